"""Builds the HIP library ``tide/_lib/libtide_hip.so`` for gfx950.

    python tide-gpr_amd/build.py            # compile (always)

Flags: -O3, no FMA contraction (results match the reference CPU
arithmetic operation for operation), packed-FP32 instructions disabled.
"""

from __future__ import annotations

import os
import shutil
import subprocess
import sys
import tempfile
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
INCLUDE = HERE.parent / "include"
OUT = HERE / "tide" / "_lib" / "libtide_hip.so"
SOURCES = ["tide_common.hip", "tide_storage.hip", "tm2d.hip", "tm2d_highorder.hip", "tm2d_launch_adj_f32.hip", "tm2d_launch_adj_f64.hip",
           "tm2d_launch_fwd_f32.hip", "tm2d_launch_fwd_f64.hip", "tm2d_launch_k_fwd_f32.hip",
           "tm2d_launch_k_fwd_f64.hip", "tm2d_launch_k_adj_f32.hip", "tm2d_launch_k_adj_f64.hip"]
OPTIONAL_SOURCES = ["em3d.hip", "tm2d_batch.hip", "tm2d_serve.hip"]
ARCH = os.environ.get("TIDE_HIP_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd: list[str], verbose: bool) -> None:
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"hipcc failed with status {proc.returncode}")


def build(verbose: bool = False) -> Path:
    """Compiles every source to an object in parallel, then links the
    shared library in place (atomically replaced)."""
    OUT.parent.mkdir(parents=True, exist_ok=True)
    srcs = [CSRC / s for s in SOURCES]
    srcs += [CSRC / s for s in OPTIONAL_SOURCES if (CSRC / s).exists()]
    flags = [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", str(CSRC), "-I", str(INCLUDE),
    ]
    tmp = OUT.with_suffix(".so.tmp")
    with tempfile.TemporaryDirectory(prefix="tide_hip_build_") as work:
        objs = [str(Path(work) / (src.stem + ".o")) for src in srcs]
        jobs = max(1, min(len(srcs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 1))))
        with ThreadPoolExecutor(max_workers=jobs) as pool:
            futures = [pool.submit(_run, [hipcc(), *flags, "-c", str(src), "-o", obj], verbose)
                       for src, obj in zip(srcs, objs)]
            for fut in futures:
                fut.result()
        _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs], verbose)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
