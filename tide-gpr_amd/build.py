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
from pathlib import Path

HERE = Path(__file__).resolve().parent
CSRC = HERE / "csrc"
INCLUDE = HERE.parent / "include"
OUT = HERE / "tide" / "_lib" / "libtide_hip.so"
SOURCES = ["tide_common.hip", "tm2d.hip"]
OPTIONAL_SOURCES = ["em3d.hip", "tm2d_batch.hip"]
ARCH = os.environ.get("TIDE_HIP_ARCH", "gfx950")


def hipcc() -> str:
    for cand in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found")


def build(verbose: bool = False) -> Path:
    OUT.parent.mkdir(parents=True, exist_ok=True)
    srcs = [str(CSRC / s) for s in SOURCES]
    srcs += [str(CSRC / s) for s in OPTIONAL_SOURCES if (CSRC / s).exists()]
    tmp = OUT.with_suffix(".so.tmp")
    cmd = [
        hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-I", str(CSRC), "-I", str(INCLUDE), "-o", str(tmp), *srcs,
    ]
    if verbose:
        print(" ".join(cmd), flush=True)
    proc = subprocess.run(cmd, capture_output=True, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout + proc.stderr)
        raise RuntimeError(f"hipcc failed with status {proc.returncode}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))
