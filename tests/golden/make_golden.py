"""Writes the golden TM2D and 3-D fixtures from the reference's own CPU library.

Run in a container where the reference sources exist:

    make -C oracle all && python tests/golden/make_golden.py

Each fixture holds the padded inputs (built by this project's host code,
``tide.maxwell.tm2d.prepare_tm2d``) and the outputs of the reference's
compiled CPU forward ``maxwell_tm_2_<dtype>_forward_cpu`` (reference
``src/tide/csrc/tm2d/maxwell_tm_inst.cpp:727-825``): the receiver traces and
the final 7 padded fields; likewise for the 3-D forward
``maxwell_3d_2_<dtype>_forward_cpu`` (reference
``src/tide/csrc/em3d/maxwell_3d_inst.cpp:446-761``, inputs from
``tide.maxwell.maxwell3d.prepare_em3d``), receivers and the six final E/H
fields.  The reference ships no golden vectors of its own
(its tests pin closed forms and cross-backend agreement), so these vectors
are produced from its compiled CPU forward path.
"""

from __future__ import annotations

import sys
from pathlib import Path

import numpy as np
import torch

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT / "tide-gpr_amd"))
sys.path.insert(0, str(ROOT))

import tide  # noqa: E402
from oracle import oracle as O  # noqa: E402
from tide.maxwell.maxwell3d import prepare_em3d  # noqa: E402
from tide.maxwell.tm2d import prepare_tm2d  # noqa: E402

OUT = Path(__file__).resolve().parent
PROFILE_NAMES = ("ay", "by", "ayh", "byh", "ax", "bx", "axh", "bxh", "ky", "kyh", "kx", "kxh")


def config1(dtype):
    """96x96, eps_r 4, sigma 0, dx 0.02, dt 4e-11, nt 300, CPML 10,
    800 MHz Ricker at (20, 48), receiver (20, 60)."""
    eps = torch.full((96, 96), 4.0, dtype=dtype)
    amp = tide.ricker(8e8, 300, 4e-11, dtype=dtype).reshape(1, 1, -1)
    return prepare_tm2d(eps, torch.zeros_like(eps), torch.ones_like(eps), 0.02, 4e-11, amp,
                        torch.tensor([[[20, 48]]]), torch.tensor([[[20, 60]]]), pml_width=10)


def heterogeneous(dtype):
    """Seeded random eps/sigma, asymmetric CPML, 3 shots x 2 sources, 10
    receivers; exercises every CPML strip and multi-source injection."""
    g = torch.Generator().manual_seed(1234)
    ny, nx, nt, shots = 40, 52, 160, 3
    eps = (1 + 8 * torch.rand(ny, nx, generator=g)).to(dtype)
    sig = (0.01 * torch.rand(ny, nx, generator=g)).to(dtype)
    dt = 0.9 * 0.01 / (2**0.5 * tide.C0)
    amp = tide.ricker(6e8, nt, dt, dtype=dtype).reshape(1, 1, -1).repeat(shots, 2, 1)
    amp[:, 1] *= 0.5
    src = torch.stack([torch.tensor([[3, 5 + 7 * s], [ny - 4, 9 + 5 * s]]) for s in range(shots)])
    rec = torch.stack([torch.tensor([[2, x] for x in range(1, nx - 1, 5)]) for _ in range(shots)])
    return prepare_tm2d(eps, sig, torch.ones_like(eps), 0.01, dt, amp, src, rec, pml_width=[5, 6, 4, 7])


def save(name, setup):
    case = O.case_from_setup(setup)
    fields, r = O.ref_tm2d_forward(case)
    arrays = {
        "ca": case.ca, "cb": case.cb, "cq": case.cq, "f": case.f,
        **{f"prof_{n}": v for n, v in zip(PROFILE_NAMES, case.profiles)},
        "src": case.src, "rec": case.rec,
        "scalars": np.array([case.rdy, case.rdx, case.dt]),
        "ints": np.array([case.nt, *case.pml], dtype=np.int64),
        "r": r,
    }
    for k, v in zip(("Ey", "Hx", "Hz", "m_Ey_x", "m_Ey_z", "m_Hx_z", "m_Hz_x"), fields):
        arrays["out_" + k] = v
    np.savez_compressed(OUT / f"{name}.npz", **arrays)
    print(name, {k: v.shape for k, v in arrays.items() if k in ("r", "ca")})


PROFILE_NAMES_3D = ("az", "bz", "azh", "bzh", "ay", "by", "ayh", "byh", "ax", "bx", "axh",
                    "bxh", "kz", "kzh", "ky", "kyh", "kx", "kxh")


def heterogeneous_3d(dtype):
    """Seeded random eps/sigma on 8x9x10, uneven CPML, 2 shots, Ex source,
    Ez receivers; exercises every CPML strip of every axis."""
    g = torch.Generator().manual_seed(4321)
    nz, ny, nx, nt, shots = 8, 9, 10, 60, 2
    eps = (1 + 8 * torch.rand(nz, ny, nx, generator=g)).to(dtype)
    sig = (0.01 * torch.rand(nz, ny, nx, generator=g)).to(dtype)
    dt = 0.9 * 0.02 / (3**0.5 * tide.C0)
    amp = tide.ricker(3e8, nt, dt, dtype=dtype).reshape(1, 1, -1).repeat(shots, 1, 1)
    src = torch.tensor([[[3, 4, 5]], [[5, 2, 6]]])
    rec = torch.tensor([[[1, 2, 3], [6, 7, 8], [4, 4, 4]], [[2, 2, 2], [4, 6, 7], [7, 8, 9]]])
    return prepare_em3d(eps, sig, torch.ones_like(eps), 0.02, dt, amp, src, rec,
                        pml_width=[2, 3, 2, 2, 3, 2], source_component="ex",
                        receiver_component="ez")


def save_3d(name, setup):
    case = O.em_case_from_setup(setup)
    fields, r = O.ref_em3d_forward(case)
    arrays = {
        "ca": case.ca, "cb": case.cb, "cq": case.cq, "f": case.f,
        **{f"prof_{n}": v for n, v in zip(PROFILE_NAMES_3D, case.profiles)},
        "src": case.src, "rec": case.rec,
        "scalars": np.array([case.rdz, case.rdy, case.rdx, case.dt]),
        "ints": np.array([case.nt, *case.pml, case.src_comp, case.rec_comp], dtype=np.int64),
        "r": r,
    }
    for k, v in zip(("Ex", "Ey", "Ez", "Hx", "Hy", "Hz"), fields[:6]):
        arrays["out_" + k] = v
    np.savez_compressed(OUT / f"{name}.npz", **arrays)
    print(name, {k: v.shape for k, v in arrays.items() if k in ("r", "ca")})


def main():
    if not O.ref_available():
        raise SystemExit("oracle/_ref/libtide_ref.so missing: run `make -C oracle ref` first")
    save("tm2d_config1_f32", config1(torch.float32))
    save("tm2d_config1_f64", config1(torch.float64))
    save("tm2d_hetero_f32", heterogeneous(torch.float32))
    save("tm2d_hetero_f64", heterogeneous(torch.float64))
    save_3d("em3d_hetero_f32", heterogeneous_3d(torch.float32))
    save_3d("em3d_hetero_f64", heterogeneous_3d(torch.float64))


if __name__ == "__main__":
    main()
