"""Shared test helpers: golden fixtures, example problems, norms."""

from __future__ import annotations

from pathlib import Path

import numpy as np
import torch

import tide
from oracle import oracle as O
from tide.maxwell.tm2d import prepare_tm2d

GOLDEN = Path(__file__).resolve().parent / "golden"
PROFILE_NAMES = ("ay", "by", "ayh", "byh", "ax", "bx", "axh", "bxh", "ky", "kyh", "kx", "kxh")
FIELD_NAMES = ("Ey", "Hx", "Hz", "m_Ey_x", "m_Ey_z", "m_Hx_z", "m_Hz_x")


def load_golden(name: str) -> tuple[O.TMCase, dict]:
    z = np.load(GOLDEN / f"{name}.npz")
    ca = z["ca"]
    n_shots = z["src"].shape[0] if z["src"].size else z["rec"].shape[0]
    ny, nx = ca.shape[-2:]
    nt, *pml = (int(v) for v in z["ints"])
    rdy, rdx, dt = (float(v) for v in z["scalars"])
    case = O.TMCase(
        ca=ca, cb=z["cb"], cq=z["cq"], f=z["f"],
        profiles=[z[f"prof_{n}"] for n in PROFILE_NAMES],
        src=z["src"], rec=z["rec"], rdy=rdy, rdx=rdx, dt=dt, nt=nt, pml=tuple(pml),
        fields=[np.zeros((n_shots, ny, nx), ca.dtype) for _ in range(7)],
    )
    expected = {"r": z["r"], **{k: z["out_" + k] for k in FIELD_NAMES}}
    return case, expected


def relative_l2(actual, reference) -> float:
    a = np.asarray(actual, dtype=np.float64)
    b = np.asarray(reference, dtype=np.float64)
    den = np.linalg.norm(b)
    assert den > 0
    return float(np.linalg.norm(a - b) / den)


def config1_setup(dtype=torch.float32, device="cpu"):
    eps = torch.full((96, 96), 4.0, dtype=dtype, device=device)
    amp = tide.ricker(8e8, 300, 4e-11, dtype=dtype, device=device).reshape(1, 1, -1)
    return prepare_tm2d(eps, torch.zeros_like(eps), torch.ones_like(eps), 0.02, 4e-11, amp,
                        torch.tensor([[[20, 48]]]), torch.tensor([[[20, 60]]]), pml_width=10)


def random_model(ny, nx, *, seed, dtype=torch.float32, device="cpu", eps_range=(1.0, 9.0), sigma_max=0.01):
    g = torch.Generator().manual_seed(seed)
    eps = eps_range[0] + (eps_range[1] - eps_range[0]) * torch.rand(ny, nx, generator=g, dtype=torch.float64)
    sig = sigma_max * torch.rand(ny, nx, generator=g, dtype=torch.float64)
    return eps.to(dtype=dtype, device=device), sig.to(dtype=dtype, device=device)


def hetero_problem(*, ny=40, nx=52, nt=160, shots=3, seed=1234, dtype=torch.float32,
                   device="cpu", pml=(5, 6, 4, 7), dx=0.01):
    eps, sig = random_model(ny, nx, seed=seed, dtype=dtype, device=device)
    dt = 0.9 * dx / (2**0.5 * tide.C0)
    amp = tide.ricker(6e8, nt, dt, dtype=dtype, device=device).reshape(1, 1, -1).repeat(shots, 2, 1)
    amp[:, 1] *= 0.5
    def col(x0, step, s, margin):  # wraps only positions that would leave the grid
        x = x0 + step * s
        return x if x < nx - 2 else x0 + (step * s) % (nx - margin)

    src = torch.stack([torch.tensor([[3, col(5, 7, s, 8)], [ny - 4, col(9, 5, s, 12)]])
                       for s in range(shots)])
    rec = torch.stack([torch.tensor([[2, x] for x in range(1, nx - 1, 5)]) for _ in range(shots)])
    return dict(epsilon=eps, sigma=sig, mu=torch.ones_like(eps), grid_spacing=dx, dt=dt,
                source_amplitude=amp, source_location=src, receiver_location=rec,
                pml_width=list(pml))


def hetero3d_problem(*, nz=10, ny=11, nx=13, nt=40, shots=2, seed=77, dtype=torch.float32,
                     device="cpu", pml=(3, 2, 4, 3, 2, 3), dx=0.02, src_comp="ey", rec_comp="ey"):
    """Small heterogeneous 3-D problem with uneven CPML for parity tests."""
    g = torch.Generator().manual_seed(seed)
    eps = 1.0 + 8.0 * torch.rand(nz, ny, nx, generator=g, dtype=torch.float64)
    sig = 0.01 * torch.rand(nz, ny, nx, generator=g, dtype=torch.float64)
    dt = 0.9 * dx / (3**0.5 * tide.C0)
    amp = tide.ricker(3e8, nt, dt, dtype=torch.float64).reshape(1, 1, -1).repeat(shots, 1, 1)
    src = torch.stack([torch.tensor([[2 + s % (nz - 4), 3, 2 + 3 * s % (nx - 4)]]) for s in range(shots)])
    # receivers on the top face within a few cells of the sources, so the
    # wave reaches them within a short run
    span = min(nx - 2, 12)
    rec = torch.stack([torch.tensor([[1, 1 + k, 1 + (2 * k + s) % span] for k in range(0, ny - 2, 2)])
                       for s in range(shots)])
    c = lambda t: t.to(dtype=dtype, device=device)  # noqa: E731
    return dict(epsilon=c(eps), sigma=c(sig), mu=c(torch.ones_like(eps)), grid_spacing=dx, dt=dt,
                source_amplitude=c(amp), source_location=src, receiver_location=rec,
                pml_width=list(pml), source_component=src_comp, receiver_component=rec_comp)


PROFILE_NAMES_3D = ("az", "bz", "azh", "bzh", "ay", "by", "ayh", "byh", "ax", "bx", "axh",
                    "bxh", "kz", "kzh", "ky", "kyh", "kx", "kxh")
FIELD_NAMES_3D = ("Ex", "Ey", "Ez", "Hx", "Hy", "Hz")


def load_golden3d(name: str) -> tuple[O.EMCase, dict]:
    z = np.load(GOLDEN / f"{name}.npz")
    ca = z["ca"]
    n_shots = z["src"].shape[0] if z["src"].size else z["rec"].shape[0]
    shape = ca.shape[-3:]
    nt, z0, y0, x0, z1, y1, x1, sc, rc = (int(v) for v in z["ints"])
    rdz, rdy, rdx, dt = (float(v) for v in z["scalars"])
    case = O.EMCase(
        ca=ca, cb=z["cb"], cq=z["cq"], f=z["f"],
        profiles=[z[f"prof_{n}"] for n in PROFILE_NAMES_3D], src=z["src"], rec=z["rec"],
        rdz=rdz, rdy=rdy, rdx=rdx, dt=dt, nt=nt, pml=(z0, y0, x0, z1, y1, x1),
        src_comp=sc, rec_comp=rc,
        fields=[np.zeros((n_shots, *shape), ca.dtype) for _ in range(18)],
    )
    expected = {"r": z["r"], **{k: z["out_" + k] for k in FIELD_NAMES_3D}}
    return case, expected


def oracle_solve(problem: dict, observed=None, *, compression=False):
    """The oracle's receivers and, given observed data, the FP64 misfit
    ``0.5 |d - d_obs|^2`` and its epsilon / sigma gradients: oracle forward
    and exact adjoint, then torch's chain rule (padding fold, coefficient
    and source-term derivatives) on the CPU.  ``problem`` holds
    ``maxwelltm`` keyword arguments on any device."""
    cpu = {k: (v.detach().cpu() if isinstance(v, torch.Tensor) else v) for k, v in problem.items()}
    want = observed is not None
    eps = cpu["epsilon"].clone().requires_grad_(want)
    sig = cpu["sigma"].clone().requires_grad_(want)
    st = prepare_tm2d(eps, sig, cpu["mu"], cpu["grid_spacing"], cpu["dt"],
                      cpu.get("source_amplitude"), cpu.get("source_location"),
                      cpu.get("receiver_location"), pml_width=cpu.get("pml_width", 20),
                      nt=cpu.get("nt"))
    case = O.case_from_setup(st)
    store = ("bf16" if compression else "full") if want else None
    _, r, es, cs = O.tm2d_forward(case, store=store)
    if not want:
        return r, None, None, None
    obs = observed.detach().cpu().double().numpy()
    resid = r.astype(np.float64) - obs
    misfit = 0.5 * float((resid**2).sum())
    gca, gcb, gf, _ = O.tm2d_backward(case, resid.astype(r.dtype), es, cs, store=store)
    dt_ = st.ca.dtype
    outs, cots = [st.ca, st.cb], [torch.from_numpy(gca).to(dt_)[None], torch.from_numpy(gcb).to(dt_)[None]]
    if st.n_src > 0:
        outs.append(st.f)
        cots.append(torch.from_numpy(gf))
    torch.autograd.backward(outs, cots)
    return r, misfit, eps.grad.numpy(), sig.grad.numpy()


def truncate_problem(problem: dict, nt: int) -> dict:
    """The same problem over its first ``nt`` steps."""
    out = dict(problem)
    out["source_amplitude"] = problem["source_amplitude"][..., :nt].contiguous()
    return out


def padded_work(problem: dict) -> float:
    """Cell-updates of one sweep of ``problem`` (stencil 2, square CPML)."""
    ny, nx = problem["epsilon"].shape[-2:]
    w = problem.get("pml_width", 20)
    w = w if isinstance(w, int) else w[0]
    shots, _, nt = problem["source_amplitude"].shape
    return float((ny + 2 * w + 1) * (nx + 2 * w + 1) * shots * nt)
