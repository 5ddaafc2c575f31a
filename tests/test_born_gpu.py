"""Born (tangent-linear) propagation on the GPU: the reference entry points
maxwell_tm_{2,4,6,8}_{float,double}_born_forward_cuda and
``op.linearize(model).jvp`` (reference ``maxwell/linearization.py:178-229``).

* Kernel: the FP64 scattered and background receivers are bit-identical
  to the oracle's tangent-linear forward (pinned against central finite
  differences and the adjoint identity in test_oracle.py) for the same
  dca / dcb / df, for every stencil order; fp32 within 1e-6.
* Operator: ``jvp`` in FP64 against central differences of the operator
  along the same direction (relative 1e-6), linear in the direction, and
  the dot-product identity <J v, w> = <v, J^T w> between ``jvp`` and
  ``vjp`` (each through its own GPU path, chain rule, padding fold and
  source term included) to 1e-9.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import tide
from helpers import hetero_problem, relative_l2
from oracle import oracle as O
from tide.maxwell.tm2d import prepare_tm2d

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    tide._native.load()


def _born_abi(case, dca, dcb, df, dtype):
    """Calls the born_forward entry point on an oracle case."""
    P = tide._native.ptr
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV, dtype)  # noqa: E731
    bg = [d(x) for x in case.fields]
    sc = [torch.zeros_like(t) for t in bg]
    r = torch.zeros(case.nt, case.n_shots, case.rec.shape[1], device=DEV, dtype=dtype)
    bg_r = torch.zeros_like(r)
    fn = getattr(tide._native.load(),
                 f"maxwell_tm_{case.stencil}_{tide._native.dtype_name(dtype)}_born_forward_cuda")
    keep = [d(case.ca), d(case.cb), d(case.cq), d(dca), d(dcb), d(case.f), d(df),
            *[d(p) for p in case.profiles]]
    src = torch.from_numpy(np.ascontiguousarray(case.src))
    rec = torch.from_numpy(np.ascontiguousarray(case.rec))
    st = fn(*[P(t) for t in keep[:7]], *[P(t) for t in bg], *[P(t) for t in sc], P(r), P(bg_r),
            *[P(t) for t in keep[7:]], P(src), P(rec), case.rdy, case.rdx, case.dt, case.nt,
            case.n_shots, case.ny, case.nx, case.src.shape[1], case.rec.shape[1], 1, False,
            False, False, 0, *case.pml, 0, 0, tide._native.stream_handle(DEV))
    tide._native.check(st)
    torch.cuda.synchronize()
    return r.cpu().numpy(), bg_r.cpu().numpy()


@pytest.mark.parametrize("stencil", [2, 4, 6, 8])
def test_born_kernel_matches_oracle_tangent(stencil):
    st = prepare_tm2d(**hetero_problem(seed=60 + stencil, dtype=torch.float64, ny=34, nx=41,
                                       nt=90), stencil=stencil)
    case = O.case_from_setup(st)
    rng = np.random.default_rng(stencil)
    dca = rng.standard_normal((case.ny, case.nx)) * 1e-3
    dcb = rng.standard_normal((case.ny, case.nx)) * 1e-3 * float(np.abs(case.cb).mean())
    df = rng.standard_normal(case.f.shape) * float(np.abs(case.f).max()) * 1e-2
    r_o, dr_o = O.tm2d_tangent(case, dca, dcb, df)
    dr, r = _born_abi(case, dca, dcb, df, torch.float64)
    print(f"stencil {stencil}: scattered {relative_l2(dr, dr_o):.2e}", flush=True)
    assert np.array_equal(r, r_o)
    assert np.array_equal(dr, dr_o)
    dr32, r32 = _born_abi(case.astype(np.float32), dca.astype(np.float32),
                          dcb.astype(np.float32), df.astype(np.float32), torch.float32)
    assert relative_l2(r32, r_o) <= 1e-5 and relative_l2(dr32, dr_o) <= 1e-5


def _operator(p, dtype):
    c = lambda t: t.to(DEV, dtype)  # noqa: E731
    return tide.MaxwellTM(
        # a fixed max velocity keeps the CPML profiles independent of the
        # model, so finite differences see only what the tangent linearizes
        tide.Discretization(spacing=p["grid_spacing"], dt=p["dt"], max_velocity=tide.C0,
                            boundary=tide.CPML(tuple(p["pml_width"]))),
        tide.Experiment(tide.Acquisition(p["source_location"].to(DEV), p["receiver_location"].to(DEV)),
                        c(p["source_amplitude"])),
    )


def test_jvp_matches_central_differences_and_is_linear():
    p = hetero_problem(seed=5, shots=2, ny=36, nx=44, nt=110, dtype=torch.float64)
    op = _operator(p, torch.float64)
    g = torch.Generator().manual_seed(9)
    de = (0.3 * torch.randn(36, 44, generator=g, dtype=torch.float64)).to(DEV)
    ds = (1e-3 * torch.randn(36, 44, generator=g, dtype=torch.float64)).to(DEV)
    model = tide.EMModel(p["epsilon"].to(DEV), p["sigma"].to(DEV), p["mu"].to(DEV))
    lin = op.linearize(model)
    t1 = lin.jvp(tide.EMDirection(epsilon=de, sigma=ds)).receiver_data
    t2 = lin.jvp(tide.EMDirection(epsilon=2 * de, sigma=2 * ds)).receiver_data
    assert relative_l2(t2.cpu().numpy(), 2 * t1.cpu().numpy()) <= 1e-12
    h = 1e-4
    with torch.no_grad():
        rp = op(tide.EMModel(model.epsilon + h * de, model.sigma + h * ds, model.mu)).receiver_data
        rm = op(tide.EMModel(model.epsilon - h * de, model.sigma - h * ds, model.mu)).receiver_data
    fd = ((rp - rm) / (2 * h)).cpu().numpy()
    err = relative_l2(t1.cpu().numpy(), fd)
    print(f"jvp vs central differences: {err:.2e}", flush=True)
    assert err <= 1e-6


def test_jvp_and_vjp_satisfy_the_dot_product_identity():
    p = hetero_problem(seed=6, shots=3, ny=36, nx=44, nt=100, dtype=torch.float64)
    op = _operator(p, torch.float64)
    g = torch.Generator().manual_seed(10)
    de = (0.3 * torch.randn(36, 44, generator=g, dtype=torch.float64)).to(DEV)
    ds = (1e-3 * torch.randn(36, 44, generator=g, dtype=torch.float64)).to(DEV)
    model = tide.EMModel(p["epsilon"].to(DEV), p["sigma"].to(DEV), p["mu"].to(DEV))
    with op.linearize(model) as lin:
        jv = lin.jvp(tide.EMDirection(epsilon=de, sigma=ds)).receiver_data
        w = torch.randn(jv.shape, generator=g, dtype=torch.float64).to(DEV)
        gt = lin.vjp(w)
    lhs = float((jv * w).sum())
    rhs = float((gt.epsilon * de).sum() + (gt.sigma * ds).sum())
    print(f"<Jv, w> = {lhs:.12e}, <v, J^T w> = {rhs:.12e}", flush=True)
    assert abs(lhs - rhs) <= 1e-9 * abs(lhs)
