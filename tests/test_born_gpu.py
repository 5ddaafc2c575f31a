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


# ---------------------------------------------------- adjoint of the Born forward
def _born_store_abi(case, dca, dcb, df, dtype):
    """born_forward_with_storage through the reference ABI: returns (r, bg_r,
    hist) with the four snapshot histories on the device."""
    P = tide._native.ptr
    d = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(DEV, dtype)  # noqa: E731
    bg = [d(x) for x in case.fields]
    sc = [torch.zeros_like(t) for t in bg]
    r = torch.zeros(case.nt, case.n_shots, case.rec.shape[1], device=DEV, dtype=dtype)
    bg_r = torch.zeros_like(r)
    hist = [torch.zeros(case.nt, case.n_shots, case.ny, case.nx, device=DEV, dtype=dtype)
            for _ in range(4)]
    fn = getattr(tide._native.load(), f"maxwell_tm_{case.stencil}_"
                 f"{tide._native.dtype_name(dtype)}_born_forward_with_storage_cuda")
    keep = [d(case.ca), d(case.cb), d(case.cq), d(dca), d(dcb), d(case.f), d(df),
            *[d(p) for p in case.profiles]]
    src = torch.from_numpy(np.ascontiguousarray(case.src))
    rec = torch.from_numpy(np.ascontiguousarray(case.rec))
    stream = tide._native.stream_handle(DEV)
    st = fn(*[P(t) for t in keep[:7]], *[P(t) for t in bg], *[P(t) for t in sc], P(r), P(bg_r),
            P(hist[0]), None, None, P(hist[1]), None, None, P(hist[2]), P(hist[3]),
            *[P(t) for t in keep[7:]], P(src), P(rec), case.rdy, case.rdx, case.dt, case.nt,
            case.n_shots, case.ny, case.nx, case.src.shape[1], case.rec.shape[1], 1, 0, 0,
            case.ny * case.nx * (8 if dtype == torch.float64 else 4), True, True, False, False,
            False, 0, *case.pml, 0, 0, stream, stream)
    tide._native.check(st)
    return r, bg_r, hist, keep


def _born_bggrad_abi(case, keep, hist, w, w_bg, dtype):
    """born_backward_bggrad through the reference ABI (outputs in the field
    type, added to zeros)."""
    P = tide._native.ptr
    n_src = case.src.shape[1]
    z = lambda *s: torch.zeros(*s, device=DEV, dtype=dtype)  # noqa: E731
    lam = [z(case.n_shots, case.ny, case.nx) for _ in range(14)]
    g_f0, g_df = z(case.nt, case.n_shots, n_src), z(case.nt, case.n_shots, n_src)
    g = [z(case.ny, case.nx) for _ in range(4)]
    scratch = z(8, case.n_shots, case.ny, case.nx)
    src = torch.from_numpy(np.ascontiguousarray(case.src))
    rec = torch.from_numpy(np.ascontiguousarray(case.rec))
    stream = tide._native.stream_handle(DEV)
    fn = getattr(tide._native.load(), f"maxwell_tm_{case.stencil}_"
                 f"{tide._native.dtype_name(dtype)}_born_backward_bggrad_cuda")
    st = fn(*[P(t) for t in keep[:7]], P(w), P(w_bg), P(hist[0]), None, None, P(hist[1]), None,
            None, P(hist[2]), P(hist[3]), None, case.n_shots, *[P(t) for t in lam[:3]],
            *[P(t) for t in lam[7:10]], P(g_f0), P(g_df), *[P(t) for t in g],
            *[P(t) for t in lam[3:7]], *[P(t) for t in lam[10:14]], None, None, None, None, None,
            None, None, P(scratch), *[P(t) for t in keep[7:]], P(src), P(rec), case.rdy,
            case.rdx, case.dt, case.nt, case.n_shots, case.ny, case.nx, n_src,
            case.rec.shape[1], 1, 0, 0, case.ny * case.nx * (8 if dtype == torch.float64 else 4),
            True, True, False, False, False, False, False, case.nt, *case.pml, 0, 0, stream,
            stream)
    tide._native.check(st)
    torch.cuda.synchronize()
    return {"ca": g[0], "cb": g[1], "dca": g[2], "dcb": g[3], "f": g_f0, "df": g_df,
            "lam": lam[:3] + lam[3:7], "eta": lam[7:10] + lam[10:14]}


@pytest.mark.parametrize("stencil", [2, 4])
def test_born_adjoint_kernels_match_oracle(stencil):
    """Born forward with its four snapshot histories, then the Born adjoint
    (reference born_backward_bggrad ABI), FP64: histories, adjoint states
    and source gradients bit-identical to the oracle's tangent adjoint
    (pinned on the CPU by its dot-product identity, the plain adjoint and
    central differences), ca / cb / dca / dcb gradients to FP64 summation
    order (1e-12)."""
    st = prepare_tm2d(**hetero_problem(seed=90 + stencil, dtype=torch.float64, ny=34, nx=41,
                                       nt=80), stencil=stencil)
    case = O.case_from_setup(st)
    rng = np.random.default_rng(40 + stencil)
    dca = rng.standard_normal((case.ny, case.nx)) * 1e-3
    dcb = rng.standard_normal((case.ny, case.nx)) * 1e-3 * float(np.abs(case.cb).mean())
    df = rng.standard_normal(case.f.shape) * float(np.abs(case.f).max()) * 1e-2
    r_o, dr_o, hist_o = O.tm2d_tangent_store(case, dca, dcb, df)
    dr, r, hist, keep = _born_store_abi(case, dca, dcb, df, torch.float64)
    assert np.array_equal(r.cpu().numpy(), r_o) and np.array_equal(dr.cpu().numpy(), dr_o)
    for a, b in zip(hist, hist_o):
        assert np.array_equal(a.cpu().numpy(), b)
    w = rng.standard_normal(dr_o.shape)
    w_bg = 0.3 * rng.standard_normal(dr_o.shape)
    go = O.tm2d_tangent_adjoint(case, dca, dcb, w, w_bg, hist_o)
    t = lambda a: torch.from_numpy(a).to(DEV)  # noqa: E731
    gg = _born_bggrad_abi(case, keep, hist, t(w), t(w_bg), torch.float64)
    for k in ("lam", "eta"):
        for a, b in zip(gg[k], go[k]):
            assert np.array_equal(a.cpu().numpy(), b), k
    for k in ("f", "df"):
        assert np.array_equal(gg[k].cpu().numpy(), go[k]), k
    for k in ("ca", "cb", "dca", "dcb"):
        err = relative_l2(gg[k].cpu().numpy(), go[k])
        print(f"stencil {stencil} grad {k}: {err:.2e}", flush=True)
        assert err < 1e-12, k


def test_second_vjp_matches_central_differences():
    """op.linearize(model).second_vjp(v, w) = d<J(m) v, w>/dm, FP64, against
    central differences of <jvp(m +- h u), w> along random model directions
    u (relative 1e-6), and zero for a zero cotangent."""
    p = hetero_problem(seed=7, shots=2, ny=34, nx=40, nt=100, dtype=torch.float64)
    op = _operator(p, torch.float64)
    g = torch.Generator().manual_seed(11)
    rnd = lambda s: (s * torch.randn(34, 40, generator=g, dtype=torch.float64)).to(DEV)  # noqa: E731
    v = tide.EMDirection(epsilon=rnd(0.3), sigma=rnd(1e-3))
    model = tide.EMModel(p["epsilon"].to(DEV), p["sigma"].to(DEV), p["mu"].to(DEV))
    with op.linearize(model) as lin:
        jv = lin.jvp(v).receiver_data
        w = torch.randn(jv.shape, generator=g, dtype=torch.float64).to(DEV)
        h2 = lin.second_vjp(v, w)
        zero = lin.second_vjp(v, torch.zeros_like(w))
    assert float(zero.epsilon.abs().max()) == 0.0 and float(zero.sigma.abs().max()) == 0.0
    h = 1e-4
    for k in range(3):
        ue, us = rnd(0.2), rnd(5e-4)

        def obj(sign):
            m = tide.EMModel(model.epsilon + sign * h * ue, model.sigma + sign * h * us, model.mu)
            with op.linearize(m) as lin_m:
                return float((lin_m.jvp(v).receiver_data * w).sum())

        fd = (obj(1) - obj(-1)) / (2 * h)
        an = float((h2.epsilon * ue).sum() + (h2.sigma * us).sum())
        err = abs(an - fd) / abs(fd)
        print(f"second_vjp direction {k}: {an:.10e} vs {fd:.10e} ({err:.2e})", flush=True)
        assert err <= 1e-6
