"""Pins the 3-D CPU oracle (test infrastructure) before it is trusted.

* forward: bit-identical to the reference's compiled CPU 3-D forward
  (``maxwell_3d_2_<dtype>_forward_cpu``) on the committed golden vectors,
  and on fresh seeded cases for every source/receiver component pairing
  when the reference library is built here;
* adjoint: the FP64 dot-product identity <J dm, w> = <dm, J^T w> against
  the tangent-linear forward, term by term (ca, cb, source), for every
  component pairing, with CPML on every face;
* tangent: central finite differences of the forward;
* closed forms: zero source gives zero; linearity in the source; the
  adjoint of a checkpointed (segmented) backward equals the one-piece one;
  BF16 snapshots keep the gradient within 1e-2 relative L2.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from helpers import FIELD_NAMES_3D, hetero3d_problem, load_golden3d, relative_l2
from oracle import oracle as O
from tide.maxwell.maxwell3d import prepare_em3d


@pytest.mark.parametrize("name", ["em3d_hetero_f32", "em3d_hetero_f64"])
def test_oracle3d_forward_bit_identical_to_reference_golden(name):
    case, expected = load_golden3d(name)
    fields, r, _ = O.em3d_forward(case)
    assert np.array_equal(r, expected["r"])
    for k, v in zip(FIELD_NAMES_3D, fields[:6]):
        assert np.array_equal(v, expected[k]), k


def _case(dtype=torch.float64, **kw):
    return O.em_case_from_setup(prepare_em3d(**hetero3d_problem(dtype=dtype, **kw)))


@pytest.mark.skipif(not O.ref_available(), reason="reference CPU library not built here")
@pytest.mark.parametrize("comps", [("ex", "ez"), ("ey", "ey"), ("ez", "ex")])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_oracle3d_forward_matches_reference_library(comps, dtype):
    case = _case(dtype, src_comp=comps[0], rec_comp=comps[1], seed=5, nt=45)
    f_o, r_o, _ = O.em3d_forward(case)
    f_r, r_r = O.ref_em3d_forward(case)
    assert np.abs(r_o).max() > 0
    assert np.array_equal(r_o, r_r)
    for a, b in zip(f_o, f_r):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("comps", [("ex", "ez"), ("ey", "ey"), ("ez", "ex")])
def test_oracle3d_adjoint_dot_product_identity(comps):
    case = _case(src_comp=comps[0], rec_comp=comps[1], nz=8, ny=9, nx=10, nt=30)
    rng = np.random.default_rng(3)
    dca = rng.standard_normal(case.shape) * float(np.abs(case.ca).mean()) * 1e-3
    dcb = rng.standard_normal(case.shape) * float(np.abs(case.cb).mean()) * 1e-3
    df = rng.standard_normal(case.f.shape) * float(np.abs(case.f).max()) * 1e-3
    w = rng.standard_normal((case.nt, case.n_shots, case.rec.shape[1]))
    _, _, stores = O.em3d_forward(case, store="full")
    gca, gcb, gf, _ = O.em3d_backward(case, w, stores)
    zero = np.zeros_like(dca)
    for a, b, c, g in ((dca, zero, 0 * df, (dca * gca).sum()),
                       (zero, dcb, 0 * df, (dcb * gcb).sum()),
                       (zero, zero, df, (df * gf).sum())):
        _, dr = O.em3d_tangent(case, a, b, c)
        lhs = float((dr * w).sum())
        assert abs(lhs) > 0
        assert abs(lhs - g) <= 1e-11 * abs(lhs), (lhs, g)


def _born_direction(case, seed):
    rng = np.random.default_rng(seed)
    dca = (rng.standard_normal(case.shape) * 1e-2).astype(case.dtype)
    dcb = (rng.standard_normal(case.shape) * float(np.abs(case.cb).mean()) * 1e-2).astype(case.dtype)
    df = (rng.standard_normal(case.f.shape) * float(np.abs(case.f).max()) * 1e-2).astype(case.dtype)
    return dca, dcb, df


@pytest.mark.skipif(not O.ref_available(), reason="reference CPU library not built here")
@pytest.mark.parametrize("comps", [("ex", "ez"), ("ey", "ey"), ("ez", "ex")])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_oracle3d_born_forward_matches_reference_library(comps, dtype):
    """The Born forward restatement is bit-identical to the reference's own
    CPU born_forward (fields, scattered fields, both receiver sets)."""
    case = _case(dtype, src_comp=comps[0], rec_comp=comps[1], seed=6, nt=40)
    dca, dcb, df = _born_direction(case, 4)
    out_o = O.em3d_born_forward(case, dca, dcb, df)
    out_r = O.ref_em3d_born_forward(case, dca, dcb, df)
    assert np.abs(out_o[2]).max() > 0 and np.abs(out_o[3]).max() > 0
    for a, b in zip(out_o[0] + out_o[1], out_r[0] + out_r[1]):
        assert np.array_equal(a, b)
    assert np.array_equal(out_o[2], out_r[2]) and np.array_equal(out_o[3], out_r[3])


def test_oracle3d_born_background_and_tangent():
    """The Born forward's background is the plain forward bit for bit, and
    its scattered receivers are the tangent-linear forward's (same terms, a
    different summation order: FP64 to 1e-13), which central differences
    pin."""
    case = _case(nt=30, nz=8, ny=9, nx=10)
    dca, dcb, df = _born_direction(case, 8)
    fields, _, dr, bg_r = O.em3d_born_forward(case, dca, dcb, df)
    f_o, r_o, _ = O.em3d_forward(case)
    assert np.array_equal(bg_r, r_o)
    for a, b in zip(fields, f_o):
        assert np.array_equal(a, b)
    _, dr_t = O.em3d_tangent(case, dca, dcb, df)
    assert np.abs(dr).max() > 0
    assert relative_l2(dr, dr_t) < 1e-13


def test_oracle3d_tangent_matches_finite_differences():
    case = _case(nt=25, nz=7, ny=8, nx=9)
    rng = np.random.default_rng(9)
    dca = rng.standard_normal(case.shape) * 1e-2
    dcb = rng.standard_normal(case.shape) * float(np.abs(case.cb).mean()) * 1e-2
    df = np.zeros_like(case.f)
    _, dr = O.em3d_tangent(case, dca, dcb, df)
    h = 1e-4

    def run(s):
        c = case.astype(np.float64)
        c.ca = np.ascontiguousarray(c.ca + s * h * dca[None])
        c.cb = np.ascontiguousarray(c.cb + s * h * dcb[None])
        return O.em3d_forward(c)[1]

    fd = (run(+1) - run(-1)) / (2 * h)
    assert relative_l2(fd, dr) < 1e-6


def test_oracle3d_zero_source_and_linearity():
    case = _case(nt=20)
    z = case.astype(case.dtype)
    z.f = np.zeros_like(case.f)
    assert not np.any(O.em3d_forward(z)[1])
    two = case.astype(case.dtype)
    two.f = np.ascontiguousarray(2.0 * case.f)
    r1, r2 = O.em3d_forward(case)[1], O.em3d_forward(two)[1]
    assert relative_l2(r2, 2.0 * r1) < 1e-13


def test_oracle3d_segmented_adjoint_equals_single_pass():
    """Backward over [t0, t1) segments with recomputed snapshots (the
    checkpoint scheme) gives the one-pass adjoint exactly."""
    case = _case(nt=24)
    rng = np.random.default_rng(4)
    w = rng.standard_normal((case.nt, case.n_shots, case.rec.shape[1]))
    _, _, stores = O.em3d_forward(case, store="full")
    g1 = O.em3d_backward(case, w, stores)
    # segmented: checkpoints every 7 steps, snapshots per segment
    seg = 7
    ckpts, state = [], [x.copy() for x in case.fields]
    for t0 in range(0, case.nt, seg):
        ckpts.append([x.copy() for x in state])
        c = case.astype(case.dtype)
        c.fields = state
        state, _, _ = O.em3d_forward(c, start_t=t0, nt=min(seg, case.nt - t0))
    lam = None
    gca = np.zeros(case.shape)
    gcb = np.zeros(case.shape)
    gf = np.zeros_like(case.f)
    for k in reversed(range(len(ckpts))):
        t0 = k * seg
        t1 = min(t0 + seg, case.nt)
        c = case.astype(case.dtype)
        c.fields = ckpts[k]
        _, _, st = O.em3d_forward(c, store="full", start_t=t0, nt=t1 - t0)
        g = O.em3d_backward(case, w, st, start_t=t1, nt=t1 - t0, lam0=lam)
        gca += g[0]
        gcb += g[1]
        gf[t0:t1] = g[2][t0:t1]
        lam = g[3]
    assert np.array_equal(gf, g1[2])
    assert np.allclose(gca, g1[0], rtol=1e-12, atol=0)
    assert np.allclose(gcb, g1[1], rtol=1e-12, atol=0)


def test_oracle3d_bf16_gradient_error_bound():
    case = _case(torch.float32, nt=40)
    rng = np.random.default_rng(2)
    w = rng.standard_normal((case.nt, case.n_shots, case.rec.shape[1])).astype(np.float32)
    _, _, s_full = O.em3d_forward(case, store="full")
    _, _, s_bf = O.em3d_forward(case, store="bf16")
    g_full = O.em3d_backward(case, w, s_full)
    g_bf = O.em3d_backward(case, w, s_bf, store="bf16")
    assert relative_l2(g_bf[0], g_full[0]) < 1e-2
    assert relative_l2(g_bf[1], g_full[1]) < 1e-2
    assert not np.array_equal(g_bf[0], g_full[0])
