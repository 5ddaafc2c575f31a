"""Pins the CPU oracle (test infrastructure) before it is trusted.

* forward: bit-identical to the reference's compiled CPU forward on the
  committed golden vectors (fp32 and fp64), and, when the reference library
  is built here, on a fresh seeded case;
* adjoint: the FP64 dot-product identity <J dm, w> = <dm, J^T w> against
  the tangent-linear forward, term by term (ca, cb, source);
* tangent: central finite differences of the forward;
* closed forms: one step of a unit source gives dt / (eps_r EP0 dx dy)
  (reference tests/numerical/test_forward.py:213-274); zero source gives
  zero; linearity and superposition; BF16 codec against torch's RNE cast.
"""

from __future__ import annotations

import math

import numpy as np
import pytest
import torch

import tide
from helpers import FIELD_NAMES, hetero_problem, load_golden, relative_l2
from oracle import oracle as O
from tide.maxwell.tm2d import prepare_tm2d


@pytest.mark.parametrize("name", ["tm2d_config1_f32", "tm2d_config1_f64", "tm2d_hetero_f32", "tm2d_hetero_f64"])
def test_oracle_forward_is_bit_identical_to_reference_golden(name):
    case, expected = load_golden(name)
    fields, r, _, _ = O.tm2d_forward(case)
    assert np.array_equal(r, expected["r"])
    for k, v in zip(FIELD_NAMES, fields):
        assert np.array_equal(v, expected[k]), k


@pytest.mark.skipif(not O.ref_available(), reason="reference CPU library not built here")
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_oracle_forward_matches_reference_library_fresh_case(dtype):
    st = prepare_tm2d(**{k: v for k, v in hetero_problem(seed=99, ny=33, nx=47, nt=90, dtype=dtype).items()})
    case = O.case_from_setup(st)
    f_o, r_o, _, _ = O.tm2d_forward(case)
    f_r, r_r = O.ref_tm2d_forward(case)
    assert np.array_equal(r_o, r_r)
    for a, b in zip(f_o, f_r):
        assert np.array_equal(a, b)


def _f64_case(seed=1, **kw):
    st = prepare_tm2d(**hetero_problem(seed=seed, dtype=torch.float64, **kw))
    return O.case_from_setup(st)


def test_adjoint_dot_product_identity():
    case = _f64_case(nt=120)
    rng = np.random.default_rng(1)
    dca = rng.standard_normal((case.ny, case.nx)) * 1e-3
    dcb = rng.standard_normal((case.ny, case.nx)) * 1e-3 * float(np.abs(case.cb).mean())
    df = rng.standard_normal(case.f.shape) * float(np.abs(case.f).max()) * 1e-2
    _, _, es, cs = O.tm2d_forward(case, store="full")
    w = rng.standard_normal((case.nt, case.n_shots, case.rec.shape[1]))
    gca, gcb, gf, _ = O.tm2d_backward(case, w, es, cs)
    zero_a, zero_b, zero_f = np.zeros_like(dca), np.zeros_like(dcb), np.zeros_like(df)
    for args, rhs in (
        ((dca, zero_b, zero_f), (gca * dca).sum()),
        ((zero_a, dcb, zero_f), (gcb * dcb).sum()),
        ((zero_a, zero_b, df), (gf * df).sum()),
    ):
        _, dr = O.tm2d_tangent(case, *args)
        lhs = (dr * w).sum()
        assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


def test_adjoint_dot_product_with_sampling_interval_one_step():
    """Gradient sampling every 3rd step scales the stored steps by 3; with a
    perturbation that is constant in time the identity still pins which
    steps are used: compare against the interval-1 gradient restricted to
    stored steps (t % 3 == 0) times 3, computed with a one-step grad_r."""
    case = _f64_case(nt=30)
    rng = np.random.default_rng(2)
    w = rng.standard_normal((case.nt, case.n_shots, case.rec.shape[1]))
    _, _, es1, cs1 = O.tm2d_forward(case, store="full", interval=1)
    _, _, es3, cs3 = O.tm2d_forward(case, store="full", interval=3)
    assert np.array_equal(es3, es1[::3])
    g1 = O.tm2d_backward(case, w, es1, cs1, interval=1)
    g3 = O.tm2d_backward(case, w, es3, cs3, interval=3)
    # Same adjoint field history; the interval only changes the imaging sum.
    assert np.array_equal(g1[2], g3[2])
    assert relative_l2(g3[0], g1[0]) < 0.5  # a coarser estimate of the same gradient


def test_tangent_matches_central_differences():
    case = _f64_case(nt=100)
    rng = np.random.default_rng(4)
    dca = rng.standard_normal((case.ny, case.nx)) * 1e-3
    zero = np.zeros_like(dca)
    _, dr = O.tm2d_tangent(case, dca, zero, np.zeros_like(case.f))
    h = 1e-6
    plus = O.TMCase(**{**case.__dict__, "ca": case.ca + h * dca[None]})
    minus = O.TMCase(**{**case.__dict__, "ca": case.ca - h * dca[None]})
    _, rp, _, _ = O.tm2d_forward(plus)
    _, rm, _, _ = O.tm2d_forward(minus)
    assert relative_l2((rp - rm) / (2 * h), dr) < 1e-6


def test_unit_source_single_step_closed_form():
    for spacing in ([0.04, 0.02], [0.02, 0.02]):
        eps = torch.full((12, 14), 4.0, dtype=torch.float64)
        st = prepare_tm2d(eps, torch.zeros_like(eps), torch.ones_like(eps), spacing, 1.0e-11,
                          torch.ones(1, 1, 1, dtype=torch.float64), torch.tensor([[[6, 7]]]), None,
                          pml_width=0, nt=1)
        case = O.case_from_setup(st)
        fields, _, _, _ = O.tm2d_forward(case)
        peak = float(np.abs(fields[0]).max())
        expected = 1.0e-11 / (4.0 * tide.EP0 * math.prod(spacing))
        assert peak == pytest.approx(expected, rel=1e-12)


def test_zero_source_linearity_superposition():
    case = _f64_case(nt=80)
    _, r, _, _ = O.tm2d_forward(case)
    zero = O.TMCase(**{**case.__dict__, "f": np.zeros_like(case.f)})
    fz, rz, _, _ = O.tm2d_forward(zero)
    assert not np.any(rz) and not any(np.any(x) for x in fz)
    scaled = O.TMCase(**{**case.__dict__, "f": 2.5 * case.f})
    _, rs, _, _ = O.tm2d_forward(scaled)
    assert relative_l2(rs, 2.5 * r) < 1e-13
    f1 = case.f.copy()
    f1[:, :, 1] = 0
    f2 = case.f - f1
    _, r1, _, _ = O.tm2d_forward(O.TMCase(**{**case.__dict__, "f": f1}))
    _, r2, _, _ = O.tm2d_forward(O.TMCase(**{**case.__dict__, "f": f2}))
    assert relative_l2(r1 + r2, r) < 1e-13


def test_bf16_codec_is_round_to_nearest_even():
    known = {1.0: 0x3F80, 1.0 + 2**-8: 0x3F80, 1.0 + 3 * 2**-8: 0x3F82, 1.0 + 2**-8 + 2**-20: 0x3F81, -2.0: 0xC000}
    for v, bits in known.items():
        assert int(O.bf16_encode(np.array([v], np.float32))[0]) == bits
    x = np.random.default_rng(0).standard_normal(10000).astype(np.float32) * 1e3
    torch_bits = torch.from_numpy(x).to(torch.bfloat16).view(torch.int16).numpy().view(np.uint16)
    assert np.array_equal(O.bf16_encode(x), torch_bits)
    assert np.array_equal(O.bf16_decode(torch_bits), torch.from_numpy(x).to(torch.bfloat16).float().numpy())


def test_bf16_snapshot_gradient_close_to_fp32():
    st = prepare_tm2d(**hetero_problem(nt=120))
    case = O.case_from_setup(st)
    rng = np.random.default_rng(6)
    _, r, es, cs = O.tm2d_forward(case, store="full")
    _, rb, eb, cb = O.tm2d_forward(case, store="bf16")
    assert np.array_equal(r, rb)  # compression never touches the forward fields
    w = rng.standard_normal(r.shape).astype(np.float32)
    g32 = O.tm2d_backward(case, w, es, cs)
    g16 = O.tm2d_backward(case, w, eb, cb, store="bf16")
    assert relative_l2(g16[0], g32[0]) < 1e-2
    assert relative_l2(g16[1], g32[1]) < 1e-2


# ------------------------------------------------------- stencils 4, 6, 8
def _stencil_case(stencil, dtype, **kw):
    p = hetero_problem(seed=40 + stencil, dtype=dtype, **kw)
    return O.case_from_setup(prepare_tm2d(**p, stencil=stencil))


@pytest.mark.skipif(not O.ref_available(), reason="reference CPU library not built here")
@pytest.mark.parametrize("stencil", [4, 6, 8])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_oracle_high_order_forward_matches_reference_library(stencil, dtype):
    """The reference's own compiled CPU forward for each order
    (maxwell_tm_{4,6,8}_*_forward_cpu) on a heterogeneous case with CPML on
    every side: fields and receivers bit-identical."""
    case = _stencil_case(stencil, dtype, ny=34, nx=41, nt=70)
    assert case.stencil == stencil
    f_o, r_o, _, _ = O.tm2d_forward(case)
    f_r, r_r = O.ref_tm2d_forward(case)
    assert np.any(r_o)
    assert np.array_equal(r_o, r_r)
    for a, b in zip(f_o, f_r):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("stencil", [4, 6, 8])
def test_high_order_adjoint_dot_product_identity(stencil):
    """<J dm, w> = <dm, J^T w> per term (ca, cb, source) in FP64."""
    case = _stencil_case(stencil, torch.float64, nt=90)
    rng = np.random.default_rng(stencil)
    dca = rng.standard_normal((case.ny, case.nx)) * 1e-3
    dcb = rng.standard_normal((case.ny, case.nx)) * 1e-3 * float(np.abs(case.cb).mean())
    df = rng.standard_normal(case.f.shape) * float(np.abs(case.f).max()) * 1e-2
    _, _, es, cs = O.tm2d_forward(case, store="full")
    w = rng.standard_normal((case.nt, case.n_shots, case.rec.shape[1]))
    gca, gcb, gf, _ = O.tm2d_backward(case, w, es, cs)
    zero_a, zero_b, zero_f = np.zeros_like(dca), np.zeros_like(dcb), np.zeros_like(df)
    for args, rhs in (
        ((dca, zero_b, zero_f), (gca * dca).sum()),
        ((zero_a, dcb, zero_f), (gcb * dcb).sum()),
        ((zero_a, zero_b, df), (gf * df).sum()),
    ):
        _, dr = O.tm2d_tangent(case, *args)
        lhs = (dr * w).sum()
        assert abs(lhs - rhs) <= 1e-12 * abs(lhs)


@pytest.mark.parametrize("stencil", [4, 8])
def test_high_order_tangent_matches_central_differences(stencil):
    case = _stencil_case(stencil, torch.float64, nt=70)
    rng = np.random.default_rng(10 + stencil)
    dcb = rng.standard_normal((case.ny, case.nx)) * 1e-3 * float(np.abs(case.cb).mean())
    zero = np.zeros_like(dcb)
    _, dr = O.tm2d_tangent(case, zero, dcb, np.zeros_like(case.f))
    h = 1e-4
    plus = O.TMCase(**{**case.__dict__, "cb": case.cb + h * dcb[None]})
    minus = O.TMCase(**{**case.__dict__, "cb": case.cb - h * dcb[None]})
    _, rp, _, _ = O.tm2d_forward(plus)
    _, rm, _, _ = O.tm2d_forward(minus)
    assert relative_l2((rp - rm) / (2 * h), dr) < 1e-6


# ------------------------------------------------------ adjoint of the tangent
def _tangent_inputs(case, seed):
    rng = np.random.default_rng(seed)
    dca = rng.standard_normal((case.ny, case.nx)) * 1e-3
    dcb = rng.standard_normal((case.ny, case.nx)) * 1e-3 * float(np.abs(case.cb).mean())
    df = rng.standard_normal(case.f.shape) * float(np.abs(case.f).max()) * 1e-2
    return dca, dcb, df


@pytest.mark.parametrize("stencil", [2, 4])
def test_tangent_adjoint_dot_product_in_the_direction(stencil):
    """<dr(dca, dcb, df), w> = <g_dca, dca> + <g_dcb, dcb> + <g_df, df>: the
    tangent is linear in the direction, so its adjoint satisfies this
    identity to FP64 rounding for any w."""
    st = prepare_tm2d(**hetero_problem(seed=70 + stencil, dtype=torch.float64, ny=30, nx=36,
                                       nt=70, shots=2), stencil=stencil)
    case = O.case_from_setup(st)
    dca, dcb, df = _tangent_inputs(case, stencil)
    _, dr, hist = O.tm2d_tangent_store(case, dca, dcb, df)
    w = np.random.default_rng(9).standard_normal(dr.shape)
    g = O.tm2d_tangent_adjoint(case, dca, dcb, w, None, hist)
    lhs = float((dr * w).sum())
    rhs = float((g["dca"] * dca).sum() + (g["dcb"] * dcb).sum() + (g["df"] * df).sum())
    assert abs(lhs - rhs) <= 1e-11 * abs(lhs)


def test_tangent_adjoint_background_part_is_the_plain_adjoint():
    """With a cotangent on the background receivers only, the tangent
    adjoint's ca/cb gradients and source gradient are the plain adjoint's
    (`backward`), and <r, w_bg> = <g_f, f> (r is linear in f)."""
    case = _f64_case(nt=80)
    zero = np.zeros((case.ny, case.nx))
    r, _, hist = O.tm2d_tangent_store(case, zero, zero, np.zeros_like(case.f))
    w_bg = np.random.default_rng(3).standard_normal(r.shape)
    g = O.tm2d_tangent_adjoint(case, zero, zero, None, w_bg, hist)
    _, _, es, cs = O.tm2d_forward(case, store="full")
    gca, gcb, gf, _ = O.tm2d_backward(case, w_bg, es, cs)
    assert relative_l2(g["ca"], gca.reshape(g["ca"].shape)) < 1e-12
    assert relative_l2(g["cb"], gcb.reshape(g["cb"].shape)) < 1e-12
    assert relative_l2(g["f"], gf) < 1e-12
    assert abs(float((r * w_bg).sum()) - float((g["f"] * case.f).sum())) <= 1e-11 * abs(float((r * w_bg).sum()))


@pytest.mark.parametrize("stencil", [2, 4])
def test_tangent_adjoint_model_gradient_matches_central_differences(stencil):
    """The second-order part: d<dr, w>/dca and d<dr, w>/dcb (through the
    background and the tangent recurrences) against central differences of
    <dr, w> along random ca and cb perturbations, and <dr, w> = <g_f, f>
    with df = 0 (dr is linear in the background source)."""
    st = prepare_tm2d(**hetero_problem(seed=80 + stencil, dtype=torch.float64, ny=30, nx=34,
                                       nt=70, shots=2), stencil=stencil)
    case = O.case_from_setup(st)
    dca, dcb, _ = _tangent_inputs(case, 10 + stencil)
    df0 = np.zeros_like(case.f)
    _, dr, hist = O.tm2d_tangent_store(case, dca, dcb, df0)
    w = np.random.default_rng(5).standard_normal(dr.shape)
    g = O.tm2d_tangent_adjoint(case, dca, dcb, w, None, hist)
    rng = np.random.default_rng(20 + stencil)
    for key, scale in (("ca", 1e-3), ("cb", 1e-3 * float(np.abs(case.cb).mean()))):
        u = rng.standard_normal((case.ny, case.nx)) * scale
        h = 1e-4

        def objective(sign):
            moved = O.TMCase(**{**case.__dict__, key: getattr(case, key) + sign * h * u[None]})
            return float((O.tm2d_tangent(moved, dca, dcb, df0)[1] * w).sum())

        fd = (objective(1) - objective(-1)) / (2 * h)
        an = float((g[key] * u).sum())
        assert abs(an - fd) <= 1e-7 * abs(fd), (key, an, fd)
    assert abs(float((dr * w).sum()) - float((g["f"] * case.f).sum())) <= 1e-11 * abs(float((dr * w).sum()))


# ------------------------------------------------------------------ Debye
def _debye_case(dtype=torch.float64, stencil=2, poles=2, seed=3, **kw):
    p = hetero_problem(seed=seed, dtype=dtype, **kw)
    g = torch.Generator().manual_seed(seed + 100)
    shape = p["epsilon"].shape
    d_eps = (2.0 + 3.0 * torch.rand((poles, *shape), generator=g, dtype=torch.float64)).to(dtype)
    tau = (torch.tensor([1e-9, 3e-10][:poles], dtype=torch.float64)
           .reshape(poles, 1, 1).expand(poles, *shape)).to(dtype)
    st = prepare_tm2d(**p, stencil=stencil, dispersion=tide.DebyeDispersion(d_eps, tau))
    deb = st.extra["debye"]
    return O.case_from_setup(st), [deb[k].numpy() for k in ("a", "b", "cp")]


@pytest.mark.skipif(not O.ref_available(), reason="reference CPU library not built here")
@pytest.mark.parametrize("dtype,stencil", [(torch.float64, 2), (torch.float32, 2), (torch.float64, 4)])
def test_debye_forward_bit_identical_to_reference_cpu(dtype, stencil):
    """The oracle's Debye forward against the reference's own CPU Debye
    forward (maxwell_tm_<s>_<dtype>_forward_cpu with has_dispersion),
    compiled from its sources: receivers, fields and polarization bitwise."""
    case, deb = _debye_case(dtype, stencil, nt=80, shots=2)
    f_o, r_o, pol_o, _ = O.tm2d_forward_debye(case, *deb)
    f_r, r_r, pol_r = O.ref_tm2d_forward_debye(case, *deb)
    assert np.abs(r_o).max() > 0
    assert np.array_equal(r_o, r_r)
    for a, b in zip(f_o, f_r):
        assert np.array_equal(a, b)
    # polarization in the E-update domain (the reference's ey_prev outside
    # it is an uninitialised buffer in its own operator; both start at zero here)
    assert np.array_equal(pol_o, pol_r)


def test_debye_coefficients_closed_forms():
    """delta_epsilon = 0 gives the plain coefficients and a zero
    polarization update; the trapezoidal pole recurrence has the static
    susceptibility b * 2 / (1 - a) = EP0 * delta_epsilon; cp = (1 - a) / den."""
    from tide.utils import compile_material_coefficients, prepare_parameters

    eps = torch.full((3, 4), 4.0, dtype=torch.float64)
    sig = torch.full_like(eps, 0.01)
    mu = torch.ones_like(eps)
    dt = 1e-11
    plain = prepare_parameters(eps, sig, mu, dt)
    zero = compile_material_coefficients(eps, sig, mu, dt, dispersion=tide.DebyeDispersion(0.0, 1e-9))
    assert torch.allclose(zero["ca"], plain[0], rtol=0, atol=1e-15)
    assert torch.allclose(zero["cb"], plain[1], rtol=1e-15, atol=0)
    assert float(zero["debye"]["b"].abs().max()) == 0.0
    c = compile_material_coefficients(eps, sig, mu, dt,
                                      dispersion=tide.DebyeDispersion(torch.tensor([2.0, 5.0]).reshape(2, 1, 1).double(),
                                                                      torch.tensor([1e-9, 2e-10]).reshape(2, 1, 1).double()))
    a, b = c["debye"]["a"], c["debye"]["b"]
    static = 2.0 * b / (1.0 - a)
    expect = tide.EP0 * torch.tensor([2.0, 5.0], dtype=torch.float64).reshape(2, 1, 1)
    assert torch.allclose(static, expect.expand_as(static), rtol=1e-12)
    den = tide.EP0 * eps + sig * dt / 2 + b.sum(0)
    assert torch.allclose(c["debye"]["cp"], (1 - a) / den, rtol=1e-14)
    with pytest.raises(ValueError, match="dt < min"):
        compile_material_coefficients(eps, sig, mu, 1e-9, dispersion=tide.DebyeDispersion(1.0, 1e-9))


def test_debye_with_zero_poles_is_the_plain_forward():
    """delta_epsilon = 0: the Debye forward is bit-identical to the plain one
    (the polarization stays zero and adds exactly nothing)."""
    p = hetero_problem(seed=4, dtype=torch.float64, nt=60)
    st = prepare_tm2d(**p, dispersion=tide.DebyeDispersion(0.0, 1e-9))
    case = O.case_from_setup(st)
    deb = [st.extra["debye"][k].numpy() for k in ("a", "b", "cp")]
    _, r_d, pol, _ = O.tm2d_forward_debye(case, *deb)
    _, r_p, _, _ = O.tm2d_forward(case)
    assert not np.any(pol)
    assert np.array_equal(r_d, r_p)
