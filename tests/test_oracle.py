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
