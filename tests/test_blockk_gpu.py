"""K-step (K = 3, 4) temporally blocked 2-D TM kernels against the oracle.

Each launch advances K steps (64 - 2K owned columns per wavefront, a
recomputed halo of K rows and columns); a remainder of the time loop runs as
two-step and single-step launches.  Every cell update keeps the single-step
operation order, so forward fields, receivers, snapshots and adjoint fields
are bit-identical to the oracle; FP64 gradients agree to summation order
(1e-12).  Cases: the small golden problems (border tiles only) and a
medium problem with interior tiles, several shots and shot groups, an nt
that is not a multiple of K, and both interior tile heights.
"""

from __future__ import annotations

import contextlib

import numpy as np
import pytest
import torch

import tide
from gpu_abi import gpu_backward, gpu_forward
from helpers import load_golden, relative_l2
from oracle import oracle as O
from tide.maxwell.tm2d import prepare_tm2d

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    tide._native.load()


@contextlib.contextmanager
def shape(k, rows, group, border):
    before = (tide._native.get_tm2d_tuning(), tide._native.get_tm2d_time_block(),
              tide._native.get_tm2d_border_rows())
    tide._native.set_tm2d_tuning(rows, group, min(rows, 4), group)
    tide._native.set_tm2d_time_block(k, k)
    tide._native.set_tm2d_border_rows(border, border)
    try:
        yield
    finally:
        tide._native.set_tm2d_tuning(*before[0])
        tide._native.set_tm2d_time_block(*before[1])
        tide._native.set_tm2d_border_rows(*before[2])


def _medium_case(dtype=torch.float32, nt=137, shots=3):
    g = torch.Generator().manual_seed(77)
    ny, nx = 190, 250
    eps = (1.0 + 8.0 * torch.rand(ny, nx, generator=g, dtype=torch.float64)).to(dtype)
    sig = (0.01 * torch.rand(ny, nx, generator=g, dtype=torch.float64)).to(dtype)
    dx = 0.01
    dt = 0.9 * dx / (2**0.5 * tide.C0)
    amp = tide.ricker(5e8, nt, dt, dtype=dtype).reshape(1, 1, -1).repeat(shots, 2, 1)
    src = torch.stack([torch.tensor([[3, 20 + 60 * s], [100, 40 + 50 * s]]) for s in range(shots)])
    rec = torch.stack([torch.tensor([[2, x] for x in range(3, nx - 3, 11)]) for _ in range(shots)])
    st = prepare_tm2d(eps, sig, torch.ones_like(eps), dx, dt, amp, src, rec,
                      pml_width=[12, 9, 14, 10])
    return O.case_from_setup(st)


def _check(case):
    f_o, r_o, es, cs = O.tm2d_forward(case, store="full")
    f_g, r_g, es_g, cs_g = gpu_forward(case, store="full")
    assert np.array_equal(r_g, r_o)
    assert np.array_equal(es_g, es) and np.array_equal(cs_g, cs)
    for a, b in zip(f_g, f_o):
        assert np.array_equal(a, b)
    w = np.random.default_rng(5).standard_normal(r_o.shape).astype(case.dtype)
    go = O.tm2d_backward(case, w, es, cs)
    gg = gpu_backward(case, w, es, cs)
    for a, b in zip(gg[3], go[3]):
        assert np.array_equal(a.reshape(b.shape), b)
    assert np.array_equal(gg[2], go[2])
    assert relative_l2(gg[0], go[0]) < 1e-12
    assert relative_l2(gg[1], go[1]) < 1e-12


@pytest.mark.parametrize("k", [3, 4])
@pytest.mark.parametrize("name", ["tm2d_hetero_f32", "tm2d_hetero_f64", "tm2d_config1_f32"])
def test_k_step_small_cases_match_oracle(k, name):
    case, _ = load_golden(name)
    print(f"K {k} {name}", flush=True)
    with shape(k, 4, 2, 0):
        _check(case)


@pytest.mark.parametrize("k,rows,group,border", [
    (3, 4, 2, 0), (4, 4, 4, 2), (4, 8, 2, 0), (3, 8, 1, 2), (4, 4, 3, 0)])
def test_k_step_medium_case_matches_oracle(k, rows, group, border):
    case = _medium_case()
    print(f"K {k} rows {rows} group {group} border {border}", flush=True)
    with shape(k, rows, group, border):
        _check(case)


@pytest.mark.parametrize("k", [3, 4])
def test_k_step_fp64_medium_case_matches_oracle(k):
    case = _medium_case(torch.float64, nt=71, shots=2)
    with shape(k, 4, 2, 2):
        _check(case)


@pytest.mark.parametrize("k", [3, 4])
@pytest.mark.parametrize("waves", [0, 4])
def test_k_step_adjoint_register_budgets_match_oracle(k, waves):
    """The fp32 K-step adjoint at the compiler's register count (0) and held
    to 128 VGPRs (4, spilling) gives the oracle's results."""
    case = _medium_case(nt=70, shots=4)
    before = tide._native.get_tm2d_adjoint_waves()
    tide._native.set_tm2d_adjoint_waves(waves)
    try:
        with shape(k, 4, 4, 2):
            _check(case)
    finally:
        tide._native.set_tm2d_adjoint_waves(before)


def test_k_step_bf16_snapshots_match_oracle():
    case = _medium_case(nt=66, shots=2)
    with shape(4, 4, 2, 0):
        f_o, r_o, es, cs = O.tm2d_forward(case, store="bf16")
        _, r_g, es_g, cs_g = gpu_forward(case, store="bf16")
        assert np.array_equal(r_g, r_o)
        assert np.array_equal(es_g, es) and np.array_equal(cs_g, cs)
        w = np.random.default_rng(6).standard_normal(r_o.shape).astype(np.float32)
        go = O.tm2d_backward(case, w, es, cs, store="bf16")
        gg = gpu_backward(case, w, es, cs, store="bf16")
        assert np.array_equal(gg[2], go[2])
        assert relative_l2(gg[0], go[0]) < 1e-12 and relative_l2(gg[1], go[1]) < 1e-12
