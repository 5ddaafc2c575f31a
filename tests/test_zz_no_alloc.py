"""The allocation-free forward entry point ``tide_tm2d_float_forward_noalloc``
against ``maxwell_tm_2_float_forward_cuda`` (same arguments), through the C
ABI: two calls in a row on the same device, the second with a smaller
problem, each bit-identical to the allocating entry point (fields and
receivers) and within 1e-6 of the oracle.  A problem whose scratch exceeds
the fixed array is rejected with a status and a message, before any device
work (the inputs stay untouched)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

import tide
from gpu_abi import gpu_forward
from helpers import hetero_problem, relative_l2
from oracle import oracle as O
from tide.maxwell.tm2d import prepare_tm2d

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    tide._native.load()


def _case(**kw):
    return O.case_from_setup(prepare_tm2d(**hetero_problem(**kw)))


def test_noalloc_forward_matches_allocating_entry_twice():
    # nt odd: both the two-step launches and the single-step tail run
    for kw in (dict(shots=4, ny=70, nx=90, nt=141, seed=21), dict(shots=2, ny=36, nx=44, nt=60, seed=22)):
        case = _case(**kw)
        print(f"case {kw}", flush=True)
        f_ref, r_ref, _, _ = gpu_forward(case)
        f_na, r_na, _, _ = gpu_forward(case, entry="tide_tm2d_float_forward_noalloc")
        assert np.array_equal(r_na, r_ref)
        for a, b in zip(f_na, f_ref):
            assert np.array_equal(a, b)
        _, r_o, _, _ = O.tm2d_forward(case)
        assert relative_l2(r_na, r_o) <= 1e-6


def test_noalloc_forward_rejects_oversized_problem():
    case = _case(shots=1, ny=36, nx=44, nt=4, seed=23)
    # 7 scratch copies of 40 x 2000 x 2000 fp32 values exceed the 64 MiB array;
    # the geometry is checked before the arrays are read, so tiny placeholder
    # buffers are never touched.
    lib = tide._native.load()
    dev = torch.device("cuda:0")
    dummy = torch.zeros(16, device=dev)
    P = tide._native.ptr
    st = lib.tide_tm2d_float_forward_noalloc(
        *[P(dummy)] * 11, None, None, None, None, None, P(dummy), 0, *[P(dummy)] * 12,
        None, None, case.rdy, case.rdx, case.dt, 4, 40, 2000, 2000, 0, 0, 1, False,
        False, False, False, 0, 20, 20, 1980, 1980, 0, 0, tide._native.stream_handle(dev))
    assert st != 0
    msg = lib.tide_last_error().decode()
    assert "scratch" in msg, msg
    assert not dummy.any()
