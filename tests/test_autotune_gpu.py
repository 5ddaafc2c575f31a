"""The launch-shape autotuner picks a candidate, installs it, caches it per
size class, and the kernels stay exact under the choice (MI355X)."""

from __future__ import annotations

import numpy as np
import pytest
import torch

from gpu_abi import gpu_forward
from helpers import load_golden
from oracle import oracle as O
from tide import _native
from tide.autotune import ADJ_CANDIDATES, FWD_CANDIDATES, tune_tm2d

pytestmark = pytest.mark.gpu


def test_autotune_installs_a_candidate_and_caches_it():
    from tide.autotune import _install, current

    before = current()
    try:
        info = tune_tm2d(180, 220, 3, nt=16)
        print(info, flush=True)
        assert info["forward"] in FWD_CANDIDATES and info["adjoint"] in ADJ_CANDIDATES
        assert current() == (info["forward"], info["adjoint"])
        again = tune_tm2d(181, 219, 3, nt=16)
        assert again["cached"] and again["forward"] == info["forward"]
        assert again["adjoint"] == info["adjoint"]
        case, expected = load_golden("tm2d_hetero_f32")
        _, r, _, _ = gpu_forward(case)
        assert np.array_equal(r, O.tm2d_forward(case)[1])
    finally:
        _install(*before)


def test_autotune_scores_a_list_of_shot_chunks():
    from tide.autotune import _install, current

    before = current()
    try:
        info = tune_tm2d(150, 170, [3, 2, 2], nt=16)
        print(info, flush=True)
        assert not info["cached"] and info["chunks"] == [3, 2, 2]
        assert info["forward"] in FWD_CANDIDATES and info["adjoint"] in ADJ_CANDIDATES
        assert current() == (info["forward"], info["adjoint"])
        again = tune_tm2d(150, 170, [2, 3, 2], nt=16)
        assert again["cached"] and again["adjoint"] == info["adjoint"]
        with pytest.raises(ValueError):
            tune_tm2d(150, 170, [], nt=16)
    finally:
        _install(*before)
