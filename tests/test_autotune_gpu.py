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
    before = _native.get_tm2d_tuning()
    before_loop = _native.get_tm2d_shot_loop()
    try:
        info = tune_tm2d(180, 220, 3, nt=16)
        print(info, flush=True)
        fr, fg, ar, ag = info["tuning"]
        fl, al = info["shot_loop"]
        assert (fr, fg, fl) in FWD_CANDIDATES and (ar, ag, al) in ADJ_CANDIDATES
        assert _native.get_tm2d_tuning() == info["tuning"]
        assert _native.get_tm2d_shot_loop() == (bool(fl), bool(al))
        again = tune_tm2d(181, 219, 3, nt=16)
        assert again["cached"] and again["tuning"] == info["tuning"]
        assert again["shot_loop"] == info["shot_loop"]
        case, expected = load_golden("tm2d_hetero_f32")
        _, r, _, _ = gpu_forward(case)
        assert np.array_equal(r, O.tm2d_forward(case)[1])
    finally:
        _native.set_tm2d_tuning(*before)
        _native.set_tm2d_shot_loop(*before_loop)
