"""Host-only checks of the launch-shape tuner (no device work): a candidate
that fails part-way leaves the installed launch shape as it was."""

from __future__ import annotations

import pytest
import torch

from tide import _native, autotune


class _LibWithFakeAdjoint:
    """The real library (host-only setters), with the adjoint entry replaced
    by a stub so that no device call is made."""

    def __init__(self, real):
        self._real = real

    def __getattr__(self, name):
        if name.endswith("_backward_f64acc"):
            return lambda *args: 0
        return getattr(self._real, name)


def test_failed_candidate_restores_the_installed_shape(monkeypatch):
    real = _native.load()
    before = autotune.current()
    monkeypatch.setattr(_native, "load", lambda: _LibWithFakeAdjoint(real))
    monkeypatch.setattr(_native, "stream_handle", lambda device: None)
    monkeypatch.setattr(_native, "get_function", lambda *a, **k: (lambda *args: 0))
    monkeypatch.setattr(torch.cuda, "synchronize", lambda *a, **k: None)
    calls = {"n": 0}

    def fake_time(fn, reps=2, setup=None):
        calls["n"] += 1
        if calls["n"] == 3:  # the first adjoint candidate fails
            raise RuntimeError("candidate failed")
        return 1.0

    monkeypatch.setattr(autotune, "_time", fake_time)
    fwd = [(4, 1, 0, 2, 0, 0), (4, 2, 0, 2, 0, 0)]
    adj = [(2, 2, 0, 1, 0, 0), (4, 4, 0, 2, 0, 2)]
    with pytest.raises(RuntimeError, match="candidate failed"):
        autotune.tune_tm2d(40, 48, 2, device="cpu", nt=16, fwd_candidates=fwd,
                           adj_candidates=adj, use_cache=False)
    assert autotune.current() == before
