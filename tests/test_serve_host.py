"""Host logic of the serving path (CPU): the FP64 closed-form model-gradient
chain rule against torch autograd through the operator's own preparation,
shot chunking under an HBM budget, the Poisson request trace and the
latency statistics."""

from __future__ import annotations

import math

import pytest
import torch

import tide
from tide.maxwell.tm2d import prepare_tm2d
from tide.serve import Ticket, job_bytes, latency_stats, model_gradient, padded_shape, plan_chunks, served_trace


@pytest.mark.parametrize("pml", [[3, 4, 2, 5], 6])
def test_model_gradient_equals_autograd_of_the_preparation(pml):
    """dJ/deps, dJ/dsigma from (dJ/dca, dJ/dcb, dJ/df) equal autograd through
    padding, coefficient compilation and source scaling (FP64)."""
    g = torch.Generator().manual_seed(0)
    ny, nx, S, nt = 17, 23, 3, 11
    eps = (1 + 8 * torch.rand(ny, nx, generator=g, dtype=torch.float64)).requires_grad_(True)
    sig = (0.01 * torch.rand(ny, nx, generator=g, dtype=torch.float64)).requires_grad_(True)
    mu = torch.ones(ny, nx, dtype=torch.float64)
    dx = 0.01
    dt = 0.9 * dx / (2**0.5 * tide.C0)
    amp = torch.randn(S, 2, nt, generator=g, dtype=torch.float64)
    src = torch.tensor([[[2, 3], [5, 7]], [[1, 1], [4, 9]], [[3, 3], [0, 0]]])
    rec = torch.tensor([[[1, 2]]] * S)
    st = prepare_tm2d(eps, sig, mu, dx, dt, amp, src, rec, pml_width=pml, nt=nt)
    gca = torch.randn(st.ny, st.nx, generator=g, dtype=torch.float64)
    gcb = torch.randn(st.ny, st.nx, generator=g, dtype=torch.float64)
    gf = torch.randn(nt, S, 2, generator=g, dtype=torch.float64)
    torch.autograd.backward([st.ca, st.cb, st.f], [gca.reshape(st.ca.shape), gcb.reshape(st.cb.shape), gf])
    st2 = prepare_tm2d(eps.detach(), sig.detach(), mu, dx, dt, amp, src, rec, pml_width=pml, nt=nt)
    ge, gs = model_gradient(gca, gcb, gf, st2, (ny, nx), amp, dx)
    assert torch.allclose(ge, eps.grad, rtol=1e-13, atol=0)
    assert torch.allclose(gs, sig.grad, rtol=1e-13, atol=0)
    # a dropped source term is caught
    ge0, _ = model_gradient(gca, gcb, None, st2, (ny, nx), amp, dx)
    assert not torch.allclose(ge0, eps.grad, rtol=1e-6, atol=0)


def test_padded_shape_and_job_bytes():
    assert padded_shape(1024, 1024, 20) == (1065, 1065)
    assert padded_shape(10, 12, [1, 2, 3, 4]) == (14, 20)
    f = job_bytes(100, 120, 4, 50, 1, 8, gradient=False)
    g = job_bytes(100, 120, 4, 50, 1, 8, gradient=True)
    h = job_bytes(100, 120, 4, 50, 1, 8, gradient=True, bf16=True)
    assert f < h < g
    # fields + memories dominate a forward job: 12 arrays per shot
    assert f >= 12 * 4 * 100 * 120 * 4
    # snapshots dominate a long gradient job: 2 x nt x shots x cells x 4 B
    assert job_bytes(100, 120, 4, 4000, 1, 8, gradient=True) >= 2 * 4000 * 4 * 100 * 120 * 4


def test_state_spec_and_pool():
    from tide.serve import DevicePool, spec_bytes, state_spec

    f = state_spec(30, 40, 3, 20, 1, 5, gradient=False)
    g = state_spec(30, 40, 3, 20, 1, 5, gradient=True, group=2)
    assert {n for n, *_ in f} == {"fields", "fields_b", "r"}
    assert {"ey_store", "lambda", "gca_groups"} <= {n for n, *_ in g}
    assert spec_bytes(f) < spec_bytes(g)
    assert spec_bytes(f) % 256 == 0 and spec_bytes(f) >= 12 * 3 * 30 * 40 * 4
    pool = DevicePool(1 << 20, torch.device("cpu"))
    a = pool.alloc(1000)
    b = pool.alloc(300_000)
    c = pool.alloc(1000)
    assert (a, b, c) == (0, 1024, 1024 + 300_032)
    assert pool.alloc(1 << 20) is None
    pool.release(b, 300_000)
    assert pool.alloc(200_000) == 1024  # first fit reuses the hole
    pool.release(a, 1000)
    pool.release(c, 1000)
    pool.release(1024, 200_000)
    assert pool.free == [(0, pool.size)] and pool.available == pool.size
    t = pool.carve(0, f)
    assert len(t["fields"]) == 7 and t["fields"][0].shape == (3, 30, 40)
    assert t["r"].shape == (20, 3, 5) and t["r"].dtype == torch.float32
    views = t["fields"] + t["fields_b"] + [t["r"]]
    spans = sorted((v.data_ptr(), v.data_ptr() + v.numel() * v.element_size()) for v in views)
    assert all(e0 <= s1 for (_, e0), (s1, _) in zip(spans, spans[1:]))  # disjoint


def test_plan_chunks():
    per_shot = lambda n: 10 * n  # noqa: E731
    assert plan_chunks(7, per_shot, 1000) == [(0, 7)]
    assert plan_chunks(7, per_shot, 40) == [(0, 2), (2, 4), (4, 6), (6, 7)]
    assert plan_chunks(3, per_shot, 15) == [(0, 1), (1, 2), (2, 3)]  # one shot may use it all
    with pytest.raises(ValueError, match="HBM budget"):
        plan_chunks(3, per_shot, 9)


def test_served_trace_is_seeded_and_in_range():
    a = served_trace(12, rate=5.0, seed=7, side_range=(128, 256), nt_range=(20, 40), shot_range=(1, 4))
    b = served_trace(12, rate=5.0, seed=7, side_range=(128, 256), nt_range=(20, 40), shot_range=(1, 4))
    times = [t for t, _ in a]
    assert times == sorted(times) and times == [t for t, _ in b]
    for (_, kw), (_, kw2) in zip(a, b):
        p = kw["problem"]
        ny, nx = p["epsilon"].shape
        assert 128 <= ny <= 256 and 128 <= nx <= 256
        nt = p["source_amplitude"].shape[-1]
        shots = p["source_amplitude"].shape[0]
        assert 20 <= nt <= 40 and 1 <= shots <= 4
        assert torch.equal(p["epsilon"], kw2["problem"]["epsilon"])
        assert p["receiver_location"].shape[:2] == (shots, min(64, nx - 2))
        if kw["kind"] == "gradient":
            assert kw["observed"].shape == (nt, shots, min(64, nx - 2))
    kinds = [kw["kind"] for _, kw in served_trace(200, rate=1.0, seed=3, side_range=(8, 9), nt_range=(2, 3))]
    assert 0.2 < kinds.count("gradient") / 200 < 0.4
    gaps = [t1 - t0 for (t0, _), (t1, _) in zip(served_trace(400, 2.0, seed=1, side_range=(8, 9),
                                                              nt_range=(2, 3))[:-1],
                                                 served_trace(400, 2.0, seed=1, side_range=(8, 9),
                                                              nt_range=(2, 3))[1:])]
    assert abs(sum(gaps) / len(gaps) - 0.5) < 0.1  # exponential gaps, mean 1/rate


def test_latency_stats():
    tickets = []
    for i in range(101):
        t = Ticket(id=i, kind="forward", arrival=float(i))
        t.completed = float(i) + 1.0 + i / 100.0
        tickets.append(t)
    s = latency_stats(tickets)
    assert s["solves"] == 101
    assert math.isclose(s["p50_s"], 1.5) and math.isclose(s["p99_s"], 1.99)
    assert math.isclose(s["span_s"], 102.0) and math.isclose(s["solves_per_s"], 101 / 102.0)


def test_head_weight_goes_to_the_oldest_request_with_running_jobs():
    """The oldest request with running jobs has all of its jobs weighted by
    head_weight and every other job by 1; the weights follow the head as it
    finishes, and weight 1 makes no engine calls."""
    from types import SimpleNamespace

    from tide.serve import TM2DServer

    calls = []

    class FakeLib:
        def tide_tm2d_engine_set_weight(self, engine, job, weight):
            calls.append((job, weight))
            return 0

    def chunk(arrival, tid):
        return SimpleNamespace(ticket=SimpleNamespace(arrival=arrival, id=tid))

    srv = object.__new__(TM2DServer)
    srv._lib, srv._engine, srv.head_weight, srv._weighted = FakeLib(), None, 4, set()
    old, new = chunk(1.0, 0), chunk(2.0, 1)
    old2 = SimpleNamespace(ticket=old.ticket)  # a second chunk of the old request
    srv._jobs = {10: new, 11: old, 12: old2}
    srv._prioritize()
    assert sorted(calls) == [(11, 4), (12, 4)]
    calls.clear()
    srv._prioritize()  # nothing changed: no calls
    assert calls == []
    del srv._jobs[11], srv._jobs[12]  # the old request finished
    srv._prioritize()
    assert calls == [(10, 4)]
    calls.clear()
    srv._jobs[13] = chunk(0.5, 2)  # an older request's chunk is admitted late
    srv._prioritize()
    assert sorted(calls) == [(10, 1), (13, 4)]
    calls.clear()
    srv.head_weight = 1
    srv._prioritize()
    assert calls == []
