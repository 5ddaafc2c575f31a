"""Served solves (tide.serve.TM2DServer) on MI355X.

The server advances every admitted job with the single-problem tile kernels
and the same shot grouping, so:
  * receivers are bitwise those of ``maxwelltm`` (forward solves and the
    forward half of gradient solves);
  * each solve's receivers, misfit and FP64 gradients are bitwise the same
    alone and in a mix (continuous batching changes nothing but timing);
  * epsilon / sigma gradients match ``maxwelltm``'s autograd within relative
    L2 1e-5 (both start from the same FP64 dJ/dca, dJ/dcb; the server applies
    the coefficient chain rule in FP64 closed form, the operator through fp32
    autograd, which alone leaves differences of about 2e-6);
  * shot chunking under a small HBM budget changes gradients only by FP64
    summation order (relative 1e-12).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

from helpers import hetero_problem, relative_l2
from tide.maxwell.tm2d import maxwelltm
from tide.serve import TM2DServer, latency_stats

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")
BUDGET = 2 << 30  # device pool of each test server

SPECS = [
    dict(kind="forward", ny=40, nx=52, nt=60, shots=3, seed=1),
    dict(kind="gradient", ny=36, nx=44, nt=50, shots=2, seed=2),
    dict(kind="forward", ny=70, nx=33, nt=45, shots=1, seed=3),
    dict(kind="gradient", ny=48, nx=64, nt=70, shots=9, seed=4),  # several shot groups
    dict(kind="gradient", ny=33, nx=40, nt=31, shots=1, seed=5, bf16=True),
    dict(kind="forward", ny=96, nx=120, nt=80, shots=6, seed=6),
]


def _problem(spec):
    p = hetero_problem(ny=spec["ny"], nx=spec["nx"], nt=spec["nt"], shots=spec["shots"], seed=spec["seed"])
    return {k: (v.to(DEV) if isinstance(v, torch.Tensor) and k not in ("source_location", "receiver_location") else v)
            for k, v in p.items()}


def _observed(spec, problem):
    g = torch.Generator().manual_seed(100 + spec["seed"])
    n_rec = problem["receiver_location"].shape[1]
    return (0.01 * torch.randn(spec["nt"], spec["shots"], n_rec, generator=g)).to(DEV)


def _submit_kwargs(spec, problem):
    kw = {"kind": spec["kind"], "storage_compression": "bf16" if spec.get("bf16") else False}
    if spec["kind"] == "gradient":
        kw["observed"] = _observed(spec, problem)
    return kw


def _reference(spec, problem):
    """The plain operator path: maxwelltm (+ autograd for gradients)."""
    if spec["kind"] == "forward":
        with torch.no_grad():
            return maxwelltm(**problem)[-1], None, None, None
    e = problem["epsilon"].clone().requires_grad_(True)
    s = problem["sigma"].clone().requires_grad_(True)
    r = maxwelltm(**dict(problem, epsilon=e, sigma=s),
                  storage_compression="bf16" if spec.get("bf16") else False)[-1]
    obs = _observed(spec, problem)
    loss = 0.5 * ((r.double() - obs.double()) ** 2).sum()
    loss.backward()
    return r.detach(), loss.detach(), e.grad, s.grad


@pytest.fixture(scope="module")
def problems():
    return [_problem(s) for s in SPECS]


@pytest.fixture(scope="module")
def mixed(problems):
    """All solves through one server, arriving 2 ms apart."""
    server = TM2DServer(DEV, hbm_budget=BUDGET, steps_per_tick=16)
    trace = [(0.002 * i, dict(problem=p, **_submit_kwargs(s, p))) for i, (s, p) in enumerate(zip(SPECS, problems))]
    tickets = server.serve(trace)
    torch.cuda.synchronize()
    stats = latency_stats(tickets)
    print(f"mixed: {stats} ticks {server.ticks}", flush=True)
    server.close()
    return tickets


def test_served_solves_match_the_operator(problems, mixed):
    for i, (spec, p, t) in enumerate(zip(SPECS, problems, mixed)):
        print(f"request {i}: {spec}", flush=True)
        assert t.done and t.latency is not None and t.latency > 0
        r, loss, ge, gs = _reference(spec, p)
        same = torch.equal(t.result.receiver_data, r)
        assert same, "receivers differ from the operator's"
        if spec["kind"] == "gradient":
            assert abs(float(t.result.misfit) - float(loss)) <= 1e-12 * abs(float(loss))
            e_eps = relative_l2(t.result.grad_epsilon.cpu().numpy(), ge.cpu().numpy())
            e_sig = relative_l2(t.result.grad_sigma.cpu().numpy(), gs.cpu().numpy())
            print(f"  gradient vs operator: eps {e_eps:.2e} sigma {e_sig:.2e}", flush=True)
            assert e_eps < 1e-5 and e_sig < 1e-5


def test_one_step_ticks_of_a_mix_match_the_operator(problems):
    """One-step ticks (no two-step launches): smaller jobs and the oldest
    request take several consecutive single steps per tick (size boost and
    head weight), each flipping the job's state buffer; receivers equal the
    operator's bitwise and gradients agree within 1e-5."""
    server = TM2DServer(DEV, hbm_budget=BUDGET, steps_per_tick=1, tick_boost=4, head_weight=4)
    tickets = [server.submit(p, **_submit_kwargs(s, p)) for s, p in zip(SPECS, problems)]
    server.drain()
    torch.cuda.synchronize()
    for i, (spec, p, t) in enumerate(zip(SPECS, problems, tickets)):
        print(f"request {i} (one-step ticks): {spec}", flush=True)
        r, loss, ge, gs = _reference(spec, p)
        same = torch.equal(t.result.receiver_data, r)
        assert same, "receivers differ from the operator's"
        if spec["kind"] == "gradient":
            assert relative_l2(t.result.grad_epsilon.cpu().numpy(), ge.cpu().numpy()) < 1e-5
            assert relative_l2(t.result.grad_sigma.cpu().numpy(), gs.cpu().numpy()) < 1e-5
    server.close()


def test_each_solve_is_the_same_alone_and_in_the_mix(problems, mixed):
    for i, (spec, p, t_mix) in enumerate(zip(SPECS, problems, mixed)):
        print(f"request {i} alone", flush=True)
        server = TM2DServer(DEV, hbm_budget=BUDGET, steps_per_tick=7)
        t = server.submit(p, **_submit_kwargs(spec, p))
        server.drain()
        torch.cuda.synchronize()
        names = ["receiver_data"] + (["misfit", "grad_epsilon", "grad_sigma"] if spec["kind"] == "gradient" else [])
        for name in names:
            a, b = getattr(t.result, name), getattr(t_mix.result, name)
            same = torch.equal(a, b)
            assert same, f"{name} differs alone vs in the mix (max |diff| {float((a - b).abs().max()):.3e})"
        server.close()


def test_small_budget_chunks_shots_and_queues(problems, mixed):
    spec, p, t_mix = SPECS[3], problems[3], mixed[3]
    from tide.serve import padded_shape, spec_bytes, state_spec

    ny_p, nx_p = padded_shape(spec["ny"], spec["nx"], p["pml_width"])
    n_rec = p["receiver_location"].shape[1]
    one = spec_bytes(state_spec(ny_p, nx_p, 2, spec["nt"], 2, n_rec, gradient=True, group=2))
    server = TM2DServer(DEV, hbm_budget=2 * one + 1024, steps_per_tick=9)
    tickets = [server.submit(p, **_submit_kwargs(spec, p)) for _ in range(2)]
    assert tickets[0].n_chunks == 5  # 9 shots in chunks of 2
    server.drain()
    torch.cuda.synchronize()
    for t in tickets:
        same = torch.equal(t.result.receiver_data, t_mix.result.receiver_data)
        assert same, "receivers differ under shot chunking"
        err = relative_l2(t.result.misfit.cpu().numpy(), t_mix.result.misfit.cpu().numpy())
        assert err < 1e-12, f"misfit {err:.2e}"
        err = relative_l2(t.result.grad_epsilon.cpu().numpy(), t_mix.result.grad_epsilon.cpu().numpy())
        assert err < 1e-12, f"grad_epsilon {err:.2e}"
        err = relative_l2(t.result.grad_sigma.cpu().numpy(), t_mix.result.grad_sigma.cpu().numpy())
        assert err < 1e-12, f"grad_sigma {err:.2e}"
    assert server.used == 0
    server.close()


def test_oversized_request_is_rejected(problems):
    server = TM2DServer(DEV, hbm_budget=1 << 16)
    with pytest.raises(ValueError, match="HBM budget"):
        server.submit(problems[0])
    server.close()


def test_engine_rejects_device_point_indices(problems):
    import ctypes

    from tide import _native
    from tide.serve import _I64

    server = TM2DServer(DEV, hbm_budget=BUDGET)
    spec, p = SPECS[0], problems[0]
    server.submit(p)
    server._admit()  # admitted normally
    ch = next(iter(server._jobs.values()))
    job = ch.job
    dev_idx = torch.zeros(4, dtype=torch.int64, device=DEV)
    job.p.sources_i = dev_idx.data_ptr()
    jid = _I64()
    rc = server._lib.tide_tm2d_engine_add(server._engine, ctypes.byref(job), ctypes.byref(jid))
    assert rc != 0 and "host arrays" in _native.load().tide_last_error().decode()
    server.drain()
    server.close()


def test_many_small_solves_share_launches():
    """Forty single-shot 24x24 solves submitted at once finish in far fewer
    ticks than one after another would need."""
    server = TM2DServer(DEV, hbm_budget=BUDGET, steps_per_tick=32)
    specs = [dict(kind="forward", ny=24, nx=24, nt=64, shots=1, seed=200 + i) for i in range(40)]
    ps = [_problem(s) for s in specs]
    tickets = [server.submit(p) for p in ps]
    server.drain()
    torch.cuda.synchronize()
    assert server.ticks <= 4
    for spec, p, t in zip(specs[:5], ps[:5], tickets[:5]):
        with torch.no_grad():
            r = maxwelltm(**p)[-1]
        same = torch.equal(t.result.receiver_data, r)
        assert same, "receivers differ from the operator's"
    server.close()
    lat = np.array([t.latency for t in tickets])
    assert np.all(lat > 0)


@pytest.mark.parametrize("tick", [1, 3])
def test_odd_and_single_step_ticks_match_the_operator(problems, tick):
    """Ticks of one step (one-step launches only) and of an odd count (the
    two-step engine rounds it down to pairs) give the operator's receivers
    bitwise and its gradients within 1e-5, for a solve of odd nt."""
    spec, p = SPECS[4], problems[4]  # nt 31, BF16 snapshots, gradient
    print(f"tick {tick}: {spec}", flush=True)
    server = TM2DServer(DEV, hbm_budget=BUDGET, steps_per_tick=tick)
    t = server.submit(p, **_submit_kwargs(spec, p))
    server.drain()
    torch.cuda.synchronize()
    r, loss, ge, gs = _reference(spec, p)
    same = torch.equal(t.result.receiver_data, r)
    assert same, "receivers differ from the operator's"
    assert relative_l2(t.result.grad_epsilon.cpu().numpy(), ge.cpu().numpy()) < 1e-5
    assert relative_l2(t.result.grad_sigma.cpu().numpy(), gs.cpu().numpy()) < 1e-5
    server.close()


def test_shared_source_cells_are_bitwise_repeatable():
    """Shots whose sources sit in the same cell: the source-term part of the
    sigma / epsilon gradient sums them without atomics, so two runs give the
    same bits."""
    spec = dict(kind="gradient", ny=40, nx=48, nt=60, shots=4, seed=8)
    p = _problem(spec)
    loc = p["source_location"].clone()
    loc[1] = loc[0]
    loc[3, 1] = loc[2, 0]
    p = dict(p, source_location=loc)
    runs = []
    for _ in range(2):
        server = TM2DServer(DEV, hbm_budget=BUDGET, steps_per_tick=16)
        t = server.submit(p, **_submit_kwargs(spec, p))
        server.drain()
        torch.cuda.synchronize()
        runs.append((t.result.grad_epsilon.clone(), t.result.grad_sigma.clone()))
        server.close()
    assert torch.equal(runs[0][0], runs[1][0])
    assert torch.equal(runs[0][1], runs[1][1])
    r, loss, ge, gs = _reference(spec, p)
    assert relative_l2(runs[0][0].cpu().numpy(), ge.cpu().numpy()) < 1e-5
    assert relative_l2(runs[0][1].cpu().numpy(), gs.cpu().numpy()) < 1e-5
