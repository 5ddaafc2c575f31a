"""Batched plans (tide.batch.TM2DBatch) against single-problem runs (MI355X).

A plan advances problems of different grid sizes, shot counts and nt with
one launch per step (or per two steps, with the two-step kernels, problems
of odd nt taking their last step alone) through the same tile kernels as
the single-problem path, so receivers are bit-identical to separate runs
and FP64 gradients agree to rounding (problems with more shots than a
workgroup holds sum per-group FP64 slices, folded after the loop).
Graph replays are checked against eager runs on every replay, with eager
work between the replays.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import tide
from gpu_abi import gpu_backward, gpu_forward
from helpers import relative_l2
from oracle import oracle as O
from tide.batch import TM2DBatch
from tide.maxwell.tm2d import maxwelltm, prepare_tm2d

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


def _problem(ny, nx, nt, shots, seed, pml=(4, 5, 3, 6)):
    g = torch.Generator().manual_seed(seed)
    eps = 1 + 8 * torch.rand(ny, nx, generator=g)
    sig = 0.01 * torch.rand(ny, nx, generator=g)
    dx = 0.01
    dt = 0.9 * dx / (2**0.5 * tide.C0)
    amp = tide.ricker(6e8, nt, dt).reshape(1, 1, -1).repeat(shots, 1, 1)
    src = torch.stack([torch.tensor([[2 + s % (ny - 4), 3 + (5 * s) % (nx - 6)]]) for s in range(shots)])
    rec = torch.stack([torch.tensor([[1, x] for x in range(1, nx - 1, 4)]) for _ in range(shots)])
    return dict(epsilon=eps.to(DEV), sigma=sig.to(DEV), mu=torch.ones(ny, nx, device=DEV),
                grid_spacing=dx, dt=dt, source_amplitude=amp.to(DEV), source_location=src.to(DEV),
                receiver_location=rec.to(DEV), pml_width=list(pml))


PROBLEMS = [
    dict(ny=40, nx=52, nt=60, shots=3, seed=1),
    dict(ny=70, nx=33, nt=45, shots=1, seed=2),
    dict(ny=96, nx=140, nt=80, shots=11, seed=3),  # two shot groups (atomics)
    dict(ny=33, nx=40, nt=31, shots=2, seed=4),
]


def _single(kw, grad_r=None):
    st = prepare_tm2d(**{k: v for k, v in kw.items()})
    case = O.case_from_setup(st)
    _, r, es, cs = gpu_forward(case, store="full")
    if grad_r is None:
        return r, None
    gca, gcb, gf, _ = gpu_backward(case, grad_r, es, cs)
    return r, (gca, gcb, gf)


@pytest.fixture(scope="module")
def problems():
    return [_problem(**p) for p in PROBLEMS]


@pytest.fixture(params=[1, 2], ids=["one_step", "two_step"])
def steps(request):
    """Time steps per launch of the plans built in the test (and of the
    single-problem runs they are compared with)."""
    before = tide._native.get_tm2d_time_block()
    tide._native.set_tm2d_time_block(request.param, request.param)
    print(f"steps per launch {request.param}", flush=True)
    yield request.param
    tide._native.set_tm2d_time_block(*before)


def test_batch_forward_and_adjoint_match_single_runs(problems, steps):
    batch = TM2DBatch(problems)
    rs = [r.cpu().numpy() for r in batch.forward()]
    rng = np.random.default_rng(0)
    grs = [rng.standard_normal(r.shape).astype(np.float32) for r in rs]
    out = batch.backward([torch.from_numpy(g).to(DEV) for g in grs])
    torch.cuda.synchronize()
    for i, (kw, r, gr, (gca, gcb, gf)) in enumerate(zip(problems, rs, grs, out)):
        print(f"problem {i}: {PROBLEMS[i]}", flush=True)
        r1, (gca1, gcb1, gf1) = _single(kw, gr)
        assert np.array_equal(r, r1)
        assert relative_l2(gca.cpu().numpy(), gca1) < 1e-12
        assert relative_l2(gcb.cpu().numpy(), gcb1) < 1e-12
        assert np.array_equal(gf.cpu().numpy(), gf1)
    batch.close()


def test_batch_graph_replays_match_eager(problems, steps):
    eager = TM2DBatch(problems, use_graph=False)
    graph = TM2DBatch(problems, use_graph=True)
    r_e = [r.clone() for r in eager.forward()]
    g_e = [tuple(t.clone() for t in x) for x in eager.backward([torch.ones_like(r) for r in r_e])]
    for replay in range(3):
        graph.reset_state()
        r_g = graph.forward()
        # unrelated eager work between the replays
        junk = torch.randn(1 << 20, device=DEV).sum()
        torch.cuda.synchronize()
        g_g = graph.backward([torch.ones_like(r) for r in r_g])
        torch.cuda.synchronize()
        assert float(junk) == float(junk)
        for a, b in zip(r_g, r_e):
            assert torch.equal(a, b), f"replay {replay}"
        for (a0, a1, a2), (b0, b1, b2) in zip(g_g, g_e):
            assert torch.allclose(a0, b0, rtol=1e-12, atol=0), f"replay {replay}"
            assert torch.allclose(a1, b1, rtol=1e-12, atol=0)
            assert torch.equal(a2, b2)
    eager.close()
    graph.close()


def test_batch_update_models_in_place(problems):
    batch = TM2DBatch(problems[:2], gradients=False)
    batch.forward()
    new = [dict(kw, epsilon=kw["epsilon"] * 1.3) for kw in problems[:2]]
    batch.update_models([kw["epsilon"] for kw in new], [kw["sigma"] for kw in new])
    batch.reset_state()
    rs = [r.cpu().numpy() for r in batch.forward()]
    torch.cuda.synchronize()
    for kw, r in zip(new, rs):
        r1, _ = _single(kw)
        assert np.array_equal(r, r1)
    batch.close()


def test_batch_gradient_matches_operator(problems, steps):
    batch = TM2DBatch(problems[:3])
    observed = []
    for kw in problems[:3]:
        with torch.no_grad():
            observed.append(maxwelltm(**dict(kw, epsilon=kw["epsilon"] * 1.05))[-1])
    misfit, ge, gs = batch.gradient(observed)
    for kw, obs, m, e_b, s_b in zip(problems[:3], observed, misfit, ge, gs):
        e = kw["epsilon"].clone().requires_grad_(True)
        s = kw["sigma"].clone().requires_grad_(True)
        r = maxwelltm(**dict(kw, epsilon=e, sigma=s))[-1]
        loss = 0.5 * ((r.double() - obs.double()) ** 2).sum()
        loss.backward()
        assert abs(float(loss) - float(m)) <= 1e-9 * abs(float(loss))
        assert relative_l2(e_b.cpu().numpy(), e.grad.cpu().numpy()) < 1e-6
        assert relative_l2(s_b.cpu().numpy(), s.grad.cpu().numpy()) < 1e-6
    batch.close()
