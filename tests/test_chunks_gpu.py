"""Shot chunks inside the public 2-D operator, the shot-separable objective
path and ``linearize().vjp``, against the CPU oracle.

A device byte limit smaller than the snapshots of all shots makes the
operator split its shots into chunks: the chunks that fit keep their
snapshots until the backward, the others are re-run there from their
initial state.  Every chunk's coefficient gradient sums into one FP64
accumulator.  Tolerances: receivers bit-identical to the unchunked
operator and within 1e-6 of the oracle; gradients within 1e-4 (fp32, the
operator-level bound) of the oracle's exact adjoint followed by the same
torch chain rule.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import tide
from helpers import hetero_problem, relative_l2
from oracle import oracle as O
from tide.maxwell.tm2d import ShotPlan, prepare_tm2d
from tide.workflow import ReceiverObjective

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    tide._native.load()


def _operator(p, dev, *, limit=None, compression=False, interval=1):
    return tide.MaxwellTM(
        tide.Discretization(spacing=p["grid_spacing"], dt=p["dt"],
                            boundary=tide.CPML(tuple(p["pml_width"]))),
        tide.Experiment(tide.Acquisition(p["source_location"].to(dev), p["receiver_location"].to(dev)),
                        p["source_amplitude"].to(dev)),
        storage=tide.StorageOptions(compression=compression, bytes_limit_device=limit),
        model_gradient_sampling_interval=interval,
    )


def _per_shot_bytes(p, interval=1, bpe=4):
    st = prepare_tm2d(p["epsilon"], p["sigma"], p["mu"], p["grid_spacing"], p["dt"],
                      p["source_amplitude"], p["source_location"], p["receiver_location"],
                      pml_width=p["pml_width"], allocate_fields=False)
    n_store = (st.nt + interval - 1) // interval
    return n_store * st.ny * st.nx * bpe * 2


def _oracle_grads(p, cot, *, store="full", interval=1):
    """Oracle forward + exact adjoint, then torch's chain rule on the CPU."""
    eps_c = p["epsilon"].clone().requires_grad_(True)
    sig_c = p["sigma"].clone().requires_grad_(True)
    st = prepare_tm2d(eps_c, sig_c, p["mu"], p["grid_spacing"], p["dt"], p["source_amplitude"],
                      p["source_location"], p["receiver_location"], pml_width=p["pml_width"])
    case = O.case_from_setup(st)
    _, r_o, es, cs = O.tm2d_forward(case, store=store, interval=interval)
    gca, gcb, gf, _ = O.tm2d_backward(case, cot, es, cs, store=store, interval=interval)
    dt_ = st.ca.dtype
    torch.autograd.backward(
        [st.ca, st.cb, st.f],
        [torch.from_numpy(gca).to(dt_)[None], torch.from_numpy(gcb).to(dt_)[None],
         torch.from_numpy(gf)],
    )
    return r_o, eps_c.grad.numpy(), sig_c.grad.numpy()


def _cotangent(r):
    return (r - 0.8 * torch.roll(r, 7, dims=0)).detach()


def test_plan_from_limit():
    """The byte limit sets the chunk plan: 7 shots, 2 per chunk, one resident."""
    p = hetero_problem(shots=7, ny=36, nx=44, nt=140)
    per = _per_shot_bytes(p)
    from tide.maxwell.tm2d import plan_shot_chunks

    plan = plan_shot_chunks(7, per, 2.5 * per)
    assert plan == ShotPlan([0, 2, 4, 6, 7], 1)


@pytest.mark.parametrize("limit_shots", [2.5, 1.2, 3.5])
def test_operator_chunks_match_unchunked_and_oracle(limit_shots):
    dev = torch.device("cuda:0")
    p = hetero_problem(shots=7, ny=36, nx=44, nt=140)
    per = _per_shot_bytes(p)
    print(f"7 shots, device limit {limit_shots} shots of snapshots", flush=True)
    grads = {}
    recs = {}
    for tag, limit in (("one", None), ("chunks", int(limit_shots * per))):
        e = p["epsilon"].to(dev).requires_grad_(True)
        s = p["sigma"].to(dev).requires_grad_(True)
        r = _operator(p, dev, limit=limit)(tide.EMModel(e, s, p["mu"].to(dev))).receiver_data
        cot = _cotangent(r)
        (r * cot).sum().backward()
        grads[tag] = (e.grad.cpu().numpy(), s.grad.cpu().numpy())
        recs[tag] = (r.detach().cpu().numpy(), cot.cpu().numpy())
    assert np.array_equal(recs["one"][0], recs["chunks"][0])
    for a, b in zip(grads["one"], grads["chunks"]):
        assert relative_l2(b, a) <= 1e-6
    r_o, ge_o, gs_o = _oracle_grads(p, recs["chunks"][1])
    print(f"receivers {relative_l2(recs['chunks'][0], r_o):.2e} grad eps "
          f"{relative_l2(grads['chunks'][0], ge_o):.2e} grad sigma "
          f"{relative_l2(grads['chunks'][1], gs_o):.2e}")
    assert relative_l2(recs["chunks"][0], r_o) <= 1e-6
    assert relative_l2(grads["chunks"][0], ge_o) <= 1e-4
    assert relative_l2(grads["chunks"][1], gs_o) <= 1e-4


def test_chunks_recompute_from_a_nonzero_state():
    """Chunks re-run in the backward start from the state the call was given."""
    dev = torch.device("cuda:0")
    p = hetero_problem(shots=5, ny=30, nx=38, nt=120)
    per = _per_shot_bytes(p)
    op_all = _operator(p, dev)
    m0 = tide.EMModel(p["epsilon"].to(dev), p["sigma"].to(dev), p["mu"].to(dev))
    with torch.no_grad():
        state = op_all(m0).final_state
    out = {}
    for tag, limit in (("one", None), ("chunks", int(1.5 * per))):
        e = p["epsilon"].to(dev).requires_grad_(True)
        r = _operator(p, dev, limit=limit)(tide.EMModel(e, p["sigma"].to(dev), p["mu"].to(dev)),
                                           state=state).receiver_data
        (r * _cotangent(r)).sum().backward()
        out[tag] = (r.detach().cpu().numpy(), e.grad.cpu().numpy())
    assert np.array_equal(out["one"][0], out["chunks"][0])
    assert relative_l2(out["chunks"][1], out["one"][1]) <= 1e-6


def test_batched_models_in_chunks():
    """Per-shot models ([N, ny, nx]) split over chunks: per-model gradients."""
    dev = torch.device("cuda:0")
    p = hetero_problem(shots=5, ny=30, nx=34, nt=90)
    g = torch.Generator().manual_seed(5)
    eps = (p["epsilon"][None] * (1 + 0.2 * torch.rand(5, 1, 1, generator=g))).to(dev)
    sig = p["sigma"][None].repeat(5, 1, 1).to(dev)
    from tide.maxwell.tm2d import maxwelltm

    per = _per_shot_bytes(p)
    res = {}
    for tag, limit in (("one", None), ("chunks", int(2.2 * per))):
        e = eps.clone().requires_grad_(True)
        r = maxwelltm(e, sig, torch.ones_like(e), p["grid_spacing"], p["dt"],
                      p["source_amplitude"].to(dev), p["source_location"].to(dev),
                      p["receiver_location"].to(dev), pml_width=p["pml_width"],
                      storage_bytes_limit_device=limit)[-1]
        (r.double() ** 2).sum().backward()
        res[tag] = (r.detach().cpu().numpy(), e.grad.cpu().numpy())
    assert np.array_equal(res["one"][0], res["chunks"][0])
    assert relative_l2(res["chunks"][1], res["one"][1]) <= 1e-6


@pytest.mark.parametrize("chunked", [False, True])
def test_separable_objective_matches_operator_and_oracle(chunked):
    """``ReceiverObjective(d_obs).gradient(op.linearize(m))`` runs chunk by
    chunk (forward, residual, adjoint) with no re-run; its gradient equals
    the operator's autograd gradient of the same least-squares misfit."""
    dev = torch.device("cuda:0")
    p = hetero_problem(shots=6, ny=36, nx=44, nt=140)
    per = _per_shot_bytes(p)
    limit = int(2.2 * per) if chunked else None
    op = _operator(p, dev, limit=limit)
    model = tide.EMModel(p["epsilon"].to(dev), p["sigma"].to(dev), p["mu"].to(dev))
    with torch.no_grad():
        eps_obs = p["epsilon"].clone()
        eps_obs[10:20, 15:25] += 1.5
        d_obs = op(tide.EMModel(eps_obs.to(dev), p["sigma"].to(dev), p["mu"].to(dev))).receiver_data
    obj = ReceiverObjective(d_obs)
    with op.linearize(model) as lin:
        value, grad = obj.value_and_gradient(lin)
        r_fast = lin.primal.receiver_data.detach().cpu().numpy()
    # the general path: primal + autograd over all shots
    e = p["epsilon"].to(dev).requires_grad_(True)
    s = p["sigma"].to(dev).requires_grad_(True)
    r = _operator(p, dev)(tide.EMModel(e, s, p["mu"].to(dev))).receiver_data
    loss = 0.5 * ((r.double() - d_obs.double()) ** 2).sum()
    loss.backward()
    assert np.array_equal(r_fast, r.detach().cpu().numpy())
    assert abs(float(value) - float(loss)) <= 1e-12 * abs(float(loss))
    assert relative_l2(grad.epsilon.cpu().numpy(), e.grad.cpu().numpy()) <= 1e-6
    assert relative_l2(grad.sigma.cpu().numpy(), s.grad.cpu().numpy()) <= 1e-6
    r_o, ge_o, gs_o = _oracle_grads(p, (r - d_obs).detach().cpu().numpy())
    assert relative_l2(grad.epsilon.cpu().numpy(), ge_o) <= 1e-4
    assert relative_l2(grad.sigma.cpu().numpy(), gs_o) <= 1e-4


def test_linearize_vjp_matches_oracle():
    """``op.linearize(model).vjp(cotangent)`` (reference
    ``maxwell/linearization.py:104-118``) against the oracle."""
    dev = torch.device("cuda:0")
    p = hetero_problem(shots=3, ny=36, nx=44, nt=140)
    op = _operator(p, dev)
    model = tide.EMModel(p["epsilon"].to(dev), p["sigma"].to(dev), p["mu"].to(dev))
    with op.linearize(model) as lin:
        r = lin.primal.receiver_data
        cot = _cotangent(r)
        g = lin.vjp(cot)
        g_eps_only = op.linearize(model, targets=("epsilon",)).vjp(cot)
    assert g.mu is None and g_eps_only.sigma is None
    r_o, ge_o, gs_o = _oracle_grads(p, cot.cpu().numpy())
    assert relative_l2(r.detach().cpu().numpy(), r_o) <= 1e-6
    assert relative_l2(g.epsilon.cpu().numpy(), ge_o) <= 1e-4
    assert relative_l2(g.sigma.cpu().numpy(), gs_o) <= 1e-4
    assert relative_l2(g_eps_only.epsilon.cpu().numpy(), ge_o) <= 1e-4
    with pytest.raises(RuntimeError):
        lin.vjp(cot)  # closed


def test_config2_size_chunked_gradient_matches_oracle():
    """1024x1024 model (1065x1065 padded, CPML 20), 4 shots in 4 chunks of
    one (one resident, three re-run), nt 300, against one oracle pass over
    all shots with FP64 gradient accumulation."""
    dev = torch.device("cuda:0")
    g = torch.Generator().manual_seed(11)
    n, nt, S = 1024, 300, 4
    eps = (1 + 24 * torch.rand(n, n, generator=g)).float()
    sig = (0.01 * torch.rand(n, n, generator=g)).float()
    dt = 0.9 * 0.01 / (2**0.5 * tide.C0)
    amp = tide.ricker(2e8, nt, dt).reshape(1, 1, -1).float().repeat(S, 1, 1)
    src = torch.tensor([[[2, 8 + 256 * s]] for s in range(S)])
    rec = torch.tensor([[[2, 2 + 4 * k] for k in range(256)]] * S)
    p = dict(epsilon=eps, sigma=sig, mu=torch.ones_like(eps), grid_spacing=0.01, dt=dt,
             source_amplitude=amp, source_location=src, receiver_location=rec, pml_width=[20] * 4)
    per = _per_shot_bytes(p)
    e = eps.to(dev).requires_grad_(True)
    s = sig.to(dev).requires_grad_(True)
    op = _operator(p, dev, limit=int(1.5 * per))
    r = op(tide.EMModel(e, s, p["mu"].to(dev))).receiver_data
    cot = _cotangent(r)
    (r * cot).sum().backward()
    print("GPU chunked gradient done; oracle pass", flush=True)
    r_o, ge_o, gs_o = _oracle_grads(p, cot.cpu().numpy())
    e_r = relative_l2(r.detach().cpu().numpy(), r_o)
    e_e = relative_l2(e.grad.cpu().numpy(), ge_o)
    e_s = relative_l2(s.grad.cpu().numpy(), gs_o)
    print(f"receivers {e_r:.2e}, grad eps {e_e:.2e}, grad sigma {e_s:.2e}")
    assert e_r <= 1e-6 and e_e <= 1e-4 and e_s <= 1e-4
