"""3-D Maxwell HIP kernels against the CPU oracle (MI355X).

The C-ABI entry points get exactly the arrays the oracle gets
(``oracle.EMCase``); the kernels keep the reference's operation order
without FMA contraction, so fields, receivers and adjoint fields are
bit-identical to the oracle and FP64 coefficient gradients agree to
rounding of the shot-sum order.  The operator tests then check
``tide.Maxwell3D`` gradients (full snapshots and checkpoint-recompute)
against the oracle plus torch's chain rule, within the stated tolerances
(fp32: receivers 1e-5, gradients 1e-4 relative L2).
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import tide
from gpu_abi import gpu_em3d_backward, gpu_em3d_forward, gpu_em3d_ref_backward, gpu_em3d_ref_forward
from helpers import FIELD_NAMES_3D, hetero3d_problem, load_golden3d, relative_l2
from oracle import oracle as O
from tide import _native
from tide.maxwell.maxwell3d import prepare_em3d

pytestmark = pytest.mark.gpu
DEV = "cuda:0"


@pytest.fixture(autouse=True)
def _default_tuning():
    _native.set_em3d_tuning(8, 4, 8)  # the library's defaults
    _native.set_em3d_forward_group(1)
    _native.set_em3d_tile_order(0)
    yield
    _native.set_em3d_tuning(8, 4, 8)
    _native.set_em3d_forward_group(1)
    _native.set_em3d_tile_order(0)


def _case(dtype=torch.float32, **kw):
    return O.em_case_from_setup(prepare_em3d(**hetero3d_problem(dtype=dtype, **kw)))


@pytest.mark.parametrize("name", ["em3d_hetero_f32", "em3d_hetero_f64"])
def test_em3d_forward_matches_reference_golden(name):
    case, expected = load_golden3d(name)
    print(f"case {name}", flush=True)
    fields, r, _ = gpu_em3d_forward(case)
    assert np.array_equal(r, expected["r"])
    for k, v in zip(FIELD_NAMES_3D, fields[:6]):
        assert np.array_equal(v, expected[k]), k


@pytest.mark.parametrize("comps", [("ex", "ez"), ("ey", "ey"), ("ez", "ex")])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_em3d_forward_bit_identical_to_oracle(comps, dtype):
    case = _case(dtype, src_comp=comps[0], rec_comp=comps[1], nz=12, ny=13, nx=70, nt=30, shots=3)
    print(f"case comps={comps} dtype={dtype} shape={case.shape}", flush=True)
    f_g, r_g, _ = gpu_em3d_forward(case)
    f_o, r_o, _ = O.em3d_forward(case)
    assert np.abs(r_o).max() > 0
    assert np.array_equal(r_g, r_o)
    for k, (a, b) in enumerate(zip(f_g, f_o)):
        assert np.array_equal(a, b), k


@pytest.mark.parametrize("store", ["full", "bf16"])
def test_em3d_storage_and_adjoint_match_oracle(store):
    case = _case(torch.float32, nz=11, ny=12, nx=66, nt=28, shots=3, src_comp="ex", rec_comp="ey")
    print(f"case store={store} shape={case.shape}", flush=True)
    f_g, r_g, s_g = gpu_em3d_forward(case, store=store)
    f_o, r_o, s_o = O.em3d_forward(case, store=store)
    assert np.array_equal(r_g, r_o)
    for a, b in zip(s_g, s_o):
        assert np.array_equal(a, b)
    rng = np.random.default_rng(11)
    w = rng.standard_normal(r_o.shape).astype(np.float32)
    gg = gpu_em3d_backward(case, w, s_o, store=store)
    go = O.em3d_backward(case, w, s_o, store=store)
    for a, b in zip(gg[3], go[3]):
        assert np.array_equal(a, b)
    assert np.array_equal(gg[2], go[2])
    assert relative_l2(gg[0], go[0]) < 1e-12
    assert relative_l2(gg[1], go[1]) < 1e-12


def test_em3d_adjoint_many_shots_and_tunings():
    """More shots than a workgroup holds (each group writes its own FP64
    gradient slice, the slices are summed in a fixed order afterwards) and
    every launch shape give the oracle's adjoint."""
    case = _case(torch.float64, nz=9, ny=10, nx=12, nt=16, shots=10)
    _, r_o, s_o = O.em3d_forward(case, store="full")
    rng = np.random.default_rng(5)
    w = rng.standard_normal(r_o.shape)
    go = O.em3d_backward(case, w, s_o)
    for fz, az, grp, order in ((8, 8, 8, 0), (4, 4, 3, 0), (16, 8, 1, 0), (8, 4, 3, 1),
                               (4, 8, 2, 1)):
        print(f"case tuning={fz},{az},{grp} tile order {order}", flush=True)
        _native.set_em3d_tuning(fz, az, grp)
        _native.set_em3d_tile_order(order)
        f_g, r_g, _ = gpu_em3d_forward(case)
        assert np.array_equal(r_g, r_o)
        gg = gpu_em3d_backward(case, w, s_o)
        for a, b in zip(gg[3], go[3]):
            assert np.array_equal(a, b)
        assert relative_l2(gg[0], go[0]) < 1e-12
        assert relative_l2(gg[1], go[1]) < 1e-12


def _operator_grads(kw, *, chunk_steps=0, compression=False, limit=None):
    dev = torch.device(DEV)
    eps = kw["epsilon"].to(dev).clone().requires_grad_(True)
    sig = kw["sigma"].to(dev).clone().requires_grad_(True)
    mu = kw["mu"].to(dev)
    out = tide.maxwell.maxwell3d.maxwell3d(
        eps, sig, mu, kw["grid_spacing"], kw["dt"], kw["source_amplitude"].to(dev),
        kw["source_location"].to(dev), kw["receiver_location"].to(dev),
        pml_width=kw["pml_width"], source_component=kw["source_component"],
        receiver_component=kw["receiver_component"], storage_chunk_steps=chunk_steps,
        storage_compression=compression, storage_bytes_limit_device=limit)
    r = out[-1]
    w = r.detach() - 0.8 * torch.roll(r.detach(), 5, dims=0)
    (r * w).sum().backward()
    return r.detach().cpu(), eps.grad.cpu(), sig.grad.cpu(), w.cpu()


def test_em3d_operator_checkpointing_equals_full_storage():
    kw = hetero3d_problem(dtype=torch.float32, nz=10, ny=11, nx=13, nt=40, shots=2,
                          src_comp="ex", rec_comp="ex")
    r1, ge1, gs1, _ = _operator_grads(kw)
    r2, ge2, gs2, _ = _operator_grads(kw, chunk_steps=7)
    assert torch.equal(r1, r2)
    assert relative_l2(ge2.numpy(), ge1.numpy()) < 1e-6
    assert relative_l2(gs2.numpy(), gs1.numpy()) < 1e-6


def test_em3d_operator_gradient_matches_oracle_chain_rule():
    kw = hetero3d_problem(dtype=torch.float32, nz=9, ny=10, nx=12, nt=36, shots=2,
                          src_comp="ez", rec_comp="ey")
    r_g, ge_g, gs_g, w = _operator_grads(kw, chunk_steps=9)
    # oracle: same forward, exact adjoint, then torch's chain rule on CPU
    eps = kw["epsilon"].clone().requires_grad_(True)
    sig = kw["sigma"].clone().requires_grad_(True)
    st = prepare_em3d(eps, sig, kw["mu"], kw["grid_spacing"], kw["dt"], kw["source_amplitude"],
                      kw["source_location"], kw["receiver_location"], pml_width=kw["pml_width"],
                      source_component=kw["source_component"],
                      receiver_component=kw["receiver_component"])
    case = O.em_case_from_setup(st)
    _, r_o, s_o = O.em3d_forward(case, store="full")
    assert relative_l2(r_g.numpy(), r_o) < 1e-5
    gca, gcb, gf, _ = O.em3d_backward(case, w.numpy(), s_o)
    torch.autograd.backward(
        [st.ca, st.cb, st.f],
        [torch.from_numpy(gca).reshape(st.ca.shape).float(),
         torch.from_numpy(gcb).reshape(st.cb.shape).float(), torch.from_numpy(gf).float()])
    assert relative_l2(ge_g.numpy(), eps.grad.numpy()) < 1e-4
    assert relative_l2(gs_g.numpy(), sig.grad.numpy()) < 1e-4


def test_em3d_operator_bf16_gradient_bound():
    kw = hetero3d_problem(dtype=torch.float32, nz=9, ny=10, nx=12, nt=36, shots=2)
    _, ge1, gs1, _ = _operator_grads(kw)
    _, ge2, gs2, _ = _operator_grads(kw, compression="bf16")
    assert relative_l2(ge2.numpy(), ge1.numpy()) < 1e-2
    cos = float((ge1 * ge2).sum() / (ge1.norm() * ge2.norm()))
    assert cos > 0.999


def _checkpoint_bytes(kw, seg):
    """Per-shot bytes of checkpoints plus one segment's snapshots."""
    from tide.maxwell.maxwell3d import _state_bytes_per_shot

    st = prepare_em3d(kw["epsilon"], kw["sigma"], kw["mu"], kw["grid_spacing"], kw["dt"],
                      kw["source_amplitude"], kw["source_location"], kw["receiver_location"],
                      pml_width=kw["pml_width"])
    numel = st.nz * st.ny * st.nx
    return -(-st.nt // seg) * _state_bytes_per_shot(st) + seg * 6 * numel * 4


def test_em3d_operator_shot_chunks_equal_one_chunk():
    """A device limit that holds the checkpoints of two of five shots: the
    shots run in chunks (one resident, the rest re-run in the backward)."""
    kw = hetero3d_problem(dtype=torch.float32, nz=10, ny=11, nx=13, nt=40, shots=5,
                          src_comp="ex", rec_comp="ez")
    per = _checkpoint_bytes(kw, 7)
    r1, ge1, gs1, _ = _operator_grads(kw, chunk_steps=7)
    r2, ge2, gs2, _ = _operator_grads(kw, chunk_steps=7, limit=int(2.3 * per))
    assert torch.equal(r1, r2)
    assert relative_l2(ge2.numpy(), ge1.numpy()) < 1e-6
    assert relative_l2(gs2.numpy(), gs1.numpy()) < 1e-6


@pytest.mark.parametrize("limit_shots", [None, 2.3])
def test_em3d_separable_objective_matches_operator(limit_shots):
    """``ReceiverObjective(d_obs).gradient(op.linearize(m))`` on a 3-D
    operator: chunk by chunk, checkpointed segments, FP64 misfit."""
    from tide.workflow import ReceiverObjective

    dev = torch.device(DEV)
    kw = hetero3d_problem(dtype=torch.float32, nz=10, ny=11, nx=13, nt=40, shots=5,
                          src_comp="ey", rec_comp="ey")
    limit = None if limit_shots is None else int(limit_shots * _checkpoint_bytes(kw, 7))
    op = tide.Maxwell3D(
        tide.Discretization(spacing=kw["grid_spacing"], dt=kw["dt"],
                            boundary=tide.CPML(tuple(kw["pml_width"]))),
        tide.Experiment(tide.Acquisition(kw["source_location"].to(dev), kw["receiver_location"].to(dev)),
                        kw["source_amplitude"].to(dev), source_component="ey",
                        receiver_component="ey"),
        storage=tide.StorageOptions(bytes_limit_device=limit, chunk_steps=7),
    )
    model = tide.EMModel(kw["epsilon"].to(dev), kw["sigma"].to(dev), kw["mu"].to(dev))
    with torch.no_grad():
        eps_obs = kw["epsilon"].clone()
        eps_obs[3:6, 4:7, 4:8] += 2.0
        d_obs = op(tide.EMModel(eps_obs.to(dev), kw["sigma"].to(dev), kw["mu"].to(dev))).receiver_data
    with op.linearize(model) as lin:
        value, grad = ReceiverObjective(d_obs).value_and_gradient(lin)
        r_fast = lin.primal.receiver_data.detach().cpu()
    e = kw["epsilon"].to(dev).requires_grad_(True)
    s = kw["sigma"].to(dev).requires_grad_(True)
    r = op(tide.EMModel(e, s, kw["mu"].to(dev))).receiver_data
    loss = 0.5 * ((r.double() - d_obs.double()) ** 2).sum()
    loss.backward()
    assert torch.equal(r_fast, r.detach().cpu())
    assert abs(float(value) - float(loss)) <= 1e-12 * abs(float(loss))
    assert relative_l2(grad.epsilon.cpu().numpy(), e.grad.cpu().numpy()) < 1e-6
    assert relative_l2(grad.sigma.cpu().numpy(), s.grad.cpu().numpy()) < 1e-6


def test_em3d_batched_models_per_shot_gradients():
    """A [2, nz, ny, nx] model (one per shot): each model's gradient equals
    the gradient of its own single-model run (per-shot accumulators)."""
    dev = torch.device(DEV)
    k0 = hetero3d_problem(dtype=torch.float32, nz=9, ny=10, nx=12, nt=36, shots=1, seed=3)
    k1 = hetero3d_problem(dtype=torch.float32, nz=9, ny=10, nx=12, nt=36, shots=1, seed=4)
    mm = tide.maxwell.maxwell3d.maxwell3d
    common = dict(grid_spacing=k0["grid_spacing"], dt=k0["dt"], pml_width=k0["pml_width"],
                  source_component="ey", receiver_component="ey")
    eps = torch.stack([k0["epsilon"], k1["epsilon"]]).to(dev).requires_grad_(True)
    sig = torch.stack([k0["sigma"], k1["sigma"]]).to(dev).requires_grad_(True)
    amp = torch.cat([k0["source_amplitude"], k1["source_amplitude"]]).to(dev)
    src = torch.cat([k0["source_location"], k1["source_location"]]).to(dev)
    rec = torch.cat([k0["receiver_location"], k1["receiver_location"]]).to(dev)
    r = mm(eps, sig, torch.ones_like(eps), source_amplitude=amp, source_location=src,
           receiver_location=rec, **common)[-1]
    (r.double() ** 2).sum().backward()
    for b, k in enumerate((k0, k1)):
        e1 = k["epsilon"].to(dev).requires_grad_(True)
        s1 = k["sigma"].to(dev).requires_grad_(True)
        r1 = mm(e1, s1, k["mu"].to(dev), source_amplitude=k["source_amplitude"].to(dev),
                source_location=k["source_location"].to(dev),
                receiver_location=k["receiver_location"].to(dev), **common)[-1]
        assert torch.equal(r[:, b], r1[:, 0])
        (r1.double() ** 2).sum().backward()
        assert relative_l2(eps.grad[b].cpu().numpy(), e1.grad.cpu().numpy()) < 1e-6
        assert relative_l2(sig.grad[b].cpu().numpy(), s1.grad.cpu().numpy()) < 1e-6


@pytest.mark.parametrize("name", ["em3d_hetero_f32", "em3d_hetero_f64"])
def test_reference_abi_3d_matches_oracle(name):
    """``maxwell_3d_2_<dtype>_{forward,forward_with_storage,backward}_cuda``
    called with the reference's arguments (compact CPML memories, 18 + 18
    pointers): fields, receivers and snapshots equal the oracle, the
    coefficient gradients match the oracle's FP64 sums, grad_eps /
    grad_sigma follow the closed-form chain rule, and zero_grad_on_entry
    clears (or keeps, and adds to) the gradient buffers."""
    case, expected = load_golden3d(name)
    print(f"case {name}", flush=True)
    f64 = case.dtype == np.float64
    tol, gtol = (1e-12, 1e-10) if f64 else (1e-6, 1e-5)
    f_o, r_o, s_o = O.em3d_forward(case, store="full")
    f_g, r_g, _ = gpu_em3d_ref_forward(case)
    assert relative_l2(r_g, r_o) <= tol
    assert relative_l2(r_g, expected["r"]) <= tol
    for a, b in zip(f_g, f_o):
        if np.any(b):
            assert relative_l2(a, b) <= tol
    f_s, r_s, s_g = gpu_em3d_ref_forward(case, store="full")
    assert np.array_equal(r_s, r_g)
    for a, b in zip(s_g, s_o):
        assert relative_l2(a, b) <= tol
    w = np.random.default_rng(5).standard_normal(r_o.shape).astype(case.dtype)
    gca_o, gcb_o, gf_o, _ = O.em3d_backward(case, w, s_o)
    gca, gcb, geps, gsig, gf = gpu_em3d_ref_backward(case, w, s_o)
    assert relative_l2(gca, gca_o) <= gtol
    assert relative_l2(gcb, gcb_o) <= gtol
    assert relative_l2(gf, gf_o) <= gtol
    ca, cb = case.ca[0].astype(np.float64), case.cb[0].astype(np.float64)
    dt = case.dt
    eps_ref = (gca * (1 - ca) * cb / dt - gcb * cb**2 / dt) * 8.8541878128e-12
    sig_ref = -0.5 * gca * (1 + ca) * cb - 0.5 * gcb * cb**2
    assert relative_l2(geps, eps_ref) <= (1e-12 if f64 else 1e-5)
    assert relative_l2(gsig, sig_ref) <= (1e-12 if f64 else 1e-5)
    # without zero_grad_on_entry the gradient is added to what the buffer held
    base = np.full_like(gca, 0.25)
    gca2, *_ = gpu_em3d_ref_backward(case, w, s_o, grad_ca0=base, zero_grad_on_entry=False)
    assert relative_l2(gca2 - base, gca) <= (1e-12 if f64 else 1e-5)


@pytest.mark.parametrize("chunk_steps", [0, 9])
def test_em3d_gradient_matches_central_differences_on_16_directions(chunk_steps):
    """The HIP gradient of J = 0.5 |d(m) - d_obs|^2 (FP64 fields, exact
    adjoint, chain rule to eps and sigma) against central differences of J
    along 16 seeded random directions (dot products <grad, v>), with whole
    snapshots and with 9-step checkpoint segments.  A fixed max velocity
    keeps the CPML profiles independent of the model, so the differences
    see only what the gradient differentiates.  Pattern of the reference's
    tests/numerical/test_gradients.py:514-635."""
    dev = torch.device(DEV)
    kw = hetero3d_problem(dtype=torch.float64, nz=14, ny=15, nx=16, nt=70, shots=2,
                          src_comp="ex", rec_comp="ez", seed=31)
    c = lambda t: t.to(dev)  # noqa: E731
    disc = tide.Discretization(spacing=kw["grid_spacing"], dt=kw["dt"], max_velocity=tide.C0,
                               boundary=tide.CPML(tuple(kw["pml_width"])))
    exp = tide.Experiment(tide.Acquisition(c(kw["source_location"]), c(kw["receiver_location"])),
                          c(kw["source_amplitude"]), source_component="ex", receiver_component="ez")
    op = tide.Maxwell3D(disc, exp, storage=tide.StorageOptions(chunk_steps=chunk_steps))
    eps, sig, mu = c(kw["epsilon"]), c(kw["sigma"]), c(kw["mu"])
    with torch.no_grad():
        observed = op(tide.EMModel(eps * 1.05, sig, mu)).receiver_data

    def misfit(e, s):
        with torch.no_grad():
            d = op(tide.EMModel(e, s, mu)).receiver_data
        return 0.5 * float(((d - observed) ** 2).sum())

    e1 = eps.clone().requires_grad_(True)
    s1 = sig.clone().requires_grad_(True)
    d = op(tide.EMModel(e1, s1, mu)).receiver_data
    (0.5 * ((d - observed) ** 2).sum()).backward()
    g = torch.Generator().manual_seed(1600 + chunk_steps)
    h = 1e-4
    worst = 0.0
    for k in range(16):
        ve = c(0.5 * torch.randn(eps.shape, generator=g, dtype=torch.float64))
        vs = c(2e-3 * torch.randn(sig.shape, generator=g, dtype=torch.float64))
        fd = (misfit(eps + h * ve, sig + h * vs) - misfit(eps - h * ve, sig - h * vs)) / (2 * h)
        an = float((e1.grad * ve).sum() + (s1.grad * vs).sum())
        err = abs(an - fd) / abs(fd)
        worst = max(worst, err)
        print(f"direction {k}: adjoint {an:.10e} central difference {fd:.10e} rel {err:.2e}",
              flush=True)
    assert worst <= 1e-6
