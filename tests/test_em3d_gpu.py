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
from gpu_abi import gpu_em3d_backward, gpu_em3d_forward
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
    yield
    _native.set_em3d_tuning(8, 4, 8)
    _native.set_em3d_forward_group(1)


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
    """More shots than a workgroup holds (FP64 atomics across groups) and
    every launch shape give the oracle's adjoint."""
    case = _case(torch.float64, nz=9, ny=10, nx=12, nt=16, shots=10)
    _, r_o, s_o = O.em3d_forward(case, store="full")
    rng = np.random.default_rng(5)
    w = rng.standard_normal(r_o.shape)
    go = O.em3d_backward(case, w, s_o)
    for fz, az, grp in ((8, 8, 8), (4, 4, 3), (16, 8, 1)):
        print(f"case tuning={fz},{az},{grp}", flush=True)
        _native.set_em3d_tuning(fz, az, grp)
        f_g, r_g, _ = gpu_em3d_forward(case)
        assert np.array_equal(r_g, r_o)
        gg = gpu_em3d_backward(case, w, s_o)
        for a, b in zip(gg[3], go[3]):
            assert np.array_equal(a, b)
        assert relative_l2(gg[0], go[0]) < 1e-12
        assert relative_l2(gg[1], go[1]) < 1e-12


def _operator_grads(kw, *, chunk_steps=0, compression=False):
    dev = torch.device(DEV)
    eps = kw["epsilon"].to(dev).clone().requires_grad_(True)
    sig = kw["sigma"].to(dev).clone().requires_grad_(True)
    mu = kw["mu"].to(dev)
    out = tide.maxwell.maxwell3d.maxwell3d(
        eps, sig, mu, kw["grid_spacing"], kw["dt"], kw["source_amplitude"].to(dev),
        kw["source_location"].to(dev), kw["receiver_location"].to(dev),
        pml_width=kw["pml_width"], source_component=kw["source_component"],
        receiver_component=kw["receiver_component"], storage_chunk_steps=chunk_steps,
        storage_compression=compression)
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
