"""One-launch 3-D forward step (H and E updates of a time step in one
kernel) against the oracle and the two-kernel step.

The kernel recomputes a halo (the row above the tile, the left column and
the plane before the march) and ping-pongs the fields and H-phase CPML
memories through a second copy; every cell update keeps the two-kernel
operation order, so fields, receivers and snapshots are bit-identical to the
oracle.  Cases cover both tile heights and march lengths on grids whose
sides are not multiples of them, several x tiles, fp32 and fp64, an odd and
an even number of steps (the final state comes back from the second copy or
not), shot groups, full and BF16 snapshots, the reference golden data, and
the operator's checkpointed and batched-model gradients (equal to the
two-kernel step's bitwise).
"""

from __future__ import annotations

import contextlib

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
SHAPES = [(2, 8), (4, 8), (2, 16), (4, 16)]


@pytest.fixture(autouse=True)
def _restore():
    before = _native.get_em3d_fused()
    yield
    _native.set_em3d_fused(*before)
    _native.set_em3d_tuning(8, 4, 8)
    _native.set_em3d_forward_group(1)


@contextlib.contextmanager
def fused(rows, planes=8):
    before = _native.get_em3d_fused()
    _native.set_em3d_fused(rows, planes)
    try:
        yield
    finally:
        _native.set_em3d_fused(*before)


def _case(dtype=torch.float32, **kw):
    return O.em_case_from_setup(prepare_em3d(**hetero3d_problem(dtype=dtype, **kw)))


@pytest.mark.parametrize("name", ["em3d_hetero_f32", "em3d_hetero_f64"])
@pytest.mark.parametrize("shape", SHAPES)
def test_fused_forward_matches_reference_golden(name, shape):
    case, expected = load_golden3d(name)
    print(f"case {name} fused {shape}", flush=True)
    with fused(*shape):
        fields, r, _ = gpu_em3d_forward(case)
    assert np.array_equal(r, expected["r"])
    for k, v in zip(FIELD_NAMES_3D, fields[:6]):
        assert np.array_equal(v, expected[k]), k


@pytest.mark.parametrize("shape", SHAPES)
@pytest.mark.parametrize("dtype,comps,nt", [(torch.float32, ("ex", "ez"), 29),
                                            (torch.float32, ("ez", "ey"), 30),
                                            (torch.float64, ("ey", "ex"), 27)])
def test_fused_forward_bit_identical_to_oracle(shape, dtype, comps, nt):
    case = _case(dtype, src_comp=comps[0], rec_comp=comps[1], nz=21, ny=13, nx=70, nt=nt, shots=3)
    print(f"case fused {shape} {dtype} {comps} nt {nt} shape {case.shape}", flush=True)
    f_o, r_o, _ = O.em3d_forward(case)
    assert np.abs(r_o).max() > 0
    for group in (1, 3):
        _native.set_em3d_forward_group(group)
        with fused(*shape):
            f_g, r_g, _ = gpu_em3d_forward(case)
        assert np.array_equal(r_g, r_o)
        for k, (a, b) in enumerate(zip(f_g, f_o)):
            assert np.array_equal(a, b), (group, k)


@pytest.mark.parametrize("store", ["full", "bf16"])
@pytest.mark.parametrize("shape", [(2, 8), (4, 16)])
def test_fused_snapshots_and_adjoint_match_oracle(store, shape):
    case = _case(torch.float32, nz=19, ny=12, nx=66, nt=28, shots=3, src_comp="ex", rec_comp="ey")
    print(f"case fused {shape} store {store}", flush=True)
    f_o, r_o, s_o = O.em3d_forward(case, store=store)
    with fused(*shape):
        f_g, r_g, s_g = gpu_em3d_forward(case, store=store)
    assert np.array_equal(r_g, r_o)
    for a, b in zip(s_g, s_o):
        assert np.array_equal(a, b)
    for a, b in zip(f_g, f_o):
        assert np.array_equal(a, b)
    w = np.random.default_rng(11).standard_normal(r_o.shape).astype(np.float32)
    gg = gpu_em3d_backward(case, w, s_g, store=store)
    go = O.em3d_backward(case, w, s_o, store=store)
    assert np.array_equal(gg[2], go[2])
    assert relative_l2(gg[0], go[0]) < 1e-12 and relative_l2(gg[1], go[1]) < 1e-12


def _operator_run(kw, chunk_steps):
    dev = torch.device(DEV)
    eps = kw["epsilon"].to(dev).clone().requires_grad_(True)
    sig = kw["sigma"].to(dev).clone().requires_grad_(True)
    r = tide.maxwell.maxwell3d.maxwell3d(
        eps, sig, kw["mu"].to(dev), kw["grid_spacing"], kw["dt"],
        kw["source_amplitude"].to(dev), kw["source_location"].to(dev),
        kw["receiver_location"].to(dev), pml_width=kw["pml_width"],
        source_component=kw["source_component"], receiver_component=kw["receiver_component"],
        storage_chunk_steps=chunk_steps)[-1]
    (r.double() ** 2).sum().backward()
    return r.detach().cpu(), eps.grad.cpu(), sig.grad.cpu()


@pytest.mark.parametrize("chunk_steps", [0, 7])
def test_fused_operator_gradient_equals_two_kernel_step(chunk_steps):
    """Full snapshots and checkpoint-recompute segments of 7 steps: the
    one-launch forward changes no bit of the receivers or gradients."""
    kw = hetero3d_problem(dtype=torch.float32, nz=18, ny=11, nx=13, nt=40, shots=2,
                          src_comp="ex", rec_comp="ex")
    with fused(0):
        r0, ge0, gs0 = _operator_run(kw, chunk_steps)
    with fused(4, 8):
        r1, ge1, gs1 = _operator_run(kw, chunk_steps)
    assert torch.equal(r0, r1)
    assert torch.equal(ge0, ge1) and torch.equal(gs0, gs1)


def test_fused_batched_models_equal_two_kernel_step():
    """Per-shot models ([2, nz, ny, nx] coefficients) through the fused step."""
    dev = torch.device(DEV)
    k0 = hetero3d_problem(dtype=torch.float32, nz=17, ny=10, nx=12, nt=30, shots=1, seed=3)
    k1 = hetero3d_problem(dtype=torch.float32, nz=17, ny=10, nx=12, nt=30, shots=1, seed=4)
    common = dict(grid_spacing=k0["grid_spacing"], dt=k0["dt"], pml_width=k0["pml_width"],
                  source_component="ey", receiver_component="ey")
    eps = torch.stack([k0["epsilon"], k1["epsilon"]]).to(dev)
    sig = torch.stack([k0["sigma"], k1["sigma"]]).to(dev)
    amp = torch.cat([k0["source_amplitude"], k1["source_amplitude"]]).to(dev)
    src = torch.cat([k0["source_location"], k1["source_location"]]).to(dev)
    rec = torch.cat([k0["receiver_location"], k1["receiver_location"]]).to(dev)
    outs = []
    for shape in ((0, 8), (2, 16)):
        with fused(*shape), torch.no_grad():
            outs.append(tide.maxwell.maxwell3d.maxwell3d(
                eps, sig, torch.ones_like(eps), source_amplitude=amp, source_location=src,
                receiver_location=rec, **common))
    for a, b in zip(*outs):
        assert torch.equal(a, b)
