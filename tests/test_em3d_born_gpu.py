"""3-D Born forward on the GPU (reference maxwell/maxwell3d_born.py,
csrc/em3d/maxwell_3d_inst.cpp:1100-1527).

* ``maxwell_3d_2_<dtype>_born_forward_cuda`` (reference argument order,
  compact CPML memories) against the oracle's Born forward, itself pinned bit
  for bit to the reference's CPU build: background and scattered fields,
  memories and both receiver sets bit-identical, fp32 and fp64, every
  source / receiver component pairing.
* ``tide.Maxwell3D(...).linearize(model).jvp(direction)`` against central
  differences of the operator's forward in FP64 along seeded directions in
  epsilon and sigma (relative L2 1e-6), and its linearity in the direction.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import tide
from gpu_abi import gpu_em3d_ref_born_forward
from helpers import hetero3d_problem, relative_l2
from oracle import oracle as O
from tide.maxwell.maxwell3d import prepare_em3d

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    tide._native.load()


def _direction(case, seed):
    rng = np.random.default_rng(seed)
    dca = (rng.standard_normal(case.shape) * 1e-2).astype(case.dtype)
    dcb = (rng.standard_normal(case.shape) * float(np.abs(case.cb).mean()) * 1e-2).astype(case.dtype)
    df = (rng.standard_normal(case.f.shape) * float(np.abs(case.f).max()) * 1e-2).astype(case.dtype)
    return dca, dcb, df


@pytest.mark.parametrize("comps", [("ex", "ez"), ("ey", "ey"), ("ez", "ex")])
@pytest.mark.parametrize("dtype", [torch.float32, torch.float64])
def test_born_forward_abi_matches_oracle(comps, dtype):
    print(f"3-D Born {comps} {dtype}", flush=True)
    case = O.em_case_from_setup(prepare_em3d(**hetero3d_problem(
        dtype=dtype, src_comp=comps[0], rec_comp=comps[1], seed=6, nt=40)))
    dca, dcb, df = _direction(case, 4)
    f_o, df_o, r_o, bg_o = O.em3d_born_forward(case, dca, dcb, df)
    f_g, df_g, r_g, bg_g = gpu_em3d_ref_born_forward(case, dca, dcb, df)
    assert np.abs(r_o).max() > 0 and np.abs(bg_o).max() > 0
    assert np.array_equal(r_g, r_o) and np.array_equal(bg_g, bg_o)
    for a, b in zip(f_g + df_g, f_o + df_o):
        assert np.array_equal(a, b)


def test_born_forward_time_chunked_calls_equal_one_call():
    """Consecutive calls over [0, 13), [13, 27), [27, 40) (``start_t``
    offsets the source and receiver rows) give the one-call results bit for
    bit, CPML memories of both fields included."""
    case = O.em_case_from_setup(prepare_em3d(**hetero3d_problem(
        dtype=torch.float32, src_comp="ez", rec_comp="ex", seed=11, nt=40)))
    dca, dcb, df = _direction(case, 5)
    one = gpu_em3d_ref_born_forward(case, dca, dcb, df)
    parts = gpu_em3d_ref_born_forward(case, dca, dcb, df, segments=(13, 14, 13))
    assert np.abs(one[2]).max() > 0
    assert np.array_equal(one[2], parts[2]) and np.array_equal(one[3], parts[3])
    for a, b in zip(one[0] + one[1], parts[0] + parts[1]):
        assert np.array_equal(a, b)


def _operator(problem, nt):
    d = tide.Discretization(spacing=problem["grid_spacing"], dt=problem["dt"],
                            boundary=tide.CPML(tuple(problem["pml_width"])))
    acq = tide.Acquisition(problem["source_location"].to(DEV),
                           problem["receiver_location"].to(DEV))
    e = tide.Experiment(acq, problem["source_amplitude"].to(DEV), nt=nt,
                        source_component=problem["source_component"],
                        receiver_component=problem["receiver_component"])
    return tide.Maxwell3D(d, e)


def test_operator_jvp_matches_central_differences():
    nt = 36
    p = hetero3d_problem(dtype=torch.float64, src_comp="ex", rec_comp="ey", seed=9, nt=nt)
    op = _operator(p, nt)
    eps, sig, mu = (p[k].to(DEV) for k in ("epsilon", "sigma", "mu"))
    g = torch.Generator().manual_seed(7100)
    de = (torch.randn(eps.shape, generator=g, dtype=torch.float64) * 0.1).to(DEV)
    ds = (torch.randn(eps.shape, generator=g, dtype=torch.float64) * 1e-3).to(DEV)
    model = tide.EMModel(eps, sig, mu)
    lin = op.linearize(model)
    res = lin.jvp(tide.EMDirection(epsilon=de, sigma=ds))
    r = res.receiver_data
    h = 1e-4

    def fwd(s):
        with torch.no_grad():
            return op(tide.EMModel(eps + s * h * de, sig + s * h * ds, mu)).receiver_data

    fd = (fwd(+1) - fwd(-1)) / (2 * h)
    err = relative_l2(r.cpu().numpy(), fd.cpu().numpy())
    print(f"jvp vs central differences: {err:.2e}", flush=True)
    assert float(fd.abs().max()) > 0
    assert err < 1e-6
    # linear in the direction
    r2 = lin.jvp(tide.EMDirection(epsilon=2.0 * de, sigma=2.0 * ds)).receiver_data
    assert relative_l2(r2.cpu().numpy(), 2.0 * r.cpu().numpy()) < 1e-12
    # the scattered final state has the layout of the operator's own state
    with torch.no_grad():
        primal = op(model).final_state
    for name in ("Ex", "Hz", "m_ey_x"):
        assert getattr(res.final_state, name).shape == getattr(primal, name).shape
