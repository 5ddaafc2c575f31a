"""Debye-dispersive 2-D TM forward on the GPU (reference native Debye path,
src/tide/csrc/tm2d/maxwell_tm_cuda_inst.cu:382-457; CPU loops
maxwell_tm_inst.cpp:214-300, 780-805).

Through ``maxwelltm(..., dispersion=DebyeDispersion(...))`` and the
``MaxwellTM`` operator with ``EMModel(dispersion=...)``: receivers and
final fields bit-identical to the oracle's Debye forward (itself
bit-identical to the reference's CPU Debye forward built from its sources,
tests/test_oracle.py), for one and two poles, fp32 and fp64, stencils 2
and 4.  Gradients with dispersion are refused, as by the reference's native
backend.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import tide
from helpers import hetero_problem
from oracle import oracle as O
from tide.maxwell.tm2d import maxwelltm, prepare_tm2d

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    tide._native.load()


def _dispersion(shape, poles, dtype, seed):
    g = torch.Generator().manual_seed(seed)
    d_eps = (2.0 + 3.0 * torch.rand((poles, *shape), generator=g, dtype=torch.float64)).to(dtype)
    tau = torch.tensor([1e-9, 3e-10][:poles], dtype=torch.float64).reshape(poles, 1, 1)
    return tide.DebyeDispersion(d_eps, tau.expand(poles, *shape).to(dtype))


@pytest.mark.parametrize("dtype,stencil,poles", [(torch.float64, 2, 2), (torch.float32, 2, 1),
                                                 (torch.float32, 4, 2), (torch.float64, 4, 1)])
def test_debye_forward_matches_oracle(dtype, stencil, poles):
    p = hetero_problem(seed=50 + stencil + poles, dtype=dtype, ny=40, nx=46, nt=90, shots=3)
    disp = _dispersion(p["epsilon"].shape, poles, dtype, seed=poles)
    dev = {k: (v.to(DEV) if torch.is_tensor(v) and k not in ("source_location", "receiver_location")
               else v) for k, v in p.items()}
    dev_disp = tide.DebyeDispersion(disp.delta_epsilon.to(DEV), disp.tau.to(DEV))
    # the oracle runs on the operator's own setup, prepared on the device
    # (torch's CPML profile arithmetic on the CPU and the GPU differs in the
    # last bits)
    st = prepare_tm2d(**dev, stencil=stencil, dispersion=dev_disp)
    case = O.case_from_setup(st)
    deb = [st.extra["debye"][k].cpu().numpy() for k in ("a", "b", "cp")]
    f_o, r_o, _, _ = O.tm2d_forward_debye(case, *deb)
    with torch.no_grad():
        out = maxwelltm(**dev, stencil=stencil, dispersion=dev_disp)
    print(f"dtype {dtype} stencil {stencil} poles {poles}: |r| {np.abs(r_o).max():.3e}", flush=True)
    assert np.abs(r_o).max() > 0
    assert np.array_equal(out[-1].cpu().numpy(), r_o)
    for a, b in zip(out[:7], f_o):
        assert np.array_equal(a.cpu().numpy(), st.unpad(torch.from_numpy(b)).numpy())


def test_debye_operator_and_gradient_refusal():
    p = hetero_problem(seed=61, dtype=torch.float32, ny=30, nx=36, nt=60, shots=2)
    disp = _dispersion(p["epsilon"].shape, 1, torch.float32, seed=9)
    disp = tide.DebyeDispersion(disp.delta_epsilon.to(DEV), disp.tau.to(DEV))
    op = tide.MaxwellTM(
        tide.Discretization(spacing=p["grid_spacing"], dt=p["dt"],
                            boundary=tide.CPML(tuple(p["pml_width"]))),
        tide.Experiment(tide.Acquisition(p["source_location"].to(DEV), p["receiver_location"].to(DEV)),
                        p["source_amplitude"].to(DEV)))
    eps, sig, mu = (p[k].to(DEV) for k in ("epsilon", "sigma", "mu"))
    with torch.no_grad():
        r_op = op(tide.EMModel(eps, sig, mu, dispersion=disp)).receiver_data
        r_fn = maxwelltm(eps, sig, mu, p["grid_spacing"], p["dt"], p["source_amplitude"].to(DEV),
                              p["source_location"].to(DEV), p["receiver_location"].to(DEV),
                              pml_width=p["pml_width"], dispersion=disp)[-1]
        r_plain = op(tide.EMModel(eps, sig, mu)).receiver_data
    assert torch.equal(r_op, r_fn)
    assert not torch.equal(r_op, r_plain)  # the poles change the response
    with pytest.raises(NotImplementedError, match="forward inference"):
        op(tide.EMModel(eps.clone().requires_grad_(True), sig, mu, dispersion=disp))
