"""Host-staged snapshot storage and wavefield callbacks of the 2-D operator.

Reference pattern: ``tests/backends/test_storage.py:230-273`` (every
storage mode gives the same gradient) and the callback chunking of
``maxwell/tm2d_autograd.py:165-261, 677-802``.  Here the snapshots are the
same bytes whichever way they travel (HBM history, pinned host history via
a three-step device ring, or a file behind a pinned host ring), so
receivers and gradients must be bit-identical across modes; the oracle
pins the device mode.  Time-chunked sweeps (callbacks) change only the
pairing of the two-step launches, so their fields and receivers are
bit-identical; their FP64 coefficient gradients differ in summation order
only (~1e-16), which can move a cell of the fp32 epsilon / sigma gradient
by one unit in the last place: held to 1e-6.
"""

from __future__ import annotations

import numpy as np
import pytest
import torch

import tide
from helpers import hetero_problem, relative_l2
from oracle import oracle as O
from tide.maxwell.tm2d import maxwelltm, prepare_tm2d

pytestmark = pytest.mark.gpu
DEV = torch.device("cuda:0")


@pytest.fixture(scope="module", autouse=True)
def _gpu():
    if not torch.cuda.is_available():
        pytest.fail("GPU tests need a HIP device")
    tide._native.load()


def _grad_run(p, tmp_path, **kw):
    e = p["epsilon"].to(DEV).requires_grad_(True)
    s = p["sigma"].to(DEV).requires_grad_(True)
    args = {k: (v.to(DEV) if isinstance(v, torch.Tensor) else v) for k, v in p.items()
            if k not in ("epsilon", "sigma")}
    out = maxwelltm(e, s, storage_path=str(tmp_path), **args, **kw)
    r = out[-1]
    cot = (r - 0.7 * torch.roll(r, 5, dims=0)).detach()
    (r * cot).sum().backward()
    torch.cuda.synchronize()
    return (r.detach().cpu().numpy(), e.grad.cpu().numpy(), s.grad.cpu().numpy(),
            [t.detach().cpu().numpy() for t in out[:-1]])


@pytest.mark.parametrize("compression", [False, "bf16"])
@pytest.mark.parametrize("interval", [1, 3])
def test_storage_modes_give_identical_gradients(tmp_path, compression, interval):
    p = hetero_problem(shots=3, ny=36, nx=44, nt=141)  # odd nt: a one-step tail
    print(f"compression {compression} interval {interval}", flush=True)
    res = {}
    for mode in ("device", "cpu", "disk"):
        res[mode] = _grad_run(p, tmp_path, storage_mode=mode, storage_compression=compression,
                              model_gradient_sampling_interval=interval)
    for mode in ("cpu", "disk"):
        for a, b in zip(res[mode][:3], res["device"][:3]):
            assert np.array_equal(a, b), mode
    assert list(tmp_path.iterdir()) == []  # the disk mode's temporary files are removed


def test_device_mode_gradient_matches_oracle(tmp_path):
    """Pins the mode the others are compared with."""
    p = hetero_problem(shots=3, ny=36, nx=44, nt=141)
    r, ge, gs, _ = _grad_run(p, tmp_path, storage_mode="cpu")
    eps_c = p["epsilon"].clone().requires_grad_(True)
    sig_c = p["sigma"].clone().requires_grad_(True)
    st = prepare_tm2d(eps_c, sig_c, p["mu"], p["grid_spacing"], p["dt"], p["source_amplitude"],
                      p["source_location"], p["receiver_location"], pml_width=p["pml_width"])
    case = O.case_from_setup(st)
    _, r_o, es, cs = O.tm2d_forward(case, store="full")
    rt = torch.from_numpy(r_o)
    cot = (rt - 0.7 * torch.roll(rt, 5, dims=0)).numpy()
    gca, gcb, gf, _ = O.tm2d_backward(case, cot, es, cs)
    torch.autograd.backward([st.ca, st.cb, st.f],
                            [torch.from_numpy(gca).float()[None], torch.from_numpy(gcb).float()[None],
                             torch.from_numpy(gf)])
    assert relative_l2(r, r_o) <= 1e-6
    assert relative_l2(ge, eps_c.grad.numpy()) <= 1e-4
    assert relative_l2(gs, sig_c.grad.numpy()) <= 1e-4


def test_auto_mode_picks_host_staging_under_a_device_limit(tmp_path):
    p = hetero_problem(shots=2, ny=30, nx=34, nt=60)
    ref = _grad_run(p, tmp_path)
    for limit_host in (None, 1):  # cpu, then disk (one byte of host budget)
        got = _grad_run(p, tmp_path, storage_mode="auto", storage_bytes_limit_device=1,
                        storage_bytes_limit_host=limit_host)
        for a, b in zip(got[:3], ref[:3]):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("every", [7, 40])
def test_callbacks_see_every_chunk_and_leave_results_unchanged(tmp_path, every):
    p = hetero_problem(shots=2, ny=36, nx=44, nt=101)
    seen_f, seen_b = [], []

    def fwd(state):
        assert not state.is_backward
        ey = state.get_wavefield("Ey", view="full")
        seen_f.append((state.step, ey.detach().clone().cpu().numpy(),
                       tuple(state.get_wavefield("Ey").shape)))

    def bwd(state):
        assert state.is_backward
        seen_b.append((state.step, state.get_gradient("epsilon", view="full").clone().cpu().numpy(),
                       state.get_wavefield("lambda_Ey", view="full").clone().cpu().numpy()))

    ref = _grad_run(p, tmp_path)
    got = _grad_run(p, tmp_path, forward_callback=fwd, backward_callback=bwd,
                    callback_frequency=every)
    nt = 101
    assert [s for s, _, _ in seen_f] == list(range(every, nt, every)) + ([nt] if nt % every else [])
    assert seen_f[-1][2] == (2, 36, 44)  # inner view: the model region
    assert np.array_equal(got[0], ref[0])  # receivers
    for a, b in zip(got[3], ref[3]):
        assert np.array_equal(a, b)  # final state
    assert relative_l2(got[1], ref[1]) <= 1e-6 and relative_l2(got[2], ref[2]) <= 1e-6
    # the forward callback saw the state of a plain run stopped at that step
    step0, ey0, _ = seen_f[0]
    part = {k: (v.to(DEV) if isinstance(v, torch.Tensor) else v) for k, v in p.items()}
    with torch.no_grad():
        out = maxwelltm(**part, nt=step0)
    st = prepare_tm2d(p["epsilon"], p["sigma"], p["mu"], p["grid_spacing"], p["dt"],
                      p["source_amplitude"], p["source_location"], p["receiver_location"],
                      pml_width=p["pml_width"], allocate_fields=False)
    pad = st.fd_pad  # the operator returns the fields without the stencil halo
    no_halo = ey0[:, pad[0]: st.ny - pad[1] if pad[1] else None,
                  pad[2]: st.nx - pad[3] if pad[3] else None]
    assert np.array_equal(no_halo, out[0].cpu().numpy())
    # the adjoint reports the step below each chunk it finished, counting
    # down (reference tm2d_autograd.py:766-768), and its accumulated
    # gradient grows to the final one
    assert [s for s, _, _ in seen_b] == [s - 1 for s in range(nt, 0, -every)]
    assert np.abs(seen_b[-1][1]).sum() > np.abs(seen_b[0][1]).sum()


def test_callbacks_with_shot_chunks_are_rejected(tmp_path):
    p = hetero_problem(shots=3, ny=30, nx=34, nt=40)
    st = prepare_tm2d(p["epsilon"], p["sigma"], p["mu"], p["grid_spacing"], p["dt"],
                      p["source_amplitude"], p["source_location"], p["receiver_location"],
                      pml_width=p["pml_width"], allocate_fields=False)
    per_shot = 2 * 40 * st.ny * st.nx * 4
    with pytest.raises(RuntimeError, match="callbacks"):
        _grad_run(p, tmp_path, storage_bytes_limit_device=int(1.5 * per_shot),
                  forward_callback=lambda s: None)


def test_operator_gradient_is_bitwise_repeatable(tmp_path):
    """Kernels, FP64 shot sums and the padding fold are all free of
    atomics: two runs of the same gradient are identical."""
    p = hetero_problem(shots=5, ny=40, nx=52, nt=90)
    a = _grad_run(p, tmp_path)
    b = _grad_run(p, tmp_path)
    for x, y in zip(a[:3], b[:3]):
        assert np.array_equal(x, y)
