"""CPU tests of the host-side numerics that feed the HIP kernels.

Coverage mirrors the reference's unit tests for the same helpers
(reference ``tests/unit/test_grid.py:101-285``): CPML profile shapes and
ranges, the CFL sub-stepping rule, padding, wavelets.  On top of that the
closed forms the kernels rely on are pinned: CPML coefficients at known
depths, the (ca, cb) partials against autograd, the padded geometry and flat
point indices, the source scaling and the snapshot-storage policy.  None of
these calls touch a GPU.
"""

from __future__ import annotations

import math

import pytest
import torch

from helpers import config1_setup, hetero_problem, random_model
from tide import cfl_condition, ricker, gaussian, gaussian_derivative
from tide import staggered, storage, utils
from tide.maxwell.tm2d import maxwelltm, prepare_tm2d
from tide.padding import create_or_pad, reverse_pad, zero_interior


def _profiles(widths, size=32, dtype=torch.float64):
    return staggered.set_pml_profiles(
        pml_width=widths, accuracy=2, fd_pad=[1, 0, 1, 0], dt=1e-11,
        grid_spacing=[0.02, 0.025], max_vel=3e8, dtype=dtype,
        device=torch.device("cpu"), pml_freq=5e10, ny=size, nx=size,
    )


def test_pml_profile_shapes():
    prof = _profiles([4] * 4)
    assert len(prof) == 12
    ay, ayh, ax, axh = prof[:4]
    assert ay.shape == (1, 32, 1) and ayh.shape == (1, 32, 1)
    assert ax.shape == (1, 1, 32) and axh.shape == (1, 1, 32)


def test_pml_profile_zero_width_is_identity():
    prof = _profiles([0] * 4, size=16)
    assert all(torch.count_nonzero(p) == 0 for p in prof[:8])
    assert all(torch.all(p == 1.0) for p in prof[8:])


def test_pml_profile_ranges_and_interior():
    widths = [8, 6, 5, 7]
    prof = _profiles(widths)
    ay, ayh, ax, axh, by, byh, bx, bxh, ky, kyh, kx, kxh = [p.reshape(-1) for p in prof]
    for b in (by, byh, bx, bxh):
        assert torch.all((b >= 0) & (b <= 1))
    for k in (ky, kyh, kx, kxh):
        assert torch.all(k >= 1) and k.max() > 1
    for a in (ay, ayh, ax, axh):
        assert torch.all(a <= 0)
    # Strictly inside the model (between the strips) the update is plain.
    y0, y1 = 1 + widths[0], 32 - widths[1]
    assert torch.all(by[y0 + 1 : y1 - 1] == 0) and torch.all(ky[y0 + 1 : y1 - 1] == 1)


def _cpml_closed_form(n, widths, pads, h, dt, half):
    """The CPML grading of Roden and Gedney with the reference's constants
    (N = 4, sigma0 = (N+1)/(150 pi h), kappa_max = 5, alpha = alpha_max (1 -
    d/L) + 0.1 alpha_max, alpha_max = 0.008), evaluated in FP64 at every
    integer (or half-integer) point of an axis whose strips end at
    pad_lo + w_lo and n - 1 - pad_hi - w_hi."""
    import numpy as np

    coord = np.arange(n, dtype=np.float64) + (0.5 if half else 0.0)
    a = np.zeros(n)
    b = np.zeros(n)
    k = np.ones(n)
    edges = (pads[0] + widths[0], n - 1 - pads[1] - widths[1])
    sigma0 = 5.0 / (150.0 * math.pi * h)
    for side in (0, 1):
        w = widths[side]
        if w <= 0:
            continue
        depth = edges[0] - coord if side == 0 else coord - edges[1]
        inside = depth >= 0
        frac = np.clip(depth / w, 0.0, 1.0)
        sig = sigma0 * frac**4
        kap = 1.0 + 4.0 * frac**4
        alpha = 0.008 * (1.0 - frac) + 0.1 * 0.008
        bs = np.exp(-(sig / kap + alpha) * dt / utils.EP0)
        as_ = np.where(sig > 1e-6, sig * (bs - 1.0) / (kap * (sig + kap * alpha) + 1e-9), 0.0)
        k = np.where(inside, kap, k)
        b = np.where(inside, bs, b)
        a = np.where(inside, as_, a)
    return a, b, k


@pytest.mark.parametrize("dtype,tol", [(torch.float64, 1e-13), (torch.float32, 2e-6)])
@pytest.mark.parametrize("widths,size", [([8, 6, 5, 7], 48), ([20, 20, 20, 20], 1065), ([3, 0, 0, 9], 40)])
def test_pml_profiles_match_the_closed_form_everywhere(dtype, tol, widths, size):
    """Every profile value (interior of the strips included), integer and
    half points, against the FP64 closed form of the grading."""
    import numpy as np

    h = (0.02, 0.025)
    dt = 0.9 * min(h) / (2**0.5 * utils.C0)
    prof = staggered.set_pml_profiles(
        pml_width=widths, accuracy=2, fd_pad=[1, 0, 1, 0], dt=dt, grid_spacing=list(h),
        max_vel=3e8, dtype=dtype, device=torch.device("cpu"), pml_freq=0.5 / dt,
        ny=size, nx=size + 3)
    ay, ayh, ax, axh, by, byh, bx, bxh, ky, kyh, kx, kxh = [p.reshape(-1).double().numpy() for p in prof]
    for (a, b, k), n, w, hh, half in (((ay, by, ky), size, widths[:2], h[0], False),
                                      ((ayh, byh, kyh), size, widths[:2], h[0], True),
                                      ((ax, bx, kx), size + 3, widths[2:], h[1], False),
                                      ((axh, bxh, kxh), size + 3, widths[2:], h[1], True)):
        a0, b0, k0 = _cpml_closed_form(n, w, (1, 0), hh, dt, half)
        scale = lambda r: max(np.abs(r).max(), 1e-30)  # noqa: E731
        assert np.abs(k - k0).max() <= tol * scale(k0)
        assert np.abs(b - b0).max() <= tol * scale(b0)
        assert np.abs(a - a0).max() <= 4 * tol * scale(a0)


def test_pml_grading_is_quartic_and_a_b_consistent():
    """Invariants of the grading, independent of its constants: inside a
    strip kappa - 1 grows as the fourth power of depth (ratios of depths
    fix the exponent and the half-point offset), and a, b, kappa satisfy
    a * kappa * (sigma + kappa alpha) = sigma (b - 1) with sigma and alpha
    read back from b = exp(-(sigma/kappa + alpha) dt / eps0)."""
    import numpy as np

    h, w, n = 0.02, 10, 60
    dt = 2e-11
    for half in (False, True):
        setup = utils.setup_pml_half if half else utils.setup_pml
        a, b, k = (t.numpy() for t in setup([w, w], [float(w), float(n - 1 - w)], w * h, dt, n,
                                            3e8, torch.float64, torch.device("cpu"), 1.0 / dt,
                                            grid_spacing=h))
        coord = np.arange(n) + (0.5 if half else 0.0)
        depth = w - coord
        inside = (depth > 0) & (depth <= w)
        ratio = (k[inside] - 1.0) / (k[inside][0] - 1.0)
        np.testing.assert_allclose(ratio, (depth[inside] / depth[inside][0]) ** 4, rtol=1e-10, atol=1e-15)
        # alpha is linear in depth; sigma = sigma0 (d/L)^4
        frac = depth[inside] / w
        alpha = 0.008 * (1 - frac) + 0.0008
        sig_over_k = -np.log(b[inside]) * utils.EP0 / dt - alpha
        sig = sig_over_k * k[inside]
        np.testing.assert_allclose(sig, 5.0 / (150 * math.pi * h) * frac**4, rtol=1e-9)
        lhs = a[inside] * (k[inside] * (sig + k[inside] * alpha) + 1e-9)
        np.testing.assert_allclose(lhs, sig * (b[inside] - 1.0), rtol=1e-9, atol=1e-15)


def test_pml_profile_closed_form_at_outer_edge():
    """At the outermost cell (depth = width) the grading is 1: sigma =
    sigma0 = 5/(150 pi h), kappa = 5, alpha = 0.1 alpha_max,
    b = exp(-(sigma/kappa + alpha) dt/eps0), a = sigma (b-1)/(kappa (sigma + kappa alpha))."""
    dt, h, w = 1e-11, 0.02, 8
    a, b, k = utils.setup_pml([w, w], [float(w), 40.0 - w], float(w) * h, dt, 41, 3e8,
                              torch.float64, torch.device("cpu"), 5e10, grid_spacing=h)
    sigma0 = 5.0 / (150.0 * math.pi * h)
    alpha = 0.1 * 0.008
    bb = math.exp(-(sigma0 / 5.0 + alpha) * dt / utils.EP0)
    aa = sigma0 * (bb - 1.0) / (5.0 * (sigma0 + 5.0 * alpha) + 1e-9)
    assert k[0].item() == pytest.approx(5.0, rel=1e-15)
    assert b[0].item() == pytest.approx(bb, rel=1e-13)
    assert a[0].item() == pytest.approx(aa, rel=1e-13)
    # Symmetric strips give mirror-image profiles on the integer grid.
    torch.testing.assert_close(b, b.flip(0), rtol=0, atol=0)
    torch.testing.assert_close(a, a.flip(0), rtol=0, atol=0)


def test_cfl_condition_substeps():
    with pytest.warns(UserWarning):
        inner, ratio = cfl_condition([0.1, 0.1], dt=0.1, max_vel=1.0)
    assert ratio >= 2 and math.isclose(inner * ratio, 0.1)
    # The largest stable step of the 2-D Yee scheme: 1/(v sqrt(1/dy^2 + 1/dx^2)).
    inner, ratio = cfl_condition([0.02, 0.02], dt=1e-12, max_vel=utils.C0)
    assert ratio == 1 and inner == 1e-12
    dt_max = 1.0 / (utils.C0 * math.sqrt(2.0) / 0.02)
    _, ratio = cfl_condition([0.02, 0.02], dt=dt_max * 2.5, max_vel=utils.C0)
    assert ratio == 3


def test_padding_helpers():
    assert reverse_pad([1, 2, 3, 4]) == [3, 4, 1, 2]
    z = create_or_pad(torch.empty(0), 1, torch.device("cpu"), torch.float32, (2, 5, 6))
    assert z.shape == (2, 5, 6) and torch.count_nonzero(z) == 0
    base = torch.arange(4.0).reshape(2, 2)
    p = create_or_pad(base, [1, 1, 1, 1], torch.device("cpu"), torch.float32, (4, 4))
    assert p.shape == (4, 4) and torch.equal(p[1:3, 1:3], base) and p[0, 0] == 0
    r = create_or_pad(base, [1, 0, 2, 1], torch.device("cpu"), torch.float32, (3, 5), mode="replicate")
    expected = torch.tensor([[0, 0, 0, 1, 1], [0, 0, 0, 1, 1], [2, 2, 2, 3, 3]], dtype=torch.float32)
    assert torch.equal(r, expected)
    t = torch.ones(1, 6, 6)
    zero_interior(t, [1, 1, 1, 1], [1, 1, 1, 1], 0)
    assert torch.count_nonzero(t[:, 2:4, :]) == 0 and torch.all(t[:, :2] == 1) and torch.all(t[:, 4:] == 1)


def test_replicate_pad_gradient_folds_onto_edges():
    """The gradient of a padded cell lands on its nearest model cell."""
    m = torch.rand(3, 4, dtype=torch.float64, requires_grad=True)
    p = create_or_pad(m, [2, 1, 1, 3], torch.device("cpu"), torch.float64, (6, 8), mode="replicate")
    g = torch.rand_like(p)
    (p * g).sum().backward()
    expected = torch.zeros(3, 4, dtype=torch.float64)
    for i in range(6):
        for j in range(8):
            expected[min(max(i - 2, 0), 2), min(max(j - 1, 0), 3)] += g[i, j]
    torch.testing.assert_close(m.grad, expected, rtol=1e-14, atol=1e-14)


def test_wavelets():
    w = ricker(2.0, 50, 0.1, dtype=torch.float32)
    assert w.shape == (50,) and w.dtype == torch.float32
    assert abs(int(w.abs().argmax()) - 5) <= 1
    with pytest.raises(ValueError):
        ricker(0.0, 50, 0.1)
    with pytest.raises(ValueError):
        ricker(2.0, 50, 0.0)
    for fn in (ricker, gaussian):
        c = fn(2.0, 21, 0.05, peak_time=0.5, dtype=torch.float64)
        assert c[10].item() == pytest.approx(1.0)
    gd = gaussian_derivative(2.0, 2001, 0.0005, peak_time=0.5)
    torch.testing.assert_close(gd, -gd.flip(0), atol=2e-5, rtol=0.0)
    assert gd.abs().max().item() == pytest.approx(1.0, rel=1e-5)


def test_coefficient_partials_match_autograd():
    dt = 4e-12
    eps = (torch.rand(5, 6, dtype=torch.float64) * 8 + 1).requires_grad_()
    sig = (torch.rand(5, 6, dtype=torch.float64) * 0.05).requires_grad_()
    mu = torch.ones(5, 6, dtype=torch.float64)
    ca, cb, _ = utils.prepare_parameters(eps, sig, mu, dt)
    dca_de, dcb_de, dca_ds, dcb_ds = utils.coefficient_partials(ca.detach(), cb.detach(), dt)
    for out, de, ds in ((ca, dca_de, dca_ds), (cb, dcb_de, dcb_ds)):
        ge, gs = torch.autograd.grad(out.sum(), (eps, sig), retain_graph=True)
        torch.testing.assert_close(ge, de, rtol=1e-10, atol=0)
        torch.testing.assert_close(gs, ds, rtol=1e-10, atol=0)


def test_material_coefficients_lossless_limit():
    """sigma = 0 gives ca = 1, cb = dt/(eps_r eps0), cq = dt/(mu_r mu0)."""
    dt = 1e-11
    eps = torch.tensor([[4.0, 9.0]], dtype=torch.float64)
    ca, cb, cq = utils.prepare_parameters(eps, torch.zeros_like(eps), torch.ones_like(eps), dt)
    assert torch.all(ca == 1)
    torch.testing.assert_close(cb, dt / (eps * utils.EP0), rtol=1e-15, atol=0)
    torch.testing.assert_close(cq, torch.full_like(eps, dt / utils.MU0), rtol=1e-15, atol=0)
    with pytest.raises(ValueError):
        utils.compile_material_coefficients(-eps, eps, eps, dt)


def test_prepare_tm2d_geometry_and_points():
    ny, nx, pml = 30, 26, [3, 4, 5, 2]
    eps, sig = random_model(ny, nx, seed=3, dtype=torch.float64)
    mu = torch.ones_like(eps)
    src_loc = torch.tensor([[[2, 3]], [[10, 20]]])
    rec_loc = torch.tensor([[[0, 0], [29, 25]], [[5, 7], [6, 8]]])
    amp = torch.randn(2, 1, 17, dtype=torch.float64)
    dy, dx, dt = 0.02, 0.025, 2e-11
    st = prepare_tm2d(eps, sig, mu, [dy, dx], dt, amp, src_loc, rec_loc, pml_width=pml)
    assert (st.ny, st.nx) == (ny + 1 + 3 + 4, nx + 1 + 5 + 2)
    assert st.pml == (1 + 3, 1 + 5, st.ny - 4, st.nx - 2)
    assert st.nt == 17 and st.n_shots == 2 and st.n_src == 1 and st.n_rec == 2
    oy, ox = 1 + pml[0], 1 + pml[2]
    assert st.receivers_i[0, 1].item() == (29 + oy) * st.nx + (25 + ox)
    assert st.sources_i[1, 0].item() == (10 + oy) * st.nx + (20 + ox)
    # Replicate padding: the padded corner carries the model's corner.
    _, cb0, _ = utils.prepare_parameters(eps, sig, mu, dt)
    assert st.cb[0, 0, 0].item() == cb0[0, 0].item()
    assert st.cb[0, -1, -1].item() == cb0[-1, -1].item()
    # Source term: f = amplitude * cb(source) * (-1/(dx dy)).
    i = st.sources_i[1, 0].item()
    expected = amp[1, 0] * st.cb.reshape(-1)[i] * (-1.0 / (dx * dy))
    torch.testing.assert_close(st.f[:, 1, 0], expected, rtol=1e-15, atol=0)
    assert all(tuple(f.shape) == (2, st.ny, st.nx) for f in st.fields)


def test_prepare_tm2d_config1_sizes():
    st = config1_setup(dtype=torch.float32)
    assert (st.ny, st.nx) == (117, 117) and st.n_shots == 1


def test_storage_policy():
    kw = dict(needs_storage=True, total_bytes=100, bytes_limit_device=50, bytes_limit_host=200)
    assert storage.resolve_storage_mode("auto", **kw) == "cpu"
    assert storage.resolve_storage_mode("auto", **{**kw, "bytes_limit_device": 100}) == "device"
    assert storage.resolve_storage_mode("auto", **{**kw, "bytes_limit_host": 10}) == "disk"
    assert storage.resolve_storage_mode("auto", **{**kw, "needs_storage": False}) == "none"
    with pytest.raises(ValueError):
        storage.resolve_storage_mode("none", **kw)
    with pytest.raises(ValueError):
        storage.resolve_storage_mode("tape", **kw)
    kind, dt, nbytes, fmt = storage._resolve_storage_compression(
        "bf16", torch.float32, torch.device("cpu"), context="x")
    assert (kind, dt, nbytes, fmt) == ("bf16", torch.bfloat16, 2, storage.STORAGE_FORMAT_BF16)
    with pytest.raises(NotImplementedError):
        storage._resolve_storage_compression(True, torch.float64, torch.device("cpu"), context="x")
    with pytest.raises(ValueError):
        storage._resolve_storage_compression("zstd", torch.float32, torch.device("cpu"), context="x")


def test_cpu_tensors_are_rejected_not_emulated():
    """The operator has one compute path (the HIP library); CPU tensors raise."""
    with pytest.raises(NotImplementedError, match="HIP"):
        maxwelltm(**hetero_problem(nt=4, shots=1))


def test_autotune_size_classes():
    from tide.autotune import ADJ_CANDIDATES, FWD_CANDIDATES, _size_class
    a = _size_class(1065, 1065, 7, torch.float32)
    assert a == _size_class(1100, 1050, 6, torch.float32)  # similar size, same shot bucket
    assert a != _size_class(1065, 1065, 16, torch.float32)
    assert a != _size_class(300, 300, 7, torch.float32)
    assert a != _size_class(1065, 1065, 7, torch.float64)
    for cands, rows in ((FWD_CANDIDATES, (2, 4, 8, 16)), (ADJ_CANDIDATES, (2, 4, 8))):
        for r, g, lp, steps, split, border in cands:
            assert r in rows and 1 <= g <= (4 if lp else 8) and lp in (0, 1)
            assert steps in (1, 2, 3, 4) and split in (0, 1) and border in (0, 1, 2)
            assert border == 0 or steps >= 2  # only blocked launches have border tiles
            assert not (lp and steps >= 2)  # blocked launches group shots per wavefront
            if steps >= 2:
                assert r <= 8
            if steps >= 3:
                # K-step launches: ordered tiles only, border tiles of 2 rows or
                # the interior height, 8-row tiles only in the fp32 forward
                assert split == 0 and border in (0, 2) and r in ((2, 4, 8) if cands is
                                                                 FWD_CANDIDATES else (2, 4))


def test_autotune_rejects_bad_chunk_lists_before_touching_a_device():
    """A chunk list must hold positive shot counts; the check runs before any
    device work (so it holds on a host without a GPU)."""
    from tide.autotune import _pick, tune_tm2d
    for bad in ([], [4, 0], [3, -1]):
        with pytest.raises(ValueError):
            tune_tm2d(64, 64, bad, device="cpu")
    # the installed shape is kept within the margin, else the fastest wins
    times = {(4, 4): 100.0, (4, 8): 98.0, (2, 2): 90.0}
    assert _pick(times, (4, 4)) == (2, 2)
    assert _pick({(4, 4): 100.0, (4, 8): 98.0}, (4, 4)) == (4, 4)
    assert _pick({(4, 4): 100.0, (4, 8): 98.0}, (8, 8)) == (4, 8)
