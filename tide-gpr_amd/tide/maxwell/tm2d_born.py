"""2-D TM Born (tangent-linear) propagation on the MI355X HIP kernels
(reference ``src/tide/maxwell/tm2d_born.py:26-336``).

``borntm`` advances the background wavefield and the scattered wavefield
``J(m) v`` together for a model direction ``v`` = (``depsilon``,
``dsigma``) or (``dca``, ``dcb``): the scattered E update adds
``dca * Ey^n + dcb * curl H^n`` of the background at every step, and with
``linearize_source`` the scattered field also receives the source term's
own tangent ``amplitude * dcb[src] * (-1/(dy dx))`` (the source is scaled
by ``cb``).  The coefficient tangents come from the closed-form partials
of ``ca, cb`` (``utils.coefficient_partials``) on the replicate-padded
direction, the same padding as the model.

Returns ``(background final state x7, scattered final state x7, background
receivers, scattered receivers)`` with the state unpadded, as the
reference.  The receivers are differentiable with respect to the
background model (``ca``, ``cb``, the coefficient tangents and both source
terms) through the native adjoint of the Born recurrence
(``BornTMPropagate``; reference ``tm2d_born_autograd.py:597-1120``): that is
what ``op.linearize(model).second_vjp`` uses.  Gradients with respect to the
final states are not propagated (as the reference).
"""

from __future__ import annotations

from collections.abc import Sequence
from typing import Any

import torch

from .. import _native
from ..grid_utils import _normalize_grid_spacing_2d
from ..padding import create_or_pad
from ..utils import coefficient_partials
from .tm2d import _reject_unsupported, prepare_tm2d


def _born_fn(stencil: int, dtype: torch.dtype, storing: bool = False):
    lib = _native.load()
    kind = "born_forward_with_storage" if storing else "born_forward"
    return getattr(lib, f"maxwell_tm_{stencil}_{_native.dtype_name(dtype)}_{kind}_cuda")


class BornTMPropagate(torch.autograd.Function):
    """Born forward kernels + their exact adjoint behind torch autograd.

    Inputs ``ca, cb`` ``[1, ny, nx]``, ``dca, dcb`` ``[ny, nx]``, source
    terms ``f0`` (background) and ``df`` (scattered) ``[nt, N, n_src]`` (or
    None), and the setup / initial states riding along in ``meta``.  Outputs
    background receivers, scattered receivers and the 7 + 7 padded final
    states.  When any input needs a gradient the forward keeps the four
    snapshot histories (background Ey^n, curl^n, scattered Ey^n, curl^n) and
    the backward runs ``tide_tm2d_<s>_<dtype>_born_backward_f64acc`` with the
    receiver cotangents, returning FP64-accumulated gradients cast once."""

    @staticmethod
    def forward(ctx, ca, cb, dca, dcb, f0, df, meta):
        setup, bg, sc = meta["setup"], meta["bg"], meta["sc"]
        dev, dtype = setup.ca.device, setup.ca.dtype
        storing = any(ctx.needs_input_grad[:6])
        r = torch.zeros(setup.nt, setup.n_shots, setup.n_rec, device=dev, dtype=dtype)
        bg_r = torch.zeros_like(r)
        hist = ([torch.empty(setup.nt, setup.n_shots, setup.ny, setup.nx, device=dev, dtype=dtype)
                 for _ in range(4)] if storing else [None] * 4)
        P = _native.ptr
        dev_idx = dev.index if dev.index is not None else torch.cuda.current_device()
        stream = _native.stream_handle(dev)
        args = [P(ca), P(cb), P(setup.cq), P(dca), P(dcb), P(f0), P(df),
                *[P(t) for t in bg], *[P(t) for t in sc], P(r), P(bg_r)]
        if storing:
            args += [P(hist[0]), None, None, P(hist[1]), None, None, P(hist[2]), P(hist[3])]
        args += [*[P(p) for p in setup.profiles], P(setup.sources_i), P(setup.receivers_i),
                 setup.rdy, setup.rdx, setup.dt, setup.nt, setup.n_shots, setup.ny, setup.nx,
                 setup.n_src, setup.n_rec, 1]
        if storing:
            args += [0, 0, setup.ny * setup.nx * dtype.itemsize, True, True]
        args += [False, False, False, 0, *setup.pml, 0, dev_idx, stream]
        if storing:
            args += [stream]
        _native.call(_born_fn(setup.stencil, dtype, storing), *args)
        ctx.meta = meta
        ctx.storing = storing
        ctx.save_for_backward(ca, cb, dca, dcb, *(hist if storing else ()))
        ctx.mark_non_differentiable(*bg, *sc)
        return (bg_r, r, *bg, *sc)

    @staticmethod
    def backward(ctx, g_bg_r, g_r, *_state_grads):
        setup = ctx.meta["setup"]
        ca, cb, dca, dcb, *hist = ctx.saved_tensors
        dev, dtype = ca.device, ca.dtype
        need = ctx.needs_input_grad
        lam = [torch.zeros(setup.n_shots, setup.ny, setup.nx, device=dev, dtype=dtype)
               for _ in range(14)]
        g_df = (torch.zeros(setup.nt, setup.n_shots, setup.n_src, device=dev, dtype=dtype)
                if need[5] and setup.n_src > 0 else None)
        g_f0 = (torch.zeros(setup.nt, setup.n_shots, setup.n_src, device=dev, dtype=dtype)
                if need[4] and setup.n_src > 0 else None)
        acc = [torch.zeros(setup.ny, setup.nx, device=dev, dtype=torch.float64) if n else None
               for n in need[:4]]
        P = _native.ptr
        w = g_r.contiguous() if g_r is not None else None
        w_bg = g_bg_r.contiguous() if g_bg_r is not None else None
        fn = getattr(_native.load(),
                     f"tide_tm2d_{setup.stencil}_{_native.dtype_name(dtype)}_born_backward_f64acc")
        dev_idx = dev.index if dev.index is not None else torch.cuda.current_device()
        _native.call(
            fn, P(ca), P(cb), P(setup.cq), P(dca), P(dcb), P(w), P(w_bg),
            *[P(h) for h in hist], *[P(t) for t in lam], P(g_df), P(g_f0),
            *[P(a) for a in acc], *[P(p) for p in setup.profiles], P(setup.sources_i),
            P(setup.receivers_i), setup.rdy, setup.rdx, setup.nt, setup.n_shots, setup.ny,
            setup.nx, setup.n_src, setup.n_rec, setup.nt, *setup.pml, dev_idx,
            _native.stream_handle(dev),
        )
        g_ca = acc[0].to(dtype)[None] if acc[0] is not None else None
        g_cb = acc[1].to(dtype)[None] if acc[1] is not None else None
        g_dca = acc[2].to(dtype) if acc[2] is not None else None
        g_dcb = acc[3].to(dtype) if acc[3] is not None else None
        return g_ca, g_cb, g_dca, g_dcb, g_f0, g_df, None


def borntm(
    epsilon: torch.Tensor,
    sigma: torch.Tensor,
    mu: torch.Tensor,
    grid_spacing: float | Sequence[float],
    dt: float,
    source_amplitude: torch.Tensor | None,
    source_location: torch.Tensor | None,
    receiver_location: torch.Tensor | None,
    bg_receiver_location: torch.Tensor | None = None,
    *,
    depsilon: torch.Tensor | None = None,
    dsigma: torch.Tensor | None = None,
    dca: torch.Tensor | None = None,
    dcb: torch.Tensor | None = None,
    stencil: int = 2,
    pml_width: int | Sequence[int] = 20,
    max_vel: float | None = None,
    Ey_0: torch.Tensor | None = None,
    Hx_0: torch.Tensor | None = None,
    Hz_0: torch.Tensor | None = None,
    m_Ey_x_0: torch.Tensor | None = None,
    m_Ey_z_0: torch.Tensor | None = None,
    m_Hx_z_0: torch.Tensor | None = None,
    m_Hz_x_0: torch.Tensor | None = None,
    dEy_0: torch.Tensor | None = None,
    dHx_0: torch.Tensor | None = None,
    dHz_0: torch.Tensor | None = None,
    dm_Ey_x_0: torch.Tensor | None = None,
    dm_Ey_z_0: torch.Tensor | None = None,
    dm_Hx_z_0: torch.Tensor | None = None,
    dm_Hz_x_0: torch.Tensor | None = None,
    nt: int | None = None,
    model_gradient_sampling_interval: int = 1,
    parameterization: str = "epsilon_sigma",
    linearize_source: bool = True,
    freq_taper_frac: float = 0.0,
    time_pad_frac: float = 0.0,
    time_taper: bool = False,
    python_backend: Any = False,
    storage_mode: str = "device",
    storage_path: str = ".",
    storage_compression: bool | str = False,
    storage_bytes_limit_device: int | None = None,
    storage_bytes_limit_host: int | None = None,
    n_threads: int | None = None,
    fallback: str = "reference",
) -> tuple[torch.Tensor, ...]:
    """Background and Born-scattered 2-D TM propagation (see module doc)."""
    del model_gradient_sampling_interval, storage_mode, storage_path, storage_compression
    del storage_bytes_limit_device, storage_bytes_limit_host, n_threads, fallback
    if epsilon.ndim != 2:
        raise NotImplementedError("borntm supports a single 2-D model")
    if freq_taper_frac or time_pad_frac or time_taper:
        raise NotImplementedError("time-axis resampling is not part of this build")
    if parameterization not in ("epsilon_sigma", "ca_cb"):
        raise ValueError("parameterization must be 'epsilon_sigma' or 'ca_cb'")
    _reject_unsupported(python_backend, stencil, None, None, None, epsilon.device)
    # the model side stays on torch's graph (ca, cb, their partials and the
    # source terms depend on epsilon and sigma); the direction is data
    setup = prepare_tm2d(
        epsilon, sigma, mu.detach(), grid_spacing, dt, source_amplitude,
        source_location, receiver_location, stencil=stencil, pml_width=pml_width, nt=nt,
        initial=(Ey_0, Hx_0, Hz_0, m_Ey_x_0, m_Ey_z_0, m_Hx_z_0, m_Hz_x_0), max_vel=max_vel,
    )
    with torch.no_grad():
        scat0 = prepare_tm2d(
            epsilon.detach(), sigma.detach(), mu.detach(), grid_spacing, dt, None,
            source_location, receiver_location, stencil=stencil, pml_width=pml_width,
            nt=setup.nt,
            initial=(dEy_0, dHx_0, dHz_0, dm_Ey_x_0, dm_Ey_z_0, dm_Hx_z_0, dm_Hz_x_0),
            max_vel=max_vel,
        ).fields
    dev, dtype = setup.ca.device, setup.ca.dtype
    pad = [a + b for a, b in zip(setup.fd_pad, setup.pml_width)]
    size = (setup.ny, setup.nx)

    def padded(t):
        return create_or_pad(t.detach().to(dtype), pad, dev, dtype, size, mode="replicate")

    if parameterization == "ca_cb":
        d_ca = padded(dca) if dca is not None else None
        d_cb = padded(dcb) if dcb is not None else None
    else:
        if dca is not None or dcb is not None:
            raise ValueError("dca / dcb need parameterization='ca_cb'")
        dca_de, dcb_de, dca_ds, dcb_ds = coefficient_partials(setup.ca[0], setup.cb[0], setup.dt)
        d_ca = d_cb = None
        if depsilon is not None:
            de = padded(depsilon)
            d_ca, d_cb = dca_de * de, dcb_de * de
        if dsigma is not None:
            ds = padded(dsigma)
            d_ca = dca_ds * ds if d_ca is None else d_ca + dca_ds * ds
            d_cb = dcb_ds * ds if d_cb is None else d_cb + dcb_ds * ds
    if (d_ca is None) != (d_cb is None):
        d_ca = torch.zeros_like(setup.ca[0]) if d_ca is None else d_ca
        d_cb = torch.zeros_like(setup.cb[0]) if d_cb is None else d_cb
    df = None
    if linearize_source and d_cb is not None and setup.n_src > 0 and setup.f.numel() > 0:
        # tangent of f = amplitude * cb[src] * (-1/(dy dx)) along dcb
        dy, dx = _normalize_grid_spacing_2d(grid_spacing)
        dcb_src = d_cb.reshape(-1)[setup.sources_i.to(dev)]  # [N, n_src]
        amp = source_amplitude.detach().permute(2, 0, 1).contiguous().to(dtype)
        df = ((amp * dcb_src.unsqueeze(0)) * (-1.0 / (dx * dy))).contiguous()
    d_ca = d_ca.contiguous() if d_ca is not None else None
    d_cb = d_cb.contiguous() if d_cb is not None else None
    meta = {"setup": setup, "bg": [t.detach().contiguous().clone() for t in setup.fields],
            "sc": [t.contiguous() for t in scat0]}
    f0 = setup.f.contiguous() if setup.f.numel() else None
    out = BornTMPropagate.apply(setup.ca.contiguous(), setup.cb.contiguous(), d_ca, d_cb, f0,
                                df, meta)
    bg_r, r = out[0], out[1]
    bg, sc = out[2:9], out[9:16]
    bg_out = [setup.unpad(t) for t in bg]
    sc_out = [setup.unpad(t) for t in sc]
    if bg_receiver_location is not None and receiver_location is not None and not torch.equal(
            bg_receiver_location.cpu(), receiver_location.cpu()):
        from .tm2d import maxwelltm

        with torch.no_grad():
            bg_r = maxwelltm(epsilon.detach(), sigma.detach(), mu.detach(), grid_spacing, dt,
                             source_amplitude, source_location, bg_receiver_location,
                             stencil=stencil, pml_width=pml_width, nt=setup.nt,
                             max_vel=max_vel)[-1]
    return (*bg_out, *sc_out, bg_r, r)


__all__ = ["borntm"]
