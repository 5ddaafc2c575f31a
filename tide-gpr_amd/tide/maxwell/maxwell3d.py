"""3-D Maxwell propagation on the MI355X HIP kernels.

``maxwell3d`` keeps the reference entry point's signature and semantics
(reference ``src/tide/maxwell/maxwell3d.py:200-951`` and the native adapter
``maxwell3d_cuda.py:27-921``), coordinates ``[z, y, x]``:

* the model is replicate-padded by ``fd_pad + pml`` per side
  (``fd_pad = [s/2, s/2-1]`` per axis), so its gradient folds back onto the
  edge cells;
* ``ca, cb, cq`` come from ``utils.prepare_parameters``; the source term is
  ``f[t, shot, k] = amplitude * cb[src] * (-1/(dz dy dx))`` on the chosen E
  component (``maxwell3d_cuda.py:504-517``), receivers read one E
  component;
* :class:`EM3DPropagate` runs the HIP forward and, in its backward, the
  exact adjoint.  When the stored snapshot history (six components per
  sampled step) does not fit the device budget, the forward keeps only
  checkpoints of the state every ``~sqrt(nt)`` steps and the backward
  recomputes each segment's snapshots from its checkpoint before running
  that segment's adjoint;
* outputs are the 18 unpadded final state tensors and receivers
  ``[nt, shots, n_rec]``.

Only the HIP device path exists: CPU tensors, the Python reference backend,
callbacks, dispersion, stencils other than 2, host-staged storage and
internal resampling (step_ratio > 1) raise ``NotImplementedError``.
"""

from __future__ import annotations

import ctypes
import math
from collections.abc import Sequence
from dataclasses import dataclass, field
from typing import Any

import torch

from .. import _native, staggered
from ..cfl import cfl_condition
from ..grid_utils import _normalize_grid_spacing_3d, _normalize_pml_width_3d
from ..padding import create_or_pad
from ..storage import (
    STORAGE_FORMAT_BF16,
    STORAGE_FORMAT_FULL,
    _resolve_storage_compression,
    resolve_storage_mode,
)
from ..utils import C0, coefficient_partials, compile_material_coefficients

STATE_NAMES = (
    "Ex", "Ey", "Ez", "Hx", "Hy", "Hz",
    "m_hz_y", "m_hy_z", "m_hx_z", "m_hz_x", "m_hy_x", "m_hx_y",
    "m_ey_z", "m_ez_y", "m_ez_x", "m_ex_z", "m_ex_y", "m_ey_x",
)
# axis (0 z, 1 y, 2 x) whose CPML strips each memory lives in
_MEMORY_AXIS = {
    "m_hz_y": 1, "m_hy_z": 0, "m_hx_z": 0, "m_hz_x": 2, "m_hy_x": 2, "m_hx_y": 1,
    "m_ey_z": 0, "m_ez_y": 1, "m_ez_x": 2, "m_ex_z": 0, "m_ex_y": 1, "m_ey_x": 2,
}
_COMPONENTS = {"ex": 0, "ey": 1, "ez": 2}


@dataclass
class EM3DSetup:
    """Padded problem handed to the native 3-D kernels."""

    ca: torch.Tensor  # [1|N, nz, ny, nx]
    cb: torch.Tensor
    cq: torch.Tensor
    f: torch.Tensor  # [nt, N, n_src] (empty when no sources)
    profiles: list[torch.Tensor]  # az bz azh bzh ay by ayh byh ax bx axh bxh kz kzh ky kyh kx kxh
    sources_i: torch.Tensor  # [N, n_src] int64, CPU
    receivers_i: torch.Tensor  # [N, n_rec] int64, CPU
    fields: list[torch.Tensor]  # 18 x [N, nz, ny, nx]
    rdz: float
    rdy: float
    rdx: float
    dt: float
    nt: int
    n_shots: int
    nz: int
    ny: int
    nx: int
    n_src: int
    n_rec: int
    pml: tuple[int, int, int, int, int, int]  # z0 y0 x0 z1 y1 x1
    fd_pad: list[int]
    pml_width: list[int]
    source_component: int
    receiver_component: int
    batched: bool
    extra: dict[str, Any] = field(default_factory=dict)

    def unpad(self, t: torch.Tensor) -> torch.Tensor:
        p = self.fd_pad
        end = lambda n, hi: n - hi if hi > 0 else None  # noqa: E731
        return t[:, p[0] : end(self.nz, p[1]), p[2] : end(self.ny, p[3]), p[4] : end(self.nx, p[5])]

    def strips(self, axis: int) -> tuple[int, int]:
        """Cells ``[0, lo)`` and ``[hi, n)`` along ``axis`` hold every CPML
        memory value (integer and half strips)."""
        n = (self.nz, self.ny, self.nx)[axis]
        lo, hi = self.pml[axis], self.pml[3 + axis]
        return lo, max(lo, min(n, hi - 1))


def _component(name: str) -> int:
    try:
        return _COMPONENTS[name.lower()]
    except (KeyError, AttributeError):
        raise ValueError(f"component must be one of 'ex', 'ey', 'ez'; got {name!r}") from None


def _flat_indices_3d(loc, pads, ny_p, nx_p):
    if loc is None or loc.numel() == 0:
        return torch.empty(0, dtype=torch.long), 0
    loc = loc.detach().to("cpu").long()
    z = loc[..., 0] + pads[0]
    y = loc[..., 1] + pads[1]
    x = loc[..., 2] + pads[2]
    return ((z * ny_p + y) * nx_p + x).contiguous(), int(loc.shape[1])


def prepare_em3d(
    epsilon: torch.Tensor,
    sigma: torch.Tensor,
    mu: torch.Tensor,
    grid_spacing: float | Sequence[float],
    dt: float,
    source_amplitude: torch.Tensor | None,
    source_location: torch.Tensor | None,
    receiver_location: torch.Tensor | None,
    *,
    stencil: int = 2,
    pml_width: int | Sequence[int] = 20,
    nt: int | None = None,
    initial: Sequence[torch.Tensor | None] = (None,) * 18,
    source_component: str = "ey",
    receiver_component: str = "ey",
    validate_material_inputs: bool = True,
) -> EM3DSetup:
    """Pad the model, compile coefficients and profiles, map the points.

    ``epsilon`` is ``[nz, ny, nx]`` (shared by all shots) or ``[N, nz, ny,
    nx]``.  Device-agnostic, so the host numerics are testable on CPU.
    """
    if epsilon.ndim not in (3, 4):
        raise RuntimeError("epsilon must be 3D or batched 4D")
    if sigma.shape != epsilon.shape or mu.shape != epsilon.shape:
        raise RuntimeError("sigma and mu must have the same shape as epsilon")
    device, dtype = epsilon.device, epsilon.dtype
    batched = epsilon.ndim == 4
    dz, dy, dx = _normalize_grid_spacing_3d(grid_spacing)
    widths = _normalize_pml_width_3d(pml_width)
    if nt is None:
        if source_amplitude is None:
            raise ValueError("Either nt or source_amplitude must be provided")
        nt = int(source_amplitude.shape[-1])
    nt = int(nt)
    if source_amplitude is not None and source_amplitude.numel() > 0:
        n_shots = int(source_amplitude.shape[0])
    elif source_location is not None and source_location.numel() > 0:
        n_shots = int(source_location.shape[0])
    elif receiver_location is not None and receiver_location.numel() > 0:
        n_shots = int(receiver_location.shape[0])
    else:
        n_shots = 1
    if batched and int(epsilon.shape[0]) != n_shots:
        raise RuntimeError("Batched model count must match the effective shot count.")

    half = stencil // 2
    fd_pad = [half, half - 1] * 3
    total = [a + b for a, b in zip(fd_pad, widths)]
    nz = int(epsilon.shape[-3]) + total[0] + total[1]
    ny = int(epsilon.shape[-2]) + total[2] + total[3]
    nx = int(epsilon.shape[-1]) + total[4] + total[5]
    size = (n_shots, nz, ny, nx) if batched else (nz, ny, nx)
    eps_p = create_or_pad(epsilon, total, device, dtype, size, mode="replicate")
    sig_p = create_or_pad(sigma, total, device, dtype, size, mode="replicate")
    mu_p = create_or_pad(mu, total, device, dtype, size, mode="replicate")
    coef = compile_material_coefficients(eps_p, sig_p, mu_p, dt, validate_inputs=validate_material_inputs)
    shape4 = (lambda t: t.contiguous()) if batched else (lambda t: t[None].contiguous())
    ca, cb, cq = shape4(coef["ca"]), shape4(coef["cb"]), shape4(coef["cq"])

    pads = (total[0], total[2], total[4])
    sources_i, n_src = _flat_indices_3d(source_location, pads, ny, nx)
    receivers_i, n_rec = _flat_indices_3d(receiver_location, pads, ny, nx)

    source_coeff = -1.0 / (dx * dy * dz)
    if n_src > 0 and source_amplitude is not None and source_amplitude.numel() > 0:
        flat_cb = cb.reshape(cb.shape[0], nz * ny * nx).expand(n_shots, -1)
        cb_src = flat_cb.gather(1, sources_i.to(device))
        amp = source_amplitude.permute(2, 0, 1).contiguous().to(dtype)
        f = (amp * cb_src.to(dtype).unsqueeze(0) * source_coeff).contiguous()
    else:
        f = torch.empty(0, device=device, dtype=dtype)

    pml = (total[0], total[2], total[4], nz - fd_pad[1] - widths[1],
           ny - fd_pad[3] - widths[3], nx - fd_pad[5] - widths[5])
    fields = []
    for k, t0 in enumerate(initial):
        if t0 is None:
            fields.append(torch.zeros(n_shots, nz, ny, nx, device=device, dtype=dtype))
            continue
        t0 = t0.to(device=device, dtype=dtype)
        if t0.ndim == 3:
            t0 = t0[None].expand(n_shots, -1, -1, -1)
        if tuple(t0.shape) != (n_shots, nz, ny, nx):
            t0 = create_or_pad(t0.detach(), fd_pad, device, dtype, (n_shots, nz, ny, nx))
        t0 = t0.detach().clone().contiguous()
        if k >= 6:  # CPML memories live only in their strips
            ax = _MEMORY_AXIS[STATE_NAMES[k]]
            lo, hi = pml[ax], max(pml[ax], pml[3 + ax] - 1)
            idx = [slice(None)] * 4
            idx[1 + ax] = slice(lo, hi)
            t0[tuple(idx)] = 0
        fields.append(t0)

    prof = staggered.set_pml_profiles_3d(
        pml_width=widths, fd_pad=fd_pad, dt=dt, grid_spacing=[dz, dy, dx], dtype=dtype,
        device=device, nz=nz, ny=ny, nx=nx,
    )
    order = ("az", "bz", "azh", "bzh", "ay", "by", "ayh", "byh", "ax", "bx", "axh", "bxh",
             "kz", "kzh", "ky", "kyh", "kx", "kxh")
    profiles = [prof[k].reshape(-1).contiguous() for k in order]
    return EM3DSetup(
        ca=ca, cb=cb, cq=cq, f=f, profiles=profiles, sources_i=sources_i,
        receivers_i=receivers_i, fields=fields, rdz=1.0 / dz, rdy=1.0 / dy, rdx=1.0 / dx,
        dt=float(dt), nt=nt, n_shots=n_shots, nz=nz, ny=ny, nx=nx, n_src=n_src, n_rec=n_rec,
        pml=pml, fd_pad=fd_pad, pml_width=widths,
        source_component=_component(source_component),
        receiver_component=_component(receiver_component), batched=batched,
        extra={"zero_init": all(t0 is None for t0 in initial)},
    )


# ------------------------------------------------------------------ native


def _call_forward(setup: EM3DSetup, ca, cb, cq, f, fields, r, *, start_t: int, nt: int,
                  interval: int, stores=None, fmt: int = STORAGE_FORMAT_FULL,
                  store_t0: int = 0) -> None:
    dev, dtype = ca.device, ca.dtype
    fn = getattr(_native.load(), f"tide_em3d_{_native.dtype_name(dtype)}_forward")
    numel = setup.nz * setup.ny * setup.nx
    bpe = 2 if fmt == STORAGE_FORMAT_BF16 else dtype.itemsize
    step_bytes = setup.n_shots * numel * bpe
    dev_idx = dev.index if dev.index is not None else torch.cuda.current_device()
    _native.call(
        fn,
        _native.ptr(ca), _native.ptr(cb), _native.ptr(cq), _native.ptr(f),
        _native.ptr_array(fields), _native.ptr(r), _native.ptr_array(setup.profiles),
        _native.ptr(setup.sources_i), _native.ptr(setup.receivers_i),
        setup.rdz, setup.rdy, setup.rdx, nt, setup.n_shots, setup.nz, setup.ny, setup.nx,
        setup.n_src, setup.n_rec, interval, setup.batched, setup.batched, setup.batched,
        start_t, _native.int64_array(setup.pml), setup.source_component,
        setup.receiver_component,
        _native.ptr_array(stores if stores is not None else [None] * 6), fmt, step_bytes,
        store_t0, dev_idx, _native.stream_handle(dev),
    )


def _call_backward(setup: EM3DSetup, ca, cb, cq, grad_r, lam, stores, *, fmt: int,
                   store_t0: int, start_t: int, nt: int, interval: int, grad_f, gca, gcb) -> None:
    dev, dtype = ca.device, ca.dtype
    fn = getattr(_native.load(), f"tide_em3d_{_native.dtype_name(dtype)}_backward_f64acc")
    numel = setup.nz * setup.ny * setup.nx
    bpe = 2 if fmt == STORAGE_FORMAT_BF16 else dtype.itemsize
    dev_idx = dev.index if dev.index is not None else torch.cuda.current_device()
    _native.call(
        fn,
        _native.ptr(ca), _native.ptr(cb), _native.ptr(cq), _native.ptr(grad_r),
        _native.ptr_array(lam), _native.ptr_array(stores if stores is not None else [None] * 6),
        fmt, setup.n_shots * numel * bpe, store_t0, _native.ptr(grad_f), _native.ptr(gca),
        _native.ptr(gcb), _native.ptr_array(setup.profiles), _native.ptr(setup.sources_i),
        _native.ptr(setup.receivers_i), setup.rdz, setup.rdy, setup.rdx, nt, setup.n_shots,
        setup.nz, setup.ny, setup.nx, setup.n_src, setup.n_rec, interval, setup.batched,
        setup.batched, setup.batched, start_t, _native.int64_array(setup.pml),
        setup.source_component, setup.receiver_component, dev_idx, _native.stream_handle(dev),
    )


def _alloc_stores(setup: EM3DSetup, n_slots: int, fmt: int, dtype, device, want_ca, want_cb):
    sdtype = torch.int16 if fmt == STORAGE_FORMAT_BF16 else dtype
    shape = (n_slots, setup.n_shots, setup.nz, setup.ny, setup.nx)
    out = []
    for k in range(6):
        need = want_ca if k < 3 else want_cb
        out.append(torch.empty(shape, device=device, dtype=sdtype) if need else None)
    return out


def _pack_state(setup: EM3DSetup, fields: list[torch.Tensor]) -> list:
    """Checkpoint: the six fields, and each CPML memory's two strips."""
    packed = [t.clone() for t in fields[:6]]
    for name, t in zip(STATE_NAMES[6:], fields[6:]):
        ax = _MEMORY_AXIS[name]
        lo, hi = setup.strips(ax)
        packed.append((t.narrow(1 + ax, 0, lo).clone(),
                       t.narrow(1 + ax, hi, t.shape[1 + ax] - hi).clone()))
    return packed


def _unpack_state(setup: EM3DSetup, packed: list, fields: list[torch.Tensor]) -> None:
    for t, p in zip(fields[:6], packed[:6]):
        t.copy_(p)
    for name, t, (lo_part, hi_part) in zip(STATE_NAMES[6:], fields[6:], packed[6:]):
        ax = _MEMORY_AXIS[name]
        lo, hi = setup.strips(ax)
        t.zero_()
        t.narrow(1 + ax, 0, lo).copy_(lo_part)
        t.narrow(1 + ax, hi, t.shape[1 + ax] - hi).copy_(hi_part)


def _chunk_setup3(setup: EM3DSetup, s0: int, s1: int, fields: list[torch.Tensor]) -> EM3DSetup:
    """The shots ``[s0, s1)`` of ``setup`` as a setup of their own (views)."""
    if s0 == 0 and s1 == setup.n_shots and fields is setup.fields:
        return setup
    sub = EM3DSetup(**{k: getattr(setup, k) for k in setup.__dataclass_fields__})
    sub.n_shots = s1 - s0
    sub.sources_i = setup.sources_i[s0:s1].contiguous() if setup.n_src > 0 else setup.sources_i
    sub.receivers_i = setup.receivers_i[s0:s1].contiguous() if setup.n_rec > 0 else setup.receivers_i
    sub.fields = fields
    sub.extra = {}
    return sub


def _shots(t: torch.Tensor | None, s0: int, s1: int, batched: bool):
    return t if (t is None or not batched) else t[s0:s1]


def _forward_chunk(sub: EM3DSetup, ca, cb, cq, f, r, *, policy: str, seg: int, interval: int,
                   fmt: int, want_ca: bool, want_cb: bool, keep: bool):
    """One chunk's forward.  ``policy`` "full" stores every sampled step,
    "checkpoint" keeps the state every ``seg`` steps; with ``keep`` False
    nothing is kept (the backward re-runs the chunk).  Returns what the
    backward needs: ``("full", stores)``, ``("checkpoint", checkpoints)`` or
    None."""
    fields = sub.fields
    if keep and policy == "full":
        n_slots = (sub.nt + interval - 1) // interval
        stores = _alloc_stores(sub, n_slots, fmt, ca.dtype, ca.device, want_ca, want_cb)
        _call_forward(sub, ca, cb, cq, f, fields, r, start_t=0, nt=sub.nt, interval=interval,
                      stores=stores, fmt=fmt)
        return ("full", stores)
    if keep and policy == "checkpoint":
        checkpoints = []
        for t0 in range(0, sub.nt, seg):
            checkpoints.append(_pack_state(sub, fields))
            _call_forward(sub, ca, cb, cq, f, fields, r, start_t=t0, nt=min(seg, sub.nt - t0),
                          interval=interval)
        return ("checkpoint", checkpoints)
    _call_forward(sub, ca, cb, cq, f, fields, r, start_t=0, nt=sub.nt, interval=interval)
    return None


def _backward_chunk(sub: EM3DSetup, ca, cb, cq, f, grad_r, lam, saved, *, seg: int,
                    interval: int, fmt: int, want_ca: bool, want_cb: bool, grad_f, gca, gcb,
                    pool: dict) -> None:
    """The exact adjoint of one chunk from its saved snapshots or
    checkpoints; ``gca``/``gcb`` are FP64 accumulators (added to)."""
    kind, data = saved
    if kind == "full":
        _call_backward(sub, ca, cb, cq, grad_r, lam, data, fmt=fmt, store_t0=0,
                       start_t=sub.nt, nt=sub.nt, interval=interval, grad_f=grad_f, gca=gca,
                       gcb=gcb)
        return
    dev, dtype = ca.device, ca.dtype
    shape = (sub.n_shots, sub.nz, sub.ny, sub.nx)
    n_slots = (seg + interval - 1) // interval
    if "work" not in pool:  # sized for the first (largest) chunk that needs it
        numel = sub.n_shots * sub.nz * sub.ny * sub.nx
        pool["work"] = [torch.empty(numel, device=dev, dtype=dtype) for _ in range(18)]
        pool["stores"] = _alloc_stores(sub, n_slots, fmt, dtype, dev, want_ca, want_cb)
        pool["r"] = torch.zeros(sub.nt * sub.n_shots * max(sub.n_rec, 1), device=dev, dtype=dtype)
    numel = sub.n_shots * sub.nz * sub.ny * sub.nx
    work = [w[:numel].view(shape) for w in pool["work"]]
    store_shape = (n_slots, *shape)
    stores = [None if t is None else t.reshape(-1)[: n_slots * numel].view(store_shape)
              for t in pool["stores"]]
    r_scratch = pool["r"][: sub.nt * sub.n_shots * sub.n_rec].view(sub.nt, sub.n_shots, sub.n_rec)
    starts = list(range(0, sub.nt, seg))
    for k in reversed(range(len(starts))):
        t0 = starts[k]
        t1 = min(t0 + seg, sub.nt)
        _unpack_state(sub, data[k], work)
        data[k] = None
        _call_forward(sub, ca, cb, cq, f, work, r_scratch, start_t=t0, nt=t1 - t0,
                      interval=interval, stores=stores, fmt=fmt, store_t0=t0)
        _call_backward(sub, ca, cb, cq, grad_r, lam, stores, fmt=fmt, store_t0=t0, start_t=t1,
                       nt=t1 - t0, interval=interval, grad_f=grad_f, gca=gca, gcb=gcb)


class EM3DPropagate(torch.autograd.Function):
    """3-D forward kernels + exact adjoint behind torch autograd.

    ``opts["policy"]`` selects the snapshot policy for a model gradient:
    ``"full"`` stores every sampled step's (E, curl H) in HBM during the
    forward; ``"checkpoint"`` keeps the state every ``segment`` steps and
    recomputes one segment's snapshots at a time in the backward (one extra
    forward sweep).  ``opts["plan"]`` splits the shots into chunks: the
    first ``resident`` chunks keep their snapshots or checkpoints until the
    backward, the others are re-run there from their initial state.  One
    FP64 accumulator per coefficient collects every chunk.  Cotangents of
    the final state are ignored, as in the reference.
    """

    @staticmethod
    def forward(ctx, ca, cb, cq, f, setup: EM3DSetup, opts: dict[str, Any]):
        dev, dtype = ca.device, ca.dtype
        fields = [t.contiguous() for t in setup.fields]
        r = torch.zeros(setup.nt, setup.n_shots, setup.n_rec, device=dev, dtype=dtype)
        f_c = f.detach().contiguous()
        want_ca, want_cb = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        interval, fmt = opts["interval"], opts["format"]
        policy = opts["policy"] if (want_ca or want_cb) else "none"
        plan = opts.get("plan") or _single(setup.n_shots)
        chunked = plan.n_chunks > 1
        saved, initial = [], []
        for i in range(plan.n_chunks):
            s0, s1 = plan.bounds[i], plan.bounds[i + 1]
            sub = _chunk_setup3(setup, s0, s1, [t[s0:s1] for t in fields] if chunked else fields)
            keep = policy != "none" and i < plan.resident
            initial.append(None if (policy == "none" or keep or setup.extra.get("zero_init"))
                           else [t.clone() for t in sub.fields])
            f_i = f_c[:, s0:s1].contiguous() if (chunked and setup.n_src > 0) else f_c
            r_i = torch.zeros(setup.nt, s1 - s0, setup.n_rec, device=dev, dtype=dtype) if chunked else r
            saved.append(_forward_chunk(sub, *(_shots(t, s0, s1, setup.batched) for t in (ca, cb, cq)),
                                        f_i, r_i, policy=policy, seg=opts["segment"],
                                        interval=interval, fmt=fmt, want_ca=want_ca,
                                        want_cb=want_cb, keep=keep))
            if chunked and setup.n_rec > 0:
                r[:, s0:s1].copy_(r_i)
        ctx.setup, ctx.opts, ctx.policy, ctx.plan = setup, opts, policy, plan
        ctx.saved_chunks, ctx.initial = saved, initial
        ctx.save_for_backward(ca, cb, cq, f_c)
        ctx.mark_non_differentiable(*fields)
        return (*fields, r)

    @staticmethod
    def backward(ctx, *grads):
        grad_r = grads[-1]
        setup: EM3DSetup = ctx.setup
        plan = ctx.plan
        ca, cb, cq, f_c = ctx.saved_tensors
        want_ca, want_cb = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        want_f = ctx.needs_input_grad[3] and setup.n_src > 0
        if (want_ca or want_cb) and ctx.policy == "none":
            raise RuntimeError("model gradients need stored snapshots; storage_mode='none' was used")
        dev, dtype = ca.device, ca.dtype
        if grad_r is None:
            grad_r = torch.zeros(setup.nt, setup.n_shots, setup.n_rec, device=dev, dtype=dtype)
        grad_r = grad_r.to(dtype)
        chunked = plan.n_chunks > 1
        max_c = max(plan.sizes)
        numel = setup.nz * setup.ny * setup.nx
        lam_all = [torch.empty(max_c * numel, device=dev, dtype=dtype) for _ in range(18)]
        grad_f = (torch.zeros(setup.nt, setup.n_shots, setup.n_src, device=dev, dtype=dtype)
                  if want_f else None)
        full_shape = (setup.n_shots, setup.nz, setup.ny, setup.nx)
        gshape = full_shape if setup.batched else full_shape[1:]
        gca = torch.zeros(gshape, device=dev, dtype=torch.float64) if want_ca else None
        gcb = torch.zeros(gshape, device=dev, dtype=torch.float64) if want_cb else None
        interval, fmt, seg = ctx.opts["interval"], ctx.opts["format"], ctx.opts["segment"]
        pool: dict = {}
        for i in range(plan.n_chunks):  # resident chunks come first
            s0, s1 = plan.bounds[i], plan.bounds[i + 1]
            n_c = s1 - s0
            cs = [_shots(t, s0, s1, setup.batched) for t in (ca, cb, cq)]
            f_i = f_c[:, s0:s1].contiguous() if (chunked and setup.n_src > 0) else f_c
            saved = ctx.saved_chunks[i]
            ctx.saved_chunks[i] = None
            if saved is None:
                # re-run this chunk's forward, keeping checkpoints only
                init = ctx.initial[i]
                ctx.initial[i] = None
                state = ([torch.zeros(n_c, setup.nz, setup.ny, setup.nx, device=dev, dtype=dtype)
                          for _ in range(18)] if init is None else init)
                sub = _chunk_setup3(setup, s0, s1, state)
                r_tmp = torch.empty(setup.nt, n_c, setup.n_rec, device=dev, dtype=dtype)
                saved = _forward_chunk(sub, *cs, f_i, r_tmp, policy="checkpoint",
                                       seg=seg if seg > 0 else _default_segment(setup.nt, interval),
                                       interval=interval, fmt=fmt, want_ca=want_ca,
                                       want_cb=want_cb, keep=True)
                seg_i = seg if seg > 0 else _default_segment(setup.nt, interval)
            else:
                sub = _chunk_setup3(setup, s0, s1, [])
                seg_i = seg
            lam = [t[: n_c * numel].view(n_c, setup.nz, setup.ny, setup.nx) for t in lam_all]
            for t in lam:
                t.zero_()
            gr_i = grad_r[:, s0:s1].contiguous() if chunked else grad_r.contiguous()
            gf_i = (torch.zeros(setup.nt, n_c, setup.n_src, device=dev, dtype=dtype)
                    if (want_f and chunked) else grad_f)
            _backward_chunk(sub, *cs, f_i, gr_i, lam, saved, seg=seg_i, interval=interval,
                            fmt=fmt, want_ca=want_ca, want_cb=want_cb, grad_f=gf_i,
                            gca=_shots(gca, s0, s1, setup.batched),
                            gcb=_shots(gcb, s0, s1, setup.batched), pool=pool)
            if want_f and chunked:
                grad_f[:, s0:s1].copy_(gf_i)
            del saved
        out_ca = gca.to(dtype).reshape(ca.shape) if want_ca else None
        out_cb = gcb.to(dtype).reshape(cb.shape) if want_cb else None
        return out_ca, out_cb, None, grad_f, None, None


def _default_segment(nt: int, interval: int) -> int:
    seg = math.ceil(math.sqrt(max(nt, 1)))
    return max(interval, math.ceil(seg / interval) * interval)


def _single(n_shots: int):
    from .tm2d import ShotPlan

    return ShotPlan([0, n_shots], 1)


def _state_bytes_per_shot(setup: EM3DSetup) -> int:
    """Bytes of one checkpoint of one shot: six fields plus the strips of
    the twelve CPML memories."""
    itemsize = setup.ca.element_size()
    numel = setup.nz * setup.ny * setup.nx
    dims = (setup.nz, setup.ny, setup.nx)
    strips = 0
    for name in STATE_NAMES[6:]:
        ax = _MEMORY_AXIS[name]
        lo, hi = setup.strips(ax)
        strips += (lo + dims[ax] - hi) * numel // dims[ax]
    return (6 * numel + strips) * itemsize


def _plan_storage(setup: EM3DSetup, interval: int, bpe: int, limit_device, chunk_steps: int,
                  device: torch.device):
    """Snapshot policy and shot chunks for a model gradient.

    ``("full", 0, plan)`` when every sampled step's six snapshot components
    of every shot fit the device budget; else ``("checkpoint", segment,
    plan)`` with a segment of about sqrt(nt) steps (a multiple of the
    sampling interval), the shots split so that one chunk's checkpoints and
    one segment's snapshots fit, and as many chunks resident as their
    checkpoints allow beside one segment's snapshots.
    """
    from .tm2d import ShotPlan, device_snapshot_budget, plan_shot_chunks

    numel = setup.nz * setup.ny * setup.nx
    n_shots = setup.n_shots
    n_slots = (setup.nt + interval - 1) // interval
    full_ps = 6 * n_slots * numel * bpe
    working = 60 * numel * setup.ca.element_size()
    budget = device_snapshot_budget(device, limit_device, working, n_shots)
    if chunk_steps <= 0 and full_ps * n_shots <= budget:
        return "full", 0, ShotPlan([0, n_shots], 1)
    seg = _default_segment(setup.nt, interval) if chunk_steps <= 0 else \
        max(interval, math.ceil(chunk_steps / interval) * interval)
    ck_ps = math.ceil(setup.nt / seg) * _state_bytes_per_shot(setup)
    seg_ps = math.ceil(seg / interval) * 6 * numel * bpe
    plan = plan_shot_chunks(n_shots, ck_ps + seg_ps, budget, group=1)
    if plan.n_chunks > 1:
        c = max(plan.sizes)
        resident = 0
        while resident < plan.n_chunks and (resident + 1) * c * ck_ps + c * seg_ps <= budget:
            resident += 1
        plan = ShotPlan(plan.bounds, max(1, resident))
    return "checkpoint", seg, plan


def _reject_unsupported(python_backend, stencil, forward_callback, backward_callback,
                        dispersion, device) -> None:
    if python_backend not in (False, "native"):
        raise NotImplementedError(
            "this build has one backend (HIP on MI355X); python_backend is not available")
    if stencil != 2:
        raise NotImplementedError(f"stencil {stencil} is not part of this build (stencil 2 is)")
    if forward_callback is not None or backward_callback is not None:
        raise NotImplementedError("wavefield callbacks are not part of this build")
    if dispersion is not None:
        raise NotImplementedError("Debye dispersion is not part of this build")
    if device.type != "cuda":
        raise NotImplementedError(
            f"the MI355X build runs on HIP devices only; got tensors on {device}")


def maxwell3d(
    epsilon: torch.Tensor,
    sigma: torch.Tensor,
    mu: torch.Tensor,
    grid_spacing: float | Sequence[float],
    dt: float,
    source_amplitude: torch.Tensor | None,
    source_location: torch.Tensor | None,
    receiver_location: torch.Tensor | None,
    stencil: int = 2,
    pml_width: int | Sequence[int] = 20,
    max_vel: float | None = None,
    Ex_0: torch.Tensor | None = None,
    Ey_0: torch.Tensor | None = None,
    Ez_0: torch.Tensor | None = None,
    Hx_0: torch.Tensor | None = None,
    Hy_0: torch.Tensor | None = None,
    Hz_0: torch.Tensor | None = None,
    m_hz_y: torch.Tensor | None = None,
    m_hy_z: torch.Tensor | None = None,
    m_hx_z: torch.Tensor | None = None,
    m_hz_x: torch.Tensor | None = None,
    m_hy_x: torch.Tensor | None = None,
    m_hx_y: torch.Tensor | None = None,
    m_ey_z: torch.Tensor | None = None,
    m_ez_y: torch.Tensor | None = None,
    m_ez_x: torch.Tensor | None = None,
    m_ex_z: torch.Tensor | None = None,
    m_ex_y: torch.Tensor | None = None,
    m_ey_x: torch.Tensor | None = None,
    nt: int | None = None,
    model_gradient_sampling_interval: int = 1,
    freq_taper_frac: float = 0.0,
    time_pad_frac: float = 0.0,
    time_taper: bool = False,
    save_snapshots: bool | None = None,
    forward_callback: Any = None,
    backward_callback: Any = None,
    callback_frequency: int = 1,
    source_component: str = "ey",
    receiver_component: str = "ey",
    python_backend: bool | str = False,
    storage_mode: str = "device",
    storage_path: str = ".",
    storage_compression: bool | str = False,
    storage_bytes_limit_device: int | None = None,
    storage_bytes_limit_host: int | None = None,
    storage_chunk_steps: int = 0,
    n_threads: int | None = None,
    dispersion: Any = None,
    fallback: str = "reference",
    initial: Sequence[torch.Tensor | None] | None = None,
):
    """3-D Maxwell solver (coordinates ``[z, y, x]``).  Returns the 18
    unpadded final state tensors (``Ex .. Hz``, then the CPML memories in
    the reference order) followed by receivers ``[nt, shots, n_rec]``.

    ``storage_chunk_steps`` > 0 forces checkpointing with segments of that
    many steps; otherwise snapshots are kept whole when they fit the device
    budget (``storage_bytes_limit_device``, default 75% of free HBM) and
    checkpointed every ~sqrt(nt) steps when they do not."""
    del freq_taper_frac, time_pad_frac, time_taper, storage_path, n_threads, fallback
    del callback_frequency
    _reject_unsupported(python_backend, stencil, forward_callback, backward_callback,
                        dispersion, epsilon.device)
    spacing = _normalize_grid_spacing_3d(grid_spacing)
    if max_vel is None:
        max_vel = float((1.0 / torch.sqrt(epsilon.detach() * mu.detach())).max().item()) * C0
    _, step_ratio = cfl_condition(spacing, dt, max_vel)
    if step_ratio > 1:
        raise NotImplementedError(
            f"dt={dt} exceeds the CFL limit (needs {step_ratio} internal steps); "
            "internal resampling is not part of this build, pass a stable dt")
    for name, t in (("Source", source_location), ("Receiver", receiver_location)):
        if t is not None and t.numel() > 0:
            hi = torch.tensor(epsilon.shape[-3:], device=t.device)
            if bool((t >= hi).any()) or bool((t < 0).any()):
                raise RuntimeError(
                    f"{name} location out of bounds for model shape {tuple(epsilon.shape[-3:])}")
    init = tuple(initial) if initial is not None else (
        Ex_0, Ey_0, Ez_0, Hx_0, Hy_0, Hz_0, m_hz_y, m_hy_z, m_hx_z, m_hz_x, m_hy_x, m_hx_y,
        m_ey_z, m_ez_y, m_ez_x, m_ex_z, m_ex_y, m_ey_x)
    setup = prepare_em3d(
        epsilon, sigma, mu, spacing, dt, source_amplitude, source_location, receiver_location,
        stencil=stencil, pml_width=pml_width, nt=nt, initial=init,
        source_component=source_component, receiver_component=receiver_component,
    )
    interval = max(1, int(model_gradient_sampling_interval))
    if setup.nt > 0:
        interval = min(interval, setup.nt)
    model_rg = epsilon.requires_grad or sigma.requires_grad
    do_save = model_rg if save_snapshots is None else bool(save_snapshots)
    kind, _, bpe, fmt = _resolve_storage_compression(
        storage_compression, epsilon.dtype, epsilon.device, context="storage_compression")
    numel = setup.nz * setup.ny * setup.nx
    n_store = (setup.nt + interval - 1) // interval
    mode = resolve_storage_mode(
        storage_mode, needs_storage=do_save and model_rg,
        total_bytes=6 * n_store * setup.n_shots * numel * bpe,
        bytes_limit_device=None, bytes_limit_host=None)
    if mode in ("cpu", "disk"):
        raise NotImplementedError(
            f"storage_mode='{mode}' (host-staged snapshots) is not part of this build; "
            "device mode checkpoints and recomputes instead")
    policy, segment, plan = "none", 0, _single(setup.n_shots)
    if mode == "device" and do_save and model_rg:
        policy, segment, plan = _plan_storage(setup, interval, bpe, storage_bytes_limit_device,
                                              int(storage_chunk_steps), epsilon.device)
    opts = {"policy": policy, "segment": segment, "interval": interval, "format": fmt,
            "kind": kind, "plan": plan}
    f = setup.f if setup.f.numel() > 0 else torch.empty(0, device=epsilon.device, dtype=epsilon.dtype)
    out = EM3DPropagate.apply(setup.ca, setup.cb, setup.cq, f, setup, opts)
    fields = [setup.unpad(t) for t in out[:18]]
    r = out[18]
    if setup.n_rec == 0:
        r = torch.empty(0, device=epsilon.device, dtype=epsilon.dtype)
    return (*fields, r)


def separable_gradient3d(
    epsilon: torch.Tensor,
    sigma: torch.Tensor,
    mu: torch.Tensor,
    grid_spacing: float | Sequence[float],
    dt: float,
    source_amplitude: torch.Tensor | None,
    source_location: torch.Tensor | None,
    receiver_location: torch.Tensor | None,
    observed: torch.Tensor,
    *,
    residual: Any,
    stencil: int = 2,
    pml_width: int | Sequence[int] = 20,
    max_vel: float | None = None,
    nt: int | None = None,
    model_gradient_sampling_interval: int = 1,
    source_component: str = "ey",
    receiver_component: str = "ey",
    storage_compression: bool | str = False,
    storage_bytes_limit_device: int | None = None,
    storage_chunk_steps: int = 0,
    targets: Sequence[str] = ("epsilon", "sigma"),
):
    """Gradient of a receiver objective that is a sum over shots (3-D).

    Each shot chunk runs its forward (snapshots, or checkpoints every
    ~sqrt(nt) steps when the snapshots do not fit), its receiver residual
    and its exact adjoint before the next chunk, so no chunk is re-run
    beyond the segment recompute; gradients of all chunks sum into one FP64
    accumulator and the chain rule is applied once by torch autograd.
    Returns ``(value, {"epsilon": g, "sigma": g}, receivers, final fields)``
    plus a dict with the policy (``"full"``/``"checkpoint"``), the segment
    and the chunk sizes.
    """
    eps = epsilon.detach().requires_grad_("epsilon" in targets)
    sig = sigma.detach().requires_grad_("sigma" in targets)
    spacing = _normalize_grid_spacing_3d(grid_spacing)
    if epsilon.ndim != 3:
        raise NotImplementedError("separable_gradient3d takes one model shared by all shots")
    if max_vel is None:
        max_vel = float((1.0 / torch.sqrt(epsilon.detach() * mu.detach())).max().item()) * C0
    if cfl_condition(spacing, dt, max_vel)[1] > 1:
        raise NotImplementedError("dt exceeds the CFL limit; pass a stable dt")
    setup = prepare_em3d(eps, sig, mu.detach(), spacing, dt, source_amplitude, source_location,
                         receiver_location, stencil=stencil, pml_width=pml_width, nt=nt,
                         source_component=source_component,
                         receiver_component=receiver_component)
    interval = max(1, min(int(model_gradient_sampling_interval), max(setup.nt, 1)))
    kind, _, bpe, fmt = _resolve_storage_compression(
        storage_compression, epsilon.dtype, epsilon.device, context="storage_compression")
    policy, seg, plan = _plan_storage(setup, interval, bpe, storage_bytes_limit_device,
                                      int(storage_chunk_steps), epsilon.device)
    # one chunk at a time: the plan's chunks need not be resident together
    if tuple(observed.shape) != (setup.nt, setup.n_shots, setup.n_rec):
        raise ValueError("observed data must be [nt, shots, receivers]")
    dev, dtype = setup.ca.device, setup.ca.dtype
    ca, cb, cq = (t.detach() for t in (setup.ca, setup.cb, setup.cq))
    f = setup.f.detach().contiguous() if setup.f.numel() > 0 else torch.empty(0, device=dev, dtype=dtype)
    fields = [t.contiguous() for t in setup.fields]
    r = torch.zeros(setup.nt, setup.n_shots, setup.n_rec, device=dev, dtype=dtype)
    numel = setup.nz * setup.ny * setup.nx
    max_c = max(plan.sizes)
    lam_all = [torch.empty(max_c * numel, device=dev, dtype=dtype) for _ in range(18)]
    gca = torch.zeros(setup.nz, setup.ny, setup.nx, device=dev, dtype=torch.float64)
    gcb = torch.zeros_like(gca)
    grad_f = torch.zeros(setup.nt, setup.n_shots, setup.n_src, device=dev, dtype=dtype)
    value = torch.zeros((), device=dev, dtype=torch.float64)
    chunked = plan.n_chunks > 1
    pool: dict = {}
    for i in range(plan.n_chunks):
        s0, s1 = plan.bounds[i], plan.bounds[i + 1]
        n_c = s1 - s0
        sub = _chunk_setup3(setup, s0, s1, [t[s0:s1] for t in fields] if chunked else fields)
        f_i = f[:, s0:s1].contiguous() if (chunked and setup.n_src > 0) else f
        r_i = torch.zeros(setup.nt, n_c, setup.n_rec, device=dev, dtype=dtype) if chunked else r
        saved = _forward_chunk(sub, ca, cb, cq, f_i, r_i, policy=policy, seg=seg,
                               interval=interval, fmt=fmt, want_ca=True, want_cb=True, keep=True)
        if chunked:
            r[:, s0:s1].copy_(r_i)
        v_i, g_i = residual(r_i, observed[:, s0:s1])
        value = value + v_i.to(torch.float64)
        lam = [t[: n_c * numel].view(n_c, setup.nz, setup.ny, setup.nx) for t in lam_all]
        for t in lam:
            t.zero_()
        gf_i = grad_f[:, s0:s1].contiguous() if chunked else grad_f
        _backward_chunk(sub, ca, cb, cq, f_i, g_i.to(dtype).contiguous(), lam, saved, seg=seg,
                        interval=interval, fmt=fmt, want_ca=True, want_cb=True,
                        grad_f=gf_i if setup.n_src > 0 else None, gca=gca, gcb=gcb, pool=pool)
        del saved
        if chunked and setup.n_src > 0:
            grad_f[:, s0:s1].copy_(gf_i)
    params = [p for p, n in ((eps, "epsilon"), (sig, "sigma")) if n in targets]
    outs = [setup.ca, setup.cb]
    grads = [gca.to(dtype)[None], gcb.to(dtype)[None]]
    if setup.n_src > 0 and setup.f.requires_grad:
        outs.append(setup.f)
        grads.append(grad_f)
    pg = torch.autograd.grad(outs, params, grads, allow_unused=True)
    names = [n for n in ("epsilon", "sigma") if n in targets]
    unpadded = [setup.unpad(t) for t in fields]
    info = {"policy": policy, "segment": seg, "shot_chunks": plan.sizes}
    return value, dict(zip(names, pg)), r, unpadded, info


def born3d(
    epsilon: torch.Tensor,
    sigma: torch.Tensor,
    mu: torch.Tensor,
    grid_spacing: float | Sequence[float],
    dt: float,
    source_amplitude: torch.Tensor | None,
    source_location: torch.Tensor | None,
    receiver_location: torch.Tensor | None,
    *,
    depsilon: torch.Tensor | None = None,
    dsigma: torch.Tensor | None = None,
    stencil: int = 2,
    pml_width: int | Sequence[int] = 20,
    max_vel: float | None = None,
    nt: int | None = None,
    source_component: str = "ey",
    receiver_component: str = "ey",
    linearize_source: bool = True,
) -> tuple[torch.Tensor, ...]:
    """Born (first-order scattered) 3-D propagation for a model direction
    (``depsilon``, ``dsigma``), the tangent of :func:`maxwell3d` (reference
    ``maxwell/maxwell3d_born.py:27-430``; native semantics
    ``csrc/em3d/maxwell_3d_inst.cpp:1100-1527``).  The background and the
    scattered field advance together on the device; the scattered E update
    adds ``dca E + dcb curl H`` of the background, with ``(dca, dcb)`` the
    closed-form coefficient partials times the padded direction, and with
    ``linearize_source`` the scattered field also takes the source term's
    tangent (the source scaling holds ``cb``).  Returns the 18 scattered
    final arrays (unpadded), the background receivers and the scattered
    receivers ``[nt, N, n_rec]``.  Not differentiable."""
    if epsilon.ndim != 3:
        raise NotImplementedError("born3d supports a single 3-D model")
    if stencil != 2:
        raise NotImplementedError("3-D Born propagation supports stencil 2")
    if epsilon.device.type != "cuda":
        raise NotImplementedError("3-D Born propagation runs on HIP devices only")
    spacing = _normalize_grid_spacing_3d(grid_spacing)
    with torch.no_grad():
        eps, sig, mu_ = epsilon.detach(), sigma.detach(), mu.detach()
        if max_vel is None:
            max_vel = float((1.0 / torch.sqrt(eps * mu_)).max().item()) * C0
        _, step_ratio = cfl_condition(spacing, dt, max_vel)
        if step_ratio > 1:
            raise NotImplementedError(
                f"dt={dt} exceeds the CFL limit (needs {step_ratio} internal steps); "
                "internal resampling is not part of this build, pass a stable dt")
        amp = source_amplitude.detach() if source_amplitude is not None else None
        setup = prepare_em3d(eps, sig, mu_, spacing, dt, amp, source_location,
                             receiver_location, stencil=stencil, pml_width=pml_width, nt=nt,
                             source_component=source_component,
                             receiver_component=receiver_component)
        dev, dtype = setup.ca.device, setup.ca.dtype
        total = [a + b for a, b in zip(setup.fd_pad, setup.pml_width)]
        shape = (setup.nz, setup.ny, setup.nx)

        def padded(t):
            return create_or_pad(t.detach().to(dtype), total, dev, dtype, shape, mode="replicate")

        dca_de, dcb_de, dca_ds, dcb_ds = coefficient_partials(setup.ca[0], setup.cb[0], setup.dt)
        d_ca = d_cb = None
        if depsilon is not None:
            de = padded(depsilon)
            d_ca, d_cb = dca_de * de, dcb_de * de
        if dsigma is not None:
            ds = padded(dsigma)
            d_ca = dca_ds * ds if d_ca is None else d_ca + dca_ds * ds
            d_cb = dcb_ds * ds if d_cb is None else d_cb + dcb_ds * ds
        d_ca = torch.zeros(shape, device=dev, dtype=dtype) if d_ca is None else d_ca.contiguous()
        d_cb = torch.zeros(shape, device=dev, dtype=dtype) if d_cb is None else d_cb.contiguous()
        numel = setup.nz * setup.ny * setup.nx
        f0 = setup.f.contiguous() if setup.f.numel() > 0 else None
        df = None
        if setup.n_src > 0 and f0 is not None:
            if linearize_source:
                # tangent of f = amplitude * cb[src] * (-1/(dz dy dx)) along dcb
                dz, dy, dx = spacing
                dcb_src = d_cb.reshape(numel)[setup.sources_i.to(dev)]  # [N, n_src]
                a = amp.permute(2, 0, 1).contiguous().to(dtype)
                df = (a * dcb_src.unsqueeze(0) * (-1.0 / (dx * dy * dz))).contiguous()
            else:
                df = torch.zeros_like(f0)
        fields = [t.contiguous().clone() for t in setup.fields]
        dfields = [torch.zeros_like(t) for t in fields]
        r = torch.zeros(setup.nt, setup.n_shots, max(setup.n_rec, 1), device=dev, dtype=dtype)
        bg_r = torch.zeros_like(r)
        fn = getattr(_native.load(), f"tide_em3d_{_native.dtype_name(dtype)}_born_forward")
        dev_idx = dev.index if dev.index is not None else torch.cuda.current_device()
        _native.call(
            fn, _native.ptr(setup.ca), _native.ptr(setup.cb), _native.ptr(setup.cq),
            _native.ptr(d_ca), _native.ptr(d_cb), _native.ptr(f0), _native.ptr(df),
            _native.ptr_array(fields), _native.ptr_array(dfields), _native.ptr(r),
            _native.ptr(bg_r), _native.ptr_array(setup.profiles),
            _native.ptr(setup.sources_i), _native.ptr(setup.receivers_i),
            setup.rdz, setup.rdy, setup.rdx, setup.nt, setup.n_shots, setup.nz, setup.ny,
            setup.nx, setup.n_src, setup.n_rec, setup.batched, setup.batched, setup.batched, 0,
            _native.int64_array(setup.pml), setup.source_component, setup.receiver_component,
            dev_idx, _native.stream_handle(dev),
        )
        r, bg_r = r[:, :, : setup.n_rec], bg_r[:, :, : setup.n_rec]
        return (*[setup.unpad(t) for t in dfields], bg_r, r)


__all__ = ["EM3DPropagate", "EM3DSetup", "STATE_NAMES", "born3d", "maxwell3d", "prepare_em3d",
           "separable_gradient3d"]
