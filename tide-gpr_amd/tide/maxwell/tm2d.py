"""2-D TM propagation on the MI355X HIP kernels.

``maxwelltm`` keeps the reference entry point's signature and semantics
(reference ``src/tide/maxwell/tm2d.py:258-745`` and the native adapter
``tm2d_cuda.py:29-823``):

* the model is replicate-padded by ``fd_pad + pml`` per side
  (``fd_pad = [s/2, s/2-1, s/2, s/2-1]``) so its gradient folds back onto
  the edge cells;
* ``ca, cb, cq`` come from ``utils.prepare_parameters``; the source term is
  ``f[t, shot, k] = amplitude * cb[src] * (-1/(dy dx))``, differentiable in
  both factors;
* :class:`TM2DPropagate` runs the forward kernels and, in its backward, the
  exact adjoint, returning ``dJ/dca``, ``dJ/dcb`` (FP64-accumulated) and
  ``dJ/df``; torch's autograd then applies the coefficient, padding and
  source-scaling chain rules exactly as in the reference;
* outputs are the unpadded final state and receivers ``[nt, shots, n_rec]``.

Snapshots stay in HBM (``storage_mode="device"``, shot chunks when they do
not fit) or are staged through pinned host memory (``"cpu"``) or a file
(``"disk"``) by the native library; ``forward_callback`` /
``backward_callback`` run the time loop in chunks of ``callback_frequency``
steps (reference ``tm2d_autograd.py:165-261, 677-802``).

Only the HIP device path exists.  CPU tensors, the Python reference
backend, dispersion and internal resampling
(step_ratio > 1) raise ``NotImplementedError``.
"""

from __future__ import annotations

import ctypes
import math
import tempfile
from collections.abc import Sequence
from dataclasses import dataclass, field
from pathlib import Path
from typing import Any

import torch

from .. import _native, staggered
from ..callbacks import CallbackState
from ..cfl import cfl_condition
from ..grid_utils import _normalize_grid_spacing_2d, _normalize_pml_width_2d
from ..padding import create_or_pad, zero_interior
from ..storage import (
    STORAGE_CPU,
    STORAGE_DEVICE,
    STORAGE_DISK,
    STORAGE_FORMAT_BF16,
    STORAGE_NONE,
    _resolve_storage_compression,
    resolve_storage_mode,
)
from ..utils import C0, compile_material_coefficients
from ..utils import validate_material_inputs as validate_debye

_FIELD_NAMES = ("Ey", "Hx", "Hz", "m_Ey_x", "m_Ey_z", "m_Hx_z", "m_Hz_x")


class KernelTimer:
    """Optional HIP-event timing of the native calls (one sweep each).

    When installed with :func:`set_kernel_timer`, every forward and adjoint
    sweep is bracketed by events on the stream the kernels run on, and the
    record ``(pass, n_steps, cells_per_step, start, end)`` is kept so a
    caller can derive per-launch kernel durations without a profiler.
    """

    def __init__(self) -> None:
        self.records: list[tuple[str, int, int, int, torch.cuda.Event, torch.cuda.Event]] = []

    def start(self) -> torch.cuda.Event:
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        return ev

    def stop(self, name: str, steps: int, cells: int, start: torch.cuda.Event,
             steps_per_launch: int = 1) -> None:
        ev = torch.cuda.Event(enable_timing=True)
        ev.record()
        self.records.append((name, steps, cells, steps_per_launch, start, ev))

    def summary(self) -> dict[str, dict[str, float]]:
        """Per pass: total ms, launches (time steps / steps per launch), time
        steps, and cell-updates (after a sync)."""
        out: dict[str, dict[str, float]] = {}
        for name, steps, cells, spl, a, b in self.records:
            d = out.setdefault(name, {"ms": 0.0, "launches": 0, "steps": 0, "cell_updates": 0.0,
                                      "steps_per_launch": spl})
            d["ms"] += a.elapsed_time(b)
            d["launches"] += (steps + spl - 1) // spl
            d["steps"] += steps
            d["cell_updates"] += float(steps) * cells
        return out


_timer: KernelTimer | None = None


def set_kernel_timer(timer: KernelTimer | None) -> None:
    global _timer
    _timer = timer


@dataclass
class TM2DSetup:
    """Everything the kernels read for one propagation, in padded layout."""

    ca: torch.Tensor  # [1|N, ny, nx]
    cb: torch.Tensor
    cq: torch.Tensor
    f: torch.Tensor  # [nt, N, n_src] (empty when no sources)
    profiles: list[torch.Tensor]  # ay, by, ayh, byh, ax, bx, axh, bxh, ky, kyh, kx, kxh
    sources_i: torch.Tensor  # [N, n_src] int64, CPU
    receivers_i: torch.Tensor  # [N, n_rec] int64, CPU
    fields: list[torch.Tensor]  # 7 x [N, ny, nx]
    rdy: float
    rdx: float
    dt: float
    nt: int
    n_shots: int
    ny: int
    nx: int
    n_src: int
    n_rec: int
    pml: tuple[int, int, int, int]  # pml_y0, pml_x0, pml_y1, pml_x1
    fd_pad: list[int]
    pml_width: list[int]
    batched: bool
    stencil: int = 2
    extra: dict[str, Any] = field(default_factory=dict)

    def unpad(self, t: torch.Tensor) -> torch.Tensor:
        p = self.fd_pad
        return t[:, p[0] : self.ny - p[1] if p[1] > 0 else None, p[2] : self.nx - p[3] if p[3] > 0 else None]


def _flat_indices(loc: torch.Tensor | None, pad_y: int, pad_x: int, nx_p: int) -> tuple[torch.Tensor, int]:
    if loc is None or loc.numel() == 0:
        return torch.empty(0, dtype=torch.long), 0
    loc = loc.detach().to("cpu")
    idx = ((loc[..., 0] + pad_y) * nx_p + (loc[..., 1] + pad_x)).long().contiguous()
    return idx, int(loc.shape[1])


def prepare_tm2d(
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
    initial: Sequence[torch.Tensor | None] = (None,) * 7,
    validate_material_inputs: bool = True,
    allocate_fields: bool = True,
    max_vel: float | None = None,
    dispersion: Any = None,
) -> TM2DSetup:
    """Pad the model, compile coefficients and profiles, map the points.

    ``epsilon`` is ``[ny, nx]`` (shared by all shots) or ``[N, ny, nx]`` (one
    model per shot).  Device-agnostic: runs on CPU tensors too, so the host
    numerics can be tested without a GPU.  ``allocate_fields=False`` leaves
    ``fields`` empty for callers that bring their own (zeroed) state.
    ``dispersion`` (Debye poles, one model): the Debye coefficients replace
    ``ca, cb`` and ``extra["debye"]`` holds the padded pole coefficients
    ``a, b, cp`` ``[n_poles, ny, nx]`` (reference ``tm2d_cuda.py:155-187``).
    """
    if epsilon.ndim not in (2, 3):
        raise RuntimeError("epsilon must be 2D or batched 3D")
    if sigma.shape != epsilon.shape or mu.shape != epsilon.shape:
        raise RuntimeError("sigma and mu must have the same shape as epsilon")
    device, dtype = epsilon.device, epsilon.dtype
    batched = epsilon.ndim == 3
    dy, dx = _normalize_grid_spacing_2d(grid_spacing)
    widths = _normalize_pml_width_2d(pml_width)
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
    fd_pad = [half, half - 1, half, half - 1]
    total = [a + b for a, b in zip(fd_pad, widths)]
    ny = int(epsilon.shape[-2]) + total[0] + total[1]
    nx = int(epsilon.shape[-1]) + total[2] + total[3]
    size = (n_shots, ny, nx) if batched else (ny, nx)
    eps_p = create_or_pad(epsilon, total, device, dtype, size, mode="replicate")
    sig_p = create_or_pad(sigma, total, device, dtype, size, mode="replicate")
    mu_p = create_or_pad(mu, total, device, dtype, size, mode="replicate")
    if dispersion is not None:
        if batched:
            raise NotImplementedError("Debye dispersion with one model per shot is not supported")
        if validate_material_inputs:
            validate_debye(epsilon, dispersion=dispersion, dt=dt)
        dispersion = _pad_debye(dispersion, tuple(epsilon.shape[-2:]), total, (ny, nx),
                                device, dtype)
    coef = compile_material_coefficients(eps_p, sig_p, mu_p, dt, dispersion=dispersion,
                                         validate_inputs=validate_material_inputs)
    shape3 = (lambda t: t.contiguous()) if batched else (lambda t: t[None].contiguous())
    ca, cb, cq = shape3(coef["ca"]), shape3(coef["cb"]), shape3(coef["cq"])

    sources_i, n_src = _flat_indices(source_location, total[0], total[2], nx)
    receivers_i, n_rec = _flat_indices(receiver_location, total[0], total[2], nx)

    source_coeff = -1.0 / (dx * dy)
    if n_src > 0 and source_amplitude is not None and source_amplitude.numel() > 0:
        flat_cb = cb.reshape(cb.shape[0], ny * nx).expand(n_shots, -1)
        cb_src = flat_cb.gather(1, sources_i.to(device))
        amp = source_amplitude.permute(2, 0, 1).contiguous().to(dtype)
        f = amp * cb_src.to(dtype).unsqueeze(0)
        f = (f * source_coeff).contiguous()
    else:
        f = torch.empty(0, device=device, dtype=dtype)

    fields = []
    for t0 in initial if allocate_fields else ():
        if t0 is None:
            fields.append(torch.zeros(n_shots, ny, nx, device=device, dtype=dtype))
            continue
        t0 = t0.to(device=device, dtype=dtype)
        if t0.ndim == 2:
            t0 = t0[None].expand(n_shots, -1, -1)
        if tuple(t0.shape) == (n_shots, ny, nx):
            fields.append(t0.detach().clone().contiguous())
        else:
            fields.append(
                create_or_pad(t0.detach(), fd_pad, device, dtype, (n_shots, ny, nx)).contiguous()
            )
    for wf, dim in zip(fields[3:], (1, 0, 0, 1)):
        zero_interior(wf, fd_pad, widths, dim)

    if max_vel is None:  # a caller's max_vel also sets the profiles (reference tm2d_cuda.py:124-130)
        max_vel = float((1.0 / torch.sqrt(epsilon.detach() * mu.detach())).max().item()) * C0
    ay, ayh, ax, axh, by, byh, bx, bxh, ky, kyh, kx, kxh = _profiles(
        tuple(widths), int(stencil), tuple(fd_pad), float(dt), (float(dy), float(dx)),
        float(max_vel), dtype, device, ny, nx)
    pml = (total[0], total[2], ny - fd_pad[1] - widths[1], nx - fd_pad[3] - widths[3])
    return TM2DSetup(
        ca=ca, cb=cb, cq=cq, f=f,
        profiles=[ay, by, ayh, byh, ax, bx, axh, bxh, ky, kyh, kx, kxh],
        sources_i=sources_i, receivers_i=receivers_i, fields=fields,
        rdy=1.0 / dy, rdx=1.0 / dx, dt=float(dt), nt=nt, n_shots=n_shots,
        ny=ny, nx=nx, n_src=n_src, n_rec=n_rec, pml=pml, fd_pad=fd_pad,
        pml_width=widths, batched=batched, stencil=int(stencil),
        extra={"zero_init": all(t0 is None for t0 in initial),
               "debye": ({k: coef["debye"][k].contiguous() for k in ("a", "b", "cp")}
                         if coef["has_dispersion"] else None),
               # padded models, for wavefield callbacks
               "models": {"epsilon": eps_p.detach(), "sigma": sig_p.detach(),
                          "mu": mu_p.detach()}},
    )


_PROFILE_CACHE: dict[tuple, list[torch.Tensor]] = {}
_PROFILE_CACHE_SIZE = 32


def _profiles(widths, stencil, fd_pad, dt, spacing, max_vel, dtype, device, ny, nx):
    """The 12 CPML profiles (flat, contiguous) of one geometry.  They are a
    pure function of these arguments and never written, so repeated calls
    on the same geometry share one set (about 50 small device operations
    fewer per call; a launch-bound small problem spends most of its host
    set-up here).  At most 32 geometries are kept, least recently used
    first out."""
    key = (widths, stencil, fd_pad, dt, spacing, max_vel, dtype, str(device), ny, nx)
    hit = _PROFILE_CACHE.pop(key, None)
    if hit is None:
        prof = staggered.set_pml_profiles(
            pml_width=list(widths), accuracy=stencil, fd_pad=list(fd_pad), dt=dt,
            grid_spacing=list(spacing), max_vel=max_vel, dtype=dtype, device=device,
            pml_freq=0.5 / dt, ny=ny, nx=nx,
        )
        hit = [p.reshape(-1).contiguous() for p in prof]
        while len(_PROFILE_CACHE) >= _PROFILE_CACHE_SIZE:
            _PROFILE_CACHE.pop(next(iter(_PROFILE_CACHE)))
    _PROFILE_CACHE[key] = hit
    return hit


def _pad_debye(dispersion, model_shape, total, padded, device, dtype):
    """Debye parameters on the padded grid: model-shaped ones (per pole)
    replicate-padded like the model, scalars kept (reference
    ``maxwell/common.py:_pad_dispersion_for_model``)."""
    from ..dispersion import DebyeDispersion

    def pad(value):
        t = torch.as_tensor(value, device=device, dtype=dtype)
        if t.ndim == 0:
            return t
        if t.ndim <= len(model_shape):
            return create_or_pad(torch.broadcast_to(t, model_shape), total, device, dtype, padded,
                                 mode="replicate")
        if t.ndim == len(model_shape) + 1:
            return torch.stack([create_or_pad(torch.broadcast_to(t[p], model_shape), total,
                                              device, dtype, padded, mode="replicate")
                                for p in range(t.shape[0])])
        raise ValueError("Debye dispersion parameters must be scalar, model-shaped, or "
                         f"[n_poles, *model_shape], but got shape {tuple(t.shape)}.")

    return DebyeDispersion(delta_epsilon=pad(dispersion.delta_epsilon), tau=pad(dispersion.tau))


def _profile_ptrs(profiles: list[torch.Tensor]) -> list[int | None]:
    return [_native.ptr(p) for p in profiles]


def _chunk_setup(setup: TM2DSetup, s0: int, s1: int, fields: list[torch.Tensor]) -> TM2DSetup:
    """The shots ``[s0, s1)`` of ``setup`` as a setup of their own (views)."""
    if s0 == 0 and s1 == setup.n_shots and fields is setup.fields:
        return setup
    sub = TM2DSetup(**{k: getattr(setup, k) for k in setup.__dataclass_fields__})
    sub.n_shots = s1 - s0
    sub.sources_i = setup.sources_i[s0:s1].contiguous() if setup.n_src > 0 else setup.sources_i
    sub.receivers_i = setup.receivers_i[s0:s1].contiguous() if setup.n_rec > 0 else setup.receivers_i
    sub.fields = fields
    # a chunk keeps the Debye pole coefficients (shared by all shots); its
    # polarization state is its own (shots are independent)
    sub.extra = {"debye": setup.extra.get("debye")}
    return sub


def _shots(t: torch.Tensor, s0: int, s1: int, batched: bool) -> torch.Tensor:
    return t[s0:s1] if batched else t


class _HostStaged:
    """Snapshots of one propagation staged outside HBM (storage_mode
    ``"cpu"`` or ``"disk"``): per snapshot stream (Ey, curl H) a device ring
    of three stored steps that the kernels read and write, pinned host
    memory behind it (the whole history for ``cpu``; a three-step ring in
    front of a file for ``disk``) and the stream the native library copies
    on (reference ``storage.py:SnapshotAllocator``,
    ``tm2d_helpers.py:_make_tm_storage_streams``)."""

    RING = 3

    def __init__(self, mode: str, want: tuple[bool, bool], n_store: int, shot_shape, sdtype,
                 device: torch.device, path: str) -> None:
        self.mode = STORAGE_CPU if mode == "cpu" else STORAGE_DISK
        self.rings = [torch.empty(self.RING, *shot_shape, device=device, dtype=sdtype) if w else None
                      for w in want]
        rows = n_store if self.mode == STORAGE_CPU else self.RING
        self.hosts = [torch.empty(rows, *shot_shape, dtype=sdtype, pin_memory=True) if w else None
                      for w in want]
        self.tmp = None
        self.names: list[Any] = [None, None]
        self._keep: list[Any] = []
        if self.mode == STORAGE_DISK:
            Path(path).mkdir(parents=True, exist_ok=True)
            self.tmp = tempfile.TemporaryDirectory(prefix="tide_tmp_", dir=path)
            for k, (w, stem) in enumerate(zip(want, ("ey", "curl"))):
                if not w:
                    continue
                buf = ctypes.create_string_buffer(str(Path(self.tmp.name) / f"{stem}.bin").encode())
                arr = (ctypes.c_char_p * 1)(ctypes.cast(buf, ctypes.c_char_p))
                self._keep += [buf, arr]
                self.names[k] = ctypes.cast(arr, ctypes.c_void_p).value
        self.stream = torch.cuda.Stream(device)  # default priority

    def args(self, k: int) -> tuple:
        """(store_1, store_3, filenames) of stream k (0 Ey, 1 curl H)."""
        return (_native.ptr(self.rings[k]), _native.ptr(self.hosts[k]), self.names[k])

    def stream_handle(self) -> int:
        return self.stream.cuda_stream

    def close(self) -> None:
        """Waits for the copies still in flight, then drops the buffers."""
        self.stream.synchronize()
        self.rings = [None, None]
        self.hosts = [None, None]
        if self.tmp is not None:
            self.tmp.cleanup()
            self.tmp = None


def _run_forward(setup: TM2DSetup, ca, cb, cq, f, fields, r, *, interval: int, fmt: int,
                 ey_store=None, curl_store=None, start_t: int = 0, nt: int | None = None,
                 staged: _HostStaged | None = None, want: tuple[bool, bool] = (False, False)) -> None:
    """One native forward over steps ``start_t .. start_t+nt-1`` (default:
    the whole run) on the caller's stream; stores snapshots when a store
    buffer is given, or stages them through ``staged`` (``want``: which
    snapshot streams)."""
    dev, dtype = ca.device, ca.dtype
    nt = setup.nt if nt is None else nt
    stream = _native.stream_handle(dev)
    dev_idx = dev.index if dev.index is not None else torch.cuda.current_device()
    tail = (
        _native.ptr(setup.sources_i), _native.ptr(setup.receivers_i),
        setup.rdy, setup.rdx, setup.dt,
        nt, setup.n_shots, setup.ny, setup.nx, setup.n_src, setup.n_rec,
    )
    storing = staged is not None or ey_store is not None or curl_store is not None
    if storing:
        fn = _native.get_function("maxwell_tm", "forward_with_storage", setup.stencil, dtype)
        shot_bytes = setup.ny * setup.nx * (2 if fmt == STORAGE_FORMAT_BF16 else dtype.itemsize)
        if staged is not None:
            mode, (e1, e3, en), (c1, c3, cn) = staged.mode, staged.args(0), staged.args(1)
            w_ca, w_cb, s_stream = want[0], want[1], staged.stream_handle()
        else:
            mode, e1, e3, en = STORAGE_DEVICE, _native.ptr(ey_store), None, None
            c1, c3, cn = _native.ptr(curl_store), None, None
            w_ca, w_cb, s_stream = ey_store is not None, curl_store is not None, stream
        _native.call(
            fn,
            _native.ptr(ca), _native.ptr(cb), _native.ptr(cq), _native.ptr(f),
            *[_native.ptr(t) for t in fields], _native.ptr(r),
            e1, e3, en, c1, c3, cn,
            *_profile_ptrs(setup.profiles), *tail, interval,
            mode, fmt, shot_bytes, w_ca, w_cb,
            setup.batched, setup.batched, setup.batched, start_t,
            *setup.pml, 0, dev_idx, stream, s_stream,
        )
    else:
        fn = _native.get_function("maxwell_tm", "forward", setup.stencil, dtype)
        debye = setup.extra.get("debye")
        if debye is not None:
            # polarization and Ey^n buffers of the Debye recurrence, kept on
            # the setup so time chunks of one run continue it
            state = setup.extra.setdefault("debye_state", {
                "pol": torch.zeros(setup.n_shots, debye["a"].shape[0], setup.ny, setup.nx,
                                   device=dev, dtype=dtype),
                "ey_prev": torch.zeros(setup.n_shots, setup.ny, setup.nx, device=dev, dtype=dtype),
            })
            deb = (_native.ptr(debye["a"]), _native.ptr(debye["b"]), _native.ptr(debye["cp"]),
                   _native.ptr(state["pol"]), _native.ptr(state["ey_prev"]))
            n_poles = int(debye["a"].shape[0])
        else:
            deb, n_poles = (None,) * 5, 0
        _native.call(
            fn,
            _native.ptr(ca), _native.ptr(cb), _native.ptr(cq), _native.ptr(f),
            *[_native.ptr(t) for t in fields],
            *deb, _native.ptr(r), n_poles,
            *_profile_ptrs(setup.profiles), *tail, interval, debye is not None,
            setup.batched, setup.batched, setup.batched, start_t,
            *setup.pml, 0, dev_idx, stream,
        )


def _run_backward(setup: TM2DSetup, ca, cb, cq, grad_r, lam, ey_store, curl_store, grad_f,
                  gca, gcb, *, interval: int, fmt: int, start_t: int | None = None,
                  nt: int | None = None, staged: _HostStaged | None = None) -> None:
    """One native exact-adjoint sweep over steps ``start_t-1`` down to
    ``start_t-nt`` (default: the whole run); ``gca``/``gcb`` are FP64
    accumulators (added to)."""
    dev, dtype = ca.device, ca.dtype
    nt = setup.nt if nt is None else nt
    start_t = setup.nt if start_t is None else start_t
    shot_bytes = setup.ny * setup.nx * (2 if fmt == STORAGE_FORMAT_BF16 else dtype.itemsize)
    fn = getattr(_native.load(), f"tide_tm2d_{_native.dtype_name(dtype)}_backward_f64acc")
    dev_idx = dev.index if dev.index is not None else torch.cuda.current_device()
    stream = _native.stream_handle(dev)
    if staged is not None:
        mode, (e1, e3, en), (c1, c3, cn) = staged.mode, staged.args(0), staged.args(1)
        s_stream = staged.stream_handle()
    else:
        mode, e1, e3, en = STORAGE_DEVICE, _native.ptr(ey_store), None, None
        c1, c3, cn = _native.ptr(curl_store), None, None
        s_stream = stream
    _native.call(
        fn,
        _native.ptr(ca), _native.ptr(cb), _native.ptr(cq), _native.ptr(grad_r),
        *[_native.ptr(t) for t in lam],
        e1 if gca is not None else None, e3, en,
        c1 if gcb is not None else None, c3, cn,
        _native.ptr(grad_f), _native.ptr(gca), _native.ptr(gcb),
        *_profile_ptrs(setup.profiles),
        _native.ptr(setup.sources_i), _native.ptr(setup.receivers_i),
        setup.rdy, setup.rdx, nt, setup.n_shots, setup.ny, setup.nx,
        setup.n_src, setup.n_rec, interval, mode, fmt, shot_bytes,
        setup.batched, setup.batched, setup.batched, start_t, *setup.pml,
        setup.stencil, dev_idx, stream, s_stream,
    )


def _time_chunks(nt: int, every: int | None, reverse: bool = False) -> list[tuple[int, int]]:
    """``(start, steps)`` of the time chunks of a sweep: one chunk, or
    chunks of ``every`` steps for callbacks (reverse: from the end, as the
    adjoint runs, with ``start`` the chunk's upper end)."""
    if not every or every >= nt:
        return [(nt if reverse else 0, nt)]
    if not reverse:
        return [(s, min(every, nt - s)) for s in range(0, nt, every)]
    return [(s, min(s, every)) for s in range(nt, 0, -every)]


def _callback_state(setup: TM2DSetup, step: int, wavefields, gradients, backward: bool):
    models = dict(setup.extra.get("models", {}))
    models.update(ca=setup.ca, cb=setup.cb, cq=setup.cq)
    return CallbackState(dt=setup.dt, step=step, nt=setup.nt, wavefields=wavefields,
                         models=models, gradients=gradients, fd_pad=list(setup.fd_pad),
                         pml_width=list(setup.pml_width), is_backward=backward)


def _model_gradients(setup: TM2DSetup, gca, gcb) -> dict[str, torch.Tensor]:
    """Gradients a backward callback sees: dJ/dca, dJ/dcb so far and their
    epsilon / sigma images through the coefficient formulas (padded grid;
    the source term and the pad fold are applied after the sweep)."""
    out: dict[str, torch.Tensor] = {}
    dtype = setup.ca.dtype
    if gca is not None:
        out["ca"] = gca.to(dtype)
    if gcb is not None:
        out["cb"] = gcb.to(dtype)
    models = setup.extra.get("models")
    if not out or not models:
        return out
    with torch.enable_grad():
        eps = models["epsilon"].detach().requires_grad_(True)
        sig = models["sigma"].detach().requires_grad_(True)
        coef = compile_material_coefficients(eps, sig, models["mu"], setup.dt, validate_inputs=False)
        outs, grads = [], []
        if gca is not None:
            outs.append(coef["ca"])
            grads.append(out["ca"].reshape(coef["ca"].shape))
        if gcb is not None:
            outs.append(coef["cb"])
            grads.append(out["cb"].reshape(coef["cb"].shape))
        ge, gs = torch.autograd.grad(outs, (eps, sig), grads, allow_unused=True)
    out["epsilon"] = ge if ge is not None else torch.zeros_like(eps)
    out["sigma"] = gs if gs is not None else torch.zeros_like(sig)
    return out


@dataclass
class ShotPlan:
    """Shot chunks of one propagation: chunk ``i`` holds shots
    ``bounds[i]:bounds[i+1]``; the first ``resident`` chunks keep their
    snapshots in device memory between the forward and the backward, the
    others are re-run (forward with storage) inside the backward."""

    bounds: list[int]
    resident: int

    @property
    def n_chunks(self) -> int:
        return len(self.bounds) - 1

    @property
    def sizes(self) -> list[int]:
        return [b - a for a, b in zip(self.bounds[:-1], self.bounds[1:])]


def plan_shot_chunks(n_shots: int, per_shot_bytes: int, budget_bytes: float,
                     group: int = 4) -> ShotPlan:
    """Split ``n_shots`` so that one chunk's snapshots fit ``budget_bytes``.

    Chunks are balanced; a chunk of at least ``group`` shots is cut to a
    multiple of ``group`` (the adjoint's shots per workgroup) so its
    workgroups are full.  As many chunks as the budget holds together stay
    resident; the rest are recomputed in the backward.  Raises when not even
    one shot's snapshots fit.
    """
    if n_shots <= 0:
        return ShotPlan([0, max(n_shots, 0)], 1)
    if per_shot_bytes <= 0 or per_shot_bytes * n_shots <= budget_bytes:
        return ShotPlan([0, n_shots], 1)
    cap = int(budget_bytes // per_shot_bytes)
    if cap < 1:
        raise RuntimeError(
            f"the snapshots of one shot ({per_shot_bytes / 1e9:.2f} GB) exceed the device "
            f"budget ({budget_bytes / 1e9:.2f} GB); use BF16 snapshot compression, a larger "
            "model_gradient_sampling_interval or storage_mode='cpu'/'disk'"
        )
    if cap >= group:
        cap = cap // group * group
    k = -(-n_shots // cap)
    base, extra = divmod(n_shots, k)
    sizes = [base + (1 if i < extra else 0) for i in range(k)]
    bounds = [0]
    for c in sizes:
        bounds.append(bounds[-1] + c)
    resident, used = 0, 0
    for c in sizes:
        if used + c * per_shot_bytes > budget_bytes:
            break
        used += c * per_shot_bytes
        resident += 1
    return ShotPlan(bounds, resident)


def device_snapshot_budget(device: torch.device, limit: int | None, working_per_shot: int,
                           n_shots: int) -> float:
    """Bytes of snapshots one call may hold on ``device``: the caller's
    limit, else 90% of the free memory (device free memory plus what torch's
    caching allocator holds unused) less the working state."""
    if limit is not None:
        return float(limit)
    free, _ = torch.cuda.mem_get_info(device)
    # memory torch's caching allocator holds but does not use is free to it
    free += torch.cuda.memory_reserved(device) - torch.cuda.memory_allocated(device)
    return max(0.0, 0.9 * free - 2.0 * working_per_shot * n_shots - (256 << 20))


class TM2DPropagate(torch.autograd.Function):
    """Forward kernels + exact adjoint behind torch autograd.

    Inputs: ``ca, cb, cq`` ``[1|N, ny, nx]``, source term ``f``
    ``[nt, N, n_src]``; the setup object and storage policy ride along.
    Returns the 7 padded final fields and receivers ``[nt, N, n_rec]``.
    Gradients flow to ``ca``, ``cb`` and ``f`` from the receiver
    cotangent; cotangents of the final state are ignored, as in the
    reference (``tm2d_autograd.py:519-547``).

    With device storage, shots run in the chunks of ``storage["plan"]``:
    chunks whose snapshots fit the device budget keep them until the
    backward, the others are re-run from their initial state in the
    backward.  With ``cpu`` / ``disk`` storage all shots run together and
    the native library stages their snapshots through host memory.  One
    FP64 accumulator per coefficient collects every chunk's gradient, cast
    to the field dtype once at the end.  Callbacks split each sweep into
    time chunks of ``storage["callback_frequency"]`` steps.
    """

    @staticmethod
    def forward(ctx, ca, cb, cq, f, setup: TM2DSetup, storage: dict[str, Any]):
        dev = ca.device
        dtype = ca.dtype
        fields = [t.contiguous() for t in setup.fields]
        r = torch.zeros(setup.nt, setup.n_shots, setup.n_rec, device=dev, dtype=dtype)
        f_c = f.detach().contiguous()
        interval = storage["interval"]
        fmt = storage["format"]
        want_ca = ctx.needs_input_grad[0]
        want_cb = ctx.needs_input_grad[1]
        mode = storage["mode"]
        store = mode in ("device", "cpu", "disk") and (want_ca or want_cb)
        staged_mode = store and mode in ("cpu", "disk")
        plan: ShotPlan = storage.get("plan") or ShotPlan([0, setup.n_shots], 1)
        n_store = (setup.nt + interval - 1) // interval
        sdtype = torch.int16 if fmt == STORAGE_FORMAT_BF16 else dtype
        chunked = plan.n_chunks > 1
        fwd_cb = storage.get("forward_callback")
        every = storage.get("callback_frequency") if fwd_cb is not None else None
        stores: list[Any] = []
        initial: list[list[torch.Tensor] | None] = []
        t0 = _timer.start() if _timer is not None else None
        for i in range(plan.n_chunks):
            s0, s1 = plan.bounds[i], plan.bounds[i + 1]
            sub = _chunk_setup(setup, s0, s1, [t[s0:s1] for t in fields] if chunked else fields)
            keep = store and i < plan.resident
            if store and not keep:
                # re-run in the backward from this initial state
                initial.append(None if setup.extra.get("zero_init") else
                               [t.clone() for t in sub.fields])
            else:
                initial.append(None)
            ey_s = curl_s = staged = None
            if keep and staged_mode:
                staged = _HostStaged(mode, (want_ca, want_cb), n_store,
                                     (s1 - s0, setup.ny, setup.nx), sdtype, dev,
                                     storage.get("path", "."))
            elif keep:
                shape = (n_store, s1 - s0, setup.ny, setup.nx)
                ey_s = torch.empty(shape, device=dev, dtype=sdtype) if want_ca else None
                curl_s = torch.empty(shape, device=dev, dtype=sdtype) if want_cb else None
            stores.append(staged if staged is not None else ((ey_s, curl_s) if keep else None))
            f_i = f_c[:, s0:s1].contiguous() if (chunked and setup.n_src > 0) else f_c
            r_i = torch.zeros(setup.nt, s1 - s0, setup.n_rec, device=dev, dtype=dtype) if chunked else r
            for start, steps in _time_chunks(setup.nt, every):
                _run_forward(sub, _shots(ca, s0, s1, setup.batched), _shots(cb, s0, s1, setup.batched),
                             _shots(cq, s0, s1, setup.batched), f_i, sub.fields, r_i,
                             interval=interval, fmt=fmt, ey_store=ey_s, curl_store=curl_s,
                             start_t=start, nt=steps, staged=staged, want=(want_ca, want_cb))
                if fwd_cb is not None:
                    fwd_cb(_callback_state(setup, start + steps, dict(zip(_FIELD_NAMES, sub.fields)),
                                           {}, False))
            if staged is not None:
                staged.stream.synchronize()  # the host history is complete
            if chunked and setup.n_rec > 0:
                r[:, s0:s1].copy_(r_i)
        if t0 is not None:
            _timer.stop("forward_store" if store else "forward", setup.nt,
                        setup.n_shots * setup.ny * setup.nx, t0, _native.get_tm2d_time_block()[0])
        ctx.setup = setup
        ctx.storage = storage
        ctx.stored = store
        ctx.plan = plan
        ctx.stores = stores
        ctx.initial = initial
        ctx.f_data = f_c if plan.resident < plan.n_chunks else None
        ctx.save_for_backward(ca, cb, cq)
        ctx.mark_non_differentiable(*fields)
        return (*fields, r)

    @staticmethod
    def backward(ctx, *grads):
        grad_r = grads[-1]
        setup: TM2DSetup = ctx.setup
        plan: ShotPlan = ctx.plan
        ca, cb, cq = ctx.saved_tensors
        want_ca = ctx.needs_input_grad[0]
        want_cb = ctx.needs_input_grad[1]
        want_f = ctx.needs_input_grad[3] and setup.n_src > 0
        if (want_ca or want_cb) and not ctx.stored:
            raise RuntimeError(
                "model gradients need stored snapshots; storage_mode='none' was used"
            )
        dev, dtype = ca.device, ca.dtype
        interval = ctx.storage["interval"]
        fmt = ctx.storage["format"]
        bwd_cb = ctx.storage.get("backward_callback")
        every = ctx.storage.get("callback_frequency") if bwd_cb is not None else None
        if grad_r is None:
            grad_r = torch.zeros(setup.nt, setup.n_shots, setup.n_rec, device=dev, dtype=dtype)
        grad_r = grad_r.to(dtype)
        chunked = plan.n_chunks > 1
        max_c = max(plan.sizes)
        lam_all = [torch.empty(max_c, setup.ny, setup.nx, device=dev, dtype=dtype) for _ in range(7)]
        grad_f = (
            torch.zeros(setup.nt, setup.n_shots, setup.n_src, device=dev, dtype=dtype)
            if want_f else None
        )
        gshape = (setup.n_shots, setup.ny, setup.nx) if setup.batched else (setup.ny, setup.nx)
        gca = torch.zeros(gshape, device=dev, dtype=torch.float64) if want_ca else None
        gcb = torch.zeros(gshape, device=dev, dtype=torch.float64) if want_cb else None
        n_store = (setup.nt + interval - 1) // interval
        sdtype = torch.int16 if fmt == STORAGE_FORMAT_BF16 else dtype
        t0 = _timer.start() if _timer is not None else None
        rerun_ey = rerun_curl = None
        for i in range(plan.n_chunks):  # resident chunks come first
            s0, s1 = plan.bounds[i], plan.bounds[i + 1]
            n_c = s1 - s0
            ca_i, cb_i, cq_i = (_shots(t, s0, s1, setup.batched) for t in (ca, cb, cq))
            held = ctx.stores[i]
            ctx.stores[i] = None
            staged = held if isinstance(held, _HostStaged) else None
            ey_s = curl_s = None
            if staged is not None or held is not None or not (want_ca or want_cb):
                if staged is None and held is not None:
                    ey_s, curl_s = held
                sub = _chunk_setup(setup, s0, s1, [])
            else:
                # re-run this chunk's forward with storage, into one buffer
                # per coefficient sized for the largest chunk
                numel = n_store * max_c * setup.ny * setup.nx
                if rerun_ey is None and rerun_curl is None:
                    rerun_ey = torch.empty(numel, device=dev, dtype=sdtype) if want_ca else None
                    rerun_curl = torch.empty(numel, device=dev, dtype=sdtype) if want_cb else None
                init = ctx.initial[i]
                state = ([torch.zeros(n_c, setup.ny, setup.nx, device=dev, dtype=dtype)
                          for _ in range(7)] if init is None else init)
                sub = _chunk_setup(setup, s0, s1, state)
                f_d = ctx.f_data
                f_i = f_d[:, s0:s1].contiguous() if setup.n_src > 0 else f_d
                r_i = torch.empty(setup.nt, n_c, setup.n_rec, device=dev, dtype=dtype)
                view = (n_store, n_c, setup.ny, setup.nx)
                used = n_store * n_c * setup.ny * setup.nx
                ey_s = rerun_ey[:used].view(view) if rerun_ey is not None else None
                curl_s = rerun_curl[:used].view(view) if rerun_curl is not None else None
                _run_forward(sub, ca_i, cb_i, cq_i, f_i, sub.fields, r_i,
                             interval=interval, fmt=fmt, ey_store=ey_s, curl_store=curl_s)
                ctx.initial[i] = None
            lam = [t[:n_c] for t in lam_all]
            for t in lam:
                t.zero_()
            gr_i = grad_r[:, s0:s1].contiguous() if chunked else grad_r.contiguous()
            gf_i = (torch.zeros(setup.nt, n_c, setup.n_src, device=dev, dtype=dtype)
                    if (want_f and chunked) else grad_f)
            ga_i = _shots(gca, s0, s1, setup.batched) if gca is not None else None
            gb_i = _shots(gcb, s0, s1, setup.batched) if gcb is not None else None
            for start, steps in _time_chunks(setup.nt, every, reverse=True):
                _run_backward(sub, ca_i, cb_i, cq_i, gr_i, lam, ey_s, curl_s, gf_i, ga_i, gb_i,
                              interval=interval, fmt=fmt, start_t=start, nt=steps, staged=staged)
                if bwd_cb is not None:
                    names = ("lambda_Ey", "lambda_Hx", "lambda_Hz", "m_lambda_Ey_x",
                             "m_lambda_Ey_z", "m_lambda_Hx_z", "m_lambda_Hz_x")
                    bwd_cb(_callback_state(setup, start - 1, dict(zip(names, lam)),
                                           _model_gradients(setup, gca, gcb), True))
            if staged is not None:
                staged.close()
            if want_f and chunked:
                grad_f[:, s0:s1].copy_(gf_i)
            del ey_s, curl_s, held, staged
        if t0 is not None:
            _timer.stop("adjoint", setup.nt, setup.n_shots * setup.ny * setup.nx, t0,
                        _native.get_tm2d_time_block()[1])
        out_ca = gca.to(dtype).reshape(ca.shape) if want_ca else None
        out_cb = gcb.to(dtype).reshape(cb.shape) if want_cb else None
        return out_ca, out_cb, None, grad_f, None, None


def _reject_unsupported(
    python_backend, stencil, forward_callback, backward_callback, dispersion, device
) -> None:
    if python_backend not in (False, "native"):
        raise NotImplementedError(
            "this build has one backend (HIP on MI355X); python_backend is not available"
        )
    if stencil not in (2, 4, 6, 8):
        raise ValueError(f"stencil must be 2, 4, 6 or 8, got {stencil}")
    if device.type != "cuda":
        raise NotImplementedError(
            f"the MI355X build runs on HIP devices only; got tensors on {device}"
        )


def _normalize_batch(epsilon, sigma, mu, shot_tensors: dict[str, torch.Tensor | None]):
    """Structured batches: ``[B, ny, nx]`` models with ``[B, S, ...]`` shot
    tensors run as ``N = B*S`` shots with per-shot models (reference
    ``maxwell/common.py:534-603``)."""
    if epsilon.ndim == 2:
        return epsilon, sigma, mu, shot_tensors, None
    B = int(epsilon.shape[0])
    S = None
    for name, t in shot_tensors.items():
        if t is None or t.numel() == 0:
            continue
        if t.ndim == 4:
            if int(t.shape[0]) != B:
                raise RuntimeError(f"{name} batch dimension must be {B}")
            s = int(t.shape[1])
        elif t.ndim == 3 and int(t.shape[0]) == B:
            s = 1
        else:
            raise RuntimeError(f"{name} must be [B, S, ...] for a batched model")
        if S is not None and s != S:
            raise RuntimeError("All shot-batched tensors must agree on S")
        S = s
    S = S or 1
    out = {}
    for name, t in shot_tensors.items():
        if t is None or t.numel() == 0:
            out[name] = t
        elif t.ndim == 4:
            out[name] = t.reshape(B * S, *t.shape[2:])
        else:
            out[name] = t
    rep = lambda m: m.repeat_interleave(S, dim=0).contiguous()  # noqa: E731
    return rep(epsilon), rep(sigma), rep(mu), out, (B, S)


def maxwelltm(
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
    Ey_0: torch.Tensor | None = None,
    Hx_0: torch.Tensor | None = None,
    Hz_0: torch.Tensor | None = None,
    m_Ey_x: torch.Tensor | None = None,
    m_Ey_z: torch.Tensor | None = None,
    m_Hx_z: torch.Tensor | None = None,
    m_Hz_x: torch.Tensor | None = None,
    nt: int | None = None,
    model_gradient_sampling_interval: int = 1,
    freq_taper_frac: float = 0.0,
    time_pad_frac: float = 0.0,
    time_taper: bool = False,
    save_snapshots: bool | None = None,
    forward_callback: Any = None,
    backward_callback: Any = None,
    callback_frequency: int = 1,
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
):
    """2-D TM solver.  Returns ``(Ey, Hx, Hz, m_Ey_x, m_Ey_z, m_Hx_z,
    m_Hz_x, receiver_amplitudes)``; fields unpadded ``[shots, ny, nx]``,
    receivers ``[nt, shots, n_rec]`` (``[nt, B, S, n_rec]`` for a batch of
    several models)."""
    call = _prepare_call(
        epsilon, sigma, mu, grid_spacing, dt, source_amplitude, source_location,
        receiver_location, stencil=stencil, pml_width=pml_width, max_vel=max_vel,
        initial=(Ey_0, Hx_0, Hz_0, m_Ey_x, m_Ey_z, m_Hx_z, m_Hz_x), nt=nt,
        model_gradient_sampling_interval=model_gradient_sampling_interval,
        save_snapshots=save_snapshots, forward_callback=forward_callback,
        backward_callback=backward_callback, python_backend=python_backend,
        storage_mode=storage_mode, storage_compression=storage_compression,
        storage_bytes_limit_device=storage_bytes_limit_device,
        storage_bytes_limit_host=storage_bytes_limit_host, dispersion=dispersion,
        storage_path=storage_path, callback_frequency=callback_frequency,
    )
    setup, storage = call.setup, call.storage
    f = setup.f if setup.f.numel() > 0 else torch.empty(0, device=setup.ca.device, dtype=setup.ca.dtype)
    out = TM2DPropagate.apply(setup.ca, setup.cb, setup.cq, f, setup, storage)
    return call.finish(list(out[:7]), out[7])


@dataclass
class _Call:
    """A validated ``maxwelltm`` call: padded setup, storage policy, and
    the structured-batch shape to restore on the outputs."""

    setup: TM2DSetup
    storage: dict[str, Any]
    bs: tuple[int, int] | None

    def finish(self, padded_fields: list[torch.Tensor], r: torch.Tensor):
        setup = self.setup
        fields = [setup.unpad(t) for t in padded_fields]
        if setup.n_rec == 0:
            r = torch.empty(0, device=setup.ca.device, dtype=setup.ca.dtype)
        if self.bs is not None and self.bs[0] > 1:
            B, S = self.bs
            fields = [t.reshape(B, S, *t.shape[1:]) for t in fields]
            if r.numel() > 0:
                r = r.reshape(r.shape[0], B, S, *r.shape[2:])
        return (*fields, r)


def _prepare_call(
    epsilon, sigma, mu, grid_spacing, dt, source_amplitude, source_location, receiver_location,
    *, stencil, pml_width, max_vel, initial, nt, model_gradient_sampling_interval,
    save_snapshots, forward_callback, backward_callback, python_backend, storage_mode,
    storage_compression, storage_bytes_limit_device, storage_bytes_limit_host, dispersion,
    force_storage: bool = False, storage_path: str = ".", callback_frequency: int = 1,
) -> _Call:
    _reject_unsupported(
        python_backend, stencil, forward_callback, backward_callback, dispersion, epsilon.device
    )
    epsilon_in, sigma_in = epsilon, sigma
    epsilon, sigma, mu, shots, bs = _normalize_batch(
        epsilon, sigma, mu,
        {
            "source_amplitude": source_amplitude,
            "source_location": source_location,
            "receiver_location": receiver_location,
        },
    )
    source_amplitude = shots["source_amplitude"]
    source_location = shots["source_location"]
    receiver_location = shots["receiver_location"]

    spacing = _normalize_grid_spacing_2d(grid_spacing)
    if max_vel is None:
        max_vel = float((1.0 / torch.sqrt(epsilon.detach() * mu.detach())).max().item()) * C0
    _, step_ratio = cfl_condition(spacing, dt, max_vel)
    if step_ratio > 1:
        raise NotImplementedError(
            f"dt={dt} exceeds the CFL limit (needs {step_ratio} internal steps); "
            "internal resampling is not part of this build, pass a stable dt"
        )
    interval = max(1, int(model_gradient_sampling_interval))
    for name, t in (("Source", source_location), ("Receiver", receiver_location)):
        if t is not None and t.numel() > 0:
            hi = torch.tensor(epsilon.shape[-2:], device=t.device)
            if bool((t >= hi).any()):
                raise RuntimeError(f"{name} location out of bounds for model shape {tuple(epsilon.shape[-2:])}")

    model_rg = force_storage or epsilon_in.requires_grad or sigma_in.requires_grad
    if dispersion is not None:
        # the reference's native Debye path is forward inference only
        # (tm2d_cuda.py:360-370); its gradients run on its Python backend
        src_rg = source_amplitude is not None and source_amplitude.requires_grad
        if model_rg or src_rg or save_snapshots or forward_callback or backward_callback:
            raise NotImplementedError(
                "Debye dispersion runs forward inference only (as the reference's native "
                "backend); gradients, snapshots and callbacks with dispersion are not part of "
                "this build")
        if bs is not None and bs[0] > 1:
            raise NotImplementedError("Debye dispersion with a batch of models is not supported")
    setup = prepare_tm2d(
        epsilon, sigma, mu, spacing, dt, source_amplitude, source_location, receiver_location,
        stencil=stencil, pml_width=pml_width, nt=nt, initial=initial, max_vel=max_vel,
        dispersion=dispersion,
    )
    if setup.nt > 0:
        interval = min(interval, setup.nt)

    do_save = model_rg if save_snapshots is None else bool(save_snapshots)
    kind, _, bpe, fmt = _resolve_storage_compression(
        storage_compression, epsilon.dtype, epsilon.device, context="storage_compression"
    )
    n_store = (setup.nt + interval - 1) // interval
    per_shot = n_store * setup.ny * setup.nx * bpe * 2
    total_bytes = per_shot * setup.n_shots
    mode = resolve_storage_mode(
        storage_mode,
        needs_storage=do_save and model_rg,
        total_bytes=total_bytes,
        bytes_limit_device=storage_bytes_limit_device,
        bytes_limit_host=storage_bytes_limit_host,
    )
    plan = ShotPlan([0, setup.n_shots], 1)
    storing = do_save and model_rg
    # Device snapshots are split into shot chunks that fit the device budget;
    # host-staged modes (cpu, disk) keep one device ring per stream and run
    # all shots together.
    if storing and mode == "device" and setup.nt > 0:
        working = 21 * setup.ny * setup.nx * epsilon.dtype.itemsize
        budget = device_snapshot_budget(epsilon.device, storage_bytes_limit_device, working,
                                        setup.n_shots)
        plan = plan_shot_chunks(setup.n_shots, per_shot, budget,
                                group=max(1, _native.get_tm2d_tuning()[3]))
    if (forward_callback is not None or backward_callback is not None) and plan.n_chunks > 1:
        raise RuntimeError(
            f"callbacks see every shot at once, but the snapshots of {setup.n_shots} shots do "
            "not fit the device budget together; use storage_mode='cpu' or 'disk', BF16 "
            "compression or a larger storage_bytes_limit_device"
        )
    freq = int(callback_frequency)
    if (forward_callback is not None or backward_callback is not None) and freq < 1:
        raise ValueError("callback_frequency must be a positive number of steps")
    storage = {"mode": mode, "format": fmt, "interval": interval, "kind": kind, "plan": plan,
               "path": storage_path, "forward_callback": forward_callback,
               "backward_callback": backward_callback, "callback_frequency": freq}
    return _Call(setup, storage, bs)


def separable_gradient(
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
    storage_compression: bool | str = False,
    storage_bytes_limit_device: int | None = None,
    targets: Sequence[str] = ("epsilon", "sigma"),
):
    """Gradient of a receiver objective that is a sum over shots.

    ``residual(pred_chunk, observed_chunk) -> (value, d value / d pred)``
    for one chunk of shots.  Each chunk of ``storage`` runs its forward
    with device snapshots, its receiver residual and its exact adjoint
    before the next chunk starts, so no chunk is ever re-run; the
    coefficient gradients of all chunks sum into one FP64 accumulator, the
    objective value is summed in FP64, and the chain rule to epsilon/sigma
    (padding fold and source term included) is applied once at the end by
    torch autograd, as in the operator path.

    Returns ``(value, {"epsilon": g, "sigma": g}, receivers, final fields,
    {"shot_chunks": sizes})`` with the receivers and final fields of every
    shot.
    """
    eps = epsilon.detach().requires_grad_("epsilon" in targets)
    sig = sigma.detach().requires_grad_("sigma" in targets)
    call = _prepare_call(
        eps, sig, mu.detach(), grid_spacing, dt, source_amplitude, source_location,
        receiver_location, stencil=stencil, pml_width=pml_width, max_vel=max_vel,
        initial=(None,) * 7, nt=nt,
        model_gradient_sampling_interval=model_gradient_sampling_interval,
        save_snapshots=True, forward_callback=None, backward_callback=None,
        python_backend=False, storage_mode="device", storage_compression=storage_compression,
        storage_bytes_limit_device=storage_bytes_limit_device, storage_bytes_limit_host=None,
        dispersion=None, force_storage=True,
    )
    if call.bs is not None and call.bs[0] > 1:
        raise NotImplementedError("separable_gradient takes one model shared by all shots")
    setup, storage = call.setup, call.storage
    plan: ShotPlan = storage["plan"]
    dev, dtype = setup.ca.device, setup.ca.dtype
    interval, fmt = storage["interval"], storage["format"]
    if tuple(observed.shape) != (setup.nt, setup.n_shots, setup.n_rec):
        raise ValueError("observed data must be [nt, shots, receivers]")
    ca, cb, cq = (t.detach() for t in (setup.ca, setup.cb, setup.cq))
    f = setup.f.detach().contiguous() if setup.f.numel() > 0 else torch.empty(0, device=dev, dtype=dtype)
    fields = [t.contiguous() for t in setup.fields]
    r = torch.zeros(setup.nt, setup.n_shots, setup.n_rec, device=dev, dtype=dtype)
    n_store = (setup.nt + interval - 1) // interval
    sdtype = torch.int16 if fmt == STORAGE_FORMAT_BF16 else dtype
    max_c = max(plan.sizes)
    numel = n_store * max_c * setup.ny * setup.nx
    ey_buf = torch.empty(numel, device=dev, dtype=sdtype)
    curl_buf = torch.empty(numel, device=dev, dtype=sdtype)
    lam_all = [torch.empty(max_c, setup.ny, setup.nx, device=dev, dtype=dtype) for _ in range(7)]
    gca = torch.zeros(setup.ny, setup.nx, device=dev, dtype=torch.float64)
    gcb = torch.zeros(setup.ny, setup.nx, device=dev, dtype=torch.float64)
    grad_f = torch.zeros(setup.nt, setup.n_shots, setup.n_src, device=dev, dtype=dtype)
    value = torch.zeros((), device=dev, dtype=torch.float64)
    chunked = plan.n_chunks > 1
    blocks = _native.get_tm2d_time_block()
    for i in range(plan.n_chunks):
        s0, s1 = plan.bounds[i], plan.bounds[i + 1]
        n_c = s1 - s0
        sub = _chunk_setup(setup, s0, s1, [t[s0:s1] for t in fields] if chunked else fields)
        view = (n_store, n_c, setup.ny, setup.nx)
        used = n_store * n_c * setup.ny * setup.nx
        ey_s, curl_s = ey_buf[:used].view(view), curl_buf[:used].view(view)
        f_i = f[:, s0:s1].contiguous() if (chunked and setup.n_src > 0) else f
        r_i = torch.zeros(setup.nt, n_c, setup.n_rec, device=dev, dtype=dtype) if chunked else r
        tf = _timer.start() if _timer is not None else None
        _run_forward(sub, ca, cb, cq, f_i, sub.fields, r_i, interval=interval, fmt=fmt,
                     ey_store=ey_s, curl_store=curl_s)
        if tf is not None:
            _timer.stop("forward_store", setup.nt, n_c * setup.ny * setup.nx, tf, blocks[0])
        if chunked:
            r[:, s0:s1].copy_(r_i)
        v_i, g_i = residual(r_i, observed[:, s0:s1])
        value = value + v_i.to(torch.float64)
        lam = [t[:n_c] for t in lam_all]
        for t in lam:
            t.zero_()
        gf_i = grad_f[:, s0:s1].contiguous() if chunked else grad_f
        ta = _timer.start() if _timer is not None else None
        _run_backward(sub, ca, cb, cq, g_i.to(dtype).contiguous(), lam, ey_s, curl_s,
                      gf_i if setup.n_src > 0 else None, gca, gcb, interval=interval, fmt=fmt)
        if ta is not None:
            _timer.stop("adjoint", setup.nt, n_c * setup.ny * setup.nx, ta, blocks[1])
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
    out = call.finish(fields, r)
    info = {"shot_chunks": plan.sizes}
    return value, dict(zip(names, pg)), out[7], list(out[:7]), info


__all__ = ["KernelTimer", "ShotPlan", "TM2DPropagate", "TM2DSetup", "maxwelltm",
           "plan_shot_chunks", "prepare_tm2d", "separable_gradient", "set_kernel_timer"]
