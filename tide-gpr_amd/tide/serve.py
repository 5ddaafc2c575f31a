"""Served 2-D TM solves: continuous batching on one GPU.

A :class:`TM2DServer` accepts forward and forward+gradient solves of any
size as they arrive and runs them on the native serving engine
(``tide_tm2d_engine_*``): every admitted job stays resident in HBM (fields,
CPML memories, snapshots, adjoint state) and each engine tick advances all
of them with one launch per time step, so solves that arrive while others
run are admitted as soon as a job slot and the HBM budget allow, and their
work overlaps on the device with whatever is already running.

Admission: a request is split into shot chunks that each fit in half the
HBM budget (the gradient of a request is the sum of its chunks' gradients,
as shots are independent); chunks wait in arrival order and are admitted
first-fit, except that a chunk bypassed ``max_bypass`` times blocks the
ones behind it (no starvation).

Streams: the engine runs on a stream of its own; host-side preparation of a
newly admitted chunk (padding, coefficients, CPML profiles, point tables,
zeroed buffers) runs on a second stream, so the host never waits for queued
engine work.  Post-processing (misfit, adjoint source, FP64 model-gradient
chain rule) is enqueued on the engine stream without host synchronisation.

Results equal the single-problem path's: the tile kernels and their shot
grouping are the same, so receivers are bitwise those of ``maxwelltm`` and
the FP64 coefficient gradients are bitwise independent of the mix; the
epsilon / sigma chain rule is the closed form of the coefficient formulas
(reference utils.py:13-41), the replicate-pad fold (reference
padding.py:93) and the source-term contribution (reference
tm2d_helpers.py:79-81), built without atomics so it too is reproducible.
"""

from __future__ import annotations

import ctypes
import math
import time
from collections import deque
from collections.abc import Sequence
from dataclasses import dataclass, field
from typing import Any

import torch

from . import _native
from .batch import _Problem
from .cfl import cfl_condition
from .grid_utils import _normalize_grid_spacing_2d, _normalize_pml_width_2d
from .padding import fold_replicate
from .maxwell.tm2d import prepare_tm2d
from .utils import C0, coefficient_partials

_P = ctypes.c_void_p
_I64 = ctypes.c_int64


class _Job(ctypes.Structure):
    """Mirror of ``tide_tm2d_job`` (include/tide_hip.h)."""

    _fields_ = [
        ("p", _Problem),
        ("fields_b", _P * 7),
        ("lambda_b", _P * 7),
        ("grad_ca_groups", _P),
        ("grad_cb_groups", _P),
        ("storage_format", _I64),
    ]


def _bind(lib: ctypes.CDLL) -> None:
    if getattr(lib, "_tide_engine_bound", False):
        return
    lib.tide_tm2d_engine_create.restype = ctypes.c_int
    lib.tide_tm2d_engine_create.argtypes = [_I64, _I64, _I64, ctypes.POINTER(_P)]
    lib.tide_tm2d_engine_set_boost.restype = ctypes.c_int
    lib.tide_tm2d_engine_set_boost.argtypes = [_P, _I64]
    lib.tide_tm2d_engine_set_weight.restype = ctypes.c_int
    lib.tide_tm2d_engine_set_weight.argtypes = [_P, _I64, _I64]
    lib.tide_tm2d_engine_group.restype = ctypes.c_int
    lib.tide_tm2d_engine_group.argtypes = [_P, ctypes.POINTER(_I64)]
    lib.tide_tm2d_engine_add.restype = ctypes.c_int
    lib.tide_tm2d_engine_add.argtypes = [_P, ctypes.POINTER(_Job), ctypes.POINTER(_I64)]
    lib.tide_tm2d_engine_tick.restype = ctypes.c_int
    lib.tide_tm2d_engine_tick.argtypes = [_P, _I64, _P, ctypes.POINTER(_I64)]
    lib.tide_tm2d_engine_finished.restype = ctypes.c_int
    lib.tide_tm2d_engine_finished.argtypes = [_P, ctypes.POINTER(_I64), _I64, ctypes.POINTER(_I64)]
    lib.tide_tm2d_engine_query.restype = ctypes.c_int
    lib.tide_tm2d_engine_query.argtypes = [_P, _I64, ctypes.POINTER(_I64), ctypes.POINTER(_I64)]
    lib.tide_tm2d_engine_begin_adjoint.restype = ctypes.c_int
    lib.tide_tm2d_engine_begin_adjoint.argtypes = [_P, _I64]
    lib.tide_tm2d_engine_remove.restype = ctypes.c_int
    lib.tide_tm2d_engine_remove.argtypes = [_P, _I64, _P]
    lib.tide_tm2d_engine_destroy.restype = ctypes.c_int
    lib.tide_tm2d_engine_destroy.argtypes = [_P]
    lib._tide_engine_bound = True


# ------------------------------------------------------------------ sizing
def padded_shape(ny: int, nx: int, pml_width: int | Sequence[int] = 20) -> tuple[int, int]:
    """Padded grid of a ``[ny, nx]`` model (stencil 2: one halo row/column
    on the low side)."""
    w = _normalize_pml_width_2d(pml_width)
    return ny + w[0] + w[1] + 1, nx + w[2] + w[3] + 1


def job_bytes(
    ny_p: int, nx_p: int, n_shots: int, nt: int, n_src: int, n_rec: int, *,
    gradient: bool, bf16: bool = False, interval: int = 1, group: int = 4,
) -> int:
    """Device bytes one engine job holds (fp32 fields, FP64 gradients)."""
    a = ny_p * nx_p
    b = 3 * a * 4 + 14 * n_shots * a * 4 + nt * n_shots * (n_src + n_rec) * 4
    b += 12 * (ny_p + nx_p) * 4
    if gradient:
        n_store = (nt + interval - 1) // interval if nt > 0 else 0
        b += 14 * n_shots * a * 4
        b += 2 * n_store * n_shots * a * (2 if bf16 else 4)
        b += nt * n_shots * (n_rec + n_src) * 4 + nt * n_shots * n_rec * 8
        b += 2 * a * 8
        groups = (n_shots + group - 1) // group
        if groups > 1:
            b += 2 * groups * a * 8
    return int(b)


def plan_chunks(n_shots: int, shot_bytes: Any, budget: int) -> list[tuple[int, int]]:
    """Split ``n_shots`` into consecutive chunks whose ``shot_bytes(count)``
    stays within half the budget (a single shot may use the whole budget).
    Raises ValueError when one shot does not fit."""
    if shot_bytes(1) > budget:
        raise ValueError(f"one shot needs {shot_bytes(1)} bytes, more than the HBM budget {budget}")
    limit = budget // 2
    per = 1
    while per < n_shots and shot_bytes(per + 1) <= limit:
        per += 1
    return [(s, min(n_shots, s + per)) for s in range(0, n_shots, per)]


_ALIGN = 256


def _aligned(n: int) -> int:
    return (int(n) + _ALIGN - 1) // _ALIGN * _ALIGN


def state_spec(
    ny_p: int, nx_p: int, n_shots: int, nt: int, n_src: int, n_rec: int, *,
    gradient: bool, bf16: bool = False, interval: int = 1, group: int = 2,
) -> list[tuple[str, torch.dtype, tuple[int, ...], int, bool]]:
    """The device arrays of one engine job carved from the server's pool:
    ``(name, dtype, shape, count, zeroed)``."""
    f32, f64 = torch.float32, torch.float64
    shape = (n_shots, ny_p, nx_p)
    spec = [
        ("fields", f32, shape, 7, True),
        ("fields_b", f32, shape, 7, True),
        ("r", f32, (nt, n_shots, max(n_rec, 1)), 1, False),
    ]
    if gradient:
        n_store = (nt + interval - 1) // interval if nt > 0 else 0
        sdt = torch.int16 if bf16 else f32
        spec += [
            ("ey_store", sdt, (n_store, *shape), 1, False),
            ("curl_store", sdt, (n_store, *shape), 1, False),
            ("lambda", f32, shape, 7, True),
            ("lambda_b", f32, shape, 7, True),
            ("grad_r", f32, (nt, n_shots, max(n_rec, 1)), 1, False),
            ("grad_f", f32, (nt, n_shots, max(n_src, 1)), 1, False),
            ("grad_ca", f64, (ny_p, nx_p), 1, True),
            ("grad_cb", f64, (ny_p, nx_p), 1, True),
        ]
        groups = (n_shots + group - 1) // group
        if groups > 1:
            spec += [("gca_groups", f64, (groups, ny_p, nx_p), 1, True),
                     ("gcb_groups", f64, (groups, ny_p, nx_p), 1, True)]
    return spec


def spec_bytes(spec) -> int:
    total = 0
    for _, dt, shape, count, _ in spec:
        n = dt.itemsize
        for d in shape:
            n *= d
        total += count * _aligned(n)
    return total


class DevicePool:
    """One device allocation made up front and carved into job state
    (first fit, 256-byte aligned, neighbours merged on release): admitting
    a solve costs no hipMalloc, whatever its size."""

    def __init__(self, nbytes: int, device: torch.device) -> None:
        self.size = _aligned(nbytes) - _ALIGN
        self.buf = torch.empty(self.size, dtype=torch.uint8, device=device)
        self.free: list[tuple[int, int]] = [(0, self.size)]

    @property
    def available(self) -> int:
        return sum(n for _, n in self.free)

    def alloc(self, nbytes: int) -> int | None:
        n = _aligned(nbytes)
        for i, (off, size) in enumerate(self.free):
            if size >= n:
                if size == n:
                    del self.free[i]
                else:
                    self.free[i] = (off + n, size - n)
                return off
        return None

    def release(self, off: int, nbytes: int) -> None:
        n = _aligned(nbytes)
        self.free.append((off, n))
        self.free.sort()
        merged: list[tuple[int, int]] = []
        for o, m in self.free:
            if merged and merged[-1][0] + merged[-1][1] == o:
                merged[-1] = (merged[-1][0], merged[-1][1] + m)
            else:
                merged.append((o, m))
        self.free = merged

    def carve(self, off: int, spec) -> dict[str, Any]:
        """Tensors for ``spec`` laid out from byte ``off`` (lists for counts > 1)."""
        out: dict[str, Any] = {}
        pos = off
        for name, dt, shape, count, _ in spec:
            n = dt.itemsize
            for d in shape:
                n *= d
            ts = []
            for _ in range(count):
                ts.append(self.buf[pos : pos + n].view(dt).view(shape))
                pos += _aligned(n)
            out[name] = ts if count > 1 else ts[0]
        return out


# ------------------------------------------------------------ chain rule
def model_gradient(
    gca: torch.Tensor, gcb: torch.Tensor, gf: torch.Tensor | None, st: Any,
    epsilon_shape: tuple[int, int], amplitude: torch.Tensor | None, grid_spacing: Any,
    source_index: torch.Tensor | None = None,
) -> tuple[torch.Tensor, torch.Tensor]:
    """dJ/d(epsilon_r), dJ/dsigma on the unpadded model from dJ/dca, dJ/dcb
    (FP64, padded grid) and dJ/df (source term), in FP64.  ``st`` is the
    chunk's setup (padded ca, cb, source indices); ``amplitude`` the chunk's
    ``[shots, n_src, nt]`` sources and ``source_index`` the flat source
    indices, both on the gradients' device (then nothing here waits for the
    device)."""
    dy, dx = _normalize_grid_spacing_2d(grid_spacing)
    ca = st.ca.reshape(st.ny, st.nx).double()
    cb = st.cb.reshape(st.ny, st.nx).double()
    g_ca = gca.reshape(st.ny, st.nx).double()
    g_cb = gcb.reshape(st.ny, st.nx).double().clone()
    if gf is not None and st.n_src > 0 and amplitude is not None and amplitude.numel() > 0:
        # f[t, s, k] = amp[s, k, t] * cb[src(s, k)] * (-1 / (dx dy))
        amp = amplitude.detach().to(device=gf.device, dtype=torch.float64).permute(2, 0, 1)
        w = (gf.double() * amp).sum(0) * (-1.0 / (dx * dy))  # [shots, n_src]
        idx = (st.sources_i if source_index is None else source_index).to(gf.device).reshape(-1)
        keep = idx >= 0  # skipped sources add zero at index 0 (no data-dependent shapes)
        idx = torch.where(keep, idx, 0)
        w = torch.where(keep, w.reshape(-1), 0.0)
        # Sources that share a cell: every entry of a cell gets the cell's
        # total (a fixed-order product with the same-cell matrix) and writes
        # the same new value, so the scatter needs no atomics and gives the
        # same bits on every run.
        same = (idx[:, None] == idx[None, :]).double()
        total = same @ w
        flat = g_cb.view(-1)
        flat.index_put_((idx,), flat[idx] + total, accumulate=False)
    dca_de, dcb_de, dca_ds, dcb_ds = coefficient_partials(ca, cb, st.dt)
    g_eps_p = g_ca * dca_de + g_cb * dcb_de
    g_sig_p = g_ca * dca_ds + g_cb * dcb_ds
    total = [a + b for a, b in zip(st.fd_pad, st.pml_width)]
    return (_fold_replicate(g_eps_p, total, epsilon_shape),
            _fold_replicate(g_sig_p, total, epsilon_shape))


def _fold_replicate(g_p: torch.Tensor, pad: Sequence[int], shape: tuple[int, int]) -> torch.Tensor:
    """Adjoint of replicate padding by ``pad = [top, bottom, left, right]``
    (deterministic, see ``tide.padding.fold_replicate``)."""
    out = fold_replicate(g_p, [int(v) for v in pad], 2)
    assert tuple(out.shape[-2:]) == tuple(shape)
    return out


# ------------------------------------------------------------- requests
@dataclass
class SolveResult:
    receiver_data: torch.Tensor  # [nt, shots, n_rec]
    misfit: torch.Tensor | None = None  # FP64 scalar, 0.5 |d - d_obs|^2
    grad_epsilon: torch.Tensor | None = None  # FP64 [ny, nx]
    grad_sigma: torch.Tensor | None = None


@dataclass
class Ticket:
    """Handle of one submitted solve; ``done`` once its results are final."""

    id: int
    kind: str
    arrival: float
    n_chunks: int = 1
    admitted: float | None = None
    completed: float | None = None
    result: SolveResult | None = None
    _parts: list = field(default_factory=list)
    _left: int = 0
    _event: Any = None

    @property
    def done(self) -> bool:
        return self.completed is not None

    @property
    def latency(self) -> float | None:
        return None if self.completed is None else self.completed - self.arrival


@dataclass
class _Chunk:
    ticket: Ticket
    index: int
    shots: tuple[int, int]
    problem: dict
    observed: torch.Tensor | None
    bytes: int
    bf16: bool
    interval: int
    inputs_ready: Any = None
    spec: Any = None
    block: int = -1
    bypassed: int = 0
    job_id: int = -1
    st: Any = None
    bufs: dict = field(default_factory=dict)
    job: Any = None


def _slice_shots(t: torch.Tensor | None, s0: int, s1: int) -> torch.Tensor | None:
    if t is None or t.numel() == 0:
        return t
    return t[s0:s1]


class TM2DServer:
    """Continuous-batching server for 2-D TM solves on one HIP device.

    Parameters
    ----------
    device:
        HIP device (default: the current one).
    hbm_budget:
        Size of the device pool the admitted jobs' state is carved from,
        allocated once at construction (default 70% of the free memory).
    max_jobs:
        Jobs resident at once.
    steps_per_tick:
        Time steps the largest running job advances per engine tick
        (admission and completion happen at tick boundaries); a smaller job
        advances proportionally more (by the ratio of cells x shots per step,
        up to ``tick_boost`` times), so it is not held to the large job's pace.
    tick_boost:
        That cap (default 4; 1 advances every job by the same steps).
    head_weight:
        Priority of the oldest request with running jobs: its jobs take
        this many times their steps per tick (default 1: the device is
        shared by size alone).  On the 4/s trace of the bench, 4 left the
        p99 latency unchanged and raised the median.
    group:
        Shots per workgroup, the same for every job of a launch (default 2:
        requests with few shots leave fewer idle wavefronts).
    max_bypass:
        Admissions that may overtake a waiting chunk before it blocks the
        queue.
    """

    def __init__(
        self,
        device: torch.device | str | None = None,
        *,
        hbm_budget: int | None = None,
        max_jobs: int = 64,
        steps_per_tick: int = 64,
        group: int | None = None,
        max_bypass: int = 8,
        tick_boost: int = 4,
        head_weight: int = 1,
    ) -> None:
        dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        if dev.type != "cuda":
            raise NotImplementedError("TM2DServer runs on HIP devices only")
        if dev.index is None:
            dev = torch.device("cuda", torch.cuda.current_device())
        self.device = dev
        self._lib = _native.load()
        _bind(self._lib)
        self.group = max(1, min(8, int(2 if group is None else group)))
        self.max_jobs = int(max_jobs)
        self.steps_per_tick = max(1, int(steps_per_tick))
        self.max_bypass = int(max_bypass)
        self.head_weight = int(head_weight)
        if not 1 <= self.head_weight <= 64:
            raise ValueError("head_weight must be in 1..64")
        self._weighted: set[int] = set()
        handle = _P()
        _native.check(self._lib.tide_tm2d_engine_create(dev.index, self.group, self.max_jobs, ctypes.byref(handle)))
        self._engine = handle
        _native.check(self._lib.tide_tm2d_engine_set_boost(self._engine, int(tick_boost)))
        self.stream = torch.cuda.Stream(dev)
        self._prep = torch.cuda.Stream(dev)
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(dev)
            hbm_budget = int(0.7 * free)
        self.budget = int(hbm_budget)
        self.pool = DevicePool(self.budget, dev)
        self._released = None  # engine-stream event after the last pool release
        self._waiting: deque[_Chunk] = deque()
        self._jobs: dict[int, _Chunk] = {}
        self._open: list[Ticket] = []
        self._inflight: deque = deque()
        self._next_id = 0
        self.ticks = 0
        # GPU-clock origin for completion times
        self._ref = torch.cuda.Event(enable_timing=True)
        with torch.cuda.stream(self.stream):
            self._ref.record(self.stream)
        self._ref.synchronize()
        self._t_ref = time.perf_counter()

    # --------------------------------------------------------------- submit
    def submit(
        self,
        problem: dict[str, Any],
        *,
        kind: str = "forward",
        observed: torch.Tensor | None = None,
        storage_compression: bool | str = False,
        model_gradient_sampling_interval: int = 1,
        arrival: float | None = None,
    ) -> Ticket:
        """Queue one solve.  ``problem`` holds ``maxwelltm`` arguments
        (``epsilon, sigma, mu`` fp32 ``[ny, nx]``, ``grid_spacing, dt,
        source_amplitude [shots, n_src, nt], source_location,
        receiver_location``, optional ``pml_width``).  ``kind`` is
        ``"forward"`` or ``"gradient"`` (misfit ``0.5 |d - observed|^2`` and
        its epsilon / sigma gradients; ``observed`` is ``[nt, shots, n_rec]``).
        ``arrival`` is the request's arrival time (``time.perf_counter()``
        clock; default now)."""
        if kind not in ("forward", "gradient"):
            raise ValueError("kind must be 'forward' or 'gradient'")
        eps = problem["epsilon"]
        if eps.ndim != 2 or eps.dtype != torch.float32:
            raise TypeError("served problems need an fp32 [ny, nx] model")
        if kind == "gradient" and observed is None:
            raise ValueError("a gradient solve needs observed data")
        amp = problem.get("source_amplitude")
        src = problem.get("source_location")
        rec = problem.get("receiver_location")
        nt = int(problem["nt"]) if problem.get("nt") is not None else int(amp.shape[-1])
        shots = next((int(t.shape[0]) for t in (amp, src, rec) if t is not None and t.numel() > 0), 1)
        n_src = int(src.shape[1]) if src is not None and src.numel() > 0 else 0
        n_rec = int(rec.shape[1]) if rec is not None and rec.numel() > 0 else 0
        ny_p, nx_p = padded_shape(int(eps.shape[0]), int(eps.shape[1]), problem.get("pml_width", 20))
        bf16 = storage_compression in (True, "bf16", "bfloat16")
        interval = max(1, int(model_gradient_sampling_interval))
        if nt > 0:
            interval = min(interval, nt)
        grad = kind == "gradient"

        def spec_of(count: int):
            return state_spec(ny_p, nx_p, count, nt, n_src, n_rec, gradient=grad, bf16=bf16,
                              interval=interval, group=self.group)

        def shot_bytes(count: int) -> int:
            return spec_bytes(spec_of(count))

        chunks = plan_chunks(shots, shot_bytes, self.pool.size)
        # inputs produced on the caller's stream are read by the preparation stream
        inputs_ready = torch.cuda.Event()
        inputs_ready.record(torch.cuda.current_stream(self.device))
        ticket = Ticket(id=self._next_id, kind=kind, n_chunks=len(chunks),
                        arrival=time.perf_counter() if arrival is None else float(arrival))
        self._next_id += 1
        ticket._left = len(chunks)
        ticket._parts = [None] * len(chunks)
        for i, (s0, s1) in enumerate(chunks):
            sub = dict(problem)
            sub["source_amplitude"] = _slice_shots(amp, s0, s1)
            sub["source_location"] = _slice_shots(src, s0, s1)
            sub["receiver_location"] = _slice_shots(rec, s0, s1)
            obs = observed[:, s0:s1] if observed is not None else None
            spec = spec_of(s1 - s0)
            self._waiting.append(_Chunk(ticket, i, (s0, s1), sub, obs, spec_bytes(spec), bf16, interval,
                                        inputs_ready, spec))
        self._open.append(ticket)
        return ticket

    # ------------------------------------------------------------ admission
    def _admit(self) -> None:
        blocked: _Chunk | None = None
        for ch in list(self._waiting):
            if len(self._jobs) >= self.max_jobs:
                break
            if blocked is not None and blocked.bypassed >= self.max_bypass:
                break
            off = self.pool.alloc(ch.bytes)
            if off is None:
                if blocked is None:
                    blocked = ch
                continue
            if blocked is not None:
                blocked.bypassed += 1
            ch.block = off
            self._waiting.remove(ch)
            self._start(ch)

    def _start(self, ch: _Chunk) -> None:
        """Prepare a chunk on the preparation stream and hand it to the engine."""
        kw = ch.problem
        dev = self.device
        f32 = torch.float32
        with torch.cuda.stream(self._prep):
            self._prep.wait_event(ch.inputs_ready)
            eps = kw["epsilon"].detach().to(dev)
            sig = kw["sigma"].detach().to(dev)
            mu = kw.get("mu")
            mu = torch.ones_like(eps) if mu is None else mu.detach().to(dev)
            spacing = _normalize_grid_spacing_2d(kw["grid_spacing"])
            max_vel = float((1.0 / torch.sqrt(eps * mu)).max().item()) * C0
            _, step_ratio = cfl_condition(spacing, kw["dt"], max_vel)
            if step_ratio > 1:
                raise NotImplementedError(f"dt={kw['dt']} exceeds the CFL limit; pass a stable dt")
            amp = kw.get("source_amplitude")
            amp = amp.detach().to(dev) if amp is not None else None
            st = prepare_tm2d(eps, sig, mu, spacing, kw["dt"], amp, kw.get("source_location"),
                              kw.get("receiver_location"), pml_width=kw.get("pml_width", 20),
                              nt=kw.get("nt"), validate_material_inputs=False, allocate_fields=False)
            # the pool block may still be in use by work queued before its release
            if self._released is not None:
                self._prep.wait_event(self._released)
            b: dict[str, Any] = self.pool.carve(ch.block, ch.spec)
            for name, _, _, _, zeroed in ch.spec:
                if zeroed:
                    for t in b[name] if isinstance(b[name], list) else [b[name]]:
                        t.zero_()
            b["ca"], b["cb"], b["cq"] = (t.detach().contiguous() for t in (st.ca, st.cb, st.cq))
            b["f"] = st.f.detach().contiguous() if st.n_src > 0 else None
            b["src"] = st.sources_i.contiguous()
            b["src_dev"] = st.sources_i.to(dev)
            b["amp"] = amp
            b["rec"] = st.receivers_i.contiguous()
            b["prof"] = list(st.profiles)
            if ch.ticket.kind == "gradient":
                obs = ch.observed
                b["observed"] = obs.detach().to(device=dev, dtype=f32) if obs is not None else None
            ready = torch.cuda.Event()
            ready.record(self._prep)
        self.stream.wait_event(ready)
        for t in _tensors(b):
            t.record_stream(self.stream)
        ch.st, ch.bufs = st, b
        ch.job = self._job_struct(ch)
        jid = _I64()
        _native.check(self._lib.tide_tm2d_engine_add(self._engine, ctypes.byref(ch.job), ctypes.byref(jid)))
        ch.job_id = int(jid.value)
        self._jobs[ch.job_id] = ch
        if ch.ticket.admitted is None:
            ch.ticket.admitted = time.perf_counter()

    def _job_struct(self, ch: _Chunk) -> _Job:
        st, b = ch.st, ch.bufs
        ptr = _native.ptr
        j = _Job()
        p = j.p
        p.ca, p.cb, p.cq = ptr(b["ca"]), ptr(b["cb"]), ptr(b["cq"])
        p.f = ptr(b["f"]) if st.n_src > 0 else None
        p.sources_i = b["src"].data_ptr() if st.n_src > 0 else None
        p.receivers_i = b["rec"].data_ptr() if st.n_rec > 0 else None
        for k, t in enumerate(b["prof"]):
            p.prof[k] = ptr(t)
        p.rdy, p.rdx = st.rdy, st.rdx
        p.nt, p.n_shots, p.ny, p.nx = st.nt, st.n_shots, st.ny, st.nx
        p.n_src, p.n_rec, p.interval = st.n_src, st.n_rec, ch.interval
        for k, v in enumerate(st.pml):
            p.pml[k] = v
        for k, t in enumerate(b["fields"]):
            p.fields[k] = ptr(t)
        for k, t in enumerate(b["fields_b"]):
            j.fields_b[k] = ptr(t)
        p.r = ptr(b["r"]) if st.n_rec > 0 else None
        if ch.ticket.kind == "gradient":
            p.ey_store, p.curl_store = ptr(b["ey_store"]), ptr(b["curl_store"])
            p.grad_r = ptr(b["grad_r"]) if st.n_rec > 0 else None
            for k in range(7):
                p.lambda_[k] = ptr(b["lambda"][k])
                j.lambda_b[k] = ptr(b["lambda_b"][k])
            p.grad_ca, p.grad_cb = ptr(b["grad_ca"]), ptr(b["grad_cb"])
            p.grad_f = ptr(b["grad_f"]) if st.n_src > 0 else None
            if "gca_groups" in b:
                j.grad_ca_groups = ptr(b["gca_groups"])
                j.grad_cb_groups = ptr(b["gcb_groups"])
        j.storage_format = 1 if ch.bf16 else 0
        return j

    # ---------------------------------------------------------------- ticks
    def step(self) -> list[Ticket]:
        """One scheduling round: admit, advance every running job by one
        tick, post-process finished phases, and return the tickets that
        completed since the last call."""
        self._admit()
        if self._jobs:
            self._prioritize()
            while len(self._inflight) >= 2:
                self._inflight.popleft().synchronize()
            n_run = _I64()
            _native.check(self._lib.tide_tm2d_engine_tick(
                self._engine, self.steps_per_tick, self.stream.cuda_stream, ctypes.byref(n_run)))
            self.ticks += 1
            self._after_tick()
            ev = torch.cuda.Event()
            ev.record(self.stream)
            self._inflight.append(ev)
        return self._collect()

    def _prioritize(self) -> None:
        """Weight the running jobs of the oldest request with running jobs
        by ``head_weight`` and every other job by 1."""
        if self.head_weight <= 1:
            return
        head = min(self._jobs.values(), key=lambda ch: (ch.ticket.arrival, ch.ticket.id)).ticket
        want = {jid for jid, ch in self._jobs.items() if ch.ticket is head}
        for jid in (self._weighted - want) & self._jobs.keys():
            _native.check(self._lib.tide_tm2d_engine_set_weight(self._engine, jid, 1))
        for jid in want - self._weighted:
            _native.check(self._lib.tide_tm2d_engine_set_weight(self._engine, jid, self.head_weight))
        self._weighted = want

    def _after_tick(self) -> None:
        cap = max(1, len(self._jobs))
        ids = (_I64 * cap)()
        n = _I64()
        _native.check(self._lib.tide_tm2d_engine_finished(self._engine, ids, cap, ctypes.byref(n)))
        phase, k = _I64(), _I64()
        with torch.cuda.stream(self.stream):
            for jid in ids[: n.value]:
                ch = self._jobs[jid]
                _native.check(self._lib.tide_tm2d_engine_query(self._engine, jid, ctypes.byref(phase), ctypes.byref(k)))
                if phase.value == 1 and ch.ticket.kind == "gradient":
                    self._begin_adjoint(ch)
                else:
                    self._finish(ch)

    def _begin_adjoint(self, ch: _Chunk) -> None:
        st, b = ch.st, ch.bufs
        if st.n_rec > 0:
            resid = b["r"][:, :, : st.n_rec].double() - b["observed"].double()
            b["misfit"] = 0.5 * (resid * resid).sum()
            b["grad_r"][:, :, : st.n_rec].copy_(resid)
        else:
            b["misfit"] = torch.zeros((), device=self.device, dtype=torch.float64)
        _native.check(self._lib.tide_tm2d_engine_begin_adjoint(self._engine, ch.job_id))

    def _finish(self, ch: _Chunk) -> None:
        st, b = ch.st, ch.bufs
        _native.check(self._lib.tide_tm2d_engine_remove(self._engine, ch.job_id, self.stream.cuda_stream))
        del self._jobs[ch.job_id]
        part: dict[str, Any] = {"r": b["r"][:, :, : st.n_rec].clone()}
        if ch.ticket.kind == "gradient":
            gf = b["grad_f"][:, :, : st.n_src] if st.n_src > 0 else None
            ge, gs = model_gradient(b["grad_ca"], b["grad_cb"], gf, st, tuple(ch.problem["epsilon"].shape),
                                    b["amp"], ch.problem["grid_spacing"], b["src_dev"])
            part.update(misfit=b["misfit"], ge=ge, gs=gs)
        # the block returns to the pool once the work queued so far is done
        self.pool.release(ch.block, ch.bytes)
        ch.block = -1
        self._released = torch.cuda.Event()
        self._released.record(self.stream)
        t = ch.ticket
        t._parts[ch.index] = part
        t._left -= 1
        ch.bufs = {}
        ch.st = None
        ch.job = None
        if t._left == 0:
            parts = t._parts
            r_all = parts[0]["r"] if len(parts) == 1 else torch.cat([p["r"] for p in parts], dim=1)
            res = SolveResult(receiver_data=r_all)
            if t.kind == "gradient":
                res.misfit = _sum([p["misfit"] for p in parts])
                res.grad_epsilon = _sum([p["ge"] for p in parts])
                res.grad_sigma = _sum([p["gs"] for p in parts])
            t.result = res
            t._parts = []
            ev = torch.cuda.Event(enable_timing=True)
            ev.record(self.stream)
            t._event = ev

    def _collect(self) -> list[Ticket]:
        done = []
        for t in list(self._open):
            if t._event is not None and t._event.query():
                t.completed = self._t_ref + self._ref.elapsed_time(t._event) / 1e3
                t._event = None
                self._open.remove(t)
                done.append(t)
        return done

    # ------------------------------------------------------------------ run
    @property
    def used(self) -> int:
        """Pool bytes held by admitted jobs."""
        return self.pool.size - self.pool.available

    @property
    def idle(self) -> bool:
        return not self._open

    def drain(self) -> list[Ticket]:
        """Run until every submitted solve has completed."""
        done = []
        while self._open:
            if not self._jobs and not self._waiting:
                torch.cuda.synchronize(self.device)
            done += self.step()
        return done

    def serve(self, requests: Sequence[tuple[float, dict[str, Any]]]) -> list[Ticket]:
        """Replay a trace: ``(arrival offset in seconds, submit kwargs)``
        pairs, submitted when the wall clock reaches each arrival; returns
        the tickets in trace order once all have completed."""
        order = sorted(range(len(requests)), key=lambda i: requests[i][0])
        t0 = time.perf_counter()
        tickets: list[Ticket | None] = [None] * len(requests)
        nxt = 0
        while nxt < len(order) or self._open:
            now = time.perf_counter()
            while nxt < len(order) and t0 + requests[order[nxt]][0] <= now:
                i = order[nxt]
                kw = dict(requests[i][1])
                problem = kw.pop("problem")
                tickets[i] = self.submit(problem, arrival=t0 + requests[i][0], **kw)
                nxt += 1
            if self._jobs or self._waiting:
                self.step()
            elif self._open:
                self._inflight.clear()
                torch.cuda.synchronize(self.device)
                self._collect()
            elif nxt < len(order):
                time.sleep(max(0.0, min(0.001, t0 + requests[order[nxt]][0] - time.perf_counter())))
        return tickets  # type: ignore[return-value]

    def close(self) -> None:
        if getattr(self, "_engine", None) is not None and self._engine.value:
            torch.cuda.synchronize(self.device)
            self._lib.tide_tm2d_engine_destroy(self._engine)
            self._engine = None
            self.pool = None

    def __del__(self) -> None:  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass


def _sum(ts: list[torch.Tensor]) -> torch.Tensor:
    out = ts[0]
    for t in ts[1:]:
        out = out + t
    return out


def _tensors(b: dict[str, Any]):
    for v in b.values():
        if isinstance(v, torch.Tensor):
            if v.is_cuda:
                yield v
        elif isinstance(v, list):
            for t in v:
                if isinstance(t, torch.Tensor) and t.is_cuda:
                    yield t


# ------------------------------------------------------------- workloads
def served_trace(
    n_requests: int,
    rate: float,
    *,
    seed: int = 2024,
    side_range: tuple[int, int] = (128, 1536),
    nt_range: tuple[int, int] = (500, 6000),
    shot_range: tuple[int, int] = (1, 32),
    gradient_fraction: float = 0.3,
    n_receivers: int = 64,
    pml_width: int = 20,
    dx: float = 0.01,
    storage_compression: bool | str = False,
    device: torch.device | str = "cpu",
) -> list[tuple[float, dict[str, Any]]]:
    """Poisson arrivals at ``rate`` per second of forward and gradient
    solves: grid sides log-uniform in ``side_range`` (independently per
    axis), nt and shot counts uniform, the seeded layered + anomaly model of
    :mod:`tide.synthetic` per request, a 200 MHz Ricker source per shot on
    the surface row and ``n_receivers`` receivers.  Gradient requests carry
    zero observed data (misfit 0.5 |d|^2)."""
    from .synthetic import cfl_dt, layered_model
    from .wavelets import ricker

    g = torch.Generator().manual_seed(seed)
    t = 0.0
    out = []
    dt = cfl_dt(dx, 2)
    lo, hi = math.log(side_range[0]), math.log(side_range[1])
    for i in range(n_requests):
        t += -math.log(1.0 - torch.rand(1, generator=g, dtype=torch.float64).item()) / rate
        ny = int(round(math.exp(lo + (hi - lo) * torch.rand(1, generator=g, dtype=torch.float64).item())))
        nx = int(round(math.exp(lo + (hi - lo) * torch.rand(1, generator=g, dtype=torch.float64).item())))
        nt = int(torch.randint(nt_range[0], nt_range[1] + 1, (1,), generator=g).item())
        shots = int(torch.randint(shot_range[0], shot_range[1] + 1, (1,), generator=g).item())
        grad = torch.rand(1, generator=g, dtype=torch.float64).item() < gradient_fraction
        eps, sig, mu = layered_model(ny, nx, seed=seed * 1000 + i, device=device)
        wav = ricker(2e8, nt, dt, dtype=torch.float32, device=device)
        xs = torch.linspace(1, nx - 2, shots).round().long()
        src = torch.stack([torch.tensor([[1, int(x)]]) for x in xs])
        nr = min(n_receivers, nx - 2)
        rx = torch.linspace(1, nx - 2, nr).round().long()
        rec = torch.stack([torch.stack([torch.ones_like(rx), rx], dim=1)] * shots)
        problem = dict(epsilon=eps, sigma=sig, mu=mu, grid_spacing=dx, dt=dt,
                       source_amplitude=wav.reshape(1, 1, -1).repeat(shots, 1, 1),
                       source_location=src, receiver_location=rec, pml_width=pml_width)
        kw: dict[str, Any] = {"problem": problem, "kind": "gradient" if grad else "forward",
                              "storage_compression": storage_compression}
        if grad:
            kw["observed"] = torch.zeros(nt, shots, nr, dtype=torch.float32, device=device)
        out.append((t, kw))
    return out


def latency_stats(tickets: Sequence[Ticket]) -> dict[str, float]:
    """Solves/s over the busy span and p50 / p99 latency in seconds."""
    lat = sorted(t.latency for t in tickets if t.latency is not None)
    if not lat:
        return {"solves": 0}
    first = min(t.arrival for t in tickets)
    last = max(t.completed for t in tickets if t.completed is not None)

    def pct(q: float) -> float:
        pos = q * (len(lat) - 1)
        lo = int(math.floor(pos))
        hi = min(lo + 1, len(lat) - 1)
        return lat[lo] + (lat[hi] - lat[lo]) * (pos - lo)

    return {
        "solves": len(lat),
        "solves_per_s": len(lat) / (last - first),
        "p50_s": pct(0.50),
        "p99_s": pct(0.99),
        "mean_s": sum(lat) / len(lat),
        "span_s": last - first,
    }


__all__ = [
    "DevicePool", "SolveResult", "TM2DServer", "Ticket", "job_bytes", "latency_stats", "model_gradient",
    "padded_shape", "plan_chunks", "served_trace", "spec_bytes", "state_spec",
]
