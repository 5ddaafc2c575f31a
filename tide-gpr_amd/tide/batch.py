"""Batched 2-D TM solves: many independent models advanced together.

A :class:`TM2DBatch` takes problems of different grid sizes, shot counts and
nt (each given as the keyword arguments of ``tide.maxwell.tm2d.maxwelltm``),
prepares them exactly as the single-problem operator does, and builds one
native plan (``tide_tm2d_plan_create``): each time step is one launch over
every problem still running, and the forward and adjoint time loops are
captured once into HIP graphs and replayed by every run.  Inputs are read
through fixed device buffers, so :meth:`TM2DBatch.update_models` rewrites
the coefficients in place and the same graphs serve the next iteration.

The gradient of each problem follows the reference semantics: the native
adjoint returns dJ/dca, dJ/dcb (FP64) and dJ/df, and torch's autograd
applies the coefficient, padding and source-scaling chain rule.
"""

from __future__ import annotations

import ctypes
from collections.abc import Sequence
from typing import Any

import torch

from . import _native
from .maxwell.tm2d import prepare_tm2d
from .storage import STORAGE_FORMAT_BF16, STORAGE_FORMAT_FULL

_P = ctypes.c_void_p


class _Problem(ctypes.Structure):
    """Mirror of ``tide_tm2d_problem`` (include/tide_hip.h)."""

    _fields_ = [
        ("ca", _P), ("cb", _P), ("cq", _P), ("f", _P),
        ("sources_i", _P), ("receivers_i", _P),
        ("prof", _P * 12),
        ("rdy", ctypes.c_float), ("rdx", ctypes.c_float),
        ("nt", ctypes.c_int64), ("n_shots", ctypes.c_int64), ("ny", ctypes.c_int64),
        ("nx", ctypes.c_int64), ("n_src", ctypes.c_int64), ("n_rec", ctypes.c_int64),
        ("interval", ctypes.c_int64), ("pml", ctypes.c_int64 * 4),
        ("fields", _P * 7), ("r", _P), ("ey_store", _P), ("curl_store", _P),
        ("grad_r", _P), ("lambda_", _P * 7), ("grad_ca", _P), ("grad_cb", _P),
        ("grad_f", _P),
    ]


def _bind(lib: ctypes.CDLL) -> None:
    if getattr(lib, "_tide_plan_bound", False):
        return
    lib.tide_tm2d_plan_create.restype = ctypes.c_int
    lib.tide_tm2d_plan_create.argtypes = [
        ctypes.POINTER(_Problem), ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
        ctypes.c_int64, ctypes.POINTER(_P)]
    for name in ("tide_tm2d_plan_forward", "tide_tm2d_plan_backward"):
        fn = getattr(lib, name)
        fn.restype = ctypes.c_int
        fn.argtypes = [_P, _P]
    lib.tide_tm2d_plan_destroy.restype = ctypes.c_int
    lib.tide_tm2d_plan_destroy.argtypes = [_P]
    lib._tide_plan_bound = True


class TM2DBatch:
    """Independent 2-D TM problems solved together on one GPU.

    Parameters
    ----------
    problems:
        One dict per problem with ``epsilon, sigma, mu, grid_spacing, dt,
        source_amplitude, source_location, receiver_location`` and optional
        ``pml_width`` (``maxwelltm`` arguments; models ``[ny, nx]``, fp32,
        on a HIP device).
    gradients:
        Keep snapshots and adjoint state so :meth:`backward` can run.
    storage_compression:
        ``"bf16"`` (or True) stores BF16 snapshots.
    model_gradient_sampling_interval:
        Snapshot / gradient sampling interval.
    use_graph:
        Replay captured HIP graphs (default) or launch every step eagerly.
    """

    def __init__(
        self,
        problems: Sequence[dict[str, Any]],
        *,
        gradients: bool = True,
        storage_compression: bool | str = False,
        model_gradient_sampling_interval: int = 1,
        use_graph: bool = True,
    ) -> None:
        if not problems:
            raise ValueError("TM2DBatch needs at least one problem")
        self._lib = _native.load()
        _bind(self._lib)
        self.problems = [dict(p) for p in problems]
        dev = self.problems[0]["epsilon"].device
        if dev.type != "cuda":
            raise NotImplementedError("TM2DBatch runs on HIP devices only")
        self.device = dev
        bf16 = storage_compression in (True, "bf16", "bfloat16")
        self.fmt = STORAGE_FORMAT_BF16 if bf16 else STORAGE_FORMAT_FULL
        self.interval = max(1, int(model_gradient_sampling_interval))
        self.gradients = gradients
        self.setups = []
        self._keep: list[Any] = []
        table = (_Problem * len(self.problems))()
        for i, kw in enumerate(self.problems):
            if kw["epsilon"].dtype != torch.float32 or kw["epsilon"].ndim != 2:
                raise TypeError("TM2DBatch problems need fp32 [ny, nx] models")
            st = prepare_tm2d(
                kw["epsilon"].detach(), kw["sigma"].detach(), kw["mu"].detach(),
                kw["grid_spacing"], kw["dt"], kw.get("source_amplitude"),
                kw.get("source_location"), kw.get("receiver_location"),
                pml_width=kw.get("pml_width", 20), nt=kw.get("nt"),
            )
            self.setups.append(st)
            self._fill(table[i], st)
        self._table = table
        handle = _P()
        _native.check(self._lib.tide_tm2d_plan_create(
            table, len(self.problems), 1 if bf16 else 0, 1 if use_graph else 0,
            dev.index if dev.index is not None else torch.cuda.current_device(),
            ctypes.byref(handle)))
        self._plan = handle

    # ------------------------------------------------------------ buffers
    def _fill(self, p: _Problem, st) -> None:
        dev, f32 = self.device, torch.float32
        shape = (st.n_shots, st.ny, st.nx)
        src = st.sources_i.to(dev)
        rec = st.receivers_i.to(dev)
        f = st.f.detach().contiguous() if st.n_src > 0 else torch.zeros(1, device=dev, dtype=f32)
        st.extra.update(
            f=f, src=src, rec=rec,
            ca=st.ca.detach().contiguous(), cb=st.cb.detach().contiguous(),
            cq=st.cq.detach().contiguous(),
            r=torch.zeros(st.nt, st.n_shots, max(st.n_rec, 1), device=dev, dtype=f32),
        )
        ex = st.extra
        p.ca, p.cb, p.cq = (_native.ptr(ex[k]) for k in ("ca", "cb", "cq"))
        p.f = _native.ptr(f) if st.n_src > 0 else None
        p.sources_i = _native.ptr(src) if st.n_src > 0 else None
        p.receivers_i = _native.ptr(rec) if st.n_rec > 0 else None
        for k, t in enumerate(st.profiles):
            p.prof[k] = _native.ptr(t)
        p.rdy, p.rdx = st.rdy, st.rdx
        p.nt, p.n_shots, p.ny, p.nx = st.nt, st.n_shots, st.ny, st.nx
        p.n_src, p.n_rec, p.interval = st.n_src, st.n_rec, self.interval
        for k, v in enumerate(st.pml):
            p.pml[k] = v
        for k, t in enumerate(st.fields):
            p.fields[k] = _native.ptr(t)
        p.r = _native.ptr(ex["r"]) if st.n_rec > 0 else None
        if self.gradients:
            n_store = (st.nt + self.interval - 1) // self.interval
            sdt = torch.int16 if self.fmt == STORAGE_FORMAT_BF16 else f32
            ex["ey_store"] = torch.empty(n_store, *shape, device=dev, dtype=sdt)
            ex["curl_store"] = torch.empty(n_store, *shape, device=dev, dtype=sdt)
            ex["grad_r"] = torch.zeros(st.nt, st.n_shots, max(st.n_rec, 1), device=dev, dtype=f32)
            ex["lambda"] = [torch.zeros(shape, device=dev, dtype=f32) for _ in range(7)]
            ex["grad_ca"] = torch.zeros(st.ny, st.nx, device=dev, dtype=torch.float64)
            ex["grad_cb"] = torch.zeros(st.ny, st.nx, device=dev, dtype=torch.float64)
            ex["grad_f"] = torch.zeros(st.nt, st.n_shots, max(st.n_src, 1), device=dev, dtype=f32)
            p.ey_store = _native.ptr(ex["ey_store"])
            p.curl_store = _native.ptr(ex["curl_store"])
            p.grad_r = _native.ptr(ex["grad_r"]) if st.n_rec > 0 else None
            for k, t in enumerate(ex["lambda"]):
                p.lambda_[k] = _native.ptr(t)
            p.grad_ca = _native.ptr(ex["grad_ca"])
            p.grad_cb = _native.ptr(ex["grad_cb"])
            p.grad_f = _native.ptr(ex["grad_f"]) if st.n_src > 0 else None

    # ------------------------------------------------------------------ runs
    def _stream(self):
        return _native.stream_handle(self.device)

    def reset_state(self) -> None:
        """Zero every problem's wavefield state (the forward's initial state)."""
        for st in self.setups:
            for t in st.fields:
                t.zero_()

    def forward(self) -> list[torch.Tensor]:
        """Run every problem forward from its current state; returns the
        receiver traces ``[nt, shots, n_rec]`` (views into plan buffers)."""
        _native.check(self._lib.tide_tm2d_plan_forward(self._plan, self._stream()))
        return [st.extra["r"][:, :, : st.n_rec] for st in self.setups]

    def backward(self, grad_r: Sequence[torch.Tensor]) -> list[tuple[torch.Tensor, torch.Tensor, torch.Tensor]]:
        """Adjoint of the last forward for receiver cotangents ``grad_r``;
        returns ``(dJ/dca, dJ/dcb, dJ/df)`` per problem (FP64, FP64, fp32)."""
        if not self.gradients:
            raise RuntimeError("TM2DBatch was built with gradients=False")
        for st, g in zip(self.setups, grad_r):
            if st.n_rec > 0:
                st.extra["grad_r"][:, :, : st.n_rec].copy_(g)
        _native.check(self._lib.tide_tm2d_plan_backward(self._plan, self._stream()))
        return [(st.extra["grad_ca"], st.extra["grad_cb"], st.extra["grad_f"][:, :, : st.n_src])
                for st in self.setups]

    def gradient(self, observed: Sequence[torch.Tensor]):
        """Forward from zero state, L2 misfit ``0.5 |d - d_obs|^2`` (FP64) and
        its epsilon / sigma gradients per problem.  Returns
        ``(misfits, eps_grads, sigma_grads)``."""
        self.reset_state()
        rs = self.forward()
        resid = [(r.double() - o.double()) for r, o in zip(rs, observed)]
        misfit = [0.5 * (x * x).sum() for x in resid]
        grads = self.backward([x.float() for x in resid])
        eps_g, sig_g = [], []
        for kw, st, (gca, gcb, gf) in zip(self.problems, self.setups, grads):
            e = kw["epsilon"].detach().clone().requires_grad_(True)
            s = kw["sigma"].detach().clone().requires_grad_(True)
            st2 = prepare_tm2d(e, s, kw["mu"].detach(), kw["grid_spacing"], kw["dt"],
                               kw.get("source_amplitude"), kw.get("source_location"),
                               kw.get("receiver_location"), pml_width=kw.get("pml_width", 20),
                               nt=kw.get("nt"))
            outs = [st2.ca, st2.cb]
            cots = [gca.to(torch.float32).reshape(st2.ca.shape), gcb.to(torch.float32).reshape(st2.cb.shape)]
            if st.n_src > 0:
                outs.append(st2.f)
                cots.append(gf)
            torch.autograd.backward(outs, cots)
            eps_g.append(e.grad)
            sig_g.append(s.grad)
        return misfit, eps_g, sig_g

    def update_models(self, epsilon: Sequence[torch.Tensor], sigma: Sequence[torch.Tensor]) -> None:
        """Replace every problem's model in place (the plan's graphs stay valid)."""
        for kw, st, e, s in zip(self.problems, self.setups, epsilon, sigma):
            kw["epsilon"], kw["sigma"] = e, s
            new = prepare_tm2d(e.detach(), s.detach(), kw["mu"].detach(), kw["grid_spacing"],
                               kw["dt"], kw.get("source_amplitude"), kw.get("source_location"),
                               kw.get("receiver_location"), pml_width=kw.get("pml_width", 20),
                               nt=kw.get("nt"))
            for k in ("ca", "cb", "cq"):
                st.extra[k].copy_(getattr(new, k))
            if st.n_src > 0:
                st.extra["f"].copy_(new.f)

    def close(self) -> None:
        if getattr(self, "_plan", None) is not None and self._plan.value:
            torch.cuda.synchronize(self.device)
            self._lib.tide_tm2d_plan_destroy(self._plan)
            self._plan = None

    def __del__(self) -> None:  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass


__all__ = ["TM2DBatch"]
