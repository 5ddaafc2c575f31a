"""Direct C-ABI calls of the HIP library on oracle-style numpy cases.

The parity tests feed the kernels exactly the arrays the oracle sees
(``oracle.TMCase``), so any difference is the kernels' own.
"""

from __future__ import annotations

import numpy as np
import torch

from oracle.oracle import TMCase
from tide import _native
from tide.storage import STORAGE_DEVICE


def _dev(a, device, dtype=None):
    t = torch.from_numpy(np.ascontiguousarray(a))
    if dtype is not None:
        t = t.to(dtype)
    return t.to(device).contiguous()


def gpu_forward(case: TMCase, *, store: str | None = None, interval: int = 1,
                device="cuda:0", nt: int | None = None, start_t: int = 0,
                fields: list | None = None, entry: str | None = None):
    """Run ``maxwell_tm_2_<dtype>_forward[_with_storage]_cuda`` (without
    storage, ``entry`` names another symbol with the same arguments).

    Returns numpy ``(fields, r, ey_store, curl_store)``."""
    dev = torch.device(device)
    dtype = torch.float64 if case.dtype == np.float64 else torch.float32
    nt = case.nt if nt is None else nt
    ca, cb, cq = (_dev(a, dev) for a in (case.ca, case.cb, case.cq))
    f = _dev(case.f, dev)
    prof = [_dev(p, dev) for p in case.profiles]
    flds = [_dev(x, dev) for x in (fields if fields is not None else case.fields)]
    N, ny, nx = case.n_shots, case.ny, case.nx
    n_src, n_rec = case.src.shape[1], case.rec.shape[1]
    r = torch.zeros(start_t + nt, N, n_rec, device=dev, dtype=dtype)
    src = torch.from_numpy(np.ascontiguousarray(case.src, np.int64))
    rec = torch.from_numpy(np.ascontiguousarray(case.rec, np.int64))
    batched = case.ca.shape[0] != 1
    stream = _native.stream_handle(dev)
    P = _native.ptr
    es = cs = None
    if store is None:
        fn = (_native.get_function("maxwell_tm", "forward", case.stencil, dtype) if entry is None
              else getattr(_native.load(), entry))
        st = fn(P(ca), P(cb), P(cq), P(f), *[P(x) for x in flds],
                None, None, None, None, None, P(r), 0, *[P(p) for p in prof],
                P(src), P(rec), case.rdy, case.rdx, case.dt, nt, N, ny, nx, n_src, n_rec,
                interval, False, batched, batched, batched, start_t, *case.pml, 0,
                dev.index or 0, stream)
    else:
        fmt = 1 if store == "bf16" else 0
        n_store = (start_t + nt + interval - 1) // interval
        sdt = torch.int16 if fmt == 1 else dtype
        es = torch.zeros(n_store, N, ny, nx, device=dev, dtype=sdt)
        cs = torch.zeros_like(es)
        shot_bytes = ny * nx * (2 if fmt == 1 else case.dtype.itemsize)
        fn = _native.get_function("maxwell_tm", "forward_with_storage", case.stencil, dtype)
        st = fn(P(ca), P(cb), P(cq), P(f), *[P(x) for x in flds], P(r),
                P(es), None, None, P(cs), None, None, *[P(p) for p in prof],
                P(src), P(rec), case.rdy, case.rdx, case.dt, nt, N, ny, nx, n_src, n_rec,
                interval, STORAGE_DEVICE, fmt, shot_bytes, True, True,
                batched, batched, batched, start_t, *case.pml, 0, dev.index or 0, stream, stream)
    _native.check(st)
    torch.cuda.synchronize(dev)
    to_np = lambda t: None if t is None else t.cpu().numpy()  # noqa: E731
    es_np, cs_np = to_np(es), to_np(cs)
    if store == "bf16":
        es_np, cs_np = es_np.view(np.uint16), cs_np.view(np.uint16)
    return [to_np(x) for x in flds], to_np(r), es_np, cs_np


def gpu_backward(case: TMCase, grad_r: np.ndarray, ey_store: np.ndarray | None,
                 curl_store: np.ndarray | None, *, store: str = "full", interval: int = 1,
                 device="cuda:0", reference_abi: bool = False, lam0=None):
    """Run the adjoint.  ``reference_abi`` selects
    ``maxwell_tm_2_<dtype>_backward_cuda`` (fp32/fp64 grad out) instead of
    the FP64-accumulator entry.  Returns ``(grad_ca, grad_cb, grad_f, lam)``."""
    dev = torch.device(device)
    dtype = torch.float64 if case.dtype == np.float64 else torch.float32
    ca, cb, cq = (_dev(a, dev) for a in (case.ca, case.cb, case.cq))
    prof = [_dev(p, dev) for p in case.profiles]
    N, ny, nx, nt = case.n_shots, case.ny, case.nx, case.nt
    n_src, n_rec = case.src.shape[1], case.rec.shape[1]
    if lam0 is None:
        lam = [torch.zeros(N, ny, nx, device=dev, dtype=dtype) for _ in range(7)]
    else:
        lam = [_dev(np.ascontiguousarray(x).reshape(N, ny, nx), dev, dtype) for x in lam0]
    gr = _dev(grad_r, dev, dtype)
    gf = torch.zeros(nt, N, n_src, device=dev, dtype=dtype)
    src = torch.from_numpy(np.ascontiguousarray(case.src, np.int64))
    rec = torch.from_numpy(np.ascontiguousarray(case.rec, np.int64))
    batched = case.ca.shape[0] != 1
    gshape = (N, ny, nx) if batched else (ny, nx)
    fmt = 1 if store == "bf16" else 0
    shot_bytes = ny * nx * (2 if fmt == 1 else case.dtype.itemsize)
    sdt = torch.int16 if fmt == 1 else dtype
    es = None if ey_store is None else torch.from_numpy(np.ascontiguousarray(ey_store).view(np.int16) if fmt == 1 else np.ascontiguousarray(ey_store)).to(dev)
    cs = None if curl_store is None else torch.from_numpy(np.ascontiguousarray(curl_store).view(np.int16) if fmt == 1 else np.ascontiguousarray(curl_store)).to(dev)
    stream = _native.stream_handle(dev)
    P = _native.ptr
    if reference_abi:
        gca = torch.zeros(gshape, device=dev, dtype=dtype)
        gcb = torch.zeros(gshape, device=dev, dtype=dtype)
        fn = _native.get_function("maxwell_tm", "backward", case.stencil, dtype)
        st = fn(P(ca), P(cb), P(cq), P(gr), *[P(x) for x in lam], None,
                P(es), None, None, P(cs), None, None, P(gf), P(gca), P(gcb), None, None,
                *[P(p) for p in prof], P(src), P(rec), case.rdy, case.rdx, case.dt,
                nt, N, ny, nx, n_src, n_rec, interval, STORAGE_DEVICE, fmt, shot_bytes,
                es is not None, cs is not None, batched, batched, batched, nt, *case.pml,
                0, dev.index or 0, stream, stream)
    else:
        gca = torch.zeros(gshape, device=dev, dtype=torch.float64)
        gcb = torch.zeros(gshape, device=dev, dtype=torch.float64)
        fn = getattr(_native.load(), f"tide_tm2d_{_native.dtype_name(dtype)}_backward_f64acc")
        st = fn(P(ca), P(cb), P(cq), P(gr), *[P(x) for x in lam], P(es), None, None,
                P(cs), None, None, P(gf),
                P(gca) if es is not None else None, P(gcb) if cs is not None else None,
                *[P(p) for p in prof], P(src), P(rec), case.rdy, case.rdx, nt, N, ny, nx,
                n_src, n_rec, interval, STORAGE_DEVICE, fmt, shot_bytes, batched, batched,
                batched, nt, *case.pml, case.stencil, dev.index or 0, stream, stream)
    _native.check(st)
    torch.cuda.synchronize(dev)
    return gca.cpu().numpy(), gcb.cpu().numpy(), gf.cpu().numpy(), [x.cpu().numpy() for x in lam]


# ---------------------------------------------------------------------- 3-D
def gpu_em3d_forward(case, *, store: str | None = None, interval: int = 1, device="cuda:0",
                     nt: int | None = None, start_t: int = 0, fields=None):
    """Run ``tide_em3d_<dtype>_forward`` on an ``oracle.EMCase``.

    Returns numpy ``(fields, r, stores)`` (stores: six arrays or None)."""
    dev = torch.device(device)
    dtype = torch.float64 if case.dtype == np.float64 else torch.float32
    nt = case.nt if nt is None else nt
    ca, cb, cq = (_dev(a, dev) for a in (case.ca, case.cb, case.cq))
    f = _dev(case.f, dev)
    prof = [_dev(p, dev) for p in case.profiles]
    flds = [_dev(x, dev) for x in (fields if fields is not None else case.fields)]
    N = case.n_shots
    nz, ny, nx = case.shape
    n_src, n_rec = case.src.shape[1], case.rec.shape[1]
    r = torch.zeros(start_t + nt, N, n_rec, device=dev, dtype=dtype)
    src = torch.from_numpy(np.ascontiguousarray(case.src, np.int64))
    rec = torch.from_numpy(np.ascontiguousarray(case.rec, np.int64))
    batched = case.ca.shape[0] != 1
    stores = None
    fmt = 1 if store == "bf16" else 0
    if store is not None:
        n_store = (start_t + nt + interval - 1) // interval
        sdt = torch.int16 if fmt == 1 else dtype
        stores = [torch.zeros(n_store, N, nz, ny, nx, device=dev, dtype=sdt) for _ in range(6)]
    step_bytes = N * nz * ny * nx * (2 if fmt == 1 else case.dtype.itemsize)
    fn = getattr(_native.load(), f"tide_em3d_{_native.dtype_name(dtype)}_forward")
    st = fn(_native.ptr(ca), _native.ptr(cb), _native.ptr(cq), _native.ptr(f),
            _native.ptr_array(flds), _native.ptr(r), _native.ptr_array(prof),
            _native.ptr(src), _native.ptr(rec), case.rdz, case.rdy, case.rdx, nt, N, nz, ny, nx,
            n_src, n_rec, interval, batched, batched, batched, start_t,
            _native.int64_array(case.pml), case.src_comp, case.rec_comp,
            _native.ptr_array(stores if stores is not None else [None] * 6), fmt, step_bytes, 0,
            dev.index or 0, _native.stream_handle(dev))
    _native.check(st)
    torch.cuda.synchronize(dev)
    out_st = None
    if stores is not None:
        out_st = [s.cpu().numpy() for s in stores]
        if fmt == 1:
            out_st = [s.view(np.uint16) for s in out_st]
    return [x.cpu().numpy() for x in flds], r.cpu().numpy(), out_st


def gpu_em3d_backward(case, grad_r, stores, *, store: str = "full", interval: int = 1,
                      device="cuda:0", lam0=None, want_ca: bool = True, want_cb: bool = True):
    """Run ``tide_em3d_<dtype>_backward_f64acc``.  Returns
    ``(grad_ca, grad_cb, grad_f, lam)`` as numpy (FP64 gradients)."""
    dev = torch.device(device)
    dtype = torch.float64 if case.dtype == np.float64 else torch.float32
    ca, cb, cq = (_dev(a, dev) for a in (case.ca, case.cb, case.cq))
    prof = [_dev(p, dev) for p in case.profiles]
    N = case.n_shots
    nz, ny, nx = case.shape
    n_src, n_rec = case.src.shape[1], case.rec.shape[1]
    shape = (N, nz, ny, nx)
    lam = ([torch.zeros(shape, device=dev, dtype=dtype) for _ in range(18)] if lam0 is None
           else [_dev(np.ascontiguousarray(x).reshape(shape), dev, dtype) for x in lam0])
    gr = _dev(grad_r, dev, dtype)
    gf = torch.zeros(case.nt, N, n_src, device=dev, dtype=dtype)
    gshape = shape if case.ca.shape[0] != 1 else shape[1:]
    gca = torch.zeros(gshape, device=dev, dtype=torch.float64) if want_ca else None
    gcb = torch.zeros(gshape, device=dev, dtype=torch.float64) if want_cb else None
    fmt = 1 if store == "bf16" else 0
    st_t = None
    if stores is not None:
        st_t = [torch.from_numpy(np.ascontiguousarray(s).view(np.int16) if fmt == 1
                                 else np.ascontiguousarray(s)).to(dev) for s in stores]
    src = torch.from_numpy(np.ascontiguousarray(case.src, np.int64))
    rec = torch.from_numpy(np.ascontiguousarray(case.rec, np.int64))
    batched = case.ca.shape[0] != 1
    step_bytes = N * nz * ny * nx * (2 if fmt == 1 else case.dtype.itemsize)
    fn = getattr(_native.load(), f"tide_em3d_{_native.dtype_name(dtype)}_backward_f64acc")
    st = fn(_native.ptr(ca), _native.ptr(cb), _native.ptr(cq), _native.ptr(gr),
            _native.ptr_array(lam), _native.ptr_array(st_t if st_t is not None else [None] * 6),
            fmt, step_bytes, 0, _native.ptr(gf), _native.ptr(gca), _native.ptr(gcb),
            _native.ptr_array(prof), _native.ptr(src), _native.ptr(rec), case.rdz, case.rdy,
            case.rdx, case.nt, N, nz, ny, nx, n_src, n_rec, interval, batched, batched, batched,
            case.nt, _native.int64_array(case.pml), case.src_comp, case.rec_comp,
            dev.index or 0, _native.stream_handle(dev))
    _native.check(st)
    torch.cuda.synchronize(dev)
    to_np = lambda t: None if t is None else t.cpu().numpy()  # noqa: E731
    return to_np(gca), to_np(gcb), gf.cpu().numpy(), [x.cpu().numpy() for x in lam]


# ------------------------------------------ 3-D, reference argument order
_MEM_AXIS_3D = (1, 0, 0, 2, 2, 1, 0, 1, 2, 0, 1, 2)  # m_hz_y .. m_ey_x


def compact_memory(full: np.ndarray, axis: int, pml) -> np.ndarray:
    """The reference's compact CPML layout of one full-size memory
    (src/tide/grid_utils.py:96-160): only [0, lo) and [max(lo, hi-1), n)
    along ``axis``."""
    n = full.shape[1 + axis]
    lo, hs = pml[axis], max(pml[axis], pml[3 + axis] - 1)
    idx = np.concatenate([np.arange(lo), np.arange(hs, n)])
    return np.ascontiguousarray(np.take(full, idx, axis=1 + axis))


def full_memory(compact: np.ndarray, axis: int, pml, shape) -> np.ndarray:
    out = np.zeros(shape, compact.dtype)
    n = shape[1 + axis]
    lo, hs = pml[axis], max(pml[axis], pml[3 + axis] - 1)
    idx = np.concatenate([np.arange(lo), np.arange(hs, n)])
    sl = [slice(None)] * 4
    sl[1 + axis] = idx
    out[tuple(sl)] = compact
    return out


def gpu_em3d_ref_forward(case, *, store: str | None = None, device="cuda:0"):
    """``maxwell_3d_2_<dtype>_forward[_with_storage]_cuda`` with the
    reference's arguments (compact CPML memories).  Returns numpy
    ``(fields with memories expanded to full size, r, stores)``."""
    dev = torch.device(device)
    dtype = torch.float64 if case.dtype == np.float64 else torch.float32
    ca, cb, cq, f = (_dev(a, dev) for a in (case.ca, case.cb, case.cq, case.f))
    prof = [_dev(p, dev) for p in case.profiles]
    N = case.n_shots
    nz, ny, nx = case.shape
    n_src, n_rec = case.src.shape[1], case.rec.shape[1]
    flds = [_dev(x, dev) for x in case.fields[:6]]
    flds += [_dev(compact_memory(x, a, case.pml), dev)
             for x, a in zip(case.fields[6:], _MEM_AXIS_3D)]
    r = torch.zeros(case.nt, N, n_rec, device=dev, dtype=dtype)
    src = torch.from_numpy(np.ascontiguousarray(case.src, np.int64)).to(dev)
    rec = torch.from_numpy(np.ascontiguousarray(case.rec, np.int64)).to(dev)
    b = case.ca.shape[0] != 1
    P = _native.ptr
    lib = _native.load()
    name = _native.dtype_name(dtype)
    tail = (case.rdz, case.rdy, case.rdx, case.dt, case.nt, N, nz, ny, nx, n_src, n_rec, 1)
    flags = (b, b, b, 0, *case.pml, case.src_comp, case.rec_comp, 0, dev.index or 0, 0)
    stream = _native.stream_handle(dev)
    stores = None
    if store is None:
        st = getattr(lib, f"maxwell_3d_2_{name}_forward_cuda")(
            P(ca), P(cb), P(cq), P(f), *[P(x) for x in flds], *([None] * 9), P(r), 0,
            *[P(p) for p in prof], P(src), P(rec), *tail, False, *flags, stream)
    else:
        fmt = 1 if store == "bf16" else 0
        sdt = torch.int16 if fmt == 1 else dtype
        stores = [torch.zeros(case.nt, N, nz, ny, nx, device=dev, dtype=sdt) for _ in range(6)]
        shot_bytes = nz * ny * nx * (2 if fmt == 1 else case.dtype.itemsize)
        storage = []
        for s in stores:
            storage += [P(s), None, None]
        st = getattr(lib, f"maxwell_3d_2_{name}_forward_with_storage_cuda")(
            P(ca), P(cb), P(cq), P(f), *[P(x) for x in flds], P(r), *storage,
            *[P(p) for p in prof], P(src), P(rec), *tail, 0, fmt, shot_bytes, True, True,
            *flags, stream, stream)
    _native.check(st)
    torch.cuda.synchronize(dev)
    shape = (N, nz, ny, nx)
    out = [x.cpu().numpy() for x in flds[:6]]
    out += [full_memory(x.cpu().numpy(), a, case.pml, shape) for x, a in zip(flds[6:], _MEM_AXIS_3D)]
    out_st = None
    if stores is not None:
        out_st = [s.cpu().numpy() for s in stores]
        if store == "bf16":
            out_st = [s.view(np.uint16) for s in out_st]
    return out, r.cpu().numpy(), out_st


def gpu_em3d_ref_backward(case, grad_r, stores, *, store: str = "full", device="cuda:0",
                          grad_ca0=None, zero_grad_on_entry=True):
    """``maxwell_3d_2_<dtype>_backward_cuda`` with the reference's
    arguments.  Returns numpy ``(grad_ca, grad_cb, grad_eps, grad_sigma,
    grad_f)`` in the field dtype."""
    dev = torch.device(device)
    dtype = torch.float64 if case.dtype == np.float64 else torch.float32
    ca, cb, cq = (_dev(a, dev) for a in (case.ca, case.cb, case.cq))
    prof = [_dev(p, dev) for p in case.profiles]
    N = case.n_shots
    nz, ny, nx = case.shape
    shape = (N, nz, ny, nx)
    n_src, n_rec = case.src.shape[1], case.rec.shape[1]
    lam = [torch.zeros(shape, device=dev, dtype=dtype) for _ in range(6)]
    # reference order of the adjoint memories: E-named first, then H-named
    names_ref = ("m_ey_z", "m_ez_y", "m_ez_x", "m_ex_z", "m_ex_y", "m_ey_x",
                 "m_hz_y", "m_hy_z", "m_hx_z", "m_hz_x", "m_hy_x", "m_hx_y")
    axis_of = dict(zip(("m_hz_y", "m_hy_z", "m_hx_z", "m_hz_x", "m_hy_x", "m_hx_y",
                        "m_ey_z", "m_ez_y", "m_ez_x", "m_ex_z", "m_ex_y", "m_ey_x"), _MEM_AXIS_3D))
    for n in names_ref:
        a = axis_of[n]
        ext = list(shape)
        lo, hs = case.pml[a], max(case.pml[a], case.pml[3 + a] - 1)
        ext[1 + a] = lo + shape[1 + a] - hs
        lam.append(torch.zeros(ext, device=dev, dtype=dtype))
    gshape = shape if case.ca.shape[0] != 1 else shape[1:]
    gca = (torch.zeros(gshape, device=dev, dtype=dtype) if grad_ca0 is None
           else _dev(grad_ca0, dev, dtype))
    gcb = torch.zeros(gshape, device=dev, dtype=dtype)
    geps = torch.full(gshape, 7.0, device=dev, dtype=dtype)
    gsig = torch.full(gshape, 7.0, device=dev, dtype=dtype)
    gf = torch.full((case.nt, N, n_src), 3.0, device=dev, dtype=dtype)
    fmt = 1 if store == "bf16" else 0
    st_t = [torch.from_numpy(np.ascontiguousarray(s).view(np.int16) if fmt == 1
                             else np.ascontiguousarray(s)).to(dev) for s in stores]
    storage = []
    for s in st_t:
        storage += [_native.ptr(s), None, None]
    src = torch.from_numpy(np.ascontiguousarray(case.src, np.int64)).to(dev)
    rec = torch.from_numpy(np.ascontiguousarray(case.rec, np.int64)).to(dev)
    b = case.ca.shape[0] != 1
    P = _native.ptr
    shot_bytes = nz * ny * nx * (2 if fmt == 1 else case.dtype.itemsize)
    fn = getattr(_native.load(), f"maxwell_3d_2_{_native.dtype_name(dtype)}_backward_cuda")
    st = fn(P(ca), P(cb), P(cq), P(_dev(grad_r, dev, dtype)), *[P(x) for x in lam], None,
            *storage, P(gf), P(gca), P(gcb), P(geps), P(gsig), None, None, zero_grad_on_entry,
            *[P(p) for p in prof], P(src), P(rec), case.rdz, case.rdy, case.rdx, case.dt,
            case.nt, N, nz, ny, nx, n_src, n_rec, 1, 0, fmt, shot_bytes, True, True, b, b, b,
            case.nt, *case.pml, case.src_comp, case.rec_comp, 0, dev.index or 0, 0,
            _native.stream_handle(dev), _native.stream_handle(dev))
    _native.check(st)
    torch.cuda.synchronize(dev)
    return tuple(t.cpu().numpy() for t in (gca, gcb, geps, gsig, gf))


def gpu_em3d_ref_born_forward(case, dca, dcb, df, *, device="cuda:0", segments=(None,)):
    """``maxwell_3d_2_<dtype>_born_forward_cuda`` with the reference's
    arguments (compact CPML memories for both fields), over the whole time
    axis or, with ``segments`` a list of step counts, as consecutive calls
    with ``start_t`` / ``nt`` (the reference's time-chunked calls).  Returns
    numpy ``(fields, dfields (memories expanded to full size), r,
    background_r)``."""
    dev = torch.device(device)
    dtype = torch.float64 if case.dtype == np.float64 else torch.float32
    ca, cb, cq, f = (_dev(a, dev) for a in (case.ca, case.cb, case.cq, case.f))
    d_ca, d_cb, d_f = (_dev(np.ascontiguousarray(a, case.dtype), dev) for a in (dca, dcb, df))
    prof = [_dev(p, dev) for p in case.profiles]
    N = case.n_shots
    nz, ny, nx = case.shape
    n_src, n_rec = case.src.shape[1], case.rec.shape[1]

    def state(arrays):
        out = [_dev(x, dev) for x in arrays[:6]]
        out += [_dev(compact_memory(x, a, case.pml), dev) for x, a in zip(arrays[6:], _MEM_AXIS_3D)]
        return out

    flds = state(case.fields)
    dflds = state([np.zeros_like(x) for x in case.fields])
    r = torch.zeros(case.nt, N, n_rec, device=dev, dtype=dtype)
    bg_r = torch.zeros_like(r)
    src = torch.from_numpy(np.ascontiguousarray(case.src, np.int64)).to(dev)
    rec = torch.from_numpy(np.ascontiguousarray(case.rec, np.int64)).to(dev)
    b = case.ca.shape[0] != 1
    P = _native.ptr
    fn = getattr(_native.load(), f"maxwell_3d_2_{_native.dtype_name(dtype)}_born_forward_cuda")
    t0 = 0
    for seg in segments:
        nt = case.nt - t0 if seg is None else seg
        st = fn(P(ca), P(cb), P(cq), P(d_ca), P(d_cb), P(f), P(d_f), *[P(x) for x in flds],
                *[P(x) for x in dflds], P(r), P(bg_r), *[P(p) for p in prof], P(src), P(rec),
                case.rdz, case.rdy, case.rdx, case.dt, nt, N, nz, ny, nx, n_src, n_rec, 1,
                b, b, b, t0, *case.pml, case.src_comp, case.rec_comp, 0, dev.index or 0, 0,
                _native.stream_handle(dev))
        _native.check(st)
        t0 += nt
    torch.cuda.synchronize(dev)
    shape = (N, nz, ny, nx)

    def back(arrays):
        out = [x.cpu().numpy() for x in arrays[:6]]
        return out + [full_memory(x.cpu().numpy(), a, case.pml, shape)
                      for x, a in zip(arrays[6:], _MEM_AXIS_3D)]

    return back(flds), back(dflds), r.cpu().numpy(), bg_r.cpu().numpy()
