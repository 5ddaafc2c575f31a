"""Python access to the CPU checkers.  TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and bench.py's cpu_baseline leg may
import this module, and only to check (or time beside) the HIP product path.

* ``liboracle.so`` (built from ``tm2d_oracle.cpp`` by ``oracle/Makefile``):
  this project's CPU restatement of the reference's TM2D forward, exact
  adjoint (FP64 gradient accumulation) and tangent-linear forward.
* ``_ref/libtide_ref.so``: the reference's own CPU propagators compiled from
  the reference sources in place (``make -C oracle ref``), called through
  their own C ABI (reference ``backend_utils.py:256-308``).
* ``_ref/libtide_ref_fast.so``: the same sources with the reference's
  Release flags (-Ofast -mavx2), for timing the CPU baseline only (fast
  math: never a parity reference).

Arrays are numpy, C-contiguous; fields are ``[shots, ny, nx]``.
"""

from __future__ import annotations

import ctypes
import subprocess
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_LIB = HERE / "_build" / "liboracle.so"
REF_LIB = HERE / "_ref" / "libtide_ref.so"
REF_FAST_LIB = HERE / "_ref" / "libtide_ref_fast.so"

_oracle: ctypes.CDLL | None = None
_ref: dict[bool, ctypes.CDLL] = {}


def build(ref_root: str | None = None) -> None:
    """Compile the checkers (the reference library only if its sources exist)."""
    subprocess.run(["make", "-s", "-C", str(HERE), "oracle"], check=True)
    root = Path(ref_root or "/root/reference")
    if (root / "src" / "tide" / "csrc").exists():
        subprocess.run(["make", "-s", "-C", str(HERE), "ref", "ref_fast", f"REF_ROOT={root}"],
                       check=True)


def oracle_lib() -> ctypes.CDLL:
    global _oracle
    if _oracle is None:
        if not ORACLE_LIB.exists():
            build()
        _oracle = ctypes.CDLL(str(ORACLE_LIB))
    return _oracle


def ref_available(fast: bool = False) -> bool:
    return (REF_FAST_LIB if fast else REF_LIB).exists()


def ref_lib(fast: bool = False) -> ctypes.CDLL:
    if fast not in _ref:
        _ref[fast] = ctypes.CDLL(str(REF_FAST_LIB if fast else REF_LIB))
    return _ref[fast]


def _p(a: np.ndarray | None):
    if a is None or a.size == 0:
        return None
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(ctypes.c_void_p)


def _parr(arrs, ctype=ctypes.c_void_p):
    return (ctype * len(arrs))(*[_p(a) for a in arrs])


@dataclass
class TMCase:
    """One TM2D propagation in padded layout (mirrors ``TM2DSetup``)."""

    ca: np.ndarray  # [1|N, ny, nx]
    cb: np.ndarray
    cq: np.ndarray
    f: np.ndarray  # [nt, N, n_src]
    profiles: list  # ay, by, ayh, byh, ax, bx, axh, bxh, ky, kyh, kx, kxh
    src: np.ndarray  # [N, n_src] int64
    rec: np.ndarray  # [N, n_rec] int64
    rdy: float
    rdx: float
    dt: float
    nt: int
    pml: tuple  # pml_y0, pml_x0, pml_y1, pml_x1
    fields: list = field(default_factory=list)  # 7 initial fields
    stencil: int = 2

    @property
    def n_shots(self) -> int:
        return int(self.fields[0].shape[0])

    @property
    def ny(self) -> int:
        return int(self.fields[0].shape[1])

    @property
    def nx(self) -> int:
        return int(self.fields[0].shape[2])

    @property
    def dtype(self):
        return self.ca.dtype

    @property
    def batched(self) -> bool:
        return self.ca.shape[0] > 1 or (self.ca.shape[0] == 1 and self.n_shots == 1 and False)

    def astype(self, dtype) -> "TMCase":
        c = lambda a: np.ascontiguousarray(a, dtype=dtype)  # noqa: E731
        return TMCase(c(self.ca), c(self.cb), c(self.cq), c(self.f),
                      [c(p) for p in self.profiles], self.src, self.rec,
                      self.rdy, self.rdx, self.dt, self.nt, self.pml,
                      [c(x) for x in self.fields], self.stencil)


def case_from_setup(setup) -> TMCase:
    """Convert a ``tide.maxwell.tm2d.TM2DSetup`` (any device) to numpy."""
    n = lambda t: np.ascontiguousarray(t.detach().cpu().numpy())  # noqa: E731
    return TMCase(
        ca=n(setup.ca), cb=n(setup.cb), cq=n(setup.cq),
        f=n(setup.f).reshape(setup.nt, setup.n_shots, setup.n_src) if setup.n_src else np.zeros((setup.nt, setup.n_shots, 0), n(setup.ca).dtype),
        profiles=[n(p) for p in setup.profiles],
        src=n(setup.sources_i).reshape(setup.n_shots, setup.n_src).astype(np.int64),
        rec=n(setup.receivers_i).reshape(setup.n_shots, setup.n_rec).astype(np.int64),
        rdy=setup.rdy, rdx=setup.rdx, dt=setup.dt, nt=setup.nt, pml=tuple(setup.pml),
        fields=[n(t) for t in setup.fields], stencil=int(getattr(setup, "stencil", 2)),
    )


def _batched_flags(case: TMCase):
    b = case.ca.shape[0] != 1 or (case.n_shots == 1 and False)
    return (ctypes.c_bool * 3)(b, b, b)


def tm2d_forward(case: TMCase, *, store: str | None = None, interval: int = 1,
                 start_t: int = 0, nt: int | None = None):
    """Oracle forward.  ``store`` None / 'full' / 'bf16'.

    Returns ``(fields, r, ey_store, curl_store)``; fields are new arrays.
    """
    lib = oracle_lib()
    dt_ = case.dtype
    f64 = dt_ == np.float64
    nt = case.nt if nt is None else nt
    fields = [np.ascontiguousarray(x.copy()) for x in case.fields]
    n_rec = case.rec.shape[1]
    r = np.zeros((start_t + nt, case.n_shots, n_rec), dt_)
    es = cs = None
    fmt = 0
    if store is not None:
        n_store = (start_t + nt + interval - 1) // interval
        sdt = np.uint16 if store == "bf16" else dt_
        fmt = 1 if store == "bf16" else 0
        es = np.zeros((n_store, case.n_shots, case.ny, case.nx), sdt)
        cs = np.zeros_like(es)
    step_bytes = case.n_shots * case.ny * case.nx * (2 if fmt == 1 else dt_.itemsize)
    ct = ctypes.c_double if f64 else ctypes.c_float
    fn = lib.oracle_tm2d_forward_f64 if f64 else lib.oracle_tm2d_forward_f32
    fn.restype = ctypes.c_int
    f = case.f if case.f.shape[0] >= start_t + nt else np.zeros((start_t + nt, case.n_shots, case.src.shape[1]), dt_)
    rc = fn(ctypes.c_int(case.stencil), _p(case.ca), _p(case.cb), _p(case.cq), _p(f),
            _parr(fields), _p(r), _parr(case.profiles), _p(case.src), _p(case.rec),
            ct(case.rdy), ct(case.rdx), ctypes.c_int64(nt), ctypes.c_int64(case.n_shots),
            ctypes.c_int64(case.ny), ctypes.c_int64(case.nx),
            ctypes.c_int64(case.src.shape[1]), ctypes.c_int64(n_rec),
            ctypes.c_int64(interval), ctypes.c_int64(start_t),
            (ctypes.c_int64 * 4)(*case.pml), _batched_flags(case),
            _p(es), _p(cs), ctypes.c_int(fmt), ctypes.c_int64(step_bytes))
    if rc != 0:
        raise RuntimeError(f"oracle forward failed ({rc})")
    return fields, r, es, cs


def tm2d_backward(case: TMCase, grad_r: np.ndarray, ey_store, curl_store, *,
                  store: str = "full", interval: int = 1, want_ca: bool = True,
                  want_cb: bool = True, lam0=None):
    """Oracle exact adjoint.  Returns ``(grad_ca, grad_cb, grad_f, lam)`` with
    FP64 coefficient gradients (``[ny, nx]``, or ``[N, ny, nx]`` batched).
    ``lam0`` optionally gives the initial adjoint state (7 arrays, order
    lambda_ey, lambda_hx, lambda_hz, m_lambda_ey_x, m_lambda_ey_z,
    m_lambda_hx_z, m_lambda_hz_x); zero otherwise."""
    lib = oracle_lib()
    dt_ = case.dtype
    f64 = dt_ == np.float64
    shape = (case.n_shots, case.ny, case.nx)
    if lam0 is None:
        lam = [np.zeros(shape, dt_) for _ in range(7)]
    else:
        lam = [np.array(x, dtype=dt_, copy=True).reshape(shape) for x in lam0]
    n_src = case.src.shape[1]
    grad_f = np.zeros((case.nt, case.n_shots, n_src), dt_)
    batched = case.ca.shape[0] != 1
    gshape = shape if batched else (case.ny, case.nx)
    gca = np.zeros(gshape, np.float64)
    gcb = np.zeros(gshape, np.float64)
    fmt = 1 if store == "bf16" else 0
    step_bytes = case.n_shots * case.ny * case.nx * (2 if fmt == 1 else dt_.itemsize)
    ct = ctypes.c_double if f64 else ctypes.c_float
    fn = lib.oracle_tm2d_backward_f64 if f64 else lib.oracle_tm2d_backward_f32
    fn.restype = ctypes.c_int
    grad_r = np.ascontiguousarray(grad_r, dtype=dt_)
    rc = fn(ctypes.c_int(case.stencil), _p(case.ca), _p(case.cb), _p(case.cq), _p(grad_r),
            _parr(lam), _p(ey_store), _p(curl_store), ctypes.c_int(fmt),
            ctypes.c_int64(step_bytes), _p(grad_f) if n_src else None, _p(gca), _p(gcb),
            _parr(case.profiles), _p(case.src), _p(case.rec), ct(case.rdy), ct(case.rdx),
            ctypes.c_int64(case.nt), ctypes.c_int64(case.n_shots), ctypes.c_int64(case.ny),
            ctypes.c_int64(case.nx), ctypes.c_int64(n_src), ctypes.c_int64(case.rec.shape[1]),
            ctypes.c_int64(interval), ctypes.c_int64(case.nt), (ctypes.c_int64 * 4)(*case.pml),
            _batched_flags(case), ctypes.c_bool(want_ca), ctypes.c_bool(want_cb))
    if rc != 0:
        raise RuntimeError(f"oracle backward failed ({rc})")
    return (gca if want_ca else None), (gcb if want_cb else None), grad_f, lam


def tm2d_tangent(case: TMCase, dca: np.ndarray, dcb: np.ndarray, df: np.ndarray):
    """FP64 tangent-linear forward: returns ``(r, dr)``."""
    lib = oracle_lib()
    c = case.astype(np.float64)
    fields = [x.copy() for x in c.fields]
    dfields = [np.zeros_like(x) for x in c.fields]
    n_rec = c.rec.shape[1]
    r = np.zeros((c.nt, c.n_shots, n_rec))
    dr = np.zeros_like(r)
    fn = lib.oracle_tm2d_tangent_f64
    fn.restype = ctypes.c_int
    rc = fn(ctypes.c_int(c.stencil), _p(c.ca), _p(c.cb), _p(c.cq), _p(c.f),
            _p(np.ascontiguousarray(dca, np.float64)), _p(np.ascontiguousarray(dcb, np.float64)),
            _p(np.ascontiguousarray(df, np.float64)), _parr(fields), _parr(dfields), _p(r), _p(dr),
            _parr(c.profiles), _p(c.src), _p(c.rec), ctypes.c_double(c.rdy), ctypes.c_double(c.rdx),
            ctypes.c_int64(c.nt), ctypes.c_int64(c.n_shots), ctypes.c_int64(c.ny),
            ctypes.c_int64(c.nx), ctypes.c_int64(c.src.shape[1]), ctypes.c_int64(n_rec),
            (ctypes.c_int64 * 4)(*c.pml))
    if rc != 0:
        raise RuntimeError(f"oracle tangent failed ({rc})")
    return r, dr


def tm2d_forward_debye(case: TMCase, debye_a: np.ndarray, debye_b: np.ndarray,
                       debye_cp: np.ndarray, *, ey_prev: np.ndarray | None = None):
    """Debye-dispersive forward (``tm2d_oracle.cpp:forward_debye``, reference
    ``csrc/tm2d/maxwell_tm_inst.cpp:214-300, 780-805``): ``case.ca``/``cb``
    are the Debye coefficients, ``debye_*`` ``[n_poles, ny, nx]``.  Returns
    ``(fields, r, polarization, ey_prev)``."""
    lib = oracle_lib()
    dt_ = case.dtype
    suffix = "f64" if dt_ == np.float64 else "f32"
    fields = [np.ascontiguousarray(x.copy()) for x in case.fields]
    n_poles = int(debye_a.shape[0])
    pol = np.zeros((case.n_shots, n_poles, case.ny, case.nx), dt_)
    prev = (np.zeros((case.n_shots, case.ny, case.nx), dt_) if ey_prev is None
            else np.ascontiguousarray(ey_prev, dt_).copy())
    n_rec = case.rec.shape[1]
    r = np.zeros((case.nt, case.n_shots, n_rec), dt_)
    ct = ctypes.c_double if dt_ == np.float64 else ctypes.c_float
    deb = [np.ascontiguousarray(x, dt_) for x in (debye_a, debye_b, debye_cp)]
    fn = getattr(lib, f"oracle_tm2d_forward_debye_{suffix}")
    fn.restype = ctypes.c_int
    rc = fn(ctypes.c_int(case.stencil), _p(case.ca), _p(case.cb), _p(case.cq), _p(case.f),
            _parr(fields), _parr(deb), _p(pol), _p(prev), ctypes.c_int64(n_poles), _p(r),
            _parr(case.profiles), _p(case.src), _p(case.rec), ct(case.rdy), ct(case.rdx),
            ctypes.c_int64(case.nt), ctypes.c_int64(case.n_shots), ctypes.c_int64(case.ny),
            ctypes.c_int64(case.nx), ctypes.c_int64(case.src.shape[1]), ctypes.c_int64(n_rec),
            ctypes.c_int64(0), (ctypes.c_int64 * 4)(*case.pml))
    if rc != 0:
        raise RuntimeError(f"oracle Debye forward failed ({rc})")
    return fields, r, pol, prev


def tm2d_tangent_store(case: TMCase, dca: np.ndarray, dcb: np.ndarray, df: np.ndarray):
    """FP64 tangent-linear forward keeping the Born adjoint's snapshots:
    returns ``(r, dr, hist)`` with ``hist`` = background Ey^n, curl^n and
    tangent dEy^n, dcurl^n, each ``[nt, shots, ny, nx]``
    (``tm2d_oracle.cpp:tangent``)."""
    lib = oracle_lib()
    c = case.astype(np.float64)
    fields = [x.copy() for x in c.fields]
    dfields = [np.zeros_like(x) for x in c.fields]
    n_rec = c.rec.shape[1]
    r = np.zeros((c.nt, c.n_shots, n_rec))
    dr = np.zeros_like(r)
    hist = [np.zeros((c.nt, c.n_shots, c.ny, c.nx)) for _ in range(4)]
    fn = lib.oracle_tm2d_tangent_store_f64
    fn.restype = ctypes.c_int
    rc = fn(ctypes.c_int(c.stencil), _p(c.ca), _p(c.cb), _p(c.cq), _p(c.f),
            _p(np.ascontiguousarray(dca, np.float64)), _p(np.ascontiguousarray(dcb, np.float64)),
            _p(np.ascontiguousarray(df, np.float64)), _parr(fields), _parr(dfields), _p(r), _p(dr),
            _parr(c.profiles), _p(c.src), _p(c.rec), ctypes.c_double(c.rdy), ctypes.c_double(c.rdx),
            ctypes.c_int64(c.nt), ctypes.c_int64(c.n_shots), ctypes.c_int64(c.ny),
            ctypes.c_int64(c.nx), ctypes.c_int64(c.src.shape[1]), ctypes.c_int64(n_rec),
            (ctypes.c_int64 * 4)(*c.pml), _parr(hist))
    if rc != 0:
        raise RuntimeError(f"oracle tangent failed ({rc})")
    return r, dr, hist


def tm2d_tangent_adjoint(case: TMCase, dca: np.ndarray, dcb: np.ndarray, w: np.ndarray | None,
                         w_bg: np.ndarray | None, hist) -> dict:
    """FP64 adjoint of the tangent-linear forward (``tm2d_oracle.cpp:
    tangent_adjoint``): gradients of ``<dr, w> + <r, w_bg>`` with respect to
    ``ca, cb, dca, dcb`` ([ny, nx]), ``f`` and ``df`` ([nt, shots, n_src]),
    plus the final adjoint states ``lam`` (tangent) and ``eta`` (background)."""
    lib = oracle_lib()
    c = case.astype(np.float64)
    n_src, n_rec = c.src.shape[1], c.rec.shape[1]
    lam = [np.zeros_like(x) for x in c.fields]
    eta = [np.zeros_like(x) for x in c.fields]
    out = {k: np.zeros((c.ny, c.nx)) for k in ("ca", "cb", "dca", "dcb")}
    out["df"] = np.zeros((c.nt, c.n_shots, n_src))
    out["f"] = np.zeros_like(out["df"])
    fn = lib.oracle_tm2d_tangent_adjoint_f64
    fn.restype = ctypes.c_int
    opt = lambda a: None if a is None else _p(np.ascontiguousarray(a, np.float64))  # noqa: E731
    rc = fn(ctypes.c_int(c.stencil), _p(c.ca), _p(c.cb), _p(c.cq),
            _p(np.ascontiguousarray(dca, np.float64)), _p(np.ascontiguousarray(dcb, np.float64)),
            opt(w), opt(w_bg), _parr(hist), _parr(lam), _parr(eta), _p(out["df"]), _p(out["f"]),
            _p(out["ca"]), _p(out["cb"]), _p(out["dca"]), _p(out["dcb"]), _parr(c.profiles),
            _p(c.src), _p(c.rec), ctypes.c_double(c.rdy), ctypes.c_double(c.rdx),
            ctypes.c_int64(c.nt), ctypes.c_int64(c.n_shots), ctypes.c_int64(c.ny),
            ctypes.c_int64(c.nx), ctypes.c_int64(n_src), ctypes.c_int64(n_rec),
            (ctypes.c_int64 * 4)(*c.pml))
    if rc != 0:
        raise RuntimeError(f"oracle tangent adjoint failed ({rc})")
    out["lam"], out["eta"] = lam, eta
    return out


def bf16_encode(x: np.ndarray) -> np.ndarray:
    """Round-to-nearest-even fp32 -> bf16 bits (same rule as the kernels)."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = u + 0x7FFF + ((u >> 16) & 1)
    return ((u >> 16) & 0xFFFF).astype(np.uint16)


def bf16_decode(h: np.ndarray) -> np.ndarray:
    return (h.astype(np.uint32) << 16).view(np.float32)


# ------------------------------------------------------------ reference lib
def ref_tm2d_forward_debye(case: TMCase, debye_a, debye_b, debye_cp):
    """The reference's own CPU forward with Debye dispersion
    (``has_dispersion``): returns ``(fields, r, polarization)``.  Its
    ``ey_prev`` buffer starts at zero, as this module's."""
    lib = ref_lib(False)
    dt_ = case.dtype
    name = "double" if dt_ == np.float64 else "float"
    fn = getattr(lib, f"maxwell_tm_{case.stencil}_{name}_forward_cpu")
    ct = ctypes.c_double if dt_ == np.float64 else ctypes.c_float
    I, P, B = ctypes.c_int64, ctypes.c_void_p, ctypes.c_bool
    fn.restype = None
    fn.argtypes = ([P] * 17 + [I] + [P] * 14 + [ct] * 3 + [I] * 7 + [B] * 4 + [I] * 7 + [P])
    fields = [np.ascontiguousarray(x.copy()) for x in case.fields]
    n_rec = case.rec.shape[1]
    n_poles = int(debye_a.shape[0])
    r = np.zeros((case.nt, case.n_shots, n_rec), dt_)
    pol = np.zeros((case.n_shots, n_poles, case.ny, case.nx), dt_)
    prev = np.zeros((case.n_shots, case.ny, case.nx), dt_)
    deb = [np.ascontiguousarray(x, dt_) for x in (debye_a, debye_b, debye_cp)]
    fn(_p(case.ca), _p(case.cb), _p(case.cq), _p(case.f), *[_p(x) for x in fields],
       *[_p(x) for x in deb], _p(pol), _p(prev), _p(r), n_poles, *[_p(p) for p in case.profiles],
       _p(case.src), _p(case.rec), case.rdy, case.rdx, case.dt, case.nt, case.n_shots,
       case.ny, case.nx, case.src.shape[1], n_rec, 1, True, False, False, False, 0,
       *case.pml, 1, 0, None)
    return fields, r, pol


def ref_tm2d_forward(case: TMCase, *, n_threads: int = 0, nt: int | None = None,
                     fast: bool = False):
    """The reference's own CPU forward (``maxwell_tm_<s>_<dtype>_forward_cpu``).

    Returns ``(fields, r)``; ``n_threads`` 0 lets OpenMP decide; ``fast``
    runs the Release-flag build."""
    lib = ref_lib(fast)
    dt_ = case.dtype
    name = "double" if dt_ == np.float64 else "float"
    fn = getattr(lib, f"maxwell_tm_{case.stencil}_{name}_forward_cpu")
    ct = ctypes.c_double if dt_ == np.float64 else ctypes.c_float
    I, P, B = ctypes.c_int64, ctypes.c_void_p, ctypes.c_bool
    fn.restype = None
    fn.argtypes = ([P] * 17 + [I] + [P] * 14 + [ct] * 3 + [I] * 7 + [B] * 4 + [I] * 7 + [P])
    nt = case.nt if nt is None else nt
    fields = [np.ascontiguousarray(x.copy()) for x in case.fields]
    n_rec = case.rec.shape[1]
    r = np.zeros((nt, case.n_shots, n_rec), dt_)
    b = case.ca.shape[0] != 1
    fn(_p(case.ca), _p(case.cb), _p(case.cq), _p(case.f), *[_p(x) for x in fields],
       None, None, None, None, None, _p(r), 0, *[_p(p) for p in case.profiles],
       _p(case.src), _p(case.rec), case.rdy, case.rdx, case.dt, nt, case.n_shots,
       case.ny, case.nx, case.src.shape[1], n_rec, 1, False, b, b, b, 0,
       *case.pml, n_threads, 0, None)
    return fields, r


def ref_tm2d_forward_with_storage_and_backward(case: TMCase, grad_r: np.ndarray, *,
                                                n_threads: int = 0, fast: bool = False):
    """The reference CPU forward_with_storage followed by its CPU backward.

    Used only to time the reference CPU path (its CPU backward re-runs the
    forward kernels on the adjoint fields and is not the exact adjoint);
    ``fast`` runs the Release-flag build."""
    lib = ref_lib(fast)
    dt_ = case.dtype
    name = "double" if dt_ == np.float64 else "float"
    ct = ctypes.c_double if dt_ == np.float64 else ctypes.c_float
    I, P, B = ctypes.c_int64, ctypes.c_void_p, ctypes.c_bool
    fws = getattr(lib, f"maxwell_tm_{case.stencil}_{name}_forward_with_storage_cpu")
    fws.restype = None
    fws.argtypes = [P] * 12 + [P] * 6 + [P] * 14 + [ct] * 3 + [I] * 7 + [I] * 3 + [B] * 2 + [B] * 3 + [I] * 7 + [P] * 2
    bwd = getattr(lib, f"maxwell_tm_{case.stencil}_{name}_backward_cpu")
    bwd.restype = None
    bwd.argtypes = [P] * 4 + [P] * 3 + [P] * 4 + [P] + [P] * 6 + [P] + [P] * 4 + [P] * 14 + [ct] * 3 + [I] * 7 + [I] * 3 + [B] * 5 + [I] * 7 + [P] * 2
    N, ny, nx, nt = case.n_shots, case.ny, case.nx, case.nt
    n_src, n_rec = case.src.shape[1], case.rec.shape[1]
    fields = [np.ascontiguousarray(x.copy()) for x in case.fields]
    r = np.zeros((nt, N, n_rec), dt_)
    es = np.zeros((nt, N, ny, nx), dt_)
    cs = np.zeros_like(es)
    shot_bytes = ny * nx * dt_.itemsize
    fws(_p(case.ca), _p(case.cb), _p(case.cq), _p(case.f), *[_p(x) for x in fields], _p(r),
        _p(es), _p(es), None, _p(cs), _p(cs), None, *[_p(p) for p in case.profiles],
        _p(case.src), _p(case.rec), case.rdy, case.rdx, case.dt, nt, N, ny, nx, n_src, n_rec, 1,
        0, 0, shot_bytes, True, True, False, False, False, 0, *case.pml, n_threads, 0, None, None)
    lam = [np.zeros((N, ny, nx), dt_) for _ in range(7)]
    scratch = np.zeros((4, N, ny, nx), dt_)
    grad_f = np.zeros((nt, N, n_src), dt_)
    gca = np.zeros((ny, nx), dt_)
    gcb = np.zeros((ny, nx), dt_)
    gcas = np.zeros((N, ny, nx), dt_)
    gcbs = np.zeros((N, ny, nx), dt_)
    grad_r = np.ascontiguousarray(grad_r, dt_)
    bwd(_p(case.ca), _p(case.cb), _p(case.cq), _p(grad_r), *[_p(x) for x in lam], _p(scratch),
        _p(es), _p(es), None, _p(cs), _p(cs), None, _p(grad_f), _p(gca), _p(gcb), _p(gcas), _p(gcbs),
        *[_p(p) for p in case.profiles], _p(case.src), _p(case.rec), case.rdy, case.rdx, case.dt,
        nt, N, ny, nx, n_src, n_rec, 1, 0, 0, shot_bytes, True, True, False, False, False, nt,
        *case.pml, n_threads, 0, None, None)
    return r, gca, gcb


# ====================================================================== 3-D
@dataclass
class EMCase:
    """One 3-D Maxwell propagation in padded layout (mirrors ``EM3DSetup``)."""

    ca: np.ndarray  # [1|N, nz, ny, nx]
    cb: np.ndarray
    cq: np.ndarray
    f: np.ndarray  # [nt, N, n_src]
    profiles: list  # az bz azh bzh ay by ayh byh ax bx axh bxh kz kzh ky kyh kx kxh
    src: np.ndarray  # [N, n_src] int64
    rec: np.ndarray  # [N, n_rec] int64
    rdz: float
    rdy: float
    rdx: float
    dt: float
    nt: int
    pml: tuple  # z0 y0 x0 z1 y1 x1
    src_comp: int = 1
    rec_comp: int = 1
    fields: list = field(default_factory=list)  # 18 initial state arrays

    @property
    def n_shots(self) -> int:
        return int(self.fields[0].shape[0])

    @property
    def shape(self) -> tuple:
        return tuple(int(v) for v in self.fields[0].shape[1:])

    @property
    def dtype(self):
        return self.ca.dtype

    def astype(self, dtype) -> "EMCase":
        c = lambda a: np.ascontiguousarray(a, dtype=dtype)  # noqa: E731
        return EMCase(c(self.ca), c(self.cb), c(self.cq), c(self.f),
                      [c(p) for p in self.profiles], self.src, self.rec, self.rdz,
                      self.rdy, self.rdx, self.dt, self.nt, self.pml, self.src_comp,
                      self.rec_comp, [c(x) for x in self.fields])


def em_case_from_setup(setup) -> EMCase:
    """Convert a ``tide.maxwell.maxwell3d.EM3DSetup`` (any device) to numpy."""
    n = lambda t: np.ascontiguousarray(t.detach().cpu().numpy())  # noqa: E731
    dt_ = n(setup.ca).dtype
    f = (n(setup.f).reshape(setup.nt, setup.n_shots, setup.n_src) if setup.n_src
         else np.zeros((setup.nt, setup.n_shots, 0), dt_))
    return EMCase(
        ca=n(setup.ca), cb=n(setup.cb), cq=n(setup.cq), f=f,
        profiles=[n(p) for p in setup.profiles],
        src=n(setup.sources_i).reshape(setup.n_shots, setup.n_src).astype(np.int64),
        rec=n(setup.receivers_i).reshape(setup.n_shots, setup.n_rec).astype(np.int64),
        rdz=setup.rdz, rdy=setup.rdy, rdx=setup.rdx, dt=setup.dt, nt=setup.nt,
        pml=tuple(setup.pml), src_comp=setup.source_component,
        rec_comp=setup.receiver_component, fields=[n(t) for t in setup.fields],
    )


def _em_common(case: EMCase):
    f64 = case.dtype == np.float64
    ct = ctypes.c_double if f64 else ctypes.c_float
    b = case.ca.shape[0] != 1
    return f64, ct, (ctypes.c_bool * 3)(b, b, b)


def em3d_forward(case: EMCase, *, store: str | None = None, interval: int = 1,
                 start_t: int = 0, nt: int | None = None):
    """Oracle 3-D forward.  ``store`` None / 'full' / 'bf16'.

    Returns ``(fields, r, stores)`` with ``stores`` = six arrays
    ``[n_stored, N, nz, ny, nx]`` (Ex, Ey, Ez, curl x, y, z) or None."""
    lib = oracle_lib()
    f64, ct, batched = _em_common(case)
    nt = case.nt if nt is None else nt
    fields = [np.ascontiguousarray(x.copy()) for x in case.fields]
    n_rec = case.rec.shape[1]
    r = np.zeros((start_t + nt, case.n_shots, n_rec), case.dtype)
    stores, fmt = None, 0
    if store is not None:
        n_store = (start_t + nt + interval - 1) // interval
        fmt = 1 if store == "bf16" else 0
        sdt = np.uint16 if fmt == 1 else case.dtype
        stores = [np.zeros((n_store, case.n_shots, *case.shape), sdt) for _ in range(6)]
    step_bytes = case.n_shots * int(np.prod(case.shape)) * (2 if fmt == 1 else case.dtype.itemsize)
    f = case.f if case.f.shape[0] >= start_t + nt else np.zeros((start_t + nt, case.n_shots, case.src.shape[1]), case.dtype)
    fn = lib.oracle_em3d_forward_f64 if f64 else lib.oracle_em3d_forward_f32
    fn.restype = ctypes.c_int
    nz, ny, nx = case.shape
    rc = fn(_p(case.ca), _p(case.cb), _p(case.cq), _p(f), _parr(fields), _p(r),
            _parr(case.profiles), _p(case.src), _p(case.rec), ct(case.rdz), ct(case.rdy),
            ct(case.rdx), ctypes.c_int64(nt), ctypes.c_int64(case.n_shots), ctypes.c_int64(nz),
            ctypes.c_int64(ny), ctypes.c_int64(nx), ctypes.c_int64(case.src.shape[1]),
            ctypes.c_int64(n_rec), ctypes.c_int64(interval), ctypes.c_int64(start_t),
            (ctypes.c_int64 * 6)(*case.pml), batched, ctypes.c_int(case.src_comp),
            ctypes.c_int(case.rec_comp),
            _parr(stores) if stores is not None else (ctypes.c_void_p * 6)(),
            ctypes.c_int(fmt), ctypes.c_int64(step_bytes))
    if rc != 0:
        raise RuntimeError(f"oracle 3-D forward failed ({rc})")
    return fields, r, stores


def em3d_backward(case: EMCase, grad_r: np.ndarray, stores, *, store: str = "full",
                  interval: int = 1, want_ca: bool = True, want_cb: bool = True, lam0=None,
                  start_t: int | None = None, nt: int | None = None):
    """Oracle exact 3-D adjoint over steps ``start_t-1`` down to
    ``start_t-nt`` (default: the whole run).  Returns ``(grad_ca, grad_cb,
    grad_f, lam)`` with FP64 coefficient gradients; ``lam0`` optionally gives
    the initial 18-array adjoint state (same slots as the forward state)."""
    start_t = case.nt if start_t is None else start_t
    nt = start_t if nt is None else nt
    lib = oracle_lib()
    f64, ct, batched = _em_common(case)
    shape = (case.n_shots, *case.shape)
    lam = ([np.zeros(shape, case.dtype) for _ in range(18)] if lam0 is None
           else [np.array(x, dtype=case.dtype, copy=True).reshape(shape) for x in lam0])
    n_src = case.src.shape[1]
    grad_f = np.zeros((case.nt, case.n_shots, n_src), case.dtype)
    gshape = shape if case.ca.shape[0] != 1 else case.shape
    gca = np.zeros(gshape, np.float64) if want_ca else None
    gcb = np.zeros(gshape, np.float64) if want_cb else None
    fmt = 1 if store == "bf16" else 0
    step_bytes = int(np.prod(shape)) * (2 if fmt == 1 else case.dtype.itemsize)
    fn = lib.oracle_em3d_backward_f64 if f64 else lib.oracle_em3d_backward_f32
    fn.restype = ctypes.c_int
    nz, ny, nx = case.shape
    grad_r = np.ascontiguousarray(grad_r, dtype=case.dtype)
    rc = fn(_p(case.ca), _p(case.cb), _p(case.cq), _p(grad_r), _parr(lam),
            _parr(stores) if stores is not None else (ctypes.c_void_p * 6)(),
            ctypes.c_int(fmt), ctypes.c_int64(step_bytes), _p(grad_f) if n_src else None,
            _p(gca), _p(gcb), _parr(case.profiles), _p(case.src), _p(case.rec),
            ct(case.rdz), ct(case.rdy), ct(case.rdx), ctypes.c_int64(nt),
            ctypes.c_int64(case.n_shots), ctypes.c_int64(nz), ctypes.c_int64(ny),
            ctypes.c_int64(nx), ctypes.c_int64(n_src), ctypes.c_int64(case.rec.shape[1]),
            ctypes.c_int64(interval), ctypes.c_int64(start_t), (ctypes.c_int64 * 6)(*case.pml),
            batched, ctypes.c_int(case.src_comp), ctypes.c_int(case.rec_comp))
    if rc != 0:
        raise RuntimeError(f"oracle 3-D backward failed ({rc})")
    return gca, gcb, grad_f, lam


def em3d_tangent(case: EMCase, dca: np.ndarray, dcb: np.ndarray, df: np.ndarray):
    """FP64 tangent-linear 3-D forward: returns ``(r, dr)``."""
    lib = oracle_lib()
    c = case.astype(np.float64)
    fields = [x.copy() for x in c.fields]
    dfields = [np.zeros_like(x) for x in c.fields]
    n_rec = c.rec.shape[1]
    r = np.zeros((c.nt, c.n_shots, n_rec))
    dr = np.zeros_like(r)
    nz, ny, nx = c.shape
    fn = lib.oracle_em3d_tangent_f64
    fn.restype = ctypes.c_int
    D = ctypes.c_double
    rc = fn(_p(c.ca), _p(c.cb), _p(c.cq), _p(c.f), _p(np.ascontiguousarray(dca, np.float64)),
            _p(np.ascontiguousarray(dcb, np.float64)), _p(np.ascontiguousarray(df, np.float64)),
            _parr(fields), _parr(dfields), _p(r), _p(dr), _parr(c.profiles), _p(c.src),
            _p(c.rec), D(c.rdz), D(c.rdy), D(c.rdx), ctypes.c_int64(c.nt),
            ctypes.c_int64(c.n_shots), ctypes.c_int64(nz), ctypes.c_int64(ny),
            ctypes.c_int64(nx), ctypes.c_int64(c.src.shape[1]), ctypes.c_int64(n_rec),
            (ctypes.c_int64 * 6)(*c.pml), ctypes.c_int(c.src_comp), ctypes.c_int(c.rec_comp))
    if rc != 0:
        raise RuntimeError(f"oracle 3-D tangent failed ({rc})")
    return r, dr


def ref_em3d_forward(case: EMCase, *, n_threads: int = 0, nt: int | None = None):
    """The reference's own CPU 3-D forward (``maxwell_3d_2_<dtype>_forward_cpu``,
    argument order of the reference ``backend_utils.py:487-503``).
    Returns ``(fields, r)``."""
    lib = ref_lib()
    dt_ = case.dtype
    name = "double" if dt_ == np.float64 else "float"
    fn = getattr(lib, f"maxwell_3d_2_{name}_forward_cpu")
    ct = ctypes.c_double if dt_ == np.float64 else ctypes.c_float
    I, P, B = ctypes.c_int64, ctypes.c_void_p, ctypes.c_bool
    fn.restype = None
    fn.argtypes = ([P] * 4 + [P] * 18 + [P] * 9 + [P] + [I] + [P] * 20 + [ct] * 4 + [I] * 8
                   + [B] + [B] * 3 + [I] + [I] * 6 + [I] * 2 + [I] * 3 + [P])
    nt = case.nt if nt is None else nt
    fields = [np.ascontiguousarray(x.copy()) for x in case.fields]
    n_rec = case.rec.shape[1]
    r = np.zeros((nt, case.n_shots, n_rec), dt_)
    b = case.ca.shape[0] != 1
    nz, ny, nx = case.shape
    fn(_p(case.ca), _p(case.cb), _p(case.cq), _p(case.f), *[_p(x) for x in fields],
       *([None] * 9), _p(r), 0, *[_p(p) for p in case.profiles], _p(case.src), _p(case.rec),
       case.rdz, case.rdy, case.rdx, case.dt, nt, case.n_shots, nz, ny, nx,
       case.src.shape[1], n_rec, 1, False, b, b, b, 0, *case.pml, case.src_comp,
       case.rec_comp, n_threads, 0, 0, None)
    return fields, r


def em3d_born_forward(case: EMCase, dca: np.ndarray, dcb: np.ndarray, df: np.ndarray | None = None,
                      dfields: list | None = None):
    """Oracle 3-D Born forward (reference ``maxwell_3d_inst.cpp:1100-1527``):
    background and scattered field in one time loop.  ``dca`` / ``dcb``
    ``[nz, ny, nx]``; ``df`` the scattered source ``[nt, N, n_src]`` (None:
    zero); ``dfields`` the 18 initial scattered arrays (None: zero).
    Returns ``(fields, dfields, r, background_r)`` with ``r`` the scattered
    receivers."""
    lib = oracle_lib()
    f64, ct, batched = _em_common(case)
    c = lambda a: np.ascontiguousarray(a, dtype=case.dtype)  # noqa: E731
    fields = [c(x).copy() for x in case.fields]
    dfd = [np.zeros_like(x) for x in fields] if dfields is None else [c(x).copy() for x in dfields]
    n_src, n_rec = case.src.shape[1], case.rec.shape[1]
    df = np.zeros((case.nt, case.n_shots, n_src), case.dtype) if df is None else c(df)
    r = np.zeros((case.nt, case.n_shots, n_rec), case.dtype)
    bg_r = np.zeros_like(r)
    nz, ny, nx = case.shape
    fn = lib.oracle_em3d_born_forward_f64 if f64 else lib.oracle_em3d_born_forward_f32
    fn.restype = ctypes.c_int
    rc = fn(_p(case.ca), _p(case.cb), _p(case.cq), _p(c(dca)), _p(c(dcb)), _p(case.f), _p(df),
            _parr(fields), _parr(dfd), _p(r), _p(bg_r), _parr(case.profiles), _p(case.src),
            _p(case.rec), ct(case.rdz), ct(case.rdy), ct(case.rdx), ctypes.c_int64(case.nt),
            ctypes.c_int64(case.n_shots), ctypes.c_int64(nz), ctypes.c_int64(ny),
            ctypes.c_int64(nx), ctypes.c_int64(n_src), ctypes.c_int64(n_rec), ctypes.c_int64(0),
            (ctypes.c_int64 * 6)(*case.pml), batched, ctypes.c_int(case.src_comp),
            ctypes.c_int(case.rec_comp))
    if rc != 0:
        raise RuntimeError(f"oracle 3-D Born forward failed ({rc})")
    return fields, dfd, r, bg_r


def ref_em3d_born_forward(case: EMCase, dca: np.ndarray, dcb: np.ndarray,
                          df: np.ndarray | None = None, *, n_threads: int = 0):
    """The reference's own CPU 3-D Born forward
    (``maxwell_3d_2_<dtype>_born_forward_cpu``, argument order of the
    reference ``backend_utils.py:556-577``).  Returns ``(fields, dfields, r,
    background_r)``."""
    lib = ref_lib()
    dt_ = case.dtype
    name = "double" if dt_ == np.float64 else "float"
    fn = getattr(lib, f"maxwell_3d_2_{name}_born_forward_cpu")
    ct = ctypes.c_double if dt_ == np.float64 else ctypes.c_float
    I, P, B = ctypes.c_int64, ctypes.c_void_p, ctypes.c_bool
    fn.restype = None
    fn.argtypes = ([P] * 7 + [P] * 18 + [P] * 18 + [P] * 2 + [P] * 18 + [P] * 2 + [ct] * 4
                   + [I] * 8 + [B] * 3 + [I] + [I] * 6 + [I] * 2 + [I] * 3 + [P])
    c = lambda a: np.ascontiguousarray(a, dtype=dt_)  # noqa: E731
    fields = [c(x).copy() for x in case.fields]
    dfd = [np.zeros_like(x) for x in fields]
    n_src, n_rec = case.src.shape[1], case.rec.shape[1]
    df = np.zeros((case.nt, case.n_shots, n_src), dt_) if df is None else c(df)
    r = np.zeros((case.nt, case.n_shots, n_rec), dt_)
    bg_r = np.zeros_like(r)
    b = case.ca.shape[0] != 1
    nz, ny, nx = case.shape
    fn(_p(case.ca), _p(case.cb), _p(case.cq), _p(c(dca)), _p(c(dcb)), _p(case.f), _p(df),
       *[_p(x) for x in fields], *[_p(x) for x in dfd], _p(r), _p(bg_r),
       *[_p(p) for p in case.profiles], _p(case.src), _p(case.rec),
       case.rdz, case.rdy, case.rdx, case.dt, case.nt, case.n_shots, nz, ny, nx,
       n_src, n_rec, 1, b, b, b, 0, *case.pml, case.src_comp, case.rec_comp, n_threads, 0, 0,
       None)
    return fields, dfd, r, bg_r
