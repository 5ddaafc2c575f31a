"""ctypes binding to the MI355X HIP library (``_lib/libtide_hip.so``).

The entry points keep the reference ABI: symbol names
``maxwell_tm_{stencil}_{dtype}_{pass}_cuda`` and argument order of the
reference's ctypes specs (reference ``src/tide/backend_utils.py:227-308``,
symbol naming ``:771-781``).  They return an ``int`` status instead of
``void``; a nonzero status is raised here as :class:`TideNativeError` with
the library's last error message.  There is no fallback: a missing library
is an error.
"""

from __future__ import annotations

import ctypes
import os
from pathlib import Path
from typing import Any

import torch

_LIB_PATH = Path(__file__).resolve().parent / "_lib" / "libtide_hip.so"
_lib: ctypes.CDLL | None = None

_P = ctypes.c_void_p
_I = ctypes.c_int64
_B = ctypes.c_bool


class TideNativeError(RuntimeError):
    """A HIP entry point returned a failure status."""


def library_path() -> Path:
    return Path(os.environ.get("TIDE_HIP_LIBRARY", _LIB_PATH))


def _tm_forward_types(f: Any) -> list[Any]:
    return [
        *([_P] * 3),  # ca, cb, cq
        _P,  # f
        *([_P] * 3),  # ey, hx, hz
        *([_P] * 4),  # m_ey_x, m_ey_z, m_hx_z, m_hz_x
        *([_P] * 5),  # debye_a, debye_b, debye_cp, polarization, ey_prev
        _P,  # r
        _I,  # n_poles
        *([_P] * 12),  # ay, by, ayh, byh, ax, bx, axh, bxh, ky, kyh, kx, kxh
        _P,
        _P,  # sources_i, receivers_i
        f,
        f,
        f,  # rdy, rdx, dt
        *([_I] * 7),  # nt, n_shots, ny, nx, n_src, n_rec, step_ratio
        _B,  # has_dispersion
        *([_B] * 3),  # ca_batched, cb_batched, cq_batched
        _I,  # start_t
        *([_I] * 4),  # pml_y0, pml_x0, pml_y1, pml_x1
        _I,
        _I,  # n_threads, device
        _P,  # compute_stream
    ]


def _tm_forward_storage_types(f: Any) -> list[Any]:
    return [
        *([_P] * 3),  # ca, cb, cq
        _P,  # f
        *([_P] * 3),  # ey, hx, hz
        *([_P] * 4),  # m_ey_x, m_ey_z, m_hx_z, m_hz_x
        _P,  # r
        *([_P] * 6),  # ey_store_1, ey_store_3, ey_filenames, curl_store_1/3, curl_filenames
        *([_P] * 12),
        _P,
        _P,
        f,
        f,
        f,
        *([_I] * 7),
        *([_I] * 3),  # storage_mode, storage_format, shot_bytes_uncomp
        _B,
        _B,  # ca_requires_grad, cb_requires_grad
        *([_B] * 3),
        _I,
        *([_I] * 4),
        _I,
        _I,
        _P,
        _P,  # compute_stream, storage_stream
    ]


def _tm_backward_types(f: Any) -> list[Any]:
    return [
        *([_P] * 3),  # ca, cb, cq
        _P,  # grad_r
        *([_P] * 3),  # lambda_ey, lambda_hx, lambda_hz
        *([_P] * 4),  # m_lambda_*
        _P,  # adjoint_memory_scratch
        *([_P] * 6),  # storage pointers / filenames
        _P,  # grad_f
        _P,
        _P,  # grad_ca, grad_cb
        _P,
        _P,  # grad_ca_shot, grad_cb_shot
        *([_P] * 12),
        _P,
        _P,
        f,
        f,
        f,
        *([_I] * 7),
        *([_I] * 3),
        _B,
        _B,
        *([_B] * 3),
        _I,
        *([_I] * 4),
        _I,
        _I,
        _P,
        _P,
    ]


def _tm_backward_f64acc_types(f: Any) -> list[Any]:
    return [
        *([_P] * 3),  # ca, cb, cq
        _P,  # grad_r
        *([_P] * 3),  # lambda fields
        *([_P] * 4),  # lambda memories
        *([_P] * 6),  # ey_store_1, ey_store_3, ey_filenames, curl_store_1/_3, curl_filenames
        _P,  # grad_f
        _P,
        _P,  # grad_ca (fp64), grad_cb (fp64)
        *([_P] * 12),
        _P,
        _P,
        f,
        f,  # rdy, rdx
        *([_I] * 7),  # nt, n_shots, ny, nx, n_src, n_rec, step_ratio
        _I,  # storage_mode
        _I,
        _I,  # storage_format, shot_bytes_uncomp
        *([_B] * 3),
        _I,  # start_t
        *([_I] * 4),
        _I,  # stencil
        _I,  # device
        _P,  # compute_stream
        _P,  # storage_stream
    ]


def _em3d_forward_types(f: Any) -> list[Any]:
    return [
        *([_P] * 4),  # ca, cb, cq, f
        _P,  # fields: void*[18]
        _P,  # r
        _P,  # profiles: void*[18]
        _P, _P,  # sources_i, receivers_i
        f, f, f,  # rdz, rdy, rdx
        *([_I] * 8),  # nt, n_shots, nz, ny, nx, n_src, n_rec, step_ratio
        *([_B] * 3),  # ca/cb/cq batched
        _I,  # start_t
        _P,  # pml: int64[6]
        _I, _I,  # source_component, receiver_component
        _P,  # stores: void*[6]
        _I, _I, _I,  # storage_format, store_step_bytes, store_t0
        _I,  # device
        _P,  # compute_stream
    ]


def _em3d_backward_types(f: Any) -> list[Any]:
    return [
        *([_P] * 4),  # ca, cb, cq, grad_r
        _P,  # lambda state: void*[18]
        _P,  # stores: void*[6]
        _I, _I, _I,  # storage_format, store_step_bytes, store_t0
        _P, _P, _P,  # grad_f, grad_ca (fp64), grad_cb (fp64)
        _P,  # profiles
        _P, _P,  # sources_i, receivers_i
        f, f, f,
        *([_I] * 8),
        *([_B] * 3),
        _I,  # start_t
        _P,  # pml
        _I, _I,  # components
        _I,  # device
        _P,  # compute_stream
    ]


def _expand_kinds(spec: str, f: Any) -> list[Any]:
    """``"P4 F2 I1"`` -> four pointers, two field floats, one int64."""
    kinds = {"P": _P, "I": _I, "B": _B, "F": f}
    out: list[Any] = []
    for tok in spec.split():
        out += [kinds[tok[0]]] * int(tok[1:])
    return out


# Argument kinds of the reference-order 3-D entry points, in the order of
# the reference's ctypes specs (src/tide/backend_utils.py:446-555).
_REF3_KINDS = {
    # ca cb cq f | 18 state | 9 Debye | r | n_poles | 18 profiles | src rec |
    # rdz rdy rdx dt | nt n_shots nz ny nx n_src n_rec step_ratio |
    # has_dispersion | 3 batched | start_t | 6 pml | 2 components |
    # n_threads device execution_backend | stream
    "forward": "P4 P18 P9 P1 I1 P18 P2 F4 I8 B1 B3 I1 I6 I2 I1 I1 I1 P1",
    # ca cb cq f | 18 state | r | 18 storage | 18 profiles | src rec | F4 | I8 |
    # storage_mode format shot_bytes | ca/cb requires_grad | flags | 2 streams
    "forward_with_storage": "P4 P18 P1 P18 P18 P2 F4 I8 I3 B2 B3 I1 I6 I2 I1 I1 I1 P2",
    # ca cb cq grad_r | 18 lambda | scratch | 18 storage | grad_f | grad_ca
    # grad_cb grad_eps grad_sigma | 2 per-shot | zero_grad_on_entry | ...
    "backward": "P4 P18 P1 P18 P1 P4 P2 B1 P18 P2 F4 I8 I3 B2 B3 I1 I6 I2 I1 I1 I1 P2",
    # ca cb cq dca dcb f0 df | 18 state | 18 scattered state | r background_r |
    # 18 profiles | src rec | F4 | I8 | 3 batched | start_t | 6 pml |
    # 2 components | n_threads device execution_backend | stream
    # (src/tide/backend_utils.py:556-577)
    "born_forward": "P7 P18 P18 P2 P18 P2 F4 I8 B3 I1 I6 I2 I1 I1 I1 P1",
}


def load() -> ctypes.CDLL:
    """Load the HIP library, raising if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    path = library_path()
    if not path.exists():
        raise TideNativeError(
            f"HIP library {path} is missing; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'`."
        )
    lib = ctypes.CDLL(str(path))
    lib.tide_last_error.restype = ctypes.c_char_p
    lib.tide_last_error.argtypes = []
    for name, ctype in (("float", ctypes.c_float), ("double", ctypes.c_double)):
        for pass_name, types in (
            ("forward", _tm_forward_types(ctype)),
            ("forward_with_storage", _tm_forward_storage_types(ctype)),
            ("backward", _tm_backward_types(ctype)),
        ):
            for stencil in (2, 4, 6, 8):
                fn = getattr(lib, f"maxwell_tm_{stencil}_{name}_{pass_name}_cuda")
                fn.restype = ctypes.c_int
                fn.argtypes = types
        fn = getattr(lib, f"tide_tm2d_{name}_backward_f64acc")
        fn.restype = ctypes.c_int
        fn.argtypes = _tm_backward_f64acc_types(ctype)
    # Born forward, reference _TM_BORN_FORWARD_SPEC (backend_utils.py:336-349):
    # 23 field/coefficient pointers, 12 profiles, 2 point arrays | rdy rdx dt
    # | nt n_shots ny nx n_src n_rec step_ratio | 3 batched | start_t |
    # 4 pml | n_threads | device | stream
    for name, ctype in (("float", ctypes.c_float), ("double", ctypes.c_double)):
        for stencil in (2, 4, 6, 8):
            fn = getattr(lib, f"maxwell_tm_{stencil}_{name}_born_forward_cuda")
            fn.restype = ctypes.c_int
            fn.argtypes = _expand_kinds("P37 F3 I7 B3 I1 I4 I1 I1 P1", ctype)
            # _TM_BORN_FORWARD_WITH_STORAGE_SPEC (backend_utils.py:351-370): the
            # same plus 8 store pointers | storage mode, format, shot bytes |
            # ca/cb requires_grad | storage stream
            fn = getattr(lib, f"maxwell_tm_{stencil}_{name}_born_forward_with_storage_cuda")
            fn.restype = ctypes.c_int
            fn.argtypes = _expand_kinds("P45 F3 I10 B5 I7 P2", ctype)
            # _TM_BORN_BACKWARD_BGGRAD_SPEC (backend_utils.py:407-444)
            fn = getattr(lib, f"maxwell_tm_{stencil}_{name}_born_backward_bggrad_cuda")
            fn.restype = ctypes.c_int
            fn.argtypes = _expand_kinds("P18 I1 P42 F3 I10 B7 I7 P2", ctype)
            # this library's Born adjoint: 27 field pointers, 4 FP64
            # accumulators, 12 profiles, 2 point arrays | rdy rdx | nt n_shots ny
            # nx n_src n_rec start_t | 4 pml | device | stream
            fn = getattr(lib, f"tide_tm2d_{stencil}_{name}_born_backward_f64acc")
            fn.restype = ctypes.c_int
            fn.argtypes = _expand_kinds("P45 F2 I12 P1", ctype)
    # the allocation-free variant of maxwell_tm_2_float_forward_cuda
    lib.tide_tm2d_float_forward_noalloc.restype = ctypes.c_int
    lib.tide_tm2d_float_forward_noalloc.argtypes = _tm_forward_types(ctypes.c_float)
    for name, ctype in (("float", ctypes.c_float), ("double", ctypes.c_double)):
        fn = getattr(lib, f"tide_em3d_{name}_forward")
        fn.restype = ctypes.c_int
        fn.argtypes = _em3d_forward_types(ctype)
        fn = getattr(lib, f"tide_em3d_{name}_backward_f64acc")
        fn.restype = ctypes.c_int
        fn.argtypes = _em3d_backward_types(ctype)
        # Born forward: 7 coefficient / source pointers, fields, dfields,
        # r, background_r, profiles, 2 point arrays | rdz rdy rdx | nt
        # n_shots nz ny nx n_src n_rec | 3 batched | start_t pml | 2
        # components device | stream
        fn = getattr(lib, f"tide_em3d_{name}_born_forward")
        fn.restype = ctypes.c_int
        fn.argtypes = _expand_kinds("P14 F3 I7 B3 I1 P1 I3 P1", ctype)
    for name, ctype in (("float", ctypes.c_float), ("double", ctypes.c_double)):
        for pass_name, kinds in _REF3_KINDS.items():
            fn = getattr(lib, f"maxwell_3d_2_{name}_{pass_name}_cuda")
            fn.restype = ctypes.c_int
            fn.argtypes = _expand_kinds(kinds, ctype)
    lib.tide_em3d_set_tuning.restype = ctypes.c_int
    lib.tide_em3d_set_tuning.argtypes = [ctypes.c_int] * 3
    lib.tide_em3d_set_forward_group.restype = ctypes.c_int
    lib.tide_em3d_set_forward_group.argtypes = [ctypes.c_int]
    lib.tide_em3d_set_fused.restype = ctypes.c_int
    lib.tide_em3d_set_fused.argtypes = [ctypes.c_int] * 2
    lib.tide_em3d_get_fused.restype = ctypes.c_int
    lib.tide_em3d_set_tile_order.restype = ctypes.c_int
    lib.tide_em3d_set_tile_order.argtypes = [ctypes.c_int]
    lib.tide_em3d_get_fused.argtypes = [ctypes.POINTER(ctypes.c_int)]
    lib.tide_tm2d_set_tuning.restype = ctypes.c_int
    lib.tide_tm2d_set_tuning.argtypes = [ctypes.c_int] * 4
    lib.tide_tm2d_get_tuning.restype = ctypes.c_int
    lib.tide_tm2d_get_tuning.argtypes = [ctypes.POINTER(ctypes.c_int)]
    lib.tide_tm2d_set_shot_loop.restype = ctypes.c_int
    lib.tide_tm2d_set_shot_loop.argtypes = [ctypes.c_int] * 2
    lib.tide_tm2d_get_shot_loop.restype = ctypes.c_int
    lib.tide_tm2d_get_shot_loop.argtypes = [ctypes.POINTER(ctypes.c_int)]
    lib.tide_tm2d_set_time_block.restype = ctypes.c_int
    lib.tide_tm2d_set_time_block.argtypes = [ctypes.c_int] * 2
    lib.tide_tm2d_get_time_block.restype = ctypes.c_int
    lib.tide_tm2d_get_time_block.argtypes = [ctypes.POINTER(ctypes.c_int)]
    lib.tide_tm2d_set_tile_split.restype = ctypes.c_int
    lib.tide_tm2d_set_tile_split.argtypes = [ctypes.c_int] * 2
    lib.tide_tm2d_get_tile_split.restype = ctypes.c_int
    lib.tide_tm2d_get_tile_split.argtypes = [ctypes.POINTER(ctypes.c_int)]
    lib.tide_tm2d_set_loop_timing.restype = ctypes.c_int
    lib.tide_tm2d_set_loop_timing.argtypes = [ctypes.c_int]
    lib.tide_tm2d_take_loop_timing.restype = ctypes.c_int
    lib.tide_tm2d_take_loop_timing.argtypes = [
        ctypes.c_int, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int64),
        ctypes.POINTER(ctypes.c_int64)]
    lib.tide_tm2d_set_adjoint_waves.restype = ctypes.c_int
    lib.tide_tm2d_set_adjoint_waves.argtypes = [ctypes.c_int]
    lib.tide_tm2d_get_adjoint_waves.restype = ctypes.c_int
    lib.tide_tm2d_get_adjoint_waves.argtypes = [ctypes.POINTER(ctypes.c_int)]
    lib.tide_tm2d_set_border_rows.restype = ctypes.c_int
    lib.tide_tm2d_set_border_rows.argtypes = [ctypes.c_int, ctypes.c_int]
    lib.tide_tm2d_get_border_rows.restype = ctypes.c_int
    lib.tide_tm2d_get_border_rows.argtypes = [ctypes.POINTER(ctypes.c_int)]
    _lib = lib
    return lib


def check(status: int) -> None:
    if status != 0:
        msg = load().tide_last_error().decode(errors="replace")
        raise TideNativeError(f"HIP propagator failed (status {status}): {msg}")


OUT_OF_MEMORY = 4  # kOutOfMemory of the C ABI


def call(fn, *args) -> None:
    """Runs a propagator entry point and raises on a failed status.

    The library allocates its scratch from HIP's stream-ordered pool, which
    cannot use blocks that torch's caching allocator holds unused: on an
    out-of-memory status the call runs once more after those blocks are
    released.  The entry points allocate all scratch before they write any
    caller array, so the retried call starts from the same state."""
    status = fn(*args)
    if (status == OUT_OF_MEMORY and torch.cuda.is_initialized()
            and torch.cuda.memory_reserved() > torch.cuda.memory_allocated()):
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
        status = fn(*args)
    check(status)


def dtype_name(dtype: torch.dtype) -> str:
    if dtype == torch.float32:
        return "float"
    if dtype == torch.float64:
        return "double"
    raise TypeError(f"unsupported field dtype {dtype}; use float32 or float64")


def get_function(propagator: str, pass_name: str, stencil: int, dtype: torch.dtype):
    """Resolve ``{propagator}_{stencil}_{dtype}_{pass}_cuda``."""
    lib = load()
    name = f"{propagator}_{stencil}_{dtype_name(dtype)}_{pass_name}_cuda"
    try:
        return getattr(lib, name)
    except AttributeError as exc:
        raise TideNativeError(
            f"{name} is not built; stencil {stencil} is not part of this build"
        ) from exc


def ptr(t: torch.Tensor | None) -> int | None:
    """Raw pointer of a contiguous tensor (None for absent/empty)."""
    if t is None or t.numel() == 0:
        return None
    if not t.is_contiguous():
        raise ValueError("native arguments must be contiguous tensors")
    return t.data_ptr()


def set_tm2d_tuning(fwd_rows: int, fwd_group: int, adj_rows: int, adj_group: int) -> None:
    """Launch shape of the TM2D step kernels (see include/tide_hip.h)."""
    check(load().tide_tm2d_set_tuning(fwd_rows, fwd_group, adj_rows, adj_group))


def set_em3d_tuning(fwd_z: int, adj_z: int, group: int) -> None:
    """z planes marched per wavefront (forward, adjoint) and shots per
    workgroup of the 3-D kernels (see include/tide_hip.h)."""
    check(load().tide_em3d_set_tuning(fwd_z, adj_z, group))


def set_em3d_forward_group(group: int) -> None:
    """Shots per workgroup of the 3-D forward kernels alone (see include/tide_hip.h)."""
    check(load().tide_em3d_set_forward_group(int(group)))


def set_em3d_fused(rows: int, planes: int = 8) -> None:
    """One-launch 3-D forward step: rows per wavefront (2 or 4; 0 runs the
    two-kernel step) and z planes per march (8 or 16)."""
    check(load().tide_em3d_set_fused(int(rows), int(planes)))


def set_tm2d_adjoint_waves(waves: int) -> None:
    """Register budget of the fp32 K-step adjoint in waves per SIMD: 4 (128
    VGPRs, the default) or 0 (the compiler's count)."""
    check(load().tide_tm2d_set_adjoint_waves(int(waves)))


def get_tm2d_adjoint_waves() -> int:
    out = ctypes.c_int()
    check(load().tide_tm2d_get_adjoint_waves(ctypes.byref(out)))
    return int(out.value)


def set_em3d_tile_order(shot_fast: int) -> None:
    """Workgroup order of the 3-D kernels: 0 shot groups outermost, 1 the
    shot groups of one tile adjacent."""
    check(load().tide_em3d_set_tile_order(int(shot_fast)))


def get_em3d_fused() -> tuple[int, int]:
    """(rows, planes) of the one-launch 3-D forward step (rows 0: off)."""
    out = (ctypes.c_int * 2)()
    check(load().tide_em3d_get_fused(out))
    return int(out[0]), int(out[1])


def ptr_array(tensors) -> ctypes.Array:
    """``void*[n]`` of tensor pointers (None for absent entries)."""
    return (ctypes.c_void_p * len(tensors))(*[ptr(t) for t in tensors])


def int64_array(values) -> ctypes.Array:
    return (ctypes.c_int64 * len(values))(*[int(v) for v in values])


def get_tm2d_tuning() -> tuple[int, int, int, int]:
    out = (ctypes.c_int * 4)()
    check(load().tide_tm2d_get_tuning(out))
    return tuple(out)  # type: ignore[return-value]


def set_tm2d_shot_loop(fwd_loop: bool, adj_loop: bool) -> None:
    """Select the shot-looping TM2D kernels per pass (see include/tide_hip.h)."""
    check(load().tide_tm2d_set_shot_loop(int(bool(fwd_loop)), int(bool(adj_loop))))


def get_tm2d_shot_loop() -> tuple[bool, bool]:
    out = (ctypes.c_int * 2)()
    check(load().tide_tm2d_get_shot_loop(out))
    return bool(out[0]), bool(out[1])


def set_tm2d_time_block(fwd_steps: int, adj_steps: int) -> None:
    """Time steps per launch (1 to 4) of the TM2D kernels per pass (see include/tide_hip.h)."""
    check(load().tide_tm2d_set_time_block(int(fwd_steps), int(adj_steps)))


def get_tm2d_time_block() -> tuple[int, int]:
    out = (ctypes.c_int * 2)()
    check(load().tide_tm2d_get_time_block(out))
    return int(out[0]), int(out[1])


def set_tm2d_tile_split(fwd_split: bool, adj_split: bool) -> None:
    """Interior / border tiles as separate two-step kernels, per pass (see include/tide_hip.h)."""
    check(load().tide_tm2d_set_tile_split(int(bool(fwd_split)), int(bool(adj_split))))


def get_tm2d_tile_split() -> tuple[bool, bool]:
    out = (ctypes.c_int * 2)()
    check(load().tide_tm2d_get_tile_split(out))
    return bool(out[0]), bool(out[1])


def set_tm2d_border_rows(fwd_rows: int, adj_rows: int) -> None:
    """Height of the two-step border tiles per pass: 1, 2 or 0 (the interior height)."""
    check(load().tide_tm2d_set_border_rows(int(fwd_rows), int(adj_rows)))


def get_tm2d_border_rows() -> tuple[int, int]:
    out = (ctypes.c_int * 2)()
    check(load().tide_tm2d_get_border_rows(out))
    return int(out[0]), int(out[1])


def set_tm2d_loop_timing(enable: bool) -> None:
    """HIP-event timing of every TM2D time loop (see include/tide_hip.h)."""
    check(load().tide_tm2d_set_loop_timing(int(bool(enable))))


def take_tm2d_loop_timing(pass_: int) -> tuple[float, int, int]:
    """(milliseconds, launches, loops) recorded for pass 0 (forward) or 1
    (adjoint) since the last call."""
    ms = ctypes.c_double()
    launches = ctypes.c_int64()
    loops = ctypes.c_int64()
    check(load().tide_tm2d_take_loop_timing(int(pass_), ctypes.byref(ms), ctypes.byref(launches),
                                            ctypes.byref(loops)))
    return float(ms.value), int(launches.value), int(loops.value)


def stream_handle(device: torch.device) -> int | None:
    """The HIP stream of torch's current stream on ``device``."""
    return torch.cuda.current_stream(device).cuda_stream


__all__ = [
    "OUT_OF_MEMORY",
    "TideNativeError",
    "call",
    "check",
    "dtype_name",
    "get_em3d_fused",
    "get_tm2d_adjoint_waves",
    "get_tm2d_border_rows",
    "get_tm2d_shot_loop",
    "get_tm2d_tile_split",
    "get_tm2d_time_block",
    "get_tm2d_tuning",
    "int64_array",
    "ptr_array",
    "set_em3d_tuning",
    "set_em3d_forward_group",
    "set_tm2d_border_rows",
    "set_tm2d_loop_timing",
    "set_tm2d_shot_loop",
    "set_tm2d_tile_split",
    "set_em3d_fused",
    "set_tm2d_adjoint_waves",
    "set_em3d_tile_order",
    "set_tm2d_time_block",
    "set_tm2d_tuning",
    "get_function",
    "library_path",
    "load",
    "ptr",
    "stream_handle",
    "take_tm2d_loop_timing",
]
