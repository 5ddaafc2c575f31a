"""The HIP library loads without a GPU and exports every C-ABI entry point
declared in include/*.h (no compute calls)."""

from __future__ import annotations

import ctypes
import json
import re
from pathlib import Path

import pytest

from tide import _native

INCLUDE = Path(__file__).resolve().parents[1] / "include"


def _declared_symbols() -> list[str]:
    names = []
    for header in sorted(INCLUDE.glob("*.h")):
        text = re.sub(r"/\*.*?\*/", "", header.read_text(), flags=re.S)
        names += re.findall(r"^\s*(?:const\s+)?\w+\s*\*?\s*(\w+)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in {"if", "while"}))


def test_header_declares_reference_entry_points():
    names = _declared_symbols()
    for d in ("float", "double"):
        for p in ("forward", "forward_with_storage", "backward"):
            assert f"maxwell_tm_2_{d}_{p}_cuda" in names
            assert f"maxwell_3d_2_{d}_{p}_cuda" in names
    assert "tide_last_error" in names


SPEC = json.loads((Path(__file__).resolve().parent / "golden" / "abi_spec.json").read_text())
REF_ENTRY = [(prop, d, p) for prop in ("tm", "3d") for d in ("float", "double")
             for p in ("forward", "forward_with_storage", "backward")]
REF_ENTRY += [("tm", d, p) for d in ("float", "double")
              for p in ("born_forward", "born_forward_with_storage", "born_backward_bggrad")]
REF_ENTRY += [("3d", d, "born_forward") for d in ("float", "double")]


def _prototype(name: str) -> list[tuple[str, str]]:
    """(kind, parameter name) of each parameter of ``name`` in the header:
    P pointer, F float/double value, I int64_t, B bool."""
    text = re.sub(r"/\*.*?\*/", "", (INCLUDE / "tide_hip.h").read_text(), flags=re.S)
    m = re.search(rf"\b{name}\s*\((.*?)\);", text, flags=re.S)
    assert m, name
    out = []
    for param in " ".join(m.group(1).split()).split(","):
        param = param.strip()
        pname = re.findall(r"(\w+)$", param)[0]
        if "*" in param:
            kind = "P"
        elif param.startswith(("float", "double")):
            kind = "F"
        elif param.startswith("int64_t"):
            kind = "I"
        elif param.startswith("bool"):
            kind = "B"
        else:
            raise AssertionError(f"{name}: unexpected parameter {param!r}")
        out.append((kind, pname))
    return out


def _spec_key(prop: str, p: str) -> str:
    return f"maxwell_{prop}_{p}"


@pytest.mark.parametrize("prop,dtype,pass_name", REF_ENTRY)
def test_prototypes_follow_the_reference_spec(prop, dtype, pass_name):
    """Each reference-compatible prototype in include/tide_hip.h has the
    parameters of the reference's ctypes spec (tests/golden/abi_spec.json,
    written from src/tide/backend_utils.py): same kinds, same names, same
    order."""
    proto = _prototype(f"maxwell_{prop}_2_{dtype}_{pass_name}_cuda")
    spec = [tuple(e.split(" ", 1)) for e in SPEC[_spec_key(prop, pass_name)]]
    assert proto == spec


@pytest.mark.parametrize("prop,dtype,pass_name", REF_ENTRY)
def test_ctypes_bindings_follow_the_reference_spec(prop, dtype, pass_name):
    """The argtypes bound by tide._native match the same spec."""
    lib = _native.load()
    fn = getattr(lib, f"maxwell_{prop}_2_{dtype}_{pass_name}_cuda")
    f = ctypes.c_float if dtype == "float" else ctypes.c_double
    kinds = {ctypes.c_void_p: "P", ctypes.c_int64: "I", ctypes.c_bool: "B", f: "F"}
    got = [kinds[t] for t in fn.argtypes]
    assert got == [e.split(" ", 1)[0] for e in SPEC[_spec_key(prop, pass_name)]]
    assert fn.restype is ctypes.c_int


def _default_args(prop: str, dtype: str, pass_name: str, **named):
    spec = [tuple(e.split(" ", 1)) for e in SPEC[_spec_key(prop, pass_name)]]
    default = {"P": None, "F": 0.0, "I": 0, "B": False}
    args = [named.pop(n, default[k]) for k, n in spec]
    assert not named, named
    return args


@pytest.mark.parametrize("prop", ["tm", "3d"])
def test_backward_rejects_mixed_coefficient_batching(prop):
    """A per-shot ca with a shared cb (both gradients wanted) would index
    the shared gradient per shot; the entry point refuses it before any
    device call."""
    lib = _native.load()
    fn = getattr(lib, f"maxwell_{prop}_2_float_backward_cuda")
    args = _default_args(prop, "float", "backward", ca_requires_grad=True,
                         cb_requires_grad=True, ca_batched=True, cb_batched=False, n_shots=2,
                         nt=4)
    assert fn(*args) == 1
    assert b"same batching" in lib.tide_last_error()


def test_library_exports_every_declared_symbol():
    lib = _native.load()
    missing = [n for n in _declared_symbols() if not hasattr(lib, n)]
    assert not missing, missing


def test_last_error_is_a_string():
    lib = _native.load()
    assert isinstance(lib.tide_last_error(), bytes)


def test_missing_library_raises(monkeypatch, tmp_path):
    monkeypatch.setattr(_native, "_lib", None)
    monkeypatch.setenv("TIDE_HIP_LIBRARY", str(tmp_path / "absent.so"))
    with pytest.raises(_native.TideNativeError):
        _native.load()
    monkeypatch.setattr(_native, "_lib", None)


def test_launch_shape_setters_round_trip_and_reject_bad_values():
    """Host-only calls: the launch-shape settings round-trip, and values the
    kernels do not support are refused with a message (state unchanged)."""
    before = (_native.get_tm2d_time_block(), _native.get_tm2d_tile_split(),
              _native.get_tm2d_border_rows())
    try:
        _native.set_tm2d_border_rows(1, 0)
        assert _native.get_tm2d_border_rows() == (1, 0)
        for bad in ((3, 0), (0, -1)):
            with pytest.raises(_native.TideNativeError, match="border tile height"):
                _native.set_tm2d_border_rows(*bad)
        assert _native.get_tm2d_border_rows() == (1, 0)
        _native.set_tm2d_time_block(2, 1)
        assert _native.get_tm2d_time_block() == (2, 1)
        _native.set_tm2d_tile_split(True, False)
        assert _native.get_tm2d_tile_split() == (True, False)
    finally:
        _native.set_tm2d_time_block(*before[0])
        _native.set_tm2d_tile_split(*before[1])
        _native.set_tm2d_border_rows(*before[2])


def test_em3d_forward_group_setter_rejects_bad_values():
    """Host-only: the 3-D forward group takes 1..8 shots per workgroup."""
    try:
        for bad in (0, 9, -1):
            with pytest.raises(_native.TideNativeError, match="forward group"):
                _native.set_em3d_forward_group(bad)
        _native.set_em3d_forward_group(2)
    finally:
        _native.set_em3d_forward_group(1)
