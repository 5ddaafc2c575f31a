"""The HIP library loads without a GPU and exports every C-ABI entry point
declared in include/*.h (no compute calls)."""

from __future__ import annotations

import ctypes
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
    assert "tide_last_error" in names


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
