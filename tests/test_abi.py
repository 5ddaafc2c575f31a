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
