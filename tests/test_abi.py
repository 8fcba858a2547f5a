"""The C-ABI library loads and exports every entry point include/monoprop_amd.h declares (no GPU calls)."""

from __future__ import annotations

import ctypes
import re
from pathlib import Path

import pytest

from monoprop_amd import _engine

HEADER = Path(__file__).resolve().parents[1] / "include" / "monoprop_amd.h"


def declared_symbols() -> list[str]:
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(mpa_[a-z_]+)\s*\(", text)))


def test_header_declares_the_binding_table():
    assert set(declared_symbols()) == set(_engine.EXPORTED_SYMBOLS)


def test_library_exports_every_symbol():
    path = _engine.library_path()
    if not path.exists():
        pytest.skip("engine library not built")
    lib = ctypes.CDLL(str(path))
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_missing_library_fails_loudly(monkeypatch, tmp_path):
    monkeypatch.setattr(_engine, "_LIB", None)
    monkeypatch.setattr(_engine, "_LIB_PATH", tmp_path / "absent.so")
    with pytest.raises(ImportError):
        _engine.load_library()
