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


def test_default_launch_reads_the_tuning_file(monkeypatch, tmp_path):
    from monoprop_amd import _engine

    f = tmp_path / "tuning.json"
    f.write_text('{"best": {"resolve_grid": 2048, "min_group_log2": 5, "ms": 1.0}}')
    monkeypatch.setattr(_engine, "TUNING_FILE", f)
    monkeypatch.setattr(_engine, "_LAUNCH", None)
    monkeypatch.setattr(_engine, "_LAUNCH_READ", False)
    assert _engine.default_launch() == (2048, 5)
    _engine.set_default_launch(None)
    assert _engine.default_launch() is None
    monkeypatch.setattr(_engine, "_LAUNCH_READ", False)
    f.write_text("not json")
    assert _engine.default_launch() is None
