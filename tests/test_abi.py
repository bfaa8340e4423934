"""The C-ABI library loads and exports every symbol include/ssx.h declares (no GPU calls)."""

import re
from pathlib import Path

import pytest

from slope_stability_petsc4py_amd import _lib

ROOT = Path(__file__).resolve().parents[1]


def _declared():
    text = (ROOT / "include" / "ssx.h").read_text()
    return sorted(set(re.findall(r"\b(ssx_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = _declared()
    assert "ssx_constitutive_3d" in names and "ssx_tangent_rows" in names and "ssx_spmv_f64" in names


def test_library_exports_every_declared_symbol():
    if not _lib.lib_path().exists():
        pytest.skip("libssx.so not built (run __graft_entry__.build())")
    lib = _lib.load()
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    assert lib.ssx_version() >= 1
    # every ctypes signature the Python side binds is declared in the header
    assert set(_lib.exported_symbols()) <= set(_declared())
