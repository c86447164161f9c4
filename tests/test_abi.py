"""The C ABI library loads and exports every entry point include/pgpu.h declares.

No compute call is made: this runs on hosts without a GPU.
"""
import ctypes
import os
import re

from proteusgpu_amd import _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "pgpu.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(pg_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared()
    assert "pg_mesh_compute" in names and "pg_sim_step" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_native.LIB_PATH), "libpgpu.so not built"
    lib = ctypes.CDLL(_native.LIB_PATH)
    missing = [n for n in declared() if not hasattr(lib, n)]
    assert not missing, f"not exported: {missing}"


def test_binding_covers_header():
    assert set(_native.exported_symbols()) == set(declared())


def test_version_string():
    assert _native.lib().pg_version().decode().startswith("pgpu")
