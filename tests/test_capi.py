"""CPU checks of the C-ABI: the library loads and exports every function that
include/mpdok_hip.h declares (no compute calls without a GPU)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "mpdok_hip.h")
LIB = os.path.join(ROOT, "mpdok_amd", "libmpdok_hip.so")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^\s*(?:const\s+)?[A-Za-z_][A-Za-z0-9_]*\s*\**\s*([A-Za-z_][A-Za-z0-9_]*)\s*\(",
                       text, flags=re.M)
    return sorted(set(n for n in names if n.startswith(("mpdok_", "py_"))))


def test_header_declares_reference_symbols():
    names = declared_functions()
    for ref in ["py_gmres_ir", "py_alloc_managed", "py_free_managed", "py_lu_ir",
                "py_lu_factor", "py_lu_solve_factored", "py_lu_alloc_managed",
                "py_lu_free_managed", "py_lu_free_factored"]:
        assert ref in names
    assert len(names) > 30


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_exports_every_declared_symbol():
    out = subprocess.run(["nm", "-D", "--defined-only", LIB], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (\w+)$", out, flags=re.M))
    missing = [n for n in declared_functions() if n not in exported]
    assert not missing, missing


@pytest.mark.skipif(not os.path.exists(LIB), reason="library not built")
def test_library_loads_and_ctypes_table_matches_header():
    import torch  # noqa: F401  (HIP runtime first)
    from mpdok_amd import _lib
    lib = _lib.load()
    for name in declared_functions():
        assert hasattr(lib, name), name
    assert set(_lib.SIGNATURES) >= set(declared_functions())
    assert lib.mpdok_last_error() == b""


def test_product_never_imports_oracle():
    pkg = os.path.join(ROOT, "mpdok_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith(".py"):
                src = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in re.findall(r"^\s*(?:from|import)\s+(\w+)", src, flags=re.M), f


def test_ooc_solver_host_validation():
    """MPDOKOOCSolver argument checks happen before any device work
    (mpdok_ooc.py:130-146, :342-343): bad store, missing path, solve first."""
    from mpdok_amd.mpdok_ooc import MPDOKOOCSolver
    s = MPDOKOOCSolver()
    assert s.tile_rows == 4096 and s.N is None and s._store is None
    with pytest.raises(ValueError, match="store must be"):
        s.build([[0.0, 0.0]], gamma=1.0, store="tape", verbose=False)
    with pytest.raises(ValueError, match="requires a path"):
        s.build([[0.0, 0.0]], gamma=1.0, store="ssd", path=None, verbose=False)
    with pytest.raises(RuntimeError, match="build"):
        s.solve([1.0])
