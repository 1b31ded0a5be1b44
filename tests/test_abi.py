"""The C-ABI library builds, loads, and exports every symbol include/tf_nr.h declares (no GPU)."""
import ctypes
import os
import re

from tissuefolding_paperversion_amd import _lib as L

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    src = open(os.path.join(ROOT, "include", "tf_nr.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(tf_[a-z_0-9]+)\s*\(", src)))


def test_header_matches_binding_list():
    assert header_symbols() == sorted(L.EXPORTS)


def test_library_exports_every_declared_symbol():
    lib = L.load()
    for name in header_symbols():
        assert hasattr(lib, name), name


def test_default_options_without_gpu():
    lib = L.load()
    o = L.SolverOpts()
    lib.tf_default_options(ctypes.byref(o))
    assert o.nr_tol == 1e-8 and o.nr_max_iter == 20 and o.krylov_rtol == 1e-12


def test_create_rejects_invalid_mesh():
    lib = L.load()
    d = L.MeshDesc()
    d.n_nodes = 0
    h = ctypes.c_void_p()
    assert lib.tf_create(ctypes.byref(d), ctypes.byref(h)) == L.TF_INVALID
    assert b"n_nodes" in lib.tf_last_error(None)
