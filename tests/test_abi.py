"""The C-ABI library loads and exports every symbol include/cs_afqmc_amd.h declares
(no compute calls: this runs without a GPU)."""

import ctypes
import os
import re

from cs_afqmc_amd import _build, _lib

HEADER = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include",
                      "cs_afqmc_amd.h")


def declared():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"AFQMC_API\s+[\w\s\*]+?\b(afqmc_\w+)\s*\(", text)))


def test_header_and_binding_agree():
    assert declared() == sorted(_lib.EXPORTS)


def test_library_exports_every_declared_symbol():
    assert os.path.exists(_build.LIB_PATH), "build first: __graft_entry__.build()"
    lib = ctypes.CDLL(_build.LIB_PATH)
    for name in declared():
        assert hasattr(lib, name), name


def test_error_path_without_gpu_work():
    L = _lib.lib()
    h = ctypes.c_void_p()
    rc = L.afqmc_ctx_create(ctypes.byref(h), None, None, None)
    assert rc == 1
    assert b"null" in L.afqmc_last_error()
    dims = _lib.Dims(7, 10, 2, 2, 5, 4, 6, 0, 0.01)  # bad walker_type
    ham = _lib.HamHost()
    rc = L.afqmc_ctx_create(ctypes.byref(h), ctypes.byref(dims), ctypes.byref(ham), None)
    assert rc == 1 and b"walker_type" in L.afqmc_last_error()
    dims = _lib.Dims(1, 300, 2, 2, 5, 4, 6, 0, 0.01)  # norb above the kernels' limit
    rc = L.afqmc_ctx_create(ctypes.byref(h), ctypes.byref(dims), ctypes.byref(ham), None)
    assert rc == 3 and b"norb" in L.afqmc_last_error()
    assert L.afqmc_propagate(None, 4, None, None, None, None, None, None, None, 0) == 1


def test_group_error_paths_without_gpu_work():
    L = _lib.lib()
    assert L.afqmc_group_create(None, 4, None) == 1
    g = ctypes.c_void_p()
    assert L.afqmc_group_create(ctypes.byref(g), 0, None) == 1
    assert b"max_fragments" in L.afqmc_last_error()
    assert L.afqmc_group_add(None, None, None, None, None, None, None, 1, None) == 1
    assert L.afqmc_group_propagate(None, 0, 0) == 1
    assert L.afqmc_group_tune(None, None) == 1
    assert L.afqmc_group_clear(None) == 1
    assert L.afqmc_group_size(None) == 0
    assert L.afqmc_group_destroy(None) == 0
