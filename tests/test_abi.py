"""C ABI checks that need no GPU: the libraries load, export every declared
symbol, and the host-only entry points behave."""
import ctypes
import os
import re

import numpy as np
import pytest

from opendihu_amd import _lib as L
from opendihu_amd.build import LIB_DIR, INCLUDE

HEADER = os.path.join(INCLUDE, "opendihu_fms.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"^\s*(?:int|void)\s+(\w+)\s*\(", text, flags=re.M)))


def _libraries():
    return sorted(f for f in os.listdir(LIB_DIR) if f.startswith("libfms_") and f.endswith(".so"))


def test_header_declares_binding_list():
    assert _declared_functions() == sorted(L.EXPORTED_SYMBOLS)


@pytest.mark.parametrize("name", _libraries() or ["<none built>"])
def test_library_exports_every_declared_symbol(name):
    if name == "<none built>":
        pytest.fail("no fibre library built; run __graft_entry__.build()")
    lib = ctypes.CDLL(os.path.join(LIB_DIR, name))
    for sym in _declared_functions():
        assert hasattr(lib, sym), f"{name} lacks {sym}"


def test_model_info_and_defaults():
    lib = L.load_library()
    assert L.model_info(lib) == (4, 9, 4, "hodgkin_huxley_1952")
    y = np.zeros(4)
    p = np.zeros(4)
    lib.fms_model_defaults(L.as_ptr(y, L._dp), L.as_ptr(p, L._dp))
    np.testing.assert_array_equal(y, [-75, 0.05, 0.6, 0.325])
    np.testing.assert_array_equal(p, [0, 120, 36, 0.3])


def test_invalid_arguments_are_reported_not_thrown():
    lib = L.load_library()
    cfg = L.FmsConfig()
    cfg.n_fibers = 0
    h = ctypes.c_void_p()
    rc = lib.fms_create(ctypes.byref(cfg), ctypes.byref(h))
    assert rc == -1 and "n_fibers" in L.last_error(lib)
    n = np.array([1], np.int32)
    e = np.zeros(1)
    pf = np.ones(1)
    cfg.n_fibers = 1
    cfg.n_nodes = L.as_ptr(n, L._ip)
    cfg.element_lengths = L.as_ptr(e, L._dp)
    cfg.prefactor = L.as_ptr(pf, L._dp)
    assert lib.fms_create(ctypes.byref(cfg), ctypes.byref(h)) == -1
    assert "at least 2 nodes" in L.last_error(lib)
    assert lib.fms_schedule(1.0, -1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 1.0, 0, None) == -1


def test_reference_abi_without_dims_reports():
    lib = L.load_library()
    env = os.environ.pop("OPENDIHU_FMS_DIMS", None)
    try:
        z = np.zeros(8)
        zi = np.zeros(8, np.int32)
        # initializeArrays before any dims were set must not crash; it records an error
        d = L.FmsReferenceDims()
        assert lib.fms_reference_set_dims(ctypes.byref(d)) == -1
    finally:
        if env is not None:
            os.environ["OPENDIHU_FMS_DIMS"] = env
    del z, zi
