"""The C-ABI boundary on the GPU box: the library the GPU tests load exports
every entry point include/scycle_hip.h declares, and a device round trip
through its memory helpers works (sc_malloc, sc_memcpy_h2d/d2h, sc_free)."""
import ctypes as C

import numpy as np
import pytest

from test_abi_io import _declared

pytestmark = pytest.mark.gpu


def test_gpu_library_exports_every_declared_symbol():
    from scycle_amd import _lib
    L = _lib.load()
    names = _declared()
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)
    major, minor = C.c_int(), C.c_int()
    assert L.sc_version(C.byref(major), C.byref(minor)) == 0
    assert (major.value, minor.value) >= (0, 2)


def test_gpu_device_memory_round_trip():
    from scycle_amd._lib import check, load
    L = load()
    x = np.arange(1000, dtype=np.float64) * 0.5
    y = np.zeros_like(x)
    p = C.c_void_p()
    check(L.sc_malloc(C.byref(p), x.nbytes))
    try:
        check(L.sc_memcpy_h2d(p, x.ctypes.data_as(C.c_void_p), x.nbytes))
        check(L.sc_memcpy_d2h(y.ctypes.data_as(C.c_void_p), p, x.nbytes))
    finally:
        check(L.sc_free(p))
    assert np.array_equal(x, y)
