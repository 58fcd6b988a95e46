"""Device buffers owned through the library's own allocator (sc_malloc)."""
import ctypes as C

import numpy as np

from ._lib import check, load


class DeviceArray:
    """A 1-D device buffer of float64 or float32."""

    def __init__(self, n, dtype=np.float64, host=None):
        self.n = int(n)
        self.dtype = np.dtype(dtype)
        self.nbytes = self.n * self.dtype.itemsize
        p = C.c_void_p()
        check(load().sc_malloc(C.byref(p), max(self.nbytes, 8)))
        self.ptr = p
        if host is not None:
            self.upload(host)
        else:
            self.upload(np.zeros(self.n, dtype=self.dtype))

    def upload(self, host):
        h = np.ascontiguousarray(host, dtype=self.dtype)
        assert h.size == self.n, (h.size, self.n)
        check(load().sc_memcpy_h2d(self.ptr, h.ctypes.data_as(C.c_void_p), self.nbytes))

    def to_host(self):
        h = np.empty(self.n, dtype=self.dtype)
        check(load().sc_memcpy_d2h(h.ctypes.data_as(C.c_void_p), self.ptr, self.nbytes))
        return h

    def free(self):
        if self.ptr is not None and self.ptr.value:
            load().sc_free(self.ptr)
        self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
