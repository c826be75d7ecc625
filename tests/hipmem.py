"""Device buffers for the GPU tests, allocated through the HIP runtime that
libmallard_amd.so itself links (symbols looked up through the library's
handle, so both share one runtime instance in the process)."""
import ctypes as C

import numpy as np

_H2D, _D2H = 1, 2


def _rt(mlr):
    L = mlr.lib()
    L.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    L.hipFree.argtypes = [C.c_void_p]
    L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    L.hipDeviceSynchronize.argtypes = []
    return L


class DeviceArray:
    """A float64 device array of a fixed shape."""

    def __init__(self, mlr, shape, host=None):
        self._L = _rt(mlr)
        self.shape = tuple(shape) if isinstance(shape, (tuple, list)) else (int(shape),)
        self.nbytes = int(np.prod(self.shape)) * 8
        p = C.c_void_p()
        if self._L.hipMalloc(C.byref(p), max(self.nbytes, 8)) != 0:
            raise RuntimeError("hipMalloc failed")
        self.ptr = p.value
        if host is not None:
            self.upload(host)
        else:
            self.upload(np.zeros(self.shape))

    def upload(self, host):
        a = np.ascontiguousarray(host, dtype=np.float64).reshape(self.shape)
        if self._L.hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, _H2D) != 0:
            raise RuntimeError("hipMemcpy to the device failed")

    def get(self):
        self._L.hipDeviceSynchronize()
        a = np.zeros(self.shape)
        if self._L.hipMemcpy(a.ctypes.data, self.ptr, self.nbytes, _D2H) != 0:
            raise RuntimeError("hipMemcpy from the device failed")
        return a

    def __del__(self):
        if getattr(self, "ptr", None):
            self._L.hipFree(self.ptr)
            self.ptr = None


def synchronize(mlr):
    _rt(mlr).hipDeviceSynchronize()
