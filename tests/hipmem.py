"""Device buffers for the GPU tests through the HIP runtime the library itself links
(libamdhip64.so.7, already loaded with libwlc_mde.so), so one runtime serves the process."""
import ctypes

import numpy as np

_hip = None


def hip():
    global _hip
    if _hip is None:
        import wlc_mde
        wlc_mde.load()
        h = ctypes.CDLL("libamdhip64.so.7")
        vp = ctypes.c_void_p
        h.hipMalloc.argtypes = [ctypes.POINTER(vp), ctypes.c_size_t]
        h.hipFree.argtypes = [vp]
        h.hipMemcpy.argtypes = [vp, vp, ctypes.c_size_t, ctypes.c_int]
        h.hipDeviceSynchronize.argtypes = []
        h.hipStreamCreate.argtypes = [ctypes.POINTER(vp)]
        h.hipStreamDestroy.argtypes = [vp]
        h.hipStreamSynchronize.argtypes = [vp]
        _hip = h
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP status {rc}")


class DeviceArray:
    """A device copy of a float64 numpy array (or an uninitialised one of a shape)."""

    def __init__(self, a=None, shape=None):
        h = hip()
        if a is not None:
            a = np.ascontiguousarray(a, dtype=np.float64)
            shape = a.shape
        self.shape = tuple(shape)
        self.nbytes = int(np.prod(self.shape)) * 8
        p = ctypes.c_void_p()
        _ok(h.hipMalloc(ctypes.byref(p), max(self.nbytes, 8)), "hipMalloc")
        self.ptr = p.value
        if a is not None:
            _ok(h.hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, 1), "hipMemcpy H2D")

    def get(self):
        out = np.empty(self.shape)
        _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
        _ok(hip().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, 2), "hipMemcpy D2H")
        return out

    def fill_nan(self):
        nan = np.full(self.shape, np.nan)
        _ok(hip().hipMemcpy(self.ptr, nan.ctypes.data, self.nbytes, 1), "hipMemcpy H2D")

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = ctypes.c_void_p()
        _ok(hip().hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        self.ptr = s.value

    def sync(self):
        _ok(hip().hipStreamSynchronize(self.ptr), "hipStreamSynchronize")

    def close(self):
        if self.ptr:
            hip().hipStreamDestroy(self.ptr)
            self.ptr = None
