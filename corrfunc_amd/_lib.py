"""Loader of the engine's shared library (libcorrfunc_amd.so, built in-tree).

The library is loaded from corrfunc_amd/lib/ only; there is no CPU fallback.
"""
import ctypes as C
import os

from . import _abi

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib", "libcorrfunc_amd.so")

_lib = None


class cf_amd_stats(C.Structure):
    _fields_ = [("nmesh", C.c_int * 3), ("refine", C.c_int * 3), ("num_cell_pairs", C.c_int64),
                ("num_work_items", C.c_int64), ("tested_pairs", C.c_double), ("kernel_ms", C.c_double),
                ("grid_ms", C.c_double), ("total_ms", C.c_double), ("engine", C.c_int), ("pad", C.c_int)]


ENGINE_TUNED = 0
ENGINE_PLAIN = 1


def lib():
    """Returns the loaded engine library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.isfile(LIB_PATH):
            raise ImportError("corrfunc_amd: %s is missing; build it with "
                              "`python -c 'import __graft_entry__ as g; g.build()'` or `make -C corrfunc_amd/csrc`"
                              % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        _abi.bind(L)
        L.cf_amd_set_engine.argtypes = [C.c_int]
        L.cf_amd_set_engine.restype = C.c_int
        L.cf_amd_get_engine.argtypes = []
        L.cf_amd_get_engine.restype = C.c_int
        L.cf_amd_last_stats.argtypes = [C.POINTER(cf_amd_stats)]
        L.cf_amd_last_stats.restype = C.c_int
        L.cf_amd_version.argtypes = []
        L.cf_amd_version.restype = C.c_char_p
        L.cf_amd_device_alloc.argtypes = [C.c_size_t]
        L.cf_amd_device_alloc.restype = C.c_void_p
        L.cf_amd_device_free.argtypes = [C.c_void_p]
        L.cf_amd_device_free.restype = C.c_int
        L.cf_amd_memcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        L.cf_amd_memcpy.restype = C.c_int
        _lib = L
    return _lib


def set_engine(kind):
    if lib().cf_amd_set_engine(int(kind)) != 0:
        raise ValueError("unknown engine kind %r" % kind)


def get_engine():
    return lib().cf_amd_get_engine()


def last_stats():
    s = cf_amd_stats()
    lib().cf_amd_last_stats(C.byref(s))
    return dict(nmesh=tuple(s.nmesh), refine=tuple(s.refine), num_cell_pairs=s.num_cell_pairs,
                num_work_items=s.num_work_items, tested_pairs=s.tested_pairs, kernel_ms=s.kernel_ms,
                grid_ms=s.grid_ms, total_ms=s.total_ms, engine=s.engine)


class DeviceArray:
    """A numpy array copied into HBM through the engine's runtime; `.ptr` is the
    device address to hand to the C ABI in place of the host pointer."""

    def __init__(self, a):
        import numpy as np
        a = np.ascontiguousarray(a)
        self.nbytes = a.nbytes
        self.dtype = a.dtype
        self.shape = a.shape
        self.ptr = lib().cf_amd_device_alloc(self.nbytes)
        if not self.ptr:
            raise MemoryError("cf_amd_device_alloc(%d) failed" % self.nbytes)
        if lib().cf_amd_memcpy(self.ptr, a.ctypes.data, self.nbytes, 0) != 0:
            raise RuntimeError("copy to device failed")

    def numpy(self):
        import numpy as np
        out = np.empty(self.shape, dtype=self.dtype)
        if lib().cf_amd_memcpy(out.ctypes.data, self.ptr, self.nbytes, 1) != 0:
            raise RuntimeError("copy from device failed")
        return out

    def free(self):
        if self.ptr:
            lib().cf_amd_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
