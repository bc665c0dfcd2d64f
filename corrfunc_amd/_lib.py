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
                ("grid_ms", C.c_double), ("total_ms", C.c_double), ("engine", C.c_int), ("sched", C.c_int),
                ("list_ms", C.c_double), ("hist_copies", C.c_int), ("autotuned", C.c_int),
                ("queue_fetch", C.c_int), ("global_hist", C.c_int), ("drain_pairs", C.c_int),
                ("workgroup_threads", C.c_int)]


ENGINE_TUNED = 0
ENGINE_PLAIN = 1

SCHED_NONE = 0
SCHED_HEURISTIC = 1
SCHED_LEARNED = 2

MODEL_FEATURES = 10
MODEL_PARAMS = 1441
MODEL_NORM = 22

_FP = C.POINTER(C.c_float)


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
        L.cf_amd_countpairs_batch.argtypes = [C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_void_p),
                                              C.POINTER(C.c_void_p), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                              C.c_int, C.c_char_p, C.POINTER(_abi.results_countpairs),
                                              C.POINTER(_abi.config_options), C.POINTER(_abi.extra_options)]
        L.cf_amd_countpairs_batch.restype = C.c_int
        for name, args, res in [
                ("cf_amd_set_autotune", [C.c_int], C.c_int),
                ("cf_amd_get_autotune", [], C.c_int),
                ("cf_amd_autotune_clear", [], C.c_int),
                ("cf_amd_set_scheduler", [C.c_int], C.c_int),
                ("cf_amd_get_scheduler", [], C.c_int),
                ("cf_amd_record_costs", [C.c_int], C.c_int),
                ("cf_amd_cost_samples", [_FP, _FP, C.c_int64], C.c_int64),
                ("cf_amd_clear_cost_samples", [], C.c_int),
                ("cf_amd_model_train", [_FP, _FP, C.c_int64, C.c_int, C.c_int, C.c_float, C.c_uint, _FP], C.c_int),
                ("cf_amd_model_get", [_FP, _FP], C.c_int),
                ("cf_amd_model_set", [_FP, _FP], C.c_int),
                ("cf_amd_model_reset", [], C.c_int),
                ("cf_amd_model_trained", [], C.c_int),
                ("cf_amd_model_predict", [_FP, C.c_int64, _FP], C.c_int),
                ("cf_amd_model_grad", [_FP, _FP, C.c_int64, _FP, _FP, _FP, _FP], C.c_int)]:
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = res
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
                grid_ms=s.grid_ms, total_ms=s.total_ms, engine=s.engine, sched=s.sched, list_ms=s.list_ms,
                hist_copies=s.hist_copies, autotuned=s.autotuned, queue_fetch=s.queue_fetch,
                global_hist=s.global_hist, drain_pairs=s.drain_pairs,
                workgroup_threads=s.workgroup_threads)


def set_autotune(on=True):
    lib().cf_amd_set_autotune(1 if on else 0)


def get_autotune():
    return bool(lib().cf_amd_get_autotune())


def autotune_clear():
    lib().cf_amd_autotune_clear()


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

    def upload(self, a):
        """Overwrites the device copy with host array `a` (same size and dtype)."""
        import numpy as np
        a = np.ascontiguousarray(a, dtype=self.dtype)
        if a.nbytes != self.nbytes:
            raise ValueError("size mismatch: %d vs %d bytes" % (a.nbytes, self.nbytes))
        if lib().cf_amd_memcpy(self.ptr, a.ctypes.data, self.nbytes, 0) != 0:
            raise RuntimeError("copy to device failed")

    def copy_from(self, other):
        """Device-to-device copy of another DeviceArray of the same size (the
        engine's own copy kernel)."""
        if other.nbytes != self.nbytes:
            raise ValueError("size mismatch: %d vs %d bytes" % (other.nbytes, self.nbytes))
        if lib().cf_amd_memcpy(self.ptr, other.ptr, self.nbytes, 2) != 0:
            raise RuntimeError("device-to-device copy failed")

    def free(self):
        if self.ptr:
            lib().cf_amd_device_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def countpairs_batch(catalogs, binfile, options=None, weights=None, nthreads=1):
    """DD(r) auto-correlations of many catalogs in one device launch.

    catalogs: sequence of (X, Y, Z) arrays (numpy, or DeviceArray for HBM-resident
    data); weights: optional sequence of per-catalog weight arrays (pair_product).
    Returns a list of dicts with the countpairs outputs of each catalog."""
    import os
    from . import _abi
    ncat = len(catalogs)
    keep = []

    def ptr(a):
        if isinstance(a, DeviceArray):
            return a.ptr
        keep.append(a)
        return a.ctypes.data

    ND = (C.c_int64 * ncat)()
    Xs, Ys, Zs, Ws = [(C.c_void_p * ncat)() for _ in range(4)]
    dtype = None
    for k, (x, y, z) in enumerate(catalogs):
        ND[k] = x.shape[0]
        Xs[k], Ys[k], Zs[k] = ptr(x), ptr(y), ptr(z)
        dtype = x.dtype
        if weights is not None:
            Ws[k] = ptr(weights[k])
    o = options if options is not None else _abi.make_options()
    o.float_type = dtype.itemsize
    extra = _abi.get_extra_options(_abi.WEIGHT_PAIR_PRODUCT if weights is not None else _abi.WEIGHT_NONE)
    res = (_abi.results_countpairs * ncat)()
    path, tmp = _abi.write_bin_file(binfile)
    try:
        st = lib().cf_amd_countpairs_batch(ncat, ND, Xs, Ys, Zs, Ws, int(nthreads), path.encode(), res,
                                           C.byref(o), C.byref(extra))
    finally:
        if tmp:
            os.remove(path)
    if st != 0:
        raise RuntimeError("cf_amd_countpairs_batch failed")
    out = []
    for k in range(ncat):
        d = _abi._collect("countpairs", res[k])
        lib().free_results(C.byref(res[k]))
        out.append(d)
    return out


class BatchGraph:
    """The batched DD(r) captured once into a HIP graph (cf_amd_batch_graph_*)
    and replayed on whatever the catalogs' device arrays hold at each launch:
    write resampled positions of the same sizes into the same DeviceArrays and
    call launch() again.

    catalogs: sequence of (X, Y, Z) DeviceArray triples; weights: optional
    sequence of DeviceArray weights (pair_product). The options must give a
    periodic box size."""

    def __init__(self, catalogs, binfile, options=None, weights=None, nthreads=1):
        import os
        from . import _abi
        L = lib()
        if not hasattr(L, "_graph_bound"):
            L.cf_amd_batch_graph_create.argtypes = [C.c_int, C.POINTER(C.c_int64), C.POINTER(C.c_void_p),
                                                    C.POINTER(C.c_void_p), C.POINTER(C.c_void_p),
                                                    C.POINTER(C.c_void_p), C.c_int, C.c_char_p,
                                                    C.POINTER(_abi.config_options), C.POINTER(_abi.extra_options),
                                                    C.POINTER(C.c_void_p)]
            L.cf_amd_batch_graph_create.restype = C.c_int
            L.cf_amd_batch_graph_launch.argtypes = [C.c_void_p, C.POINTER(_abi.results_countpairs)]
            L.cf_amd_batch_graph_launch.restype = C.c_int
            L.cf_amd_batch_graph_destroy.argtypes = [C.c_void_p]
            L.cf_amd_batch_graph_destroy.restype = C.c_int
            L._graph_bound = True
        self.ncat = ncat = len(catalogs)
        self.catalogs, self.weights = catalogs, weights  # keep the device arrays alive
        ND = (C.c_int64 * ncat)()
        Xs, Ys, Zs, Ws = [(C.c_void_p * ncat)() for _ in range(4)]
        for k, (x, y, z) in enumerate(catalogs):
            if not isinstance(x, DeviceArray):
                raise TypeError("a graph batch reads its catalogs from DeviceArrays")
            ND[k] = x.shape[0]
            Xs[k], Ys[k], Zs[k] = x.ptr, y.ptr, z.ptr
            if weights is not None:
                Ws[k] = weights[k].ptr
        o = options if options is not None else _abi.make_options()
        o.float_type = catalogs[0][0].dtype.itemsize
        extra = _abi.get_extra_options(_abi.WEIGHT_PAIR_PRODUCT if weights is not None else _abi.WEIGHT_NONE)
        path, tmp = _abi.write_bin_file(binfile)
        self.handle = C.c_void_p()
        try:
            st = L.cf_amd_batch_graph_create(ncat, ND, Xs, Ys, Zs, Ws, int(nthreads), path.encode(), C.byref(o),
                                             C.byref(extra), C.byref(self.handle))
        finally:
            if tmp:
                os.remove(path)
        if st != 0:
            raise RuntimeError("cf_amd_batch_graph_create failed")

    def launch(self):
        """Replays the graph; returns a list of dicts with the countpairs outputs of each catalog."""
        from . import _abi
        res = (_abi.results_countpairs * self.ncat)()
        if lib().cf_amd_batch_graph_launch(self.handle, res) != 0:
            raise RuntimeError("cf_amd_batch_graph_launch failed")
        out = []
        for k in range(self.ncat):
            out.append(_abi._collect("countpairs", res[k]))
            lib().free_results(C.byref(res[k]))
        return out

    def close(self):
        if self.handle:
            lib().cf_amd_batch_graph_destroy(self.handle)
            self.handle = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
