"""cf_amd_countpairs_noalloc (no device allocation; fixed scratch array) against
the allocating entry point countpairs() on the same device-resident inputs,
through the C ABI: npairs bit-exact, ravg and weightavg within 1e-10 relative.
Two calls in a row, the second on a smaller input; a weighted cross-correlation;
the clear errors for an input larger than the scratch array and for host
pointers, after which the entry point still works."""
import ctypes as C

import numpy as np
import pytest

from corrfunc_amd import _abi, _lib
from helpers import rel_err, uniform, weights

pytestmark = pytest.mark.gpu


def _noalloc(lib):
    f = lib.cf_amd_countpairs_noalloc
    f.argtypes = [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p,
                  C.c_int, C.c_int, C.c_char_p, C.POINTER(_abi.results_countpairs), C.POINTER(_abi.config_options),
                  C.POINTER(_abi.extra_options)]
    f.restype = C.c_int
    return f


class Results:
    """Caller-owned results arrays of `cap` entries."""

    def __init__(self, cap):
        self.np = np.zeros(cap, np.uint64)
        self.rupp = np.zeros(cap)
        self.ravg = np.zeros(cap)
        self.wavg = np.zeros(cap)
        self.r = _abi.results_countpairs()
        self.r.nbin = cap
        self.r.npairs = self.np.ctypes.data_as(C.POINTER(C.c_uint64))
        self.r.rupp = self.rupp.ctypes.data_as(C.POINTER(C.c_double))
        self.r.rpavg = self.ravg.ctypes.data_as(C.POINTER(C.c_double))
        self.r.weightavg = self.wavg.ctypes.data_as(C.POINTER(C.c_double))

    def out(self):
        n = self.r.nbin
        return dict(nbin=n, npairs=self.np[:n].copy(), rupp=self.rupp[:n].copy(), ravg=self.ravg[:n].copy(),
                    weightavg=self.wavg[:n].copy())


def _device(arrs):
    return [_lib.DeviceArray(a) for a in arrs]


def _call(lib, fn, d1, n1, d2, n2, w1, w2, autocorr, binpath, res=None):
    o = _abi.make_options(periodic=True, boxsize=300.0, need_avg_sep=True, isa="fallback", float_type=8)
    e = _abi.get_extra_options(_abi.WEIGHT_PAIR_PRODUCT if w1 is not None else _abi.WEIGHT_NONE)
    if w1 is not None:
        e.weights0.weights[0] = w1.ptr
        if w2 is not None:
            e.weights1.weights[0] = w2.ptr
    q = [C.c_void_p(d.ptr) for d in d2] if d2 else [None, None, None]
    args = [n1] + [C.c_void_p(d.ptr) for d in d1] + [n2] + q + [4, autocorr, binpath.encode()]
    if res is None:
        r = _abi.results_countpairs()
        st = lib.countpairs(*args, C.byref(r), C.byref(o), C.byref(e))
        assert st == 0
        out = _abi._collect("countpairs", r)
        lib.free_results(C.byref(r))
        return out
    st = fn(*args, C.byref(res.r), C.byref(o), C.byref(e))
    return res.out() if st == 0 else None


def _same(got, want):
    s = _abi.reported_slots("countpairs", want)
    assert got["nbin"] == want["nbin"]
    assert np.array_equal(got["rupp"], want["rupp"])
    assert np.array_equal(got["npairs"][s], want["npairs"][s])
    assert rel_err(got["ravg"][s], want["ravg"][s]) <= 1e-10
    assert rel_err(got["weightavg"][s], want["weightavg"][s]) <= 1e-10
    assert want["npairs"][s].sum() > 0


def test_noalloc_matches_countpairs():
    lib = _lib.lib()
    fn = _noalloc(lib)
    _lib.set_engine(_lib.ENGINE_TUNED)
    L = 300.0
    # the scratch holds ~70 B per particle plus 64 B per neighbour-cell pair:
    # rmax = 30 in a 300 box gives a 20 x 20 x 10 grid (~150k cell pairs)
    bins = np.logspace(-1, np.log10(30.0), 16)
    binpath, _ = _abi.write_bin_file(bins)
    # first call: 300k points; second: 40k points and fewer bins (a smaller input)
    X = uniform(300000, L, 5)
    d = _device(X)
    res = Results(64)
    got = _call(lib, fn, d, 300000, None, 0, None, None, 1, binpath, res)
    want = _call(lib, fn, d, 300000, None, 0, None, None, 1, binpath)
    assert got is not None
    _same(got, want)
    bins2 = np.logspace(0, np.log10(25.0), 9)
    binpath2, _ = _abi.write_bin_file(bins2)
    X2 = uniform(40000, L, 6)
    d2 = _device(X2)
    res2 = Results(64)
    got2 = _call(lib, fn, d2, 40000, None, 0, None, None, 1, binpath2, res2)
    want2 = _call(lib, fn, d2, 40000, None, 0, None, None, 1, binpath2)
    assert got2 is not None and got2["nbin"] == 9
    _same(got2, want2)


def test_noalloc_weighted_cross_and_errors():
    lib = _lib.lib()
    fn = _noalloc(lib)
    _lib.set_engine(_lib.ENGINE_TUNED)
    L = 300.0
    bins = np.concatenate([[0.0], np.logspace(-1, np.log10(30.0), 10)])
    binpath, _ = _abi.write_bin_file(bins)
    A, B = uniform(60000, L, 7), uniform(50000, L, 8)
    wa, wb = weights(60000, 9, np.float64), weights(50000, 10, np.float64)
    da, db = _device(A), _device(B)
    dwa, dwb = _lib.DeviceArray(wa), _lib.DeviceArray(wb)
    got = _call(lib, fn, da, 60000, db, 50000, dwa, dwb, 0, binpath, Results(16))
    want = _call(lib, fn, da, 60000, db, 50000, dwa, dwb, 0, binpath)
    assert got is not None
    _same(got, want)
    # an input larger than the scratch array: a clear error, nothing allocated
    big = _device(uniform(3000000, L, 11))
    assert _call(lib, fn, big, 3000000, None, 0, None, None, 1, binpath, Results(16)) is None
    # results arrays too small for the bins
    assert _call(lib, fn, da, 60000, None, 0, None, None, 1, binpath, Results(4)) is None
    # host pointers are refused
    o = _abi.make_options(periodic=True, boxsize=L, isa="fallback", float_type=8)
    e = _abi.get_extra_options(_abi.WEIGHT_NONE)
    r = Results(16)
    st = fn(60000, A[0].ctypes.data, A[1].ctypes.data, A[2].ctypes.data, 0, None, None, None, 4, 1,
            binpath.encode(), C.byref(r.r), C.byref(o), C.byref(e))
    assert st != 0
    # and the entry point still works after the errors
    again = _call(lib, fn, da, 60000, db, 50000, dwa, dwb, 0, binpath, Results(16))
    _same(again, want)
