"""ctypes mirror of the C ABI declared in include/corrfunc_theory.h.

The structures reproduce the byte layout of the reference's public structs
(utils/defs.h:37-45 api_cell_timings, :53-156 config_options, :353-419
weight_struct / extra_options; theory/<mode>/countpairs*.h results structs).
`call_mode` drives any shared library that exports the reference entry points,
optionally under a symbol prefix, with the same option handling as the
reference's CPython extension (theory/python_bindings/_countpairs.c:1136-1441).
"""
import ctypes as C
import os
import re
import tempfile

import numpy as np

API_VERSION = b'"2.5.3"'  # quotes included, as the reference stores it
BOXSIZE_NOTGIVEN = -2.0
NLATMAX = 100
BINNING_DFL = 0x0
BINNING_CUST = 0x1
MAX_NUM_WEIGHTS = 10

ISA = {"fallback": 0, "sse42": 6, "avx": 7, "avx2": 8, "avx512f": 9, "arm64": 10, "fastest": -1}
WEIGHT_NONE = -42
WEIGHT_PAIR_PRODUCT = 0


class api_cell_timings(C.Structure):
    _fields_ = [("N1", C.c_int64), ("N2", C.c_int64), ("time_in_ns", C.c_int64),
                ("first_cellindex", C.c_int), ("second_cellindex", C.c_int), ("tid", C.c_int)]


class config_options(C.Structure):
    _fields_ = [
        ("boxsize_x", C.c_double), ("boxsize_y", C.c_double), ("boxsize_z", C.c_double),
        ("OMEGA_M", C.c_double), ("OMEGA_B", C.c_double), ("OMEGA_L", C.c_double),
        ("HUBBLE", C.c_double), ("LITTLE_H", C.c_double), ("SIGMA_8", C.c_double), ("NS", C.c_double),
        ("c_api_time", C.c_double),
        ("cell_timings", C.POINTER(api_cell_timings)),
        ("totncells_timings", C.c_int64),
        ("float_type", C.c_size_t),
        ("instruction_set", C.c_int32),
        ("version", C.c_char * 32),
        ("verbose", C.c_uint8), ("c_api_timer", C.c_uint8), ("c_cell_timer", C.c_uint8),
        ("need_avg_sep", C.c_uint8), ("autocorr", C.c_uint8), ("periodic", C.c_uint8),
        ("sort_on_z", C.c_uint8), ("is_comoving_dist", C.c_uint8), ("link_in_dec", C.c_uint8),
        ("link_in_ra", C.c_uint8), ("fast_divide_and_NR_steps", C.c_uint8), ("fast_acos", C.c_uint8),
        ("enable_min_sep_opt", C.c_uint8),
        ("bin_refine_factors", C.c_int8 * 3),
        ("max_cells_per_dim", C.c_uint16),
        ("copy_particles", C.c_uint8), ("use_heap_sort", C.c_uint8),
        ("binning_flags", C.c_uint32),
        ("reserved", C.c_uint8 * (1024 - 172)),
    ]


class weight_struct(C.Structure):
    _fields_ = [("weights", C.c_void_p * MAX_NUM_WEIGHTS), ("num_weights", C.c_int64)]


class extra_options(C.Structure):
    _fields_ = [("weights0", weight_struct), ("weights1", weight_struct), ("weight_method", C.c_int),
                ("reserved", C.c_uint8 * (1024 - 2 * 88 - 4))]


class results_countpairs(C.Structure):
    _fields_ = [("npairs", C.POINTER(C.c_uint64)), ("rupp", C.POINTER(C.c_double)),
                ("rpavg", C.POINTER(C.c_double)), ("weightavg", C.POINTER(C.c_double)), ("nbin", C.c_int)]


class results_countpairs_rp_pi(C.Structure):
    _fields_ = [("npairs", C.POINTER(C.c_uint64)), ("rupp", C.POINTER(C.c_double)),
                ("rpavg", C.POINTER(C.c_double)), ("weightavg", C.POINTER(C.c_double)),
                ("pimax", C.c_double), ("nbin", C.c_int), ("npibin", C.c_int)]


class results_countpairs_s_mu(C.Structure):
    _fields_ = [("npairs", C.POINTER(C.c_uint64)), ("supp", C.POINTER(C.c_double)),
                ("savg", C.POINTER(C.c_double)), ("mu_max", C.c_double), ("mu_min", C.c_double),
                ("weightavg", C.POINTER(C.c_double)), ("nsbin", C.c_int), ("nmu_bins", C.c_int)]


class results_countpairs_xi(C.Structure):
    _fields_ = [("npairs", C.POINTER(C.c_uint64)), ("xi", C.POINTER(C.c_double)),
                ("rupp", C.POINTER(C.c_double)), ("ravg", C.POINTER(C.c_double)),
                ("weightavg", C.POINTER(C.c_double)), ("nbin", C.c_int)]


class results_countpairs_wp(C.Structure):
    _fields_ = [("npairs", C.POINTER(C.c_uint64)), ("wp", C.POINTER(C.c_double)),
                ("rupp", C.POINTER(C.c_double)), ("rpavg", C.POINTER(C.c_double)),
                ("weightavg", C.POINTER(C.c_double)), ("pimax", C.c_double), ("nbin", C.c_int)]


assert C.sizeof(config_options) == 1024
assert C.sizeof(extra_options) == 1024
assert config_options.instruction_set.offset == 112
assert config_options.binning_flags.offset == 168
assert extra_options.weight_method.offset == 176


def get_config_options():
    """Defaults of reference utils/defs.h:236-346 (theory.options build)."""
    o = config_options()
    C.memset(C.byref(o), 0, C.sizeof(o))
    o.version = API_VERSION
    o.boxsize_x = o.boxsize_y = o.boxsize_z = BOXSIZE_NOTGIVEN
    o.float_type = 4
    o.verbose = 1
    o.periodic = 1
    o.instruction_set = -1
    o.enable_min_sep_opt = 1
    o.copy_particles = 1
    o.max_cells_per_dim = NLATMAX
    o.bin_refine_factors[0] = 2
    o.bin_refine_factors[1] = 2
    o.bin_refine_factors[2] = 1
    o.binning_flags = BINNING_DFL
    return o


def get_extra_options(method):
    e = extra_options()
    C.memset(C.byref(e), 0, C.sizeof(e))
    e.weight_method = method
    n = 1 if method == WEIGHT_PAIR_PRODUCT else 0
    e.weights0.num_weights = n
    e.weights1.num_weights = n
    return e


_SIGS = {
    "countpairs": [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                   C.c_void_p, C.c_int, C.c_int, C.c_char_p, C.POINTER(results_countpairs),
                   C.POINTER(config_options), C.POINTER(extra_options)],
    "countpairs_rp_pi": [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                         C.c_void_p, C.c_int, C.c_int, C.c_char_p, C.c_double,
                         C.POINTER(results_countpairs_rp_pi), C.POINTER(config_options),
                         C.POINTER(extra_options)],
    "countpairs_s_mu": [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                        C.c_void_p, C.c_int, C.c_int, C.c_char_p, C.c_double, C.c_int,
                        C.POINTER(results_countpairs_s_mu), C.POINTER(config_options),
                        C.POINTER(extra_options)],
    "countpairs_xi": [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_char_p,
                      C.POINTER(results_countpairs_xi), C.POINTER(config_options),
                      C.POINTER(extra_options)],
    "countpairs_wp": [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_double, C.c_int, C.c_char_p,
                      C.c_double, C.POINTER(results_countpairs_wp), C.POINTER(config_options),
                      C.POINTER(extra_options)],
}
class results_countpairs_mocks(C.Structure):
    _fields_ = [("npairs", C.POINTER(C.c_uint64)), ("rupp", C.POINTER(C.c_double)),
                ("rpavg", C.POINTER(C.c_double)), ("weightavg", C.POINTER(C.c_double)),
                ("pimax", C.c_double), ("nbin", C.c_int), ("npibin", C.c_int)]


class results_countpairs_mocks_s_mu(C.Structure):
    _fields_ = [("npairs", C.POINTER(C.c_uint64)), ("supp", C.POINTER(C.c_double)),
                ("savg", C.POINTER(C.c_double)), ("mu_max", C.c_double), ("mu_min", C.c_double),
                ("weightavg", C.POINTER(C.c_double)), ("nsbin", C.c_int), ("nmu_bins", C.c_int)]


_MOCK_SIGS = {
    "countpairs_mocks": [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                         C.c_void_p, C.c_int, C.c_int, C.c_char_p, C.c_double, C.c_int,
                         C.POINTER(results_countpairs_mocks), C.POINTER(config_options),
                         C.POINTER(extra_options)],
    "countpairs_mocks_s_mu": [C.c_int64, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_void_p,
                              C.c_void_p, C.c_int, C.c_int, C.c_char_p, C.c_double, C.c_int, C.c_int,
                              C.POINTER(results_countpairs_mocks_s_mu), C.POINTER(config_options),
                              C.POINTER(extra_options)],
}
_RES = {"countpairs": results_countpairs, "countpairs_rp_pi": results_countpairs_rp_pi,
        "countpairs_s_mu": results_countpairs_s_mu, "countpairs_xi": results_countpairs_xi,
        "countpairs_wp": results_countpairs_wp, "countpairs_mocks": results_countpairs_mocks,
        "countpairs_mocks_s_mu": results_countpairs_mocks_s_mu}
_FREE = {"countpairs": "free_results", "countpairs_rp_pi": "free_results_rp_pi",
         "countpairs_s_mu": "free_results_s_mu", "countpairs_xi": "free_results_xi",
         "countpairs_wp": "free_results_wp", "countpairs_mocks": "free_results_mocks",
         "countpairs_mocks_s_mu": "free_results_mocks_s_mu"}
ENTRY_POINTS = tuple(_SIGS) + tuple(_MOCK_SIGS) + tuple(_FREE.values())


def bind(lib, prefix=""):
    """Attach argtypes to the entry points (and free functions) of `lib`: the
    five theory counters and, where the library has them, the two on-sky ones."""
    sigs = dict(_SIGS)
    sigs.update({k: v for k, v in _MOCK_SIGS.items() if hasattr(lib, prefix + k)})
    for name, sig in sigs.items():
        f = getattr(lib, prefix + name)
        f.argtypes = sig
        f.restype = C.c_int
        g = getattr(lib, prefix + _FREE[name])
        g.argtypes = [C.POINTER(_RES[name])]
        g.restype = None
    return lib


def write_bin_file(binfile):
    """Array of edges -> temporary `lo hi` file, exactly as reference
    Corrfunc/utils.py:324-381 (return_file_with_rbins) writes it."""
    if isinstance(binfile, (str, bytes, os.PathLike)):
        path = os.fsdecode(binfile)
        if not os.path.isfile(path):
            raise IOError("Could not find file = `{0}`".format(path))
        return path, False
    edges = np.asarray(binfile, dtype=np.float64)
    if edges.ndim != 1 or len(edges) < 2:
        raise TypeError("Input `binfile` was not a valid array of bin edges")
    edges = sorted(edges.tolist())
    fd, path = tempfile.mkstemp(suffix=".bins", text=True)
    with os.fdopen(fd, "w") as f:
        for lo, hi in zip(edges[:-1], edges[1:]):
            f.write("{0} {1}\n".format(lo, hi))
    return path, True


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _prep_positions(arrs):
    """All position arrays must share one floating dtype (float32/float64)."""
    out = []
    dt = None
    for a in arrs:
        if a is None:
            out.append(None)
            continue
        a = np.ascontiguousarray(a)
        if a.dtype.byteorder not in ("=", "|"):
            a = a.astype(a.dtype.newbyteorder("="))
        if a.dtype not in (np.float32, np.float64):
            raise TypeError("positions must be float32 or float64 arrays, got %s" % a.dtype)
        if dt is None:
            dt = a.dtype
        elif a.dtype != dt:
            raise TypeError("The two arrays must have the same data-type")
        out.append(a)
    return out, dt


def _prep_weights(w, n, dt):
    if w is None:
        return None
    w = np.atleast_1d(np.asarray(w))
    if w.ndim == 1:
        w = w.reshape(1, -1)
    if w.shape[-1] != n:
        raise ValueError("weights must have shape (n_weights, N)")
    return np.ascontiguousarray(w, dtype=dt)


def highest_host_isa():
    """The x86 ISA the reference's 'fastest' resolves to on this host
    (reference utils/cpu_features.c:74 get_max_usable_isa)."""
    try:
        flags = open("/proc/cpuinfo").read()
    except OSError:
        return ISA["fallback"]
    for name, key in (("avx512f", "avx512f"), ("avx2", "avx2"), ("avx", "avx"), ("sse4_2", "sse42")):
        if re.search(r"\b%s\b" % name, flags):
            return ISA[key]
    return ISA["fallback"]


def make_options(periodic=True, boxsize=None, verbose=False, need_avg_sep=False, refine=(2, 2, 1),
                 max_cells_per_dim=100, copy_particles=True, enable_min_sep_opt=True,
                 c_api_timer=False, isa="fastest", float_type=8, c_cell_timer=False,
                 fast_divide_and_NR_steps=0):
    """config_options exactly as the reference CPython binding fills them
    (_countpairs.c:1151-1250): custom refine factors set BINNING_CUST."""
    o = get_config_options()
    o.verbose = int(bool(verbose))
    o.periodic = int(bool(periodic))
    o.need_avg_sep = int(bool(need_avg_sep))
    o.c_api_timer = int(bool(c_api_timer))
    o.c_cell_timer = int(bool(c_cell_timer))
    o.copy_particles = int(bool(copy_particles))
    o.enable_min_sep_opt = int(bool(enable_min_sep_opt))
    o.max_cells_per_dim = int(max_cells_per_dim)
    o.fast_divide_and_NR_steps = int(fast_divide_and_NR_steps)
    o.float_type = int(float_type)
    if isinstance(isa, str):
        key = isa.lower()
        if key not in ISA:
            raise ValueError("Do not know instruction type %s" % isa)
        o.instruction_set = ISA[key]
    else:
        o.instruction_set = int(isa)
    if o.instruction_set == -1:
        # as the reference's CPython binding does before calling C
        o.instruction_set = highest_host_isa()
    if boxsize is not None:
        b = np.atleast_1d(np.asarray(boxsize, dtype=np.float64))
        if len(b) == 1:
            b = np.array([b[0]] * 3)
        o.boxsize_x, o.boxsize_y, o.boxsize_z = float(b[0]), float(b[1]), float(b[2])
    if tuple(refine) != (2, 2, 1):
        for i in range(3):
            o.bin_refine_factors[i] = int(refine[i])
        o.binning_flags = (o.binning_flags & ~0xF) | BINNING_CUST
    return o


def call_mode(lib, mode, prefix="", *, autocorr=1, nthreads=1, binfile=None, X1=None, Y1=None, Z1=None,
              weights1=None, X2=None, Y2=None, Z2=None, weights2=None, weight_type=None,
              pimax=None, mu_max=None, nmu_bins=None, boxsize=None, options=None, raw=False):
    """Run entry point `mode` of `lib` and return a dict of numpy arrays
    (full padded arrays, copied before the library frees them)."""
    (X1, Y1, Z1, X2, Y2, Z2), dt = _prep_positions([X1, Y1, Z1, X2, Y2, Z2])
    n1 = len(X1)
    n2 = len(X2) if X2 is not None else 0
    method = WEIGHT_NONE
    if weight_type is not None:
        if weight_type not in ("pair_product", "p"):
            raise ValueError("unknown weight_type %r" % weight_type)
        method = WEIGHT_PAIR_PRODUCT
    extra = get_extra_options(method)
    w1 = w2 = None
    if method != WEIGHT_NONE:
        w1 = _prep_weights(weights1, n1, dt)
        if w1 is None:
            raise ValueError("weight_type given but no weights1")
        extra.weights0.weights[0] = w1.ctypes.data
        if not autocorr:
            w2 = _prep_weights(weights2, n2, dt)
            if w2 is None:
                raise ValueError("cross-correlation with weights needs weights2")
            extra.weights1.weights[0] = w2.ctypes.data
    o = options if options is not None else make_options()
    o.float_type = dt.itemsize
    path, tmp = write_bin_file(binfile)
    res = _RES[mode]()
    fn = getattr(lib, prefix + mode)
    try:
        if mode == "countpairs":
            st = fn(n1, _ptr(X1), _ptr(Y1), _ptr(Z1), n2, _ptr(X2), _ptr(Y2), _ptr(Z2), int(nthreads),
                    int(autocorr), path.encode(), C.byref(res), C.byref(o), C.byref(extra))
        elif mode == "countpairs_rp_pi":
            st = fn(n1, _ptr(X1), _ptr(Y1), _ptr(Z1), n2, _ptr(X2), _ptr(Y2), _ptr(Z2), int(nthreads),
                    int(autocorr), path.encode(), float(pimax), C.byref(res), C.byref(o), C.byref(extra))
        elif mode == "countpairs_s_mu":
            st = fn(n1, _ptr(X1), _ptr(Y1), _ptr(Z1), n2, _ptr(X2), _ptr(Y2), _ptr(Z2), int(nthreads),
                    int(autocorr), path.encode(), float(mu_max), int(nmu_bins), C.byref(res), C.byref(o),
                    C.byref(extra))
        elif mode == "countpairs_xi":
            st = fn(n1, _ptr(X1), _ptr(Y1), _ptr(Z1), float(boxsize), int(nthreads), path.encode(),
                    C.byref(res), C.byref(o), C.byref(extra))
        elif mode == "countpairs_wp":
            st = fn(n1, _ptr(X1), _ptr(Y1), _ptr(Z1), float(boxsize), int(nthreads), path.encode(),
                    float(pimax), C.byref(res), C.byref(o), C.byref(extra))
        else:
            raise ValueError(mode)
    finally:
        if tmp:
            os.remove(path)
    if st != 0:
        return None
    out = _collect(mode, res)
    out["api_time"] = o.c_api_time
    getattr(lib, prefix + _FREE[mode])(C.byref(res))
    if o.c_cell_timer:
        out["cell_timings"] = take_cell_timings(o)
    return out


CELL_TIMINGS_DTYPE = np.dtype([("N1", np.int64), ("N2", np.int64), ("time_in_ns", np.int64),
                               ("cellidx1", np.int32), ("cellidx2", np.int32), ("tid", np.int32)])


def take_cell_timings(o):
    """Copies options->cell_timings (totncells_timings rows) into a structured
    array with the fields of the reference's _convert_cell_timer
    (Corrfunc/theory/wp.py:244-280) and releases it as free_cell_timings does."""
    n = int(o.totncells_timings)
    out = np.zeros(max(n, 0), dtype=CELL_TIMINGS_DTYPE)
    if n > 0 and o.cell_timings:
        raw_dt = np.dtype({"names": list(CELL_TIMINGS_DTYPE.names),
                           "formats": [np.int64, np.int64, np.int64, np.int32, np.int32, np.int32],
                           "offsets": [0, 8, 16, 24, 28, 32], "itemsize": C.sizeof(api_cell_timings)})
        raw = np.ctypeslib.as_array(C.cast(o.cell_timings, C.POINTER(C.c_uint8)),
                                    shape=(n * C.sizeof(api_cell_timings),))
        rows = raw.view(raw_dt)
        for name in CELL_TIMINGS_DTYPE.names:
            out[name] = rows[name]
        C.CDLL(None).free(C.cast(o.cell_timings, C.c_void_p))
    o.cell_timings = None
    o.totncells_timings = 0
    return out


def _arr(p, n, dtype):
    if not p:
        return np.zeros(n, dtype=dtype)
    return np.ctypeslib.as_array(p, shape=(n,)).astype(dtype, copy=True)


def _collect(mode, r):
    if mode == "countpairs":
        n = r.nbin
        return dict(nbin=n, npairs=_arr(r.npairs, n, np.uint64), rupp=_arr(r.rupp, n, np.float64),
                    ravg=_arr(r.rpavg, n, np.float64), weightavg=_arr(r.weightavg, n, np.float64))
    if mode == "countpairs_rp_pi":
        n, m = r.nbin, r.npibin
        tot = (n + 1) * (m + 1)
        return dict(nbin=n, npibin=m, pimax=r.pimax, npairs=_arr(r.npairs, tot, np.uint64),
                    rupp=_arr(r.rupp, n, np.float64), ravg=_arr(r.rpavg, tot, np.float64),
                    weightavg=_arr(r.weightavg, tot, np.float64))
    if mode == "countpairs_s_mu":
        n, m = r.nsbin, r.nmu_bins
        tot = (n + 1) * (m + 1)
        return dict(nbin=n, nmu_bins=m, mu_max=r.mu_max, npairs=_arr(r.npairs, tot, np.uint64),
                    rupp=_arr(r.supp, n, np.float64), ravg=_arr(r.savg, tot, np.float64),
                    weightavg=_arr(r.weightavg, tot, np.float64))
    if mode == "countpairs_xi":
        n = r.nbin
        return dict(nbin=n, npairs=_arr(r.npairs, n, np.uint64), xi=_arr(r.xi, n, np.float64),
                    rupp=_arr(r.rupp, n, np.float64), ravg=_arr(r.ravg, n, np.float64),
                    weightavg=_arr(r.weightavg, n, np.float64))
    if mode == "countpairs_wp":
        n = r.nbin
        return dict(nbin=n, pimax=r.pimax, npairs=_arr(r.npairs, n, np.uint64), wp=_arr(r.wp, n, np.float64),
                    rupp=_arr(r.rupp, n, np.float64), ravg=_arr(r.rpavg, n, np.float64),
                    weightavg=_arr(r.weightavg, n, np.float64))
    raise ValueError(mode)


def reported_slots(mode, out):
    """Indices of the slots the reference's callers read (1..nbin-1, and
    pi/mu columns 0..n2-1)."""
    n = out["nbin"]
    if mode in ("countpairs", "countpairs_xi", "countpairs_wp"):
        return np.arange(1, n)
    m = out["npibin"] if mode in ("countpairs_rp_pi", "countpairs_mocks") else out["nmu_bins"]
    return np.array([i * (m + 1) + j for i in range(1, n) for j in range(m)], dtype=np.int64)


def call_mocks(lib, mode, prefix="", *, autocorr=1, nthreads=1, binfile=None, RA1=None, DEC1=None, CZ1=None,
               weights1=None, RA2=None, DEC2=None, CZ2=None, weights2=None, weight_type=None, pimax=None,
               mu_max=None, nmu_bins=None, cosmology=1, options=None):
    """Run on-sky entry point `mode` ("countpairs_mocks" or "countpairs_mocks_s_mu")
    of `lib`; returns a dict of numpy arrays as call_mode does, or None on failure.
    The caller's arrays are not changed: where the library would fix a set in
    place (RA < 0, DEC > 90, or cz read as z), that set is handed over as a copy."""
    (RA1, DEC1, CZ1, RA2, DEC2, CZ2), dt = _prep_positions([RA1, DEC1, CZ1, RA2, DEC2, CZ2])
    comoving = bool(options.is_comoving_dist) if options is not None else False

    def _private(ra, dec, cz):
        if len(ra) and (ra.min() < 0 or dec.max() > 90 or (not comoving and cz.max() < 10)):
            return [a.copy() for a in (ra, dec, cz)]
        return [ra, dec, cz]
    RA1, DEC1, CZ1 = _private(RA1, DEC1, CZ1)
    if RA2 is not None:
        RA2, DEC2, CZ2 = _private(RA2, DEC2, CZ2)
    n1 = len(RA1)
    n2 = len(RA2) if RA2 is not None else 0
    method = WEIGHT_NONE
    if weight_type is not None:
        if weight_type not in ("pair_product", "p"):
            raise ValueError("unknown weight_type %r" % weight_type)
        method = WEIGHT_PAIR_PRODUCT
    extra = get_extra_options(method)
    if method != WEIGHT_NONE:
        w1 = _prep_weights(weights1, n1, dt)
        extra.weights0.weights[0] = w1.ctypes.data
        if not autocorr:
            w2 = _prep_weights(weights2, n2, dt)
            extra.weights1.weights[0] = w2.ctypes.data
    o = options if options is not None else make_options(periodic=False)
    o.float_type = dt.itemsize
    path, tmp = write_bin_file(binfile)
    res = _RES[mode]()
    fn = getattr(lib, prefix + mode)
    try:
        if mode == "countpairs_mocks":
            st = fn(n1, _ptr(RA1), _ptr(DEC1), _ptr(CZ1), n2, _ptr(RA2), _ptr(DEC2), _ptr(CZ2), int(nthreads),
                    int(autocorr), path.encode(), float(pimax), int(cosmology), C.byref(res), C.byref(o),
                    C.byref(extra))
        elif mode == "countpairs_mocks_s_mu":
            st = fn(n1, _ptr(RA1), _ptr(DEC1), _ptr(CZ1), n2, _ptr(RA2), _ptr(DEC2), _ptr(CZ2), int(nthreads),
                    int(autocorr), path.encode(), float(mu_max), int(nmu_bins), int(cosmology), C.byref(res),
                    C.byref(o), C.byref(extra))
        else:
            raise ValueError(mode)
    finally:
        if tmp:
            os.remove(path)
    if st != 0:
        return None
    if mode == "countpairs_mocks":
        n, m = res.nbin, res.npibin
        tot = (n + 1) * (m + 1)
        out = dict(nbin=n, npibin=m, pimax=res.pimax, npairs=_arr(res.npairs, tot, np.uint64),
                   rupp=_arr(res.rupp, n, np.float64), ravg=_arr(res.rpavg, tot, np.float64),
                   weightavg=_arr(res.weightavg, tot, np.float64))
    else:
        n, m = res.nsbin, res.nmu_bins
        tot = (n + 1) * (m + 1)
        out = dict(nbin=n, nmu_bins=m, mu_max=res.mu_max, npairs=_arr(res.npairs, tot, np.uint64),
                   rupp=_arr(res.supp, n, np.float64), ravg=_arr(res.savg, tot, np.float64),
                   weightavg=_arr(res.weightavg, tot, np.float64))
    out["api_time"] = o.c_api_time
    getattr(lib, prefix + _FREE[mode])(C.byref(res))
    return out
