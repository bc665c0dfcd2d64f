"""Shared argument handling of the corrfunc_amd.theory functions.

Follows the reference Python layer (Corrfunc/utils.py:464-518
translate_isa_string_to_enum, :967-1027 process_weights, :324-381
return_file_with_rbins) and the option filling of the reference CPython
extension (theory/python_bindings/_countpairs.c:1151-1250).
"""
import numpy as np

from .. import _abi
from .._lib import lib

VALID_ISA = ("FALLBACK", "AVX512F", "AVX2", "AVX", "SSE42", "FASTEST")


def translate_isa_string_to_enum(isa):
    if not isinstance(isa, str):
        raise TypeError("Input to translate_isa_string_to_enum must be of string type. "
                        "Found type = {0}".format(type(isa)))
    key = isa.upper()
    if key not in VALID_ISA:
        raise ValueError("Do not know instruction type = {0}. Valid instruction sets are {1}".format(isa, VALID_ISA))
    return _abi.ISA[key.lower()]


def process_weights(weights1, weights2, X1, X2, weight_type, autocorr):
    if weight_type is None:
        return weights1, weights2

    def prep(w, x):
        if w is None:
            return w
        if isinstance(w, float):
            w = np.array(w, dtype=x.dtype)
        w = np.atleast_1d(w)
        if w.shape[-1] == 1:
            w = np.tile(w, len(x))
        return np.atleast_2d(w)

    weights1 = prep(weights1, X1)
    if not autocorr:
        weights2 = prep(weights2, X2)
        if (weights1 is None) != (weights2 is None):
            if weight_type != "pair_product":
                raise ValueError("If using a weight_type other than 'pair_product', you must provide "
                                 "both weight arrays.")
        if weights1 is None and weights2 is not None:
            weights1 = np.ones((len(weights2), len(X1)), dtype=X1.dtype)
        if weights2 is None and weights1 is not None:
            weights2 = np.ones((len(weights1), len(X2)), dtype=X2.dtype)
    return weights1, weights2


def boxsize_tuple(boxsize):
    if boxsize is None:
        return None
    b = np.atleast_1d(boxsize)
    if len(b) == 1:
        b = (b[0], b[0], b[0])
    return tuple(float(v) for v in b)


def run(mode, *, autocorr, nthreads, binfile, X1, Y1, Z1, weights1=None, X2=None, Y2=None, Z2=None,
        weights2=None, weight_type=None, periodic=True, boxsize=None, verbose=False, need_avg=False,
        refine=(2, 2, 1), max_cells_per_dim=100, copy_particles=True, enable_min_sep_opt=True,
        c_api_timer=False, isa="fastest", pimax=None, mu_max=None, nmu_bins=None,
        fast_divide_and_NR_steps=0, c_cell_timer=False):
    integer_isa = translate_isa_string_to_enum(isa)
    if weight_type is not None and weight_type not in ("pair_product", "p"):
        raise ValueError('unknown weight_type "{0}"'.format(weight_type))
    if weight_type is not None:
        w1 = np.atleast_2d(weights1) if weights1 is not None else None
        if w1 is not None and w1.shape[0] != 1:
            raise ValueError("specified weighting method {0} which requires 1 weight(s)-per-particle, but found "
                             "{1} weight(s) instead!".format(weight_type, w1.shape[0]))
    opts = _abi.make_options(periodic=periodic, boxsize=boxsize, verbose=verbose, need_avg_sep=need_avg,
                             refine=refine, max_cells_per_dim=max_cells_per_dim, copy_particles=copy_particles,
                             enable_min_sep_opt=enable_min_sep_opt, c_api_timer=c_api_timer, isa=integer_isa,
                             fast_divide_and_NR_steps=fast_divide_and_NR_steps, c_cell_timer=c_cell_timer)
    if boxsize is None and mode not in ("countpairs_xi", "countpairs_wp"):
        opts.boxsize_x = opts.boxsize_y = opts.boxsize_z = _abi.BOXSIZE_NOTGIVEN
    xb = None
    if mode in ("countpairs_xi", "countpairs_wp"):
        xb = float(np.atleast_1d(boxsize)[0])
    out = _abi.call_mode(lib(), mode, autocorr=autocorr, nthreads=nthreads, binfile=binfile, X1=X1, Y1=Y1, Z1=Z1,
                         weights1=weights1, X2=X2, Y2=Y2, Z2=Z2, weights2=weights2, weight_type=weight_type,
                         pimax=pimax, mu_max=mu_max, nmu_bins=nmu_bins, boxsize=xb, options=opts)
    if out is None:
        raise RuntimeError("RuntimeError occurred")
    return out
