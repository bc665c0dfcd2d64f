"""The tuned kernel's work queue hands out 4 items per atomic when the work list
has many small items (>= 100 per resident wave, <= 8 per cell pair), as a
uniform catalog of a million points gives. At that size the tuned kernel is
checked against the plain FP64 kernel (which matches the oracle case by case
elsewhere): npairs equal, averages within 1e-10."""
import os

import numpy as np
import pytest

from corrfunc_amd import _abi, _lib
from helpers import rel_err, run, uniform

pytestmark = pytest.mark.gpu


def _call(engine, kind, X, bins, L):
    os.environ["CORRFUNC_AMD_DENSE"] = "0"
    _lib.set_engine(kind)
    o = _abi.make_options(periodic=True, boxsize=L, need_avg_sep=True, isa="avx512f")
    r = run(engine, "countpairs", options=o, nthreads=4, binfile=bins, X1=X[0], Y1=X[1], Z1=X[2], autocorr=1)
    return r, _lib.last_stats()


def test_many_small_items_tuned_equals_plain(engine):
    L = 420.0
    X = uniform(1048576, L, 42, np.float64)
    bins = np.logspace(-1, np.log10(25.0), 21)
    got, st = _call(engine, _lib.ENGINE_TUNED, X, bins, L)
    try:
        want, _ = _call(engine, _lib.ENGINE_PLAIN, X, bins, L)
    finally:
        _lib.set_engine(_lib.ENGINE_TUNED)
    # the list this catalog gives is in the batched regime
    assert st["num_work_items"] <= 8 * st["num_cell_pairs"]
    assert st["num_work_items"] >= 100 * 256 * 24  # 24 resident waves per CU at most
    s = _abi.reported_slots("countpairs", want)
    assert np.array_equal(got["npairs"][s], want["npairs"][s])
    assert want["npairs"][s].sum() > 5e8
    assert rel_err(got["ravg"][s], want["ravg"][s]) <= 1e-10
