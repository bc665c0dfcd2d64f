"""The tuned kernel's work queue hands out 4 items per atomic when the work list
has many small items (>= 100 per resident wave, <= 8 per cell pair), as a
uniform catalog of a million points gives, and 1 where heavy cell pairs were
split into many items (a clustered catalog). At the million-point size the
tuned kernel is checked against the plain FP64 kernel (which matches the oracle
case by case elsewhere): npairs equal, averages within 1e-10; the clustered
case is checked against the oracle."""
import numpy as np
import pytest

from corrfunc_amd import _abi, _lib
from helpers import assert_same, clustered, mode_kwargs, rel_err, run, uniform, weights

pytestmark = pytest.mark.gpu


def _call(engine, kind, X, bins, L):
    _lib.set_engine(kind)
    o = _abi.make_options(periodic=True, boxsize=L, need_avg_sep=True, isa="avx512f")
    r = run(engine, "countpairs", options=o, nthreads=4, binfile=bins, X1=X[0], Y1=X[1], Z1=X[2], autocorr=1)
    return r, _lib.last_stats()


def test_many_small_items_tuned_equals_plain(engine, monkeypatch):
    monkeypatch.setenv("CORRFUNC_AMD_DENSE", "0")
    monkeypatch.delenv("CORRFUNC_AMD_FETCH", raising=False)
    L = 420.0
    X = uniform(1048576, L, 42, np.float64)
    bins = np.logspace(-1, np.log10(25.0), 21)
    got, st = _call(engine, _lib.ENGINE_TUNED, X, bins, L)
    try:
        want, _ = _call(engine, _lib.ENGINE_PLAIN, X, bins, L)
    finally:
        _lib.set_engine(_lib.ENGINE_TUNED)
    # the list this catalog gives is in the batched regime, and the kernel took 4 per atomic
    assert st["num_work_items"] <= 8 * st["num_cell_pairs"]
    assert st["queue_fetch"] == 4, st
    s = _abi.reported_slots("countpairs", want)
    assert np.array_equal(got["npairs"][s], want["npairs"][s])
    assert want["npairs"][s].sum() > 5e8
    assert rel_err(got["ravg"][s], want["ravg"][s]) <= 1e-10


def test_split_heavy_cell_pairs_fetch_one(engine, oracle, monkeypatch):
    """A clustered catalog whose heavy cell pairs are split into many items:
    the queue hands out one item per atomic."""
    monkeypatch.setenv("CORRFUNC_AMD_DENSE", "0")
    monkeypatch.delenv("CORRFUNC_AMD_FETCH", raising=False)
    from corrfunc_amd import scheduler
    scheduler.set_scheduler("heuristic")
    try:
        _lib.set_engine(_lib.ENGINE_TUNED)
        L = 300.0
        X = clustered(200000, L, 13, nparent=40, sigma=2.0, dtype=np.float32)
        W = weights(200000, 14, np.float32)
        bins = np.logspace(-1, np.log10(20.0), 13)
        kw = mode_kwargs("countpairs_rp_pi", X, W, autocorr=1, bins=bins, L=L, pimax=20.0)
        o = dict(periodic=True, boxsize=L, need_avg_sep=True, isa="avx512f")
        got = run(engine, "countpairs_rp_pi", options=_abi.make_options(**o), **kw)
        st = _lib.last_stats()
    finally:
        scheduler.set_scheduler("learned")
    assert st["num_work_items"] > 8 * st["num_cell_pairs"], st
    assert st["queue_fetch"] == 1, st
    want = run(oracle, "countpairs_rp_pi", "oracle_", options=_abi.make_options(**o), **kw)
    assert_same("countpairs_rp_pi", got, want)
