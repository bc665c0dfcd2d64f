"""Histogram handling of the GPU engine against the FP64 oracle.

* histograms of any size: (nbin+1)*(n2+1) slots as the reference allocates
  (theory/DDrppi/countpairs_rp_pi_impl.c.src:207, theory/DDsmu/countpairs_s_mu_impl.c.src:226),
  in LDS where one copy fits the workgroup, else in global memory;
* the sums of separations and weight products: wide-range, negative and zero
  weights, separations far below their bin's upper edge, coincident points;
* the survivor queue's window-relative indices (windows forced short);
* wp's c_api_time in nanoseconds (theory/wp/countpairs_wp_impl.c.src:669-673);
* concurrent calls from two host threads.

npairs bit-exact; averages within 1e-10 relative unless a test states otherwise.
"""
import os
import threading

import numpy as np
import pytest

from corrfunc_amd import _abi, _lib
from helpers import assert_same, clustered, mode_kwargs, rel_err, run, uniform, weights

pytestmark = pytest.mark.gpu

# (pair kernel, dense mode, survivors per lane per drain or None for the engine's choice)
ENGINES = {"tuned": (_lib.ENGINE_TUNED, "0", None), "dense": (_lib.ENGINE_TUNED, "1", None),
           "sparse": (_lib.ENGINE_TUNED, "2", None), "sparse_two": (_lib.ENGINE_TUNED, "2", "2"),
           "sparse_one": (_lib.ENGINE_TUNED, "2", "1"), "plain": (_lib.ENGINE_PLAIN, "0", None)}


@pytest.fixture
def select(monkeypatch):
    def sel(eng):
        kind, dense, drain = ENGINES[eng]
        monkeypatch.setenv("CORRFUNC_AMD_DENSE", dense)
        if drain:
            monkeypatch.setenv("CORRFUNC_AMD_DRAIN_PAIRS", drain)
        else:
            monkeypatch.delenv("CORRFUNC_AMD_DRAIN_PAIRS", raising=False)
        _lib.set_engine(kind)
    yield sel
    _lib.set_engine(_lib.ENGINE_TUNED)


def _opts(**kw):
    return _abi.make_options(**kw)


# (mode, bins, pimax / nmu_bins, where the tuned kernel's histogram lives)
LARGE = [
    # 41 edges x pimax 100: 42 x 101 = 4,242 slots (101 KiB: one LDS copy)
    ("countpairs_rp_pi", np.logspace(-1, np.log10(30.0), 41), 100.0, "lds"),
    # 101 edges x 100 mu bins: 102 x 101 = 10,302 slots (241 KiB: global memory)
    ("countpairs_s_mu", np.linspace(0.0, 30.0, 101), 100, "global"),
    # 3,001 edges of a 1-D histogram (70 KiB: one LDS copy, the plain kernel's global)
    ("countpairs", np.logspace(-1, np.log10(30.0), 3001), None, "lds"),
    # 10,001 edges of a 1-D histogram: 235 KiB of slots, global memory
    ("countpairs_xi", np.logspace(-1, np.log10(30.0), 10001), None, "global"),
    # 20,001 edges: the edges alone (8 B each) exceed a workgroup's LDS next to
    # the queues: the call falls back to the plain kernel
    ("countpairs", np.logspace(-1, np.log10(30.0), 20001), None, "plain"),
]


@pytest.mark.parametrize("eng", list(ENGINES))
@pytest.mark.parametrize("case", LARGE, ids=[c[0] + "_" + c[3] for c in LARGE])
def test_large_histograms(engine, oracle, select, eng, case):
    mode, bins, extra, where = case
    select(eng)
    # pimax = 100 must stay below half the box; a smaller uniform catalog where
    # the oracle's top-down search walks 20,000 edges
    L = 250.0
    if where == "plain":
        X = uniform(6000, L, 43)
    else:
        X = clustered(40000, L, 41, nparent=30, sigma=3.0, dtype=np.float64)
    W = weights(len(X[0]), 42, np.float64)
    kw = mode_kwargs(mode, X, W, autocorr=1, bins=bins, L=L,
                     pimax=extra if mode == "countpairs_rp_pi" else None,
                     mu_max=1.0, nmu_bins=extra if mode == "countpairs_s_mu" else 10)
    o = dict(periodic=True, boxsize=L, need_avg_sep=True, isa="fallback")
    got = run(engine, mode, options=_opts(**o), **kw)
    st = _lib.last_stats()
    want = run(oracle, mode, "oracle_", options=_opts(**o), **kw)
    assert_same(mode, got, want)
    assert want["npairs"].sum() > 0
    if eng == "plain" or where == "plain":
        assert st["engine"] == _lib.ENGINE_PLAIN
    else:
        assert st["engine"] == _lib.ENGINE_TUNED
        assert st["global_hist"] == (1 if where == "global" else 0)
        assert st["hist_copies"] == (0 if where == "global" else 1)


@pytest.mark.parametrize("eng", ["tuned", "dense", "sparse", "sparse_two"])
@pytest.mark.parametrize("mode", ["countpairs", "countpairs_rp_pi", "countpairs_s_mu"])
@pytest.mark.parametrize("wkind", ["wide", "signed", "zero", "outlier"])
def test_fixed_point_weight_sums(engine, oracle, select, eng, mode, wkind):
    """Weight products of any range and sign. With signed weights a bin's sum
    can cancel, so it is compared against 1e-10 of the sum of |w_i w_j|."""
    select(eng)
    L = 150.0
    n = 20000
    X = clustered(n, L, 51, nparent=20, sigma=2.0, dtype=np.float64)
    rng = np.random.default_rng(52)
    if wkind == "wide":
        W = np.exp(rng.uniform(np.log(1e-3), np.log(1e3), n))
    elif wkind == "signed":
        W = rng.uniform(-1.5, 1.5, n)
    elif wkind == "zero":
        W = np.zeros(n)
    else:
        W = rng.uniform(0.5, 1.5, n)
        W[123] = 1e6
    bins = np.concatenate([[0.0], np.logspace(-1, np.log10(12.0), 10)])
    kw = mode_kwargs(mode, X, W, autocorr=1, bins=bins, L=L, pimax=12.0, mu_max=1.0, nmu_bins=6)
    o = dict(periodic=True, boxsize=L, need_avg_sep=True, isa="fallback")
    got = run(engine, mode, options=_opts(**o), **kw)
    want = run(oracle, mode, "oracle_", options=_opts(**o), **kw)
    s = _abi.reported_slots(mode, want)
    assert np.array_equal(got["npairs"][s], want["npairs"][s])
    assert rel_err(got["ravg"][s], want["ravg"][s]) <= 1e-10
    kwa = dict(kw, weights1=np.abs(W))
    absw = run(oracle, mode, "oracle_", options=_opts(**o), **kwa)
    bound = 1e-10 * absw["weightavg"][s] + 1e-300
    assert np.all(np.abs(got["weightavg"][s] - want["weightavg"][s]) <= bound), wkind
    if wkind != "signed":
        assert rel_err(got["weightavg"][s], want["weightavg"][s]) <= 1e-10


@pytest.mark.parametrize("mode", ["countpairs", "countpairs_s_mu"])
def test_separations_far_below_their_bin_edge(engine, oracle, select, mode):
    """A first bin [0, 5) holding pairs at separations of 1e-7 .. 1e-3 next to
    ordinary ones, and coincident points."""
    select("tuned")
    L = 100.0
    rng = np.random.default_rng(61)
    base = rng.uniform(0, L, (4000, 3))
    d = np.exp(rng.uniform(np.log(1e-7), np.log(1e-3), (4000, 1))) * rng.normal(size=(4000, 3))
    pos = np.concatenate([base, (base + d) % L, base[:500]])
    X = [pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy()]
    W = weights(len(pos), 62, np.float64)
    bins = np.linspace(0.0, 20.0, 5)
    kw = mode_kwargs(mode, X, W, autocorr=1, bins=bins, L=L, mu_max=1.0, nmu_bins=4)
    o = dict(periodic=True, boxsize=L, need_avg_sep=True, isa="fallback")
    got = run(engine, mode, options=_opts(**o), **kw)
    want = run(oracle, mode, "oracle_", options=_opts(**o), **kw)
    assert_same(mode, got, want)


@pytest.mark.parametrize("knobs", [{"CORRFUNC_AMD_MAX_WINDOW": "8"},
                                   {"CORRFUNC_AMD_MAX_WINDOW": "100", "CORRFUNC_AMD_DENSE": "2"},
                                   {"CORRFUNC_AMD_MAX_WINDOW": "4096", "CORRFUNC_AMD_DENSE": "2"},
                                   {"CORRFUNC_AMD_MAX_WINDOW": "100", "CORRFUNC_AMD_DENSE": "2",
                                    "CORRFUNC_AMD_DRAIN_PAIRS": "2"},
                                   {"CORRFUNC_AMD_MAX_WINDOW": "8", "CORRFUNC_AMD_DRAIN_PAIRS": "1"}],
                         ids=["window_8", "window_100_sparse", "window_4096_sparse", "window_100_two_per_lane",
                              "window_8_one_per_lane"])
@pytest.mark.parametrize("mode", ["countpairs", "countpairs_rp_pi", "countpairs_wp", "countpairs_s_mu"])
def test_short_windows(engine, oracle, select, monkeypatch, knobs, mode):
    """Secondary windows scanned in short sub-windows (each with its own queue
    base, drained at its end) keep the counts and sums."""
    select("tuned")
    for k, v in knobs.items():
        monkeypatch.setenv(k, v)
    L = 150.0
    X = clustered(30000, L, 71, nparent=25, sigma=2.5, dtype=np.float32)
    W = weights(30000, 72, np.float32)
    bins = np.logspace(-1, np.log10(14.0), 12)
    kw = mode_kwargs(mode, X, W, autocorr=1, bins=bins, L=L, pimax=14.0)
    o = dict(periodic=True, boxsize=L, need_avg_sep=True, isa="avx512f")
    got = run(engine, mode, options=_opts(**o), **kw)
    want = run(oracle, mode, "oracle_", options=_opts(**o), **kw)
    assert_same(mode, got, want, what=str(knobs))


def test_wp_api_time_in_nanoseconds(engine):
    """wp reports c_api_time in ns, the other modes in seconds, as the reference."""
    L = 200.0
    X = uniform(50000, L, 81)
    bins = np.logspace(-1, np.log10(20.0), 10)
    o = _opts(periodic=True, boxsize=L, c_api_timer=True, isa="fallback")
    wp = run(engine, "countpairs_wp", options=o, **mode_kwargs("countpairs_wp", X, bins=bins, L=L, pimax=20.0))
    o = _opts(periodic=True, boxsize=L, c_api_timer=True, isa="fallback")
    dd = run(engine, "countpairs", options=o, **mode_kwargs("countpairs", X, bins=bins, L=L))
    assert 1e3 < wp["api_time"] < 1e12, wp["api_time"]
    assert 0.0 < dd["api_time"] < 100.0, dd["api_time"]
    assert wp["api_time"] > 1e3 * dd["api_time"]


def test_two_threads(engine, oracle):
    """Calls from two host threads at once: each equals the oracle (the entry
    points serialise on the engine)."""
    L = 150.0
    jobs = []
    for k, mode in enumerate(["countpairs", "countpairs_rp_pi", "countpairs_s_mu", "countpairs_xi"] * 2):
        X = clustered(25000 + 1000 * k, L, 90 + k, nparent=20, sigma=2.0, dtype=np.float64)
        W = weights(len(X[0]), 190 + k, np.float64)
        bins = np.logspace(-1, np.log10(12.0), 9)
        kw = mode_kwargs(mode, X, W, autocorr=1, bins=bins, L=L, pimax=12.0, nmu_bins=5)
        jobs.append((mode, kw))
    o = dict(periodic=True, boxsize=L, need_avg_sep=True, isa="fallback")
    got = [None] * len(jobs)

    def worker(idx):
        for i in idx:
            mode, kw = jobs[i]
            got[i] = run(engine, mode, options=_opts(**o), **kw)

    threads = [threading.Thread(target=worker, args=(range(t, len(jobs), 2),)) for t in range(2)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    for (mode, kw), g in zip(jobs, got):
        want = run(oracle, mode, "oracle_", options=_opts(**o), **kw)
        assert_same(mode, g, want)
