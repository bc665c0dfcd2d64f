"""The corrfunc_amd.theory entry points on the GPU against the reference's own
doctest tables (Corrfunc/theory/DD.py:166-202, DDrppi.py:183-249,
DDsmu.py:199-265, xi.py:148-182, wp.py:437-472): the same inputs
(RandomState(42) uniform in a 420 box, the reference's `bins` file, unit
weights), the same calls and the printed columns, plus what the tables do not
show: result dtypes, the pimax / mu_max upper-edge columns, the
(results, api_time) tuples, wp's cell-timer return, scalar and per-axis
boxsizes, and find_fastest_wp_bin_refs."""
import json
import os

import numpy as np
import pytest

from corrfunc_amd import _abi, _lib
from corrfunc_amd.theory import DD, DDrppi, DDsmu, wp, xi
from corrfunc_amd.theory.wp import find_fastest_wp_bin_refs
from helpers import run

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DOCTEST = json.load(open(os.path.join(HERE, "golden", "doctest_tables.json")))
BINS = os.path.join(HERE, "golden", "bins")


def _inputs(N):
    np.random.seed(42)
    X = np.random.uniform(0, 420.0, N)
    Y = np.random.uniform(0, 420.0, N)
    Z = np.random.uniform(0, 420.0, N)
    return X, Y, Z


def _compare(res, name):
    spec = DOCTEST[name]
    rows = res[spec["first_row"]:]
    assert len(rows) == len(spec["rows"]), name
    for r, want in zip(rows, spec["rows"]):
        for col, w in zip(spec["cols"], want):
            g = r[col]
            if col == "npairs":
                assert int(g) == w, (name, col, r, want)
            else:
                # the tables print 6 decimals (%10.6f) or 1 (pimax, mu_max columns)
                tol = 0.051 if col in ("pimax", "mu_max") and name in ("DDrppi", "DDsmu") else 5.1e-7
                assert abs(g - w) <= tol + 1e-9 * abs(w), (name, col, r, want)


def _fields(res, names):
    assert res.dtype.names == tuple(names)
    for n in names:
        assert res.dtype[n] == (np.uint64 if n == "npairs" else np.float64), n


@pytest.fixture(autouse=True)
def _tuned():
    _lib.set_engine(_lib.ENGINE_TUNED)


def test_DD_doctest(engine):
    X, Y, Z = _inputs(DOCTEST["DD"]["N"])
    res = DD(1, 4, BINS, X, Y, Z, weights1=np.ones_like(X), weight_type="pair_product", output_ravg=True,
             boxsize=420.0, periodic=True)
    _fields(res, ("rmin", "rmax", "ravg", "npairs", "weightavg"))
    _compare(res, "DD")
    # per-axis boxsize tuple and the api timer (seconds)
    res2, t = DD(1, 4, BINS, X, Y, Z, weights1=np.ones_like(X), weight_type="pair_product", output_ravg=True,
                 boxsize=(420.0, 420.0, 420.0), periodic=True, c_api_timer=True)
    assert np.array_equal(res2["npairs"], res["npairs"])
    assert 0.0 < t < 100.0


def test_DDrppi_doctest(engine):
    X, Y, Z = _inputs(DOCTEST["DDrppi"]["N"])
    pimax = DOCTEST["DDrppi"]["pimax"]
    res = DDrppi(1, 4, pimax, BINS, X, Y, Z, weights1=np.ones_like(X), weight_type="pair_product",
                 output_rpavg=True, boxsize=420.0, periodic=True)
    _fields(res, ("rmin", "rmax", "rpavg", "pimax", "npairs", "weightavg"))
    _compare(res, "DDrppi")
    # the pimax column holds the upper edge (j+1)*dpi of each pi bin, dpi = pimax/int(pimax)
    npibin = int(pimax)
    assert np.array_equal(res["pimax"][:npibin], (np.arange(npibin) + 1) * (pimax / npibin))
    assert len(res) == (len(np.loadtxt(BINS))) * npibin


def test_DDsmu_doctest(engine):
    X, Y, Z = _inputs(DOCTEST["DDsmu"]["N"])
    spec = DOCTEST["DDsmu"]
    res = DDsmu(1, 4, BINS, spec["mu_max"], spec["nmu_bins"], X, Y, Z, weights1=np.ones_like(X),
                weight_type="pair_product", output_savg=True, boxsize=420.0, periodic=True)
    _fields(res, ("smin", "smax", "savg", "mu_max", "npairs", "weightavg"))
    _compare(res, "DDsmu")
    m = spec["nmu_bins"]
    assert np.array_equal(res["mu_max"][:m], (np.arange(m) + 1) * (spec["mu_max"] / m))


def test_xi_doctest(engine):
    X, Y, Z = _inputs(DOCTEST["xi"]["N"])
    res = xi(420.0, 4, BINS, X, Y, Z, weights=np.ones_like(X), weight_type="pair_product", output_ravg=True)
    _fields(res, ("rmin", "rmax", "ravg", "xi", "npairs", "weightavg"))
    _compare(res, "xi")


def test_wp_doctest_and_timers(engine):
    X, Y, Z = _inputs(DOCTEST["wp"]["N"])
    pimax = DOCTEST["wp"]["pimax"]
    res = wp(420.0, pimax, 4, BINS, X, Y, Z, weights=np.ones_like(X), weight_type="pair_product")
    _fields(res, ("rmin", "rmax", "rpavg", "wp", "npairs", "weightavg"))
    _compare(res, "wp")
    # (results, api_time, cell_time): api time in ns, one timing row per cell pair
    r2, t_ns, cells = wp(420.0, pimax, 4, BINS, X, Y, Z, weights=np.ones_like(X), weight_type="pair_product",
                         c_api_timer=True, c_cell_timer=True)
    assert np.array_equal(r2["npairs"], res["npairs"])
    assert t_ns > 1e3
    st = _lib.last_stats()
    assert len(cells) == st["num_cell_pairs"] > 0
    assert cells.dtype.names == ("N1", "N2", "time_in_ns", "cellidx1", "cellidx2", "tid")
    r3, cells3 = wp(420.0, pimax, 4, BINS, X, Y, Z, c_cell_timer=True)
    assert np.array_equal(r3["npairs"], res["npairs"]) and len(cells3) == len(cells)


def test_python_entry_points_match_the_oracle(engine, oracle):
    """Cross-correlation and non-periodic calls through the Python layer equal
    the oracle's C-ABI results (the layer adds no arithmetic of its own)."""
    rng = np.random.default_rng(5)
    L = 250.0
    X1 = [rng.uniform(0, L, 30000) for _ in range(3)]
    X2 = [rng.uniform(0, L, 20000) for _ in range(3)]
    w1, w2 = rng.uniform(0.5, 1.5, 30000), rng.uniform(0.5, 1.5, 20000)
    bins = np.logspace(-1, np.log10(20.0), 12)
    res = DDsmu(0, 4, bins, 0.7, 7, *X1, weights1=w1, X2=X2[0], Y2=X2[1], Z2=X2[2], weights2=w2,
                weight_type="pair_product", output_savg=True, periodic=False, isa="fallback")
    want = run(oracle, "countpairs_s_mu", "oracle_",
               options=_abi.make_options(periodic=False, need_avg_sep=True, isa="fallback"),
               autocorr=0, nthreads=2, binfile=bins, X1=X1[0], Y1=X1[1], Z1=X1[2], X2=X2[0], Y2=X2[1],
               Z2=X2[2], weights1=w1, weights2=w2, weight_type="pair_product", mu_max=0.7, nmu_bins=7)
    m = 7
    idx = np.array([i * (m + 1) + j for i in range(1, len(bins)) for j in range(m)])
    assert np.array_equal(res["npairs"], want["npairs"][idx])
    assert np.allclose(res["savg"], want["ravg"][idx], rtol=1e-10, atol=0)
    assert np.allclose(res["weightavg"], want["weightavg"][idx], rtol=1e-10, atol=0)


def test_find_fastest_wp_bin_refs(engine):
    X, Y, Z = _inputs(50000)
    best, runtimes = find_fastest_wp_bin_refs(420.0, 40.0, 4, BINS, X, Y, Z, maxbinref=2, nrepeats=2,
                                              return_runtimes=True)
    assert len(best) == 3 and all(1 <= v <= 2 for v in best)
    assert runtimes.dtype.names == ("nx", "ny", "nz", "avg_time", "sigma_time")
    assert len(runtimes) == 8 and np.all(np.isfinite(runtimes["avg_time"]))
    assert (runtimes["nx"][0], runtimes["ny"][0], runtimes["nz"][0]) == best
    assert np.all(np.diff(runtimes["avg_time"]) >= 0)
    assert find_fastest_wp_bin_refs(420.0, 40.0, 4, BINS, X, Y, Z, maxbinref=1, nrepeats=1) == (1, 1, 1)
    # every triple counts the same pairs
    ref = wp(420.0, 40.0, 4, BINS, X, Y, Z)
    for nx, ny, nz in [(1, 1, 1), (2, 2, 2), best]:
        r = wp(420.0, 40.0, 4, BINS, X, Y, Z, xbin_refine_factor=nx, ybin_refine_factor=ny, zbin_refine_factor=nz)
        assert np.array_equal(r["npairs"], ref["npairs"])
