"""Pins of the CPU oracle (oracle/cf_oracle.c) -- no GPU needed.

1. golden vectors of the reference library compiled from its own sources
   (tests/golden/ref_cases.json, written by tests/golden/make_golden.py);
2. the known-answer tables of the reference's doctests
   (tests/golden/doctest_tables.json; Corrfunc/theory/*.py docstrings);
3. brute force over all N^2 pairs (restates Corrfunc/tests/test_theory.py:197-286);
4. explicit answers (test_theory.py:115-194);
5. invariants (cross(D, D) == auto(D) for rmin = 0, refine-factor invariance,
   periodic box larger than the data == non-periodic);
6. the compiled reference itself on fresh random inputs, where it is built here.
"""
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

from corrfunc_amd import _abi  # noqa: E402
from helpers import MODES, assert_same, mode_kwargs, rel_err, run, uniform, weights  # noqa: E402
import make_golden  # noqa: E402

GOLDEN = json.load(open(os.path.join(HERE, "golden", "ref_cases.json")))
DOCTEST = json.load(open(os.path.join(HERE, "golden", "doctest_tables.json")))
BINS = os.path.join(HERE, "golden", "bins")


def check_golden(out, rec, what):
    s = np.array(rec["slots"])
    got = out["npairs"][s]
    want = np.array(rec["npairs"], dtype=np.uint64)
    assert np.array_equal(got, want), (what, np.nonzero(got != want)[0][:5])
    for k in ("ravg", "weightavg"):
        w = np.array([float.fromhex(v) for v in rec[k]])
        assert rel_err(out[k][s], w) <= 1e-10, (what, k)
    for k in ("xi", "wp"):
        if k in rec:
            w = np.array([float.fromhex(v) for v in rec[k]])
            scale = 1.0 if k == "xi" else 2.0 * rec["case"]["pimax"]
            assert np.all(np.abs(out[k][1:] - w) <= 1e-10 * (scale + np.abs(w))), (what, k)


@pytest.mark.parametrize("idx", range(len(GOLDEN)))
def test_oracle_golden(oracle, idx):
    rec = GOLDEN[idx]
    out = make_golden.call(oracle, rec["case"], prefix="oracle_")
    check_golden(out, rec, rec["case"])


def _doctest_inputs(N):
    np.random.seed(42)
    X = np.random.uniform(0, 420.0, N)
    Y = np.random.uniform(0, 420.0, N)
    Z = np.random.uniform(0, 420.0, N)
    return X, Y, Z


DOCTEST_MODE = {"DD": "countpairs", "DDrppi": "countpairs_rp_pi", "DDsmu": "countpairs_s_mu",
                "xi": "countpairs_xi", "wp": "countpairs_wp"}


def doctest_rows(lib, prefix, name):
    """Runs the doctest call of `name` and returns its printed rows (same columns)."""
    spec = DOCTEST[name]
    X, Y, Z = _doctest_inputs(spec["N"])
    mode = DOCTEST_MODE[name]
    kw = dict(nthreads=4, binfile=BINS, X1=X, Y1=Y, Z1=Z, weights1=np.ones_like(X), weight_type="pair_product")
    if mode in ("countpairs_xi", "countpairs_wp"):
        kw["boxsize"] = 420.0
    else:
        kw["autocorr"] = 1
    if "pimax" in spec:
        kw["pimax"] = spec["pimax"]
    if mode == "countpairs_s_mu":
        kw.update(mu_max=spec["mu_max"], nmu_bins=spec["nmu_bins"])
    o = _abi.make_options(periodic=True, boxsize=420.0, need_avg_sep=spec["output_ravg"], isa="fallback")
    out = run(lib, mode, prefix, options=o, **kw)
    n = out["nbin"]
    rupp = out["rupp"]
    rows = []
    if mode in ("countpairs", "countpairs_xi", "countpairs_wp"):
        for i in range(1, n):
            r = [rupp[i - 1], rupp[i], out["ravg"][i]]
            if mode == "countpairs_xi":
                r.append(out["xi"][i])
            if mode == "countpairs_wp":
                r.append(out["wp"][i])
            rows.append(r + [int(out["npairs"][i]), out["weightavg"][i]])
    else:
        m = out["npibin"] if mode == "countpairs_rp_pi" else out["nmu_bins"]
        d = (out["pimax"] / m) if mode == "countpairs_rp_pi" else (out["mu_max"] / m)
        for i in range(1, n):
            for j in range(m):
                idx = i * (m + 1) + j
                rows.append([rupp[i - 1], rupp[i], out["ravg"][idx], (j + 1) * d, int(out["npairs"][idx]),
                             out["weightavg"][idx]])
    return rows[spec["first_row"]:]


def compare_doctest(rows, spec, what):
    assert len(rows) == len(spec["rows"]), what
    for got, want in zip(rows, spec["rows"]):
        for g, w, col in zip(got, want, spec["cols"]):
            if col == "npairs":
                assert int(g) == w, (what, col, got, want)
            else:
                assert abs(g - w) <= 5.1e-7 + 1e-9 * abs(w), (what, col, got, want)


@pytest.mark.parametrize("name", ["DD", "DDrppi", "DDsmu", "xi", "wp"])
def test_oracle_doctest_tables(oracle, name):
    compare_doctest(doctest_rows(oracle, "oracle_", name), DOCTEST[name], name)


# ------------------------------------------------------------ brute force
def brute_counts(mode, pos, boxsize, periodic, bins, pimax, mu_max, nmu_bins):
    pdiff = np.abs(pos[:, np.newaxis] - pos)
    if periodic:
        mask = pdiff >= boxsize / 2
        pdiff -= mask * boxsize
    if mode == "countpairs_rp_pi":
        sqr_rp = (pdiff[:, :, :2] ** 2).sum(axis=-1).reshape(-1)
        pi = np.abs(pdiff[:, :, 2]).reshape(-1)
        pibins = np.linspace(0.0, pimax, int(pimax) + 1)
        c, _, _ = np.histogram2d(sqr_rp, pi, bins=(bins ** 2, pibins))
        return c.reshape(-1)
    if mode == "countpairs_s_mu":
        s = np.sqrt((pdiff ** 2).sum(axis=-1).reshape(-1))
        s[s == 0.0] = np.inf
        mu = np.abs(pdiff[:, :, 2]).reshape(-1) / s
        c, _, _ = np.histogram2d(s, mu, bins=(bins, np.linspace(0, mu_max, nmu_bins + 1)))
        return c.reshape(-1)
    c, _ = np.histogram((pdiff ** 2).sum(axis=-1).reshape(-1), bins=bins ** 2)
    return c


def brute_case(lib, prefix, mode, autocorr, binref, min_sep_opt, maxcells, boxsize, periodic):
    np.random.seed(1234)
    npts = 100
    eps = 0.2
    boxsize = np.array(boxsize)
    bins = np.linspace(0.01, 0.49 * boxsize.min(), 20) if periodic else np.linspace(0.01, 2 * boxsize.max(), 20)
    pimax = np.floor(0.49 * boxsize.min())
    mu_max, nmu_bins = 0.5, 10
    pos = np.random.uniform(low=-eps, high=eps, size=(npts, 3)) * boxsize
    pos[npts // 2:] += boxsize / 2.0
    pos %= boxsize
    want = brute_counts(mode, pos, boxsize, periodic, bins, pimax, mu_max, nmu_bins)
    o = _abi.make_options(periodic=periodic, boxsize=tuple(np.broadcast_to(boxsize, 3)), isa="fallback",
                          refine=(binref, binref, binref), max_cells_per_dim=maxcells,
                          enable_min_sep_opt=min_sep_opt)
    kw = dict(autocorr=autocorr, nthreads=2, binfile=bins, X1=pos[:, 0].copy(), Y1=pos[:, 1].copy(),
              Z1=pos[:, 2].copy())
    if not autocorr:
        kw.update(X2=pos[:, 0].copy(), Y2=pos[:, 1].copy(), Z2=pos[:, 2].copy())
    if mode == "countpairs_rp_pi":
        kw["pimax"] = float(pimax)
    if mode == "countpairs_s_mu":
        kw.update(mu_max=mu_max, nmu_bins=nmu_bins)
    out = run(lib, mode, prefix, options=o, **kw)
    assert np.any(want > 0)
    if mode == "countpairs":
        got = out["npairs"][1:]
    else:
        m = out["npibin"] if mode == "countpairs_rp_pi" else out["nmu_bins"]
        got = np.array([out["npairs"][i * (m + 1) + j] for i in range(1, out["nbin"]) for j in range(m)])
    return got, want


BRUTE = [(m, a, r, mso, mc, b, p) for m in ("countpairs", "countpairs_rp_pi", "countpairs_s_mu") for a in (0, 1)
         for r in (1, 2, 3) for mso in (False, True) for mc in (1, 2, 3) for b in (123.0, (51.0, 75.0, 123.0))
         for p in (False, True)]


@pytest.mark.parametrize("mode,autocorr,binref,mso,maxcells,boxsize,periodic", BRUTE[::3])
def test_oracle_brute(oracle, mode, autocorr, binref, mso, maxcells, boxsize, periodic):
    got, want = brute_case(oracle, "oracle_", mode, autocorr, binref, mso, maxcells, boxsize, periodic)
    assert np.array_equal(got, want)


def explicit_cases(lib, prefix):
    """Corrfunc/tests/test_theory.py:115-194 (narrow extent, pairs across the wrap)."""
    out = []
    # narrow extent: N=2 -> [2, 0]; N=1 -> [0, 0]
    for N, pos, want in ((2, np.array([[0., 0.], [0., 0.], [0., 0.5]]), [2, 0]),
                         (1, np.array([[0.1], [0.2], [0.3]]), [0, 0])):
        o = _abi.make_options(periodic=True, boxsize=(3.0, 3.0, 3.0), isa="fallback")
        r = run(lib, "countpairs", prefix, options=o, autocorr=1, nthreads=1, binfile=np.array([0.2, 0.6, 1.0]),
                X1=pos[0], Y1=pos[1], Z1=pos[2])
        out.append((list(r["npairs"][1:]), want, "narrow N=%d" % N))
    for autocorr in (0, 1):
        for binref in (1, 2, 3):
            for mso in (False, True):
                for maxcells in (1, 2, 3):
                    L = 432.0
                    for bins, pos, want in ((np.array([0.01, 0.4]) * L, np.array([[0.02, 0.98], [0., 0.], [0., 0.]]) * L, [2]),
                                            (np.array([0.2, 0.3, 0.49]) * L, np.array([[0., 0.], [0., 0.], [0., 0.48]]) * L, [0, 2])):
                        o = _abi.make_options(periodic=True, boxsize=L, isa="fallback", max_cells_per_dim=maxcells,
                                              refine=(binref, binref, binref), enable_min_sep_opt=mso)
                        r = run(lib, "countpairs", prefix, options=o, autocorr=autocorr, nthreads=1, binfile=bins,
                                X1=pos[0].copy(), Y1=pos[1].copy(), Z1=pos[2].copy(), X2=pos[0].copy(),
                                Y2=pos[1].copy(), Z2=pos[2].copy())
                        out.append((list(r["npairs"][1:]), want, "dup a%d r%d m%d c%d" % (autocorr, binref, mso, maxcells)))
    return out


def test_oracle_explicit(oracle):
    for got, want, what in explicit_cases(oracle, "oracle_"):
        assert [int(v) for v in got] == want, what


# --------------------------------------------------------------- invariants
@pytest.mark.parametrize("mode", ["countpairs", "countpairs_rp_pi", "countpairs_s_mu"])
def test_oracle_cross_self_equals_auto(oracle, mode):
    """countpairs_impl.c.src:625-633: with rmin = 0, cross(D, D) == auto(D) (self pairs included)."""
    L = 60.0
    X = uniform(1500, L, 7)
    W = weights(1500, 8, np.float64)
    bins = np.array([0.0, 1.0, 2.0, 4.0, 8.0])
    o = dict(periodic=True, boxsize=L, need_avg_sep=True, isa="fallback")
    a = run(oracle, mode, "oracle_", options=_abi.make_options(**o),
            **mode_kwargs(mode, X, W, autocorr=1, bins=bins, pimax=8.0, mu_max=1.0, nmu_bins=4))
    c = run(oracle, mode, "oracle_", options=_abi.make_options(**o),
            **mode_kwargs(mode, X, W, X, W, autocorr=0, bins=bins, pimax=8.0, mu_max=1.0, nmu_bins=4))
    s = _abi.reported_slots(mode, a)
    if mode == "countpairs_s_mu":
        # DDsmu: the cross call drops the N zero-separation self pairs through its
        # mu cut (dz^2 >= s^2 mu_max^2 holds for s = 0), the auto call adds them to
        # the first valid slot (s_mu_impl:628-636)
        diff = a["npairs"].astype(np.int64) - c["npairs"].astype(np.int64)
        expect = np.zeros_like(diff)
        expect[a["nmu_bins"] + 1] = 1500
        assert np.array_equal(diff[s], expect[s])
        return
    assert np.array_equal(a["npairs"][s], c["npairs"][s])
    assert rel_err(a["ravg"][s], c["ravg"][s]) < 1e-12


@pytest.mark.parametrize("mode", MODES)
def test_oracle_refine_invariance(oracle, mode):
    """tests_common.h:72-137: results do not depend on bin refine factors / min-sep option."""
    L = 80.0
    X = uniform(3000, L, 9)
    bins = np.logspace(-0.5, np.log10(15.0), 8)
    base = None
    for refine in ((1, 1, 1), (2, 2, 1), (3, 2, 3)):
        for mso in (True, False):
            o = _abi.make_options(periodic=True, boxsize=L, need_avg_sep=True, isa="fallback", refine=refine,
                                  enable_min_sep_opt=mso)
            r = run(oracle, mode, "oracle_", options=o,
                    **mode_kwargs(mode, X, bins=bins, L=L, pimax=15.0, mu_max=0.8, nmu_bins=6))
            if base is None:
                base = r
            else:
                assert_same(mode, r, base, tol=1e-12)


@pytest.mark.parametrize("mode", ["countpairs", "countpairs_rp_pi", "countpairs_s_mu"])
def test_oracle_big_box_equals_nonperiodic(oracle, mode):
    """test_theory.py:109-112: a periodic box much larger than the data equals no wrapping."""
    X = uniform(2000, 50.0, 10)
    bins = np.logspace(-0.5, np.log10(12.0), 8)
    kw = mode_kwargs(mode, X, bins=bins, autocorr=1, pimax=12.0, mu_max=1.0, nmu_bins=5)
    a = run(oracle, mode, "oracle_", options=_abi.make_options(periodic=True, boxsize=500.0, isa="fallback"), **kw)
    b = run(oracle, mode, "oracle_", options=_abi.make_options(periodic=False, isa="fallback"), **kw)
    assert_same(mode, a, b)


def test_oracle_error_paths(oracle):
    """Reference failure triggers return EXIT_FAILURE (None here)."""
    X = uniform(100, 10.0, 1)
    kw = dict(autocorr=1, nthreads=1, X1=X[0], Y1=X[1], Z1=X[2])
    # rmax >= L/2 with periodic wrap (gridlink_utils.c.src:37-41)
    assert run(oracle, "countpairs", "oracle_", options=_abi.make_options(periodic=True, boxsize=10.0),
               binfile=np.array([0.1, 6.0]), **kw) is None
    # periodic without a boxsize (countpairs_impl.c.src:231-235)
    assert run(oracle, "countpairs", "oracle_", options=_abi.make_options(periodic=True, boxsize=None),
               binfile=np.array([0.1, 1.0]), **kw) is None
    # DDsmu mu_max outside (0, 1] and nmu_bins < 1 (s_mu_impl:214-223)
    for mu, nmu in ((0.0, 5), (1.5, 5), (0.5, 0)):
        assert run(oracle, "countpairs_s_mu", "oracle_", options=_abi.make_options(periodic=True, boxsize=10.0),
                   binfile=np.array([0.1, 1.0]), mu_max=mu, nmu_bins=nmu, **kw) is None
    # wrong version string (countpairs.c:52-55)
    o = _abi.make_options(periodic=True, boxsize=10.0)
    o.version = b"1.0"
    assert run(oracle, "countpairs", "oracle_", options=o, binfile=np.array([0.1, 1.0]), **kw) is None
    # xi: particles outside [0, boxsize] (gridlink_impl.c.src:183)
    assert run(oracle, "countpairs_xi", "oracle_", options=_abi.make_options(), nthreads=1,
               binfile=np.array([0.1, 1.0]), boxsize=5.0, X1=X[0], Y1=X[1], Z1=X[2]) is None


# ---------------------------------------------------- compiled reference
@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("isa", ["fallback", "avx512f", "avx"])
def test_oracle_vs_compiled_reference(oracle, reference, mode, isa):
    L = 90.0
    for dtype, seed in ((np.float64, 41), (np.float32, 42)):
        X = uniform(2500, L, seed, dtype)
        W = weights(2500, seed + 1, dtype)
        X2 = uniform(2000, L, seed + 2, dtype)
        W2 = weights(2000, seed + 3, dtype)
        bins = np.concatenate([[0.0], np.logspace(-0.5, np.log10(20.0), 9)])
        for autocorr in ((1, 0) if mode not in ("countpairs_xi", "countpairs_wp") else (1,)):
            kw = mode_kwargs(mode, X, W, X2, W2, autocorr=autocorr, bins=bins, L=L, pimax=17.0, mu_max=0.75,
                             nmu_bins=9)
            kw64 = {k: (v.astype(np.float64) if hasattr(v, "dtype") and v.dtype == np.float32 else v)
                    for k, v in kw.items()}
            o = dict(periodic=True, boxsize=L, need_avg_sep=True, isa=isa)
            want = run(reference, mode, options=_abi.make_options(**o), **kw64)
            got = run(oracle, mode, "oracle_", options=_abi.make_options(**o), **kw)
            assert_same(mode, got, want, what="%s autocorr=%d" % (dtype.__name__, autocorr))
