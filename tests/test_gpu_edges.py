"""Edge inputs of the GPU engine against the oracle (which matches the compiled
reference on each of them): empty and single-particle sets, coincident points
with and without rmin = 0, particles on the box faces, a one-cell grid, large
coordinate offsets, and the error paths (both return failure)."""
import os

import numpy as np
import pytest

from corrfunc_amd import _abi, _lib
from helpers import MODES, assert_same, mode_kwargs, run

pytestmark = pytest.mark.gpu

L = 50.0
BINS = np.logspace(-1, 1, 6)
BINS0 = np.concatenate([[0.0], BINS])


def _select(eng):
    os.environ["CORRFUNC_AMD_DENSE"] = "0"
    _lib.set_engine(_lib.ENGINE_TUNED if eng == "tuned" else _lib.ENGINE_PLAIN)


def _both(engine, oracle, mode, X, X2=None, W=None, W2=None, bins=BINS, opts=None, **extra):
    kw = mode_kwargs(mode, X, W, X2, W2, autocorr=1 if X2 is None else 0, bins=bins, L=L, pimax=10.0, mu_max=1.0,
                     nmu_bins=4)
    kw.update(extra)
    o = opts or dict(periodic=True, boxsize=L, need_avg_sep=True, isa="fallback")
    got = run(engine, mode, options=_abi.make_options(**o), **kw)
    want = run(oracle, mode, "oracle_", options=_abi.make_options(**o), **kw)
    return got, want


def _pts(*rows):
    a = np.array(rows, dtype=np.float64)
    return [a[:, 0].copy(), a[:, 1].copy(), a[:, 2].copy()]


@pytest.mark.parametrize("eng", ["tuned", "plain"])
@pytest.mark.parametrize("mode", MODES)
def test_tiny_sets(engine, oracle, eng, mode):
    _select(eng)
    empty = [np.zeros(0), np.zeros(0), np.zeros(0)]
    one = _pts((1.0, 2.0, 3.0))
    two = _pts((1.0, 2.0, 3.0), (1.0, 2.0, 3.0))
    many = _pts(*([(7.5, 8.5, 9.5)] * 50 + [(7.5, 8.5, 9.6)] * 3))
    for name, X, bins in (("empty", empty, BINS), ("one", one, BINS), ("two coincident, rmin 0", two, BINS0),
                          ("53 near-coincident, rmin 0", many, BINS0), ("53 near-coincident", many, BINS)):
        got, want = _both(engine, oracle, mode, X, bins=bins)
        assert_same(mode, got, want, what=name)
    if mode not in ("countpairs_xi", "countpairs_wp"):
        rng = np.random.default_rng(1)
        X = [rng.uniform(0, L, 300) for _ in range(3)]
        got, want = _both(engine, oracle, mode, X, X2=empty)
        assert_same(mode, got, want, what="cross with an empty set")
        got, want = _both(engine, oracle, mode, empty, X2=X)
        assert_same(mode, got, want, what="cross from an empty set")


@pytest.mark.parametrize("eng", ["tuned", "plain"])
@pytest.mark.parametrize("mode", MODES)
def test_box_faces_and_one_cell(engine, oracle, eng, mode):
    _select(eng)
    rng = np.random.default_rng(2)
    X = [rng.uniform(0, L, 2000) for _ in range(3)]
    for d in range(3):
        X[d][:40] = 0.0
        X[d][40:80] = L if mode not in ("countpairs_xi", "countpairs_wp") else np.nextafter(L, 0)
    W = rng.uniform(0.5, 1.5, 2000)
    got, want = _both(engine, oracle, mode, X, W=W, weight_type="pair_product")
    assert_same(mode, got, want, what="faces")
    o = dict(periodic=True, boxsize=L, need_avg_sep=True, isa="fallback", max_cells_per_dim=1)
    got, want = _both(engine, oracle, mode, X, W=W, weight_type="pair_product", opts=o)
    assert_same(mode, got, want, what="max_cells_per_dim=1")


@pytest.mark.parametrize("eng", ["tuned", "plain"])
@pytest.mark.parametrize("mode", ["countpairs", "countpairs_rp_pi", "countpairs_s_mu"])
def test_large_offsets_non_periodic(engine, oracle, eng, mode):
    _select(eng)
    rng = np.random.default_rng(3)
    X = [1.0e6 + rng.uniform(0, 40, 3000), -5.0e5 + rng.uniform(0, 40, 3000), rng.uniform(-20, 20, 3000)]
    o = dict(periodic=False, need_avg_sep=True, isa="fallback")
    got, want = _both(engine, oracle, mode, X, opts=o)
    assert_same(mode, got, want, what="offsets")
    assert want["npairs"].sum() > 0


@pytest.mark.parametrize("mode", MODES)
def test_error_paths(engine, oracle, mode, capfd):
    _select("tuned")
    rng = np.random.default_rng(4)
    X = [rng.uniform(0, L, 500) for _ in range(3)]
    # periodic with rmax >= L/2
    got, want = _both(engine, oracle, mode, X, bins=np.array([1.0, 30.0]))
    assert got is None and want is None
    # bins with rmin >= rmax are rejected
    got, want = _both(engine, oracle, mode, X, bins=np.array([5.0, 5.0]))
    assert got is None and want is None
    if mode in ("countpairs_xi", "countpairs_wp"):
        Xo = [a.copy() for a in X]
        Xo[0][3] = L + 10.0  # outside [0, boxsize]
        got, want = _both(engine, oracle, mode, Xo)
        assert got is None and want is None
    if mode == "countpairs_s_mu":
        for mu in (0.0, 1.5):
            kw = mode_kwargs(mode, X, bins=BINS, mu_max=mu, nmu_bins=4)
            o = _abi.make_options(periodic=True, boxsize=L, isa="fallback")
            assert run(engine, mode, options=o, **kw) is None
            assert run(oracle, mode, "oracle_", options=_abi.make_options(periodic=True, boxsize=L,
                                                                         isa="fallback"), **kw) is None
    capfd.readouterr()
