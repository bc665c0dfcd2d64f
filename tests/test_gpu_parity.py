"""GPU engine vs the FP64 oracle on identical seeded inputs, through the C ABI.

npairs must be bit-exact; ravg / savg / weightavg / xi / wp within 1e-10 relative.
"""
import os

import numpy as np
import pytest

from corrfunc_amd import _abi, _lib
from helpers import MODES, assert_same, clustered, mode_kwargs, run, uniform, weights

pytestmark = pytest.mark.gpu

# engine variants: (kernel, CORRFUNC_AMD_DENSE) -- the tuned kernel picks its
# dense or sparse path per work item; 1 / 2 force one path for every item
ENGINES = {"tuned": (_lib.ENGINE_TUNED, "0"), "dense": (_lib.ENGINE_TUNED, "1"),
           "sparse": (_lib.ENGINE_TUNED, "2"), "plain": (_lib.ENGINE_PLAIN, "0")}


def _select(eng):
    kind, dense = ENGINES[eng]
    os.environ["CORRFUNC_AMD_DENSE"] = dense
    _lib.set_engine(kind)


def _opts(**kw):
    return _abi.make_options(**kw)


@pytest.mark.parametrize("eng", list(ENGINES))
@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
@pytest.mark.parametrize("isa", ["fallback", "avx", "avx2", "avx512f"])
def test_modes_uniform(engine, oracle, eng, mode, dtype, isa):
    """Every mode under every ISA's FP64 rounding order (xi under AVX/AVX2 sums
    dx^2 + (dy^2 + dz^2), wp under AVX2 uses an FMA; the rest as fallback or AVX512F)."""
    _select(eng)
    L = 200.0
    X = uniform(20000, L, 11, dtype)
    W = weights(20000, 12, dtype)
    bins = np.logspace(np.log10(0.5), np.log10(25.0), 15)
    for autocorr in ((1, 0) if mode not in ("countpairs_xi", "countpairs_wp") else (1,)):
        X2 = uniform(15000, L, 13, dtype)
        W2 = weights(15000, 14, dtype)
        kw = mode_kwargs(mode, X, W, X2, W2, autocorr=autocorr, bins=bins, L=L, pimax=20.0, mu_max=0.8, nmu_bins=8)
        o = dict(periodic=True, boxsize=L, need_avg_sep=True, isa=isa)
        got = run(engine, mode, options=_opts(**o), **kw)
        want = run(oracle, mode, "oracle_", options=_opts(**o), **kw)
        assert_same(mode, got, want, what="autocorr=%d" % autocorr)


@pytest.mark.parametrize("eng", list(ENGINES))
@pytest.mark.parametrize("mode", ["countpairs", "countpairs_rp_pi", "countpairs_s_mu"])
@pytest.mark.parametrize("periodic", [True, False])
@pytest.mark.parametrize("refine", [(2, 2, 1), (1, 1, 1), (3, 3, 2)])
def test_options_clustered(engine, oracle, eng, mode, periodic, refine):
    _select(eng)
    L = 300.0
    X = clustered(30000, L, 21, nparent=40, sigma=3.0)
    W = weights(30000, 22, np.float32)
    bins = np.concatenate([[0.0], np.logspace(-1, np.log10(30.0), 12)])
    kw = mode_kwargs(mode, X, W, autocorr=1, bins=bins, L=L, pimax=35.0, mu_max=1.0, nmu_bins=12)
    o = dict(periodic=periodic, boxsize=L if periodic else None, need_avg_sep=True, refine=refine, isa="fallback")
    got = run(engine, mode, options=_opts(**o), **kw)
    want = run(oracle, mode, "oracle_", options=_opts(**o), **kw)
    assert_same(mode, got, want)


@pytest.mark.parametrize("eng", list(ENGINES))
@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_device_resident_inputs(engine, oracle, eng, dtype):
    """Positions and weights already in HBM are used in place (same result as host inputs)."""
    import ctypes as C
    import os
    _select(eng)
    L = 150.0
    X = uniform(10000, L, 31, dtype)
    W = weights(10000, 32, dtype)
    bins = np.logspace(-0.5, np.log10(20.0), 10)
    dev = [_lib.DeviceArray(a) for a in X + [W]]
    o = _opts(periodic=True, boxsize=L, need_avg_sep=True, isa="fallback", float_type=np.dtype(dtype).itemsize)
    e = _abi.get_extra_options(_abi.WEIGHT_PAIR_PRODUCT)
    e.weights0.weights[0] = dev[3].ptr
    res = _abi.results_countpairs()
    path, tmp = _abi.write_bin_file(bins)
    try:
        st = engine.countpairs(len(X[0]), dev[0].ptr, dev[1].ptr, dev[2].ptr, 0, None, None, None, 1, 1,
                               path.encode(), C.byref(res), C.byref(o), C.byref(e))
    finally:
        os.remove(path)
    assert st == 0
    got = _abi._collect("countpairs", res)
    engine.free_results(C.byref(res))
    for a, d in zip(X + [W], dev):
        assert np.array_equal(a, d.numpy())  # inputs untouched
        d.free()
    want = run(oracle, "countpairs", "oracle_", options=_opts(periodic=True, boxsize=L, need_avg_sep=True,
                                                               isa="fallback"),
               autocorr=1, nthreads=2, binfile=bins, X1=X[0], Y1=X[1], Z1=X[2], weights1=W,
               weight_type="pair_product")
    assert_same("countpairs", got, want)


@pytest.mark.parametrize("drain", ["auto", "1", "2"])
@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_batch_matches_single_calls_and_oracle(engine, oracle, monkeypatch, weighted, dtype, drain):
    """cf_amd_countpairs_batch: each catalog equals countpairs() on it and the oracle
    (with either form of the batched kernel's survivor drain)."""
    _select("tuned")
    if drain != "auto":
        monkeypatch.setenv("CORRFUNC_AMD_DRAIN_PAIRS", drain)
    L = 120.0
    rng = np.random.default_rng(77)
    sizes = [int(v) for v in np.exp(rng.uniform(np.log(300), np.log(20000), 12))]
    cats = [uniform(n, L, 1000 + k, dtype) for k, n in enumerate(sizes)]
    cats[3] = [np.ascontiguousarray(a) for a in clustered(sizes[3], L, 5, nparent=6, sigma=1.5, dtype=dtype)]
    ws = [weights(n, 2000 + k, dtype) for k, n in enumerate(sizes)] if weighted else None
    bins = np.logspace(np.log10(0.5), np.log10(18.0), 12)
    o = dict(periodic=True, boxsize=L, need_avg_sep=True, isa="fallback")
    outs = _lib.countpairs_batch([tuple(c) for c in cats], bins, options=_opts(**o), weights=ws)
    if drain != "auto":
        assert _lib.last_stats()["drain_pairs"] == int(drain)
    for k, c in enumerate(cats):
        kw = dict(autocorr=1, nthreads=2, binfile=bins, X1=c[0], Y1=c[1], Z1=c[2])
        if weighted:
            kw.update(weights1=ws[k], weight_type="pair_product")
        single = run(engine, "countpairs", options=_opts(**o), **kw)
        want = run(oracle, "countpairs", "oracle_", options=_opts(**o), **kw)
        assert_same("countpairs", outs[k], want, what="catalog %d" % k)
        assert np.array_equal(outs[k]["npairs"], single["npairs"])


@pytest.mark.parametrize("nbytes", [1, 15, 16, 4096 + 7, (1 << 22) + 3])
def test_device_copy(engine, nbytes):
    """Device-to-device copies (the engine's copy kernel, with its byte tail) are exact."""
    src = np.random.default_rng(nbytes).integers(0, 256, nbytes, dtype=np.uint8)
    a = _lib.DeviceArray(src)
    b = _lib.DeviceArray(np.zeros(nbytes, np.uint8))
    b.copy_from(a)
    assert np.array_equal(b.numpy(), src)
    a.free()
    b.free()


# Engine variables of the tuned kernel's work list and paths (read on every call).
KNOBS = [
    {"CORRFUNC_AMD_FETCH": "4"},
    {"CORRFUNC_AMD_FETCH": "3", "CORRFUNC_AMD_HEAVY": "1e9"},
    {"CORRFUNC_AMD_DENSE_FRAC": "0"},
    {"CORRFUNC_AMD_DENSE_FRAC": "0", "CORRFUNC_AMD_STRIDE": "1"},
    {"CORRFUNC_AMD_DENSE_FRAC": "2", "CORRFUNC_AMD_PRIM_CACHE": "0"},
    {"CORRFUNC_AMD_PRIM_CACHE": "1", "CORRFUNC_AMD_HIST_COPIES": "1"},
    {"CORRFUNC_AMD_SPLIT": "64", "CORRFUNC_AMD_HEAVY": "1"},
    {"CORRFUNC_AMD_SPLIT": "0", "CORRFUNC_AMD_DENSE": "2"},
    {"CORRFUNC_AMD_QBLOCK_LOG2": "5"},
    {"CORRFUNC_AMD_QBLOCK_LOG2": "3", "CORRFUNC_AMD_FETCH": "4", "CORRFUNC_AMD_HEAVY": "1e9"},
    {"CORRFUNC_AMD_QBLOCK_LOG2": "12", "CORRFUNC_AMD_SPLIT": "64", "CORRFUNC_AMD_HEAVY": "1"},
    {"CORRFUNC_AMD_DRAIN_PAIRS": "2"},
    {"CORRFUNC_AMD_DRAIN_PAIRS": "2", "CORRFUNC_AMD_DENSE_FRAC": "2", "CORRFUNC_AMD_PRIM_CACHE": "0"},
    {"CORRFUNC_AMD_DRAIN_PAIRS": "2", "CORRFUNC_AMD_PRIM_CACHE": "1", "CORRFUNC_AMD_FETCH": "4"},
    {"CORRFUNC_AMD_DRAIN_PAIRS": "1", "CORRFUNC_AMD_HIST_COPIES": "1"},
    {"CORRFUNC_AMD_THREADS": "1024"},
    {"CORRFUNC_AMD_THREADS": "1024", "CORRFUNC_AMD_DRAIN_PAIRS": "1", "CORRFUNC_AMD_HIST_COPIES": "2"},
    {"CORRFUNC_AMD_THREADS": "1024", "CORRFUNC_AMD_DRAIN_PAIRS": "2", "CORRFUNC_AMD_DENSE_FRAC": "2"},
    {"CORRFUNC_AMD_THREADS": "256", "CORRFUNC_AMD_DRAIN_PAIRS": "2"},
]


@pytest.mark.parametrize("knobs", KNOBS, ids=lambda k: ",".join("%s=%s" % (a[13:], b) for a, b in k.items()))
@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("sched", ["none", "heuristic"])
def test_engine_variables_keep_counts(engine, oracle, knobs, mode, sched):
    """Fetch batches, sub-queue stealing, dense/sparse choice, the primaries cache,
    histogram replication, workgroup size, heavy-first order and window splitting
    never change a count."""
    from corrfunc_amd import scheduler
    _select("tuned")
    L = 160.0
    X = clustered(40000, L, 91, nparent=25, sigma=2.5, dtype=np.float64)
    W = weights(40000, 92, np.float64)
    bins = np.concatenate([[0.0], np.logspace(np.log10(0.2), np.log10(15.0), 11)])
    kw = mode_kwargs(mode, X, W, autocorr=1, bins=bins, L=L, pimax=12.0, mu_max=1.0, nmu_bins=9)
    o = dict(periodic=True, boxsize=L, need_avg_sep=True, isa="avx512f")
    scheduler.set_scheduler(sched)
    old = {k: os.environ.get(k) for k in knobs}
    try:
        os.environ.update(knobs)
        got = run(engine, mode, options=_opts(**o), **kw)
        if "CORRFUNC_AMD_DRAIN_PAIRS" in knobs:
            assert _lib.last_stats()["drain_pairs"] == int(knobs["CORRFUNC_AMD_DRAIN_PAIRS"])
        if "CORRFUNC_AMD_THREADS" in knobs:
            assert _lib.last_stats()["workgroup_threads"] == int(knobs["CORRFUNC_AMD_THREADS"])
    finally:
        for k, v in old.items():
            if v is None:
                os.environ.pop(k, None)
            else:
                os.environ[k] = v
        scheduler.set_scheduler("learned")
    want = run(oracle, mode, "oracle_", options=_opts(**o), **kw)
    assert_same(mode, got, want, what=str(knobs))


def _lattice(n_side, spacing, offset, jitter, seed):
    """Points on a cubic lattice (pairs exactly at multiples of the spacing),
    optionally displaced by a tiny jitter, shifted by `offset`."""
    g = np.arange(n_side, dtype=np.float64) * spacing
    X = np.array(np.meshgrid(g, g, g, indexing="ij")).reshape(3, -1) + offset
    if jitter:
        X = X + np.random.default_rng(seed).uniform(-jitter, jitter, X.shape)
    return X


# (name, points, bins, box, periodic, pimax, max_cells, refine): each case puts
# many pairs at the quantized filter's thresholds or stretches its quantum
FILTER_CASES = [
    # pairs exactly at rmax, at the bin edges and at pimax (spacing 2.5)
    ("edges_exact", lambda: _lattice(16, 2.5, 0.0, 0.0, 0), np.arange(1, 8) * 2.5, 40.0, True, 5.0, 100, (2, 2, 1)),
    # the same, a few ulps off the edges, far from the origin (non-periodic)
    ("edges_jitter_far", lambda: _lattice(16, 2.5, 1.0e4, 1e-12 * 1e4, 1), np.arange(1, 9) * 2.5, None, False,
     5.0, 100, (2, 2, 1)),
    # lower bound far from zero: only a thin shell counts
    ("thin_shell", lambda: uniform(20000, 100.0, 3), np.array([9.0, 9.5, 10.0]), 100.0, True, 10.0, 100,
     (2, 2, 1)),
    # cells much larger than rmax (coarse quantum)
    ("big_cells", lambda: uniform(20000, 100.0, 4), np.logspace(-1.5, np.log10(0.8), 6), 100.0, True, 3.0, 4,
     (1, 1, 1)),
    # rmax close to half the box, cells smaller than rmax
    ("small_cells", lambda: uniform(6000, 50.0, 5), np.linspace(1.0, 24.0, 7), 50.0, True, 24.0, 100, (3, 3, 3)),
]


@pytest.mark.parametrize("eng", ["tuned", "sparse"])
@pytest.mark.parametrize("mode", ["countpairs", "countpairs_rp_pi", "countpairs_s_mu"])
@pytest.mark.parametrize("case", FILTER_CASES, ids=[c[0] for c in FILTER_CASES])
def test_filter_boundaries(engine, oracle, eng, mode, case):
    """npairs bit-exact where pairs sit on the separation thresholds of the
    integer pair filter (rmax, rmin, pimax) and where its quantum is coarse."""
    _select(eng)
    name, make, bins, box, periodic, pimax, max_cells, refine = case
    X = np.asarray(make())
    n = X.shape[1]
    W = weights(n, 7, np.float64)
    kw = mode_kwargs(mode, X, W, autocorr=1, bins=bins, L=box, pimax=pimax, mu_max=1.0, nmu_bins=5)
    o = dict(periodic=periodic, boxsize=box if periodic else None, need_avg_sep=True, isa="avx512f",
             max_cells_per_dim=max_cells, refine=refine)
    got = run(engine, mode, options=_opts(**o), **kw)
    want = run(oracle, mode, "oracle_", options=_opts(**o), **kw)
    assert_same(mode, got, want, what=name)
    if name == "edges_exact":
        assert want["npairs"].sum() > 0
