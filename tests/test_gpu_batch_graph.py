"""The graph-captured batch (cf_amd_batch_graph_*): captured once, replayed on
new positions of the same catalog sizes written into the same device arrays
(bootstrap-like redraws, jackknife-like moves of a region), with eager engine
calls and synchronisations between the replays. Every replay equals a fresh
cf_amd_countpairs_batch call on the same arrays and the oracle (npairs
bit-exact, averages within 1e-10). Positions the capture's capacities cannot
hold make the launch fail with a message, and the next launch works again."""
import numpy as np
import pytest

from corrfunc_amd import _abi, _lib
from helpers import assert_same, clustered, run, uniform, weights

pytestmark = pytest.mark.gpu

L = 120.0
BINS = np.logspace(np.log10(0.5), np.log10(18.0), 12)


def _cats(sizes, seed, dtype):
    cats = []
    for k, n in enumerate(sizes):
        if k % 3 == 1:
            cats.append([np.ascontiguousarray(a) for a in clustered(n, L, seed + k, nparent=6, sigma=1.5,
                                                                     dtype=dtype)])
        else:
            cats.append(uniform(n, L, seed + k, dtype))
    return cats


def _check(engine, oracle, graph_out, dev, ws, cats, wts, o):
    fresh = _lib.countpairs_batch([tuple(d) for d in dev], BINS, options=_abi.make_options(**o), weights=ws)
    for k, c in enumerate(cats):
        assert np.array_equal(graph_out[k]["npairs"], fresh[k]["npairs"]), k
        kw = dict(autocorr=1, nthreads=2, binfile=BINS, X1=c[0], Y1=c[1], Z1=c[2])
        if wts is not None:
            kw.update(weights1=wts[k], weight_type="pair_product")
        want = run(oracle, "countpairs", "oracle_", options=_abi.make_options(**o), **kw)
        assert_same("countpairs", graph_out[k], want, what="catalog %d" % k)


@pytest.mark.parametrize("weighted", [False, True])
@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_graph_replays_resamples(engine, oracle, weighted, dtype):
    _lib.set_engine(_lib.ENGINE_TUNED)
    rng = np.random.default_rng(7)
    sizes = [int(v) for v in np.exp(rng.uniform(np.log(500), np.log(30000), 10))]
    cats = _cats(sizes, 100, dtype)
    wts = [weights(n, 300 + k, dtype) for k, n in enumerate(sizes)] if weighted else None
    dev = [[_lib.DeviceArray(a) for a in c] for c in cats]
    dws = [_lib.DeviceArray(w) for w in wts] if weighted else None
    o = dict(periodic=True, boxsize=L, need_avg_sep=True, isa="fallback")
    g = _lib.BatchGraph([tuple(d) for d in dev], BINS, options=_abi.make_options(**o), weights=dws)
    try:
        _check(engine, oracle, g.launch(), dev, dws, cats, wts, o)
        for it in range(3):
            # bootstrap-like: every catalog redrawn at its size; jackknife-like:
            # one region's points moved elsewhere (sizes kept)
            new = _cats(sizes, 1000 + 50 * it, dtype)
            if it == 1:
                for c in new:
                    sel = (c[0] < L / 4) & (c[1] < L / 4)
                    c[0][sel] += np.asarray(L / 2, dtype)
            neww = [weights(n, 2000 + 50 * it + k, dtype) for k, n in enumerate(sizes)] if weighted else None
            for d, c in zip(dev, new):
                for da, a in zip(d, c):
                    da.upload(a)
            if weighted:
                for dw, w in zip(dws, neww):
                    dw.upload(w)
            # eager engine work between the replays
            run(engine, "countpairs", options=_abi.make_options(**o), autocorr=1, nthreads=2, binfile=BINS,
                X1=new[0][0], Y1=new[0][1], Z1=new[0][2])
            out = g.launch()
            _check(engine, oracle, out, dev, dws, new, neww, o)
            again = g.launch()
            for a, b in zip(out, again):
                assert np.array_equal(a["npairs"], b["npairs"])
    finally:
        g.close()


def test_graph_rejects_what_it_cannot_hold(engine, oracle):
    _lib.set_engine(_lib.ENGINE_TUNED)
    sizes = [3000, 8000, 15000]
    cats = _cats(sizes, 40, np.float64)
    dev = [[_lib.DeviceArray(a) for a in c] for c in cats]
    o = dict(periodic=True, boxsize=L, need_avg_sep=True, isa="avx512f")
    g = _lib.BatchGraph([tuple(d) for d in dev], BINS, options=_abi.make_options(**o))
    try:
        _check(engine, oracle, g.launch(), dev, None, cats, None, o)
        # positions spread over three boxes: more cells than the capture holds
        wide = [[a * 3.0 for a in c] for c in cats]
        for d, c in zip(dev, wide):
            for da, a in zip(d, c):
                da.upload(a)
        with pytest.raises(RuntimeError):
            g.launch()
        # every catalog collapsed into one clump: more work items than captured
        tight = [[np.full(n, 5.0) + np.random.default_rng(k).uniform(0, 0.1, n) for _ in range(3)]
                 for k, n in enumerate(sizes)]
        for d, c in zip(dev, tight):
            for da, a in zip(d, c):
                da.upload(a)
        try:
            out = g.launch()
            _check(engine, oracle, out, dev, None, tight, None, o)
        except RuntimeError:
            pass
        # back to positions it holds: counted again
        for d, c in zip(dev, cats):
            for da, a in zip(d, c):
                da.upload(a)
        _check(engine, oracle, g.launch(), dev, None, cats, None, o)
    finally:
        g.close()
    # a non-periodic batch cannot be captured
    with pytest.raises(RuntimeError):
        _lib.BatchGraph([tuple(d) for d in dev], BINS, options=_abi.make_options(periodic=False))
