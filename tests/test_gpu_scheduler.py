"""Work-list schedulers and the learned cost model on the GPU.

* the MLP's forward and backward kernels against the FP64 CPU oracle
  (oracle/cost_model.py), FP32 tolerance;
* training on the GPU lowers the loss;
* counts are bit-identical under every scheduler (none / heuristic / learned,
  including a deliberately bad model that reorders and splits at random);
* the A/B: a model trained on per-cell-pair timings the engine recorded on
  clustered catalogs predicts the cost of held-out catalogs better than the
  N1*N2 heuristic;
* wp's c_cell_timer rows.
"""
import os
import sys

import numpy as np
import pytest

from corrfunc_amd import _abi, _lib, catalogs, scheduler
from helpers import MODES, assert_same, clustered, mode_kwargs, run, weights

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import cost_model as cm  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _restore():
    _lib.set_engine(_lib.ENGINE_TUNED)
    os.environ["CORRFUNC_AMD_DENSE"] = "0"
    yield
    scheduler.record(False)
    scheduler.reset()
    scheduler.set_scheduler("learned")


def _problem(n, seed):
    rng = np.random.default_rng(seed)
    X = (rng.normal(0, 1, (n, cm.F)) * rng.uniform(0.5, 3, cm.F) + rng.uniform(-2, 5, cm.F)).astype(np.float32)
    y = rng.normal(10, 3, n).astype(np.float32)
    p = rng.normal(0, 0.3, cm.NPARAM).astype(np.float32)
    return p, cm.make_norm(X, y).astype(np.float32), X, y


@pytest.mark.parametrize("n", [1, 63, 64, 65, 1000, 4099])
def test_model_gradient_matches_oracle(engine, n):
    p, norm, X, y = _problem(n, n)
    loss, g = scheduler.grad(X, y, p, norm)
    L, gw = cm.loss_and_grad(p.astype(np.float64), norm.astype(np.float64), X, y)
    assert loss == pytest.approx(L, rel=2e-5)
    scale = max(np.abs(gw).max(), 1e-30)
    assert np.abs(g - gw).max() <= 2e-5 * scale, np.abs(g - gw).argmax()


def test_model_predict_matches_oracle(engine):
    p, norm, X, y = _problem(2000, 7)
    scheduler.set_params(p, norm)
    got = scheduler.predict(X)
    want = cm.predict(p.astype(np.float64), norm.astype(np.float64), X)
    assert np.abs(got - want).max() <= 1e-4 * max(1.0, np.abs(want).max())
    q, nm = scheduler.params()
    assert np.array_equal(q, p) and np.array_equal(nm, norm)


@pytest.mark.parametrize("step", [1, 2, 37])
def test_adam_step_matches_oracle(engine, step):
    """One k_adam update (the training kernel's) against the FP64 oracle's
    adam_step, FP32 tolerance."""
    rng = np.random.default_rng(step)
    p = rng.normal(0, 0.5, cm.NPARAM).astype(np.float32)
    g = (rng.normal(0, 1, cm.NPARAM) * np.exp(rng.uniform(-6, 2, cm.NPARAM))).astype(np.float32)
    m = (rng.normal(0, 0.1, cm.NPARAM) if step > 1 else np.zeros(cm.NPARAM)).astype(np.float32)
    v = (rng.uniform(0, 0.1, cm.NPARAM) if step > 1 else np.zeros(cm.NPARAM)).astype(np.float32)
    gp, gm, gv = scheduler.adam_step(p, g, m, v, step, 3e-3)
    p64, g64, m64, v64 = (a.astype(np.float64) for a in (p, g, m, v))
    wp, wm, wv = cm.adam_step(p64, g64, m64, v64, step, 3e-3)
    # moments: each constant and each product rounded once to FP32, then the sum
    # (the terms may cancel, so the bound is on the terms, not on the result)
    u = 2.0 ** -24
    assert np.all(np.abs(gm - wm) <= 4 * u * (0.9 * np.abs(m64) + 0.1 * np.abs(g64)) + 1e-30)
    assert np.all(np.abs(gv - wv) <= 5 * u * (0.999 * v64 + 0.001 * g64 * g64) + 1e-30)
    # the update from the kernel's own moments: FP32 rounding of the bias
    # corrections, products, square root, divide (about 8 roundings) and of the
    # new parameter
    mh = gm.astype(np.float64) / (1 - 0.9 ** step)
    vh = gv.astype(np.float64) / (1 - 0.999 ** step)
    upd = 3e-3 * mh / (np.sqrt(vh) + 1e-8)
    assert np.all(np.abs((p64 - gp) - upd) <= 12 * u * np.abs(upd) + 2 * u * np.abs(p64))
    # where the moments do not cancel, the parameters follow the oracle's update
    ok = np.abs(wm) > 0.5 * (0.9 * np.abs(m64) + 0.1 * np.abs(g64))
    assert ok.sum() > cm.NPARAM // 2
    assert np.all(np.abs(gp - wp)[ok] <= 24 * u * np.abs(p64 - wp)[ok] + 2 * u * np.abs(p64)[ok])


def test_training_lowers_the_loss(engine):
    rng = np.random.default_rng(11)
    n = 20000
    X = rng.uniform(0, 10, (n, cm.F)).astype(np.float32)
    y = (0.7 * X[:, 0] + 0.9 * X[:, 1] + 2.0 * np.maximum(X[:, 7] - 5, 0) + rng.normal(0, 0.1, n)).astype(np.float32)
    loss = scheduler.train(X, y, epochs=30, batch=256, lr=3e-3, seed=1)
    # standardised loss: 0.5 for the best constant predictor
    assert loss < 0.01
    p, nm = scheduler.params()
    assert cm.loss(p, nm, X, y) == pytest.approx(loss, rel=1e-3)
    pred = scheduler.predict(X)
    assert np.sqrt(np.mean((pred - y) ** 2)) < 0.5


def _random_model(seed):
    p, norm, _, _ = _problem(10, seed)
    return p, norm


@pytest.mark.parametrize("mode", MODES)
def test_counts_identical_under_every_scheduler(engine, oracle, mode):
    L = 120.0
    X = clustered(30000, L, 41, nparent=12, sigma=1.5, dtype=np.float64)
    W = weights(30000, 42, np.float64)
    bins = np.logspace(np.log10(0.3), np.log10(18.0), 12)
    kw = mode_kwargs(mode, X, W, autocorr=1, bins=bins, L=L, pimax=15.0, mu_max=1.0, nmu_bins=8)
    o = dict(periodic=True, boxsize=L, need_avg_sep=True, isa="avx512f")
    outs = {}
    for name in ("none", "heuristic", "learned-random", "learned-trained"):
        if name.startswith("learned"):
            scheduler.set_scheduler("learned")
            if name == "learned-random":
                scheduler.set_params(*_random_model(5))
        else:
            scheduler.set_scheduler(name)
        outs[name] = run(engine, mode, options=_abi.make_options(**o), **kw)
        assert _lib.last_stats()["sched"] == {"none": 0, "heuristic": 1}.get(name, 2), name
        if name == "learned-random":
            # train a model on what this call records, for the last variant
            scheduler.record(True)
            scheduler.clear()
            run(engine, mode, options=_abi.make_options(**o), **kw)
            scheduler.record(False)
            Xs, ys = scheduler.samples()
            assert len(ys) == _lib.last_stats()["num_cell_pairs"] or len(ys) > 0
            scheduler.train(Xs, ys, epochs=5, batch=128)
    want = run(oracle, mode, "oracle_", options=_abi.make_options(**o), **kw)
    for name, got in outs.items():
        assert_same(mode, got, want, what=name)
        assert np.array_equal(got["npairs"], outs["none"]["npairs"]), name


def _record_xi(cat, bins, L):
    scheduler.clear()
    scheduler.record(True)
    r = run(_lib.lib(), "countpairs_xi", nthreads=1, binfile=bins, X1=cat[0], Y1=cat[1], Z1=cat[2], boxsize=L)
    scheduler.record(False)
    return r, scheduler.samples()


def test_learned_cost_model_beats_the_heuristic(engine):
    """Record per-cell-pair timings on three clustered (NFW halo) catalogs, train
    on the GPU, and predict the timings of a fourth, held-out catalog."""
    L = 300.0
    bins = np.logspace(np.log10(0.1), np.log10(20.0), 16)
    trainX, trainy = [], []
    for seed in (101, 102, 103):
        cat = catalogs.nfw_halos(n=400000, L=L, nhalo=1500, seed=seed)
        _, (Xs, ys) = _record_xi(cat, bins, L)
        trainX.append(Xs)
        trainy.append(ys)
    Xt, yt = np.concatenate(trainX), np.concatenate(trainy)
    loss = scheduler.train(Xt, yt, epochs=10, batch=512, lr=3e-3, seed=0)
    held = catalogs.nfw_halos(n=400000, L=L, nhalo=1500, seed=104)
    _, (Xh, yh) = _record_xi(held, bins, L)
    ev = scheduler.evaluate(Xh, yh, offset_from=(Xt, yt), workers=2048)
    print("\ncost model A/B on a held-out catalog (%d cell pairs, train loss %.4f): %s" % (len(yh), loss, ev))
    assert ev["rmse_learned"] < ev["rmse_heuristic"]
    # simulated makespan of the engine's work items (heavy-first order and
    # splitting) with each estimate: both orders far better than list order,
    # the model's no worse than the heuristic's
    assert ev["items_imbalance_heuristic"] < 0.5 * ev["items_imbalance_none"]
    assert ev["items_imbalance_learned"] <= ev["items_imbalance_heuristic"] + 0.01
    # kernel time under each scheduler on the held-out catalog (median of 3), same counts
    res = {}
    for name in ("none", "heuristic", "learned"):
        scheduler.set_scheduler(name)
        ms = []
        for _ in range(3):
            r = run(_lib.lib(), "countpairs_xi", nthreads=1, binfile=bins, X1=held[0], Y1=held[1], Z1=held[2],
                    boxsize=L)
            ms.append(_lib.last_stats()["kernel_ms"])
        res[name] = (r["npairs"], float(np.median(ms)))
    print("kernel ms by scheduler: %s" % {k: round(v[1], 3) for k, v in res.items()})
    assert np.array_equal(res["none"][0], res["learned"][0])
    assert np.array_equal(res["none"][0], res["heuristic"][0])
    # measured: list order about 2.1-2.6 ms, either cost order about 1.5 ms
    assert res["heuristic"][1] < 0.9 * res["none"][1]
    assert res["learned"][1] < 0.9 * res["none"][1]


def test_wp_cell_timer(engine):
    L = 200.0
    X = clustered(50000, L, 77, nparent=20, sigma=2.0)
    bins = np.logspace(-1, np.log10(15.0), 10)
    o = _abi.make_options(c_cell_timer=True, isa="fallback")
    out = _abi.call_mode(_lib.lib(), "countpairs_wp", nthreads=1, binfile=bins, X1=X[0], Y1=X[1], Z1=X[2],
                         pimax=20.0, boxsize=L, options=o)
    st = _lib.last_stats()
    t = out["cell_timings"]
    assert len(t) == st["num_cell_pairs"] > 0
    same = t["cellidx1"] == t["cellidx2"]
    n1, n2 = t["N1"].astype(np.float64), t["N2"].astype(np.float64)
    assert np.sum(np.where(same, n1 * (n1 - 1) / 2, n1 * n2)) == st["tested_pairs"]
    ncells = st["nmesh"][0] * st["nmesh"][1] * st["nmesh"][2]
    assert t["cellidx1"].min() >= 0 and t["cellidx1"].max() < ncells and t["cellidx2"].max() < ncells
    assert (t["time_in_ns"] >= 0).all() and (t["time_in_ns"] > 0).mean() > 0.9
    assert ((t["tid"] >= 0) | (t["time_in_ns"] == 0)).all()
    # the npairs do not depend on the timer
    plain = _abi.call_mode(_lib.lib(), "countpairs_wp", nthreads=1, binfile=bins, X1=X[0], Y1=X[1], Z1=X[2],
                           pimax=20.0, boxsize=L, options=_abi.make_options(isa="fallback"))
    assert np.array_equal(plain["npairs"], out["npairs"])
    assert o.totncells_timings == 0  # released


@pytest.mark.parametrize("mode", ["countpairs", "countpairs_rp_pi", "countpairs_xi"])
def test_autotuner_choice_keeps_counts(engine, oracle, mode, tmp_path):
    """The autotuner's refine factors give the oracle's counts for those factors
    (and the default factors' counts); the second call reuses the cached choice,
    also from the cache file in a fresh in-process cache."""
    L = 150.0
    X = clustered(40000, L, 61, nparent=30, sigma=2.0, dtype=np.float64)
    W = weights(40000, 62, np.float64)
    bins = np.logspace(np.log10(0.3), np.log10(12.0), 10)
    kw = mode_kwargs(mode, X, W, autocorr=1, bins=bins, L=L, pimax=10.0)
    o = dict(periodic=True, boxsize=L, need_avg_sep=True, isa="fallback")
    cache = str(tmp_path / "tune.txt")
    os.environ["CORRFUNC_AMD_TUNE_CACHE"] = cache
    try:
        _lib.autotune_clear()
        _lib.set_autotune(True)
        got = run(engine, mode, options=_abi.make_options(**o), **kw)
        st = _lib.last_stats()
        assert st["autotuned"] == 2
        chosen = st["refine"]
        got2 = run(engine, mode, options=_abi.make_options(**o), **kw)
        assert _lib.last_stats()["autotuned"] == 1 and _lib.last_stats()["refine"] == chosen
        _lib.autotune_clear()
        got3 = run(engine, mode, options=_abi.make_options(**o), **kw)
        assert _lib.last_stats()["autotuned"] == 1 and _lib.last_stats()["refine"] == chosen
        assert len(open(cache).read().strip().splitlines()) == 1
    finally:
        _lib.set_autotune(False)
        _lib.autotune_clear()
        del os.environ["CORRFUNC_AMD_TUNE_CACHE"]
    oc = _abi.make_options(**o)
    for i in range(3):
        oc.bin_refine_factors[i] = chosen[i]
    oc.binning_flags = (oc.binning_flags & ~0xF) | 0x1  # BINNING_CUST: exactly these factors
    want = run(oracle, mode, "oracle_", options=oc, **kw)
    want_dflt = run(oracle, mode, "oracle_", options=_abi.make_options(**o), **kw)
    for g in (got, got2, got3):
        assert_same(mode, g, want, what="refine %s" % (chosen,))
        assert_same(mode, g, want_dflt, what="default refine")


def test_autotuned_batch(engine, oracle):
    L = 200.0
    cats = [clustered(n, L, 70 + k, nparent=20, sigma=3.0) for k, n in enumerate([3000, 20000, 9000, 60000])]
    bins = np.logspace(np.log10(0.5), np.log10(20.0), 12)
    o = dict(periodic=True, boxsize=L, need_avg_sep=True, isa="avx512f")
    _lib.autotune_clear()
    _lib.set_autotune(True)
    try:
        outs = _lib.countpairs_batch(cats, bins, options=_abi.make_options(**o))
        assert _lib.last_stats()["autotuned"] == 2
    finally:
        _lib.set_autotune(False)
        _lib.autotune_clear()
    for k, c in enumerate(cats):
        want = run(oracle, "countpairs", "oracle_", options=_abi.make_options(**o), autocorr=1, nthreads=2,
                   binfile=bins, X1=c[0], Y1=c[1], Z1=c[2])
        assert_same("countpairs", outs[k], want, what="catalog %d" % k)
