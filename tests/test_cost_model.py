"""The cost model's CPU oracle (oracle/cost_model.py), pinned by mathematics, and
the scheduler's host-side evaluation logic (no GPU)."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import cost_model as cm  # noqa: E402

from corrfunc_amd import scheduler  # noqa: E402


def _problem(n=40, seed=0):
    rng = np.random.default_rng(seed)
    X = rng.normal(0, 1, (n, cm.F)) * rng.uniform(0.5, 3, cm.F) + rng.uniform(-2, 5, cm.F)
    y = rng.normal(10, 3, n)
    p = rng.normal(0, 0.4, cm.NPARAM)
    return p, cm.make_norm(X, y), X, y


def test_gradient_matches_central_differences():
    """Every component of the reverse-mode gradient equals the central finite
    difference of the loss (FP64, h = 1e-6)."""
    p, norm, X, y = _problem()
    L, g = cm.loss_and_grad(p, norm, X, y)
    assert L == pytest.approx(cm.loss(p, norm, X, y), rel=1e-14)
    h = 1e-6
    fd = np.empty(cm.NPARAM)
    for k in range(cm.NPARAM):
        e = np.zeros(cm.NPARAM)
        e[k] = h
        fd[k] = (cm.loss(p + e, norm, X, y) - cm.loss(p - e, norm, X, y)) / (2 * h)
    assert np.max(np.abs(fd - g)) <= 1e-7 * max(1.0, np.max(np.abs(g)))
    # every parameter block is exercised
    for a, b in [(cm.OFF_W1, cm.OFF_B1), (cm.OFF_B1, cm.OFF_W2), (cm.OFF_W2, cm.OFF_B2), (cm.OFF_B2, cm.OFF_W3),
                 (cm.OFF_W3, cm.OFF_B3), (cm.OFF_B3, cm.NPARAM)]:
        assert np.any(g[a:b] != 0)


def test_linear_network_closed_form():
    """With W2 = I, b = 0 and inputs kept positive, the network is linear:
    o = W3 . (W1 x~), so its least-squares loss has the textbook value."""
    rng = np.random.default_rng(3)
    X = rng.uniform(1, 2, (30, cm.F))
    norm = np.concatenate([np.zeros(cm.F), np.ones(cm.F), [0.0, 1.0]])
    p = np.zeros(cm.NPARAM)
    W1 = np.abs(rng.normal(0, 1, (cm.H, cm.F)))
    p[cm.OFF_W1:cm.OFF_B1] = W1.ravel()
    p[cm.OFF_W2:cm.OFF_B2] = np.eye(cm.H).ravel()
    w3 = rng.normal(0, 1, cm.H)
    p[cm.OFF_W3:cm.OFF_B3] = w3
    y = rng.normal(0, 1, 30)
    o = X @ W1.T @ w3
    assert cm.predict(p, norm, X) == pytest.approx(o, rel=1e-12)
    assert cm.loss(p, norm, X, y) == pytest.approx(0.5 * np.mean((o - y) ** 2), rel=1e-12)


def test_adam_first_step_moves_by_lr_times_sign():
    """Bias correction makes Adam's first update -lr * g / (|g| + eps)."""
    g = np.array([3.0, -0.5, 1e-3, 0.0])
    p, m, v = cm.adam_step(np.zeros(4), g, np.zeros(4), np.zeros(4), 1, 0.01)
    assert p == pytest.approx(-0.01 * g / (np.abs(g) + 1e-8), rel=1e-12, abs=1e-15)
    assert m == pytest.approx(0.1 * g) and v == pytest.approx(0.001 * g * g)


def test_adam_second_step_by_hand():
    """Moments and bias correction of step 2, worked by hand: g = 2 then -1."""
    lr = 0.1
    p1, m1, v1 = cm.adam_step(np.zeros(1), np.array([2.0]), np.zeros(1), np.zeros(1), 1, lr)
    p2, m2, v2 = cm.adam_step(p1, np.array([-1.0]), m1, v1, 2, lr)
    # m1 = 0.2, v1 = 0.004, m2 = 0.18 - 0.1 = 0.08, v2 = 0.003996 + 0.001 = 0.004996
    assert m2[0] == pytest.approx(0.08, rel=1e-12) and v2[0] == pytest.approx(0.004996, rel=1e-12)
    mh, vh = 0.08 / (1 - 0.81), 0.004996 / (1 - 0.998001)
    assert p2[0] == pytest.approx(p1[0] - lr * mh / (np.sqrt(vh) + 1e-8), rel=1e-12)


def test_adam_minimises_a_quadratic():
    """Adam on f(p) = |p - c|^2 / 2 (gradient p - c) ends near c."""
    c = np.array([1.5, -2.0, 0.25])
    p, m, v = np.zeros(3), np.zeros(3), np.zeros(3)
    for step in range(1, 3001):
        p, m, v = cm.adam_step(p, p - c, m, v, step, 0.01)
    assert np.abs(p - c).max() < 1e-3


def test_make_norm_standardises():
    p, norm, X, y = _problem(200, 5)
    xs = (X - norm[:cm.F]) * norm[cm.F:2 * cm.F]
    assert np.abs(xs.mean(0)).max() < 1e-12
    assert np.abs(xs.std(0) - 1).max() < 1e-12
    Xc = X.copy()
    Xc[:, 4] = 1.0  # a constant feature is switched off, not divided by zero
    assert cm.make_norm(Xc, y)[cm.F + 4] == 0.0


def test_heuristic_cost_from_features():
    n1 = np.array([1, 5, 64, 1000, 7])
    n2 = np.array([9, 5, 100, 1, 7])
    same = np.array([0, 0, 0, 0, 1])
    X = np.zeros((5, scheduler.FEATURES))
    X[:, 0] = np.log2(1 + n1)
    X[:, 1] = np.log2(1 + n2)
    X[:, 3] = same
    want = np.log2(1 + np.where(same, n1 * (n1 - 1) / 2, n1 * n2))
    assert scheduler.heuristic_log2_cost(X) == pytest.approx(want, rel=1e-12)


def test_list_schedule_known_answers():
    # greedy in order (3, 3, 2, 2, 2) on 2 workers: 3|3, 5|5, 7 -> 7 (optimum 6)
    assert scheduler.list_schedule_makespan([3, 3, 2, 2, 2], [0, 1, 2, 3, 4], 2) == 7
    # heavy item last on 3 workers: 1,1,1,1,1,1 then 6 -> 2 + 6 = 8
    c = [1] * 6 + [6]
    assert scheduler.list_schedule_makespan(c, range(7), 3) == 8
    # heavy first: 6 | 1,1,1 | 1,1,1 -> 6
    assert scheduler.list_schedule_makespan(c, [6, 0, 1, 2, 3, 4, 5], 3) == 6


def test_evaluate_prefers_the_better_estimate():
    rng = np.random.default_rng(1)
    n = 3000
    X = np.zeros((n, scheduler.FEATURES))
    n1 = rng.integers(1, 500, n)
    n2 = rng.integers(1, 500, n)
    X[:, 0], X[:, 1] = np.log2(1 + n1), np.log2(1 + n2)
    y = np.log2(n1 * n2 * rng.uniform(0.2, 5.0, n))  # cost = N1*N2 * a factor the heuristic cannot see
    ev = scheduler.evaluate(X, y, pred=y + rng.normal(0, 0.05, n), workers=64)
    assert ev["rmse_learned"] < 0.1 < ev["rmse_heuristic"]
    assert ev["imbalance_learned"] <= ev["imbalance_heuristic"] + 1e-9
    assert ev["imbalance_oracle_order"] <= ev["imbalance_none"]


def test_engine_items_order_and_splits_by_hand():
    """One heavy cell pair (1000 of a total 1007, 4x the mean is 503.5) goes
    first and is split into ceil(1000 / (1007 / (4 x 2 workers))) = 8 items;
    the light ones keep list order, one item each (one 64-primary chunk)."""
    X = np.zeros((8, scheduler.FEATURES))
    est = np.log2([1.0, 1.0, 1000.0, 1.0, 1.0, 1.0, 1.0, 1.0])
    order, nitems = scheduler.engine_items(X, est, workers=2)
    assert list(order) == [2, 0, 1, 3, 4, 5, 6, 7]
    assert list(nitems) == [8, 1, 1, 1, 1, 1, 1, 1]
    # 8 items of 125 alternate over the 2 workers (500 each), then 7 of 1
    true = np.exp2(est)
    assert scheduler.item_makespan(true, order, nitems, 2) == 504.0
    # a cell pair of 3 chunks splits each chunk: 3 x ceil(333.3 / 125.9) = 9 items
    X[2, 2] = np.log2(3)
    _, nitems = scheduler.engine_items(X, est, workers=2)
    assert nitems[0] == 9


def test_item_level_imbalance_prefers_the_better_estimate():
    rng = np.random.default_rng(2)
    n = 4000
    X = np.zeros((n, scheduler.FEATURES))
    n1 = rng.integers(1, 300, n)
    n2 = rng.integers(1, 300, n)
    X[:, 0], X[:, 1] = np.log2(1 + n1), np.log2(1 + n2)
    X[:, 2] = np.log2(np.maximum(1, np.ceil(n1 / 64)))
    # a few cell pairs far more expensive than N1*N2 says (what the model can learn)
    f = np.where(rng.uniform(size=n) < 0.01, 200.0, 1.0)
    y = np.log2(n1 * n2 * f)
    ev = scheduler.evaluate(X, y, pred=y + rng.normal(0, 0.05, n), workers=256)
    assert ev["items_imbalance_oracle_order"] < ev["items_imbalance_none"]
    assert ev["items_imbalance_learned"] < ev["items_imbalance_heuristic"]
