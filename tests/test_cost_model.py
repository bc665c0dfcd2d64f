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
