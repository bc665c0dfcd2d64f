"""cost_model.py -- CPU restatement (numpy, FP64) of the engine's work-list cost model.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and nothing else) as the checker
of the GPU kernels k_mlp_predict, k_mlp_grad + k_mlp_reduce and k_adam. The
model has no counterpart in the reference; what fixes it is mathematics:

  network   x~ = (x - mean) * inv_sd;  h1 = relu(W1 x~ + b1);
            h2 = relu(W2 h1 + b2);     o = W3 . h2 + b3;
            cost = y_mean + y_sd * o
  loss      L = 1/(2n) sum_s (o_s - (y_s - y_mean) / y_sd)^2
  gradient  reverse-mode differentiation of L (checked here against central
            finite differences in FP64, tests/test_cost_model.py)
  Adam      Kingma & Ba, "Adam: A Method for Stochastic Optimization" (2015),
            Algorithm 1 with bias-corrected moments

Parameter vector layout (as the engine's):
  W1[H][F] | b1[H] | W2[H][H] | b2[H] | W3[H] | b3,   H = 32, F = 10
Normalisation vector: mean[F] | inv_sd[F] | y_mean | y_sd.
"""
import numpy as np

F = 10
H = 32
OFF_W1 = 0
OFF_B1 = OFF_W1 + H * F
OFF_W2 = OFF_B1 + H
OFF_B2 = OFF_W2 + H * H
OFF_W3 = OFF_B2 + H
OFF_B3 = OFF_W3 + H
NPARAM = OFF_B3 + 1
NNORM = 2 * F + 2


def unpack(p):
    p = np.asarray(p, dtype=np.float64)
    return (p[OFF_W1:OFF_B1].reshape(H, F), p[OFF_B1:OFF_W2], p[OFF_W2:OFF_B2].reshape(H, H),
            p[OFF_B2:OFF_W3], p[OFF_W3:OFF_B3], p[OFF_B3])


def _standardise(X, norm):
    norm = np.asarray(norm, dtype=np.float64)
    return (np.asarray(X, dtype=np.float64) - norm[:F]) * norm[F:2 * F]


def forward(p, norm, X):
    """Network output o (standardised units) and the activations."""
    W1, b1, W2, b2, W3, b3 = unpack(p)
    xs = _standardise(X, norm)
    z1 = xs @ W1.T + b1
    h1 = np.maximum(z1, 0.0)
    z2 = h1 @ W2.T + b2
    h2 = np.maximum(z2, 0.0)
    o = h2 @ W3 + b3
    return o, (xs, z1, h1, z2, h2)


def predict(p, norm, X):
    """Predicted log2 cost."""
    norm = np.asarray(norm, dtype=np.float64)
    o, _ = forward(p, norm, X)
    return norm[2 * F] + norm[2 * F + 1] * o


def loss(p, norm, X, y):
    norm = np.asarray(norm, dtype=np.float64)
    o, _ = forward(p, norm, X)
    yt = (np.asarray(y, dtype=np.float64) - norm[2 * F]) / norm[2 * F + 1]
    return 0.5 * np.mean((o - yt) ** 2)


def loss_and_grad(p, norm, X, y):
    """L and dL/dp by reverse-mode differentiation."""
    norm = np.asarray(norm, dtype=np.float64)
    W1, b1, W2, b2, W3, b3 = unpack(p)
    o, (xs, z1, h1, z2, h2) = forward(p, norm, X)
    n = len(o)
    yt = (np.asarray(y, dtype=np.float64) - norm[2 * F]) / norm[2 * F + 1]
    err = o - yt
    L = 0.5 * np.mean(err ** 2)
    do = err / n
    d2 = np.outer(do, W3) * (z2 > 0)
    d1 = (d2 @ W2) * (z1 > 0)
    g = np.zeros(NPARAM)
    g[OFF_W1:OFF_B1] = (d1.T @ xs).ravel()
    g[OFF_B1:OFF_W2] = d1.sum(0)
    g[OFF_W2:OFF_B2] = (d2.T @ h1).ravel()
    g[OFF_B2:OFF_W3] = d2.sum(0)
    g[OFF_W3:OFF_B3] = do @ h2
    g[OFF_B3] = do.sum()
    return L, g


def adam_step(p, g, m, v, step, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One bias-corrected Adam update; returns (p, m, v)."""
    m = beta1 * m + (1 - beta1) * g
    v = beta2 * v + (1 - beta2) * g * g
    mh = m / (1 - beta1 ** step)
    vh = v / (1 - beta2 ** step)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def make_norm(X, y):
    """Standardisation from samples, as the engine computes it (sd <= 1e-6:
    inv_sd = 0 for a feature, y_sd = 1 for the target)."""
    X = np.asarray(X, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    mu = X.mean(0)
    sd = X.std(0)
    inv = np.where(sd > 1e-6, 1.0 / np.where(sd > 1e-6, sd, 1.0), 0.0)
    ysd = y.std()
    return np.concatenate([mu, inv, [y.mean(), ysd if ysd > 1e-6 else 1.0]])

