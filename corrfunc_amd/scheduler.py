"""Work-list scheduling of the tuned pair kernel, and its learned cost model.

The tuned kernel's persistent waves pull work items (<= 64 primaries of a cell
pair, or a slice of their secondary window) from a queue in list order. The
scheduler orders the cell pairs by descending predicted cost and splits the
heaviest into slices, so that the last items to start are short:

  "none"       cell-pair list order, no splitting
  "heuristic"  cost = N1*N2 (N1(N1-1)/2 for a cell with itself)
  "learned"    cost = the cost model's prediction (falls back to "heuristic"
               while no model is trained or loaded)

The cost model is a small MLP (10 features -> 32 -> 32 -> 1, ReLU) whose
inference, gradient and Adam kernels are hand-written HIP in the engine. It is
trained on per-cell-pair timings that the engine records itself:

    from corrfunc_amd import scheduler
    scheduler.record(True)            # time every cell pair of the next calls
    ... run countpairs_* calls on representative catalogs ...
    scheduler.record(False)
    X, y = scheduler.samples()        # features, log2 wall-clock ticks
    scheduler.train()                 # on the recorded samples, on the GPU
    scheduler.save("cost_model.npz")  # later: scheduler.load(...)

Counts are identical under every scheduler; only the kernel time changes.
"""
import ctypes as C

import numpy as np

from . import _lib

NONE = _lib.SCHED_NONE
HEURISTIC = _lib.SCHED_HEURISTIC
LEARNED = _lib.SCHED_LEARNED
_KINDS = {"none": NONE, "heuristic": HEURISTIC, "learned": LEARNED}
FEATURES = _lib.MODEL_FEATURES
NPARAM = _lib.MODEL_PARAMS
NNORM = _lib.MODEL_NORM
FEATURE_NAMES = ("log2(1+N1)", "log2(1+N2)", "log2(chunks)", "same_cell", "dense", "min_sep2/rmax2",
                 "max_sep2/rmax2", "z_overlap", "projected", "log2(1+N2*z_overlap)")


def _fp(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def set_scheduler(kind):
    """kind: "none" | "heuristic" | "learned" (or the SCHED_* integer)."""
    k = _KINDS.get(kind, kind) if isinstance(kind, str) else kind
    if k not in (NONE, HEURISTIC, LEARNED) or _lib.lib().cf_amd_set_scheduler(int(k)) != 0:
        raise ValueError("unknown scheduler %r" % (kind,))


def get_scheduler():
    k = _lib.lib().cf_amd_get_scheduler()
    return {v: n for n, v in _KINDS.items()}[k]


def record(on=True):
    """Record (features, log2 ticks) of every cell pair of the following tuned calls."""
    if _lib.lib().cf_amd_record_costs(1 if on else 0) != 0:
        raise RuntimeError("cf_amd_record_costs failed")


def clear():
    if _lib.lib().cf_amd_clear_cost_samples() != 0:
        raise RuntimeError("cf_amd_clear_cost_samples failed")


def samples():
    """The recorded samples: X [n, FEATURES] float32, y [n] float32 (log2 ticks)."""
    L = _lib.lib()
    n = L.cf_amd_cost_samples(None, None, 0)
    if n < 0:
        raise RuntimeError("cf_amd_cost_samples failed")
    X = np.zeros((n, FEATURES), np.float32)
    y = np.zeros(n, np.float32)
    if n and L.cf_amd_cost_samples(_fp(X), _fp(y), n) < 0:
        raise RuntimeError("cf_amd_cost_samples failed")
    return X, y


def train(X=None, y=None, epochs=40, batch=256, lr=3e-3, seed=0):
    """Trains the model from fresh (seeded) weights on the GPU; X = None trains
    on the recorded samples. Returns the final loss (standardised MSE / 2)."""
    loss = np.zeros(1, np.float32)
    if X is None:
        st = _lib.lib().cf_amd_model_train(None, None, 0, int(epochs), int(batch), float(lr), int(seed), _fp(loss))
    else:
        X = np.ascontiguousarray(X, np.float32)
        y = np.ascontiguousarray(y, np.float32)
        if X.ndim != 2 or X.shape[1] != FEATURES or y.shape != (X.shape[0],):
            raise ValueError("X must be [n, %d] and y [n]" % FEATURES)
        st = _lib.lib().cf_amd_model_train(_fp(X), _fp(y), X.shape[0], int(epochs), int(batch), float(lr),
                                           int(seed), _fp(loss))
    if st != 0:
        raise RuntimeError("cost model training failed")
    return float(loss[0])


def trained():
    return bool(_lib.lib().cf_amd_model_trained())


def reset():
    _lib.lib().cf_amd_model_reset()


def params():
    """(params [NPARAM], norm [NNORM]) of the installed model."""
    p = np.zeros(NPARAM, np.float32)
    nm = np.zeros(NNORM, np.float32)
    if _lib.lib().cf_amd_model_get(_fp(p), _fp(nm)) != 0:
        raise RuntimeError("no cost model is installed")
    return p, nm


def set_params(p, norm):
    p = np.ascontiguousarray(p, np.float32)
    nm = np.ascontiguousarray(norm, np.float32)
    if p.shape != (NPARAM,) or nm.shape != (NNORM,):
        raise ValueError("params must have %d and norm %d entries" % (NPARAM, NNORM))
    if _lib.lib().cf_amd_model_set(_fp(p), _fp(nm)) != 0:
        raise RuntimeError("cf_amd_model_set failed")


def save(path):
    p, nm = params()
    np.savez(path, params=p, norm=nm)


def load(path):
    with np.load(path, allow_pickle=False) as f:
        set_params(f["params"], f["norm"])


def predict(X):
    """log2 cost predicted by the installed model (GPU kernel)."""
    X = np.ascontiguousarray(X, np.float32)
    out = np.zeros(X.shape[0], np.float32)
    if X.shape[0] and _lib.lib().cf_amd_model_predict(_fp(X), X.shape[0], _fp(out)) != 0:
        raise RuntimeError("cf_amd_model_predict failed (is a model installed?)")
    return out


def grad(X, y, p, norm):
    """(loss, gradient) of the model at parameters p on (X, y), computed by the
    GPU's forward/backward kernels."""
    X = np.ascontiguousarray(X, np.float32)
    y = np.ascontiguousarray(y, np.float32)
    p = np.ascontiguousarray(p, np.float32)
    nm = np.ascontiguousarray(norm, np.float32)
    g = np.zeros(NPARAM, np.float32)
    loss = np.zeros(1, np.float32)
    if _lib.lib().cf_amd_model_grad(_fp(X), _fp(y), X.shape[0], _fp(p), _fp(nm), _fp(g), _fp(loss)) != 0:
        raise RuntimeError("cf_amd_model_grad failed")
    return float(loss[0]), g


def adam_step(p, g, m, v, step, lr):
    """One Adam step of the GPU training kernel on host arrays; returns (p, m, v)."""
    p = np.array(p, np.float32)
    g = np.ascontiguousarray(g, np.float32)
    m = np.array(m, np.float32)
    v = np.array(v, np.float32)
    L = _lib.lib()
    fp = C.POINTER(C.c_float)
    L.cf_amd_model_adam_step.argtypes = [fp, fp, fp, fp, C.c_int, C.c_float]
    L.cf_amd_model_adam_step.restype = C.c_int
    if L.cf_amd_model_adam_step(_fp(p), _fp(g), _fp(m), _fp(v), int(step), float(lr)) != 0:
        raise RuntimeError("cf_amd_model_adam_step failed")
    return p, m, v


# ------------------------------------------------------------- evaluation
def heuristic_log2_cost(X):
    """The N1*N2 heuristic from the features: log2(1 + N1*N2), or
    log2(1 + N1(N1-1)/2) for a cell with itself."""
    X = np.asarray(X, np.float64)
    n1 = np.exp2(X[:, 0]) - 1
    n2 = np.exp2(X[:, 1]) - 1
    pairs = np.where(X[:, 3] > 0.5, 0.5 * n1 * (n1 - 1), n1 * n2)
    return np.log2(1 + pairs)


def list_schedule_makespan(true_cost, order, workers):
    """Makespan of greedy list scheduling: items start in `order`, each on the
    worker that becomes free first (what the persistent waves do)."""
    import heapq
    free = [0.0] * workers
    heapq.heapify(free)
    for k in order:
        t = heapq.heappop(free)
        heapq.heappush(free, t + true_cost[k])
    return max(free)


def engine_items(X, est, workers, heavy=4.0, split=4, max_parts=32):
    """The work items the engine makes from cell pairs whose log2 cost estimate
    is `est` (cf_sched.hip, k_count_chunks): the order (cell pairs with an
    estimate >= heavy x the mean go first, heaviest first in quarter-octave
    buckets; the rest keep list order), and per cell pair its 64-primary chunks,
    each split into ceil(chunk estimate / cap) parts (at most max_parts) when
    above cap = total estimate / (split x workers).

    Returns (cell pair index, number of items) per cell pair in launch order."""
    X = np.asarray(X, np.float64)
    est = np.asarray(est, np.float64)
    n = len(est)
    c = np.exp2(est)
    total = c.sum()
    light = np.log2(max(total / max(n, 1) * heavy, 1.0))
    hv = est >= light
    bucket = np.floor((48.0 - est[hv]) * 4.0)
    hidx = np.nonzero(hv)[0]
    order = np.concatenate([hidx[np.argsort(bucket, kind="stable")], np.nonzero(~hv)[0]])
    nch = np.maximum(1.0, np.round(np.exp2(X[:, 2])))
    cap = total / (split * workers)
    per_chunk = c / nch
    parts = np.where(per_chunk > cap, np.minimum(max_parts, np.ceil(per_chunk / cap)), 1.0)
    return order, (nch * parts)[order].astype(np.int64)


def item_makespan(true_cost, order, nitems, workers):
    """Makespan of greedy list scheduling of the items of `engine_items` (each
    cell pair's true cost spread evenly over its items)."""
    import heapq
    free = [0.0] * workers
    heapq.heapify(free)
    for k, m in zip(order, nitems):
        each = true_cost[k] / m
        for _ in range(int(m)):
            t = heapq.heappop(free)
            heapq.heappush(free, t + each)
    return max(free)


def evaluate(X, y, pred=None, offset_from=None, workers=2048):
    """Prediction error and load imbalance of the model and of the heuristic on
    recorded samples (X, y = log2 ticks).

    rmse_*: RMS error in log2 units; the heuristic gets its best constant
    offset (fitted on `offset_from` = (X, y) when given, else on (X, y)).
    imbalance_*: makespan / (total / workers) - 1 of greedy list scheduling of
    whole cell pairs in the order each cost estimate gives (descending),
    measured with the true costs; "none" keeps the recorded order.
    items_imbalance_*: the same for the work items the engine makes with each
    estimate (its heavy-first order and splitting, `engine_items`); "none" is
    the list order with 64-primary chunks and no splitting, as the engine runs
    without a scheduler."""
    X = np.asarray(X, np.float64)
    y = np.asarray(y, np.float64)
    h = heuristic_log2_cost(X)
    if offset_from is not None:
        off = float(np.mean(np.asarray(offset_from[1], np.float64) - heuristic_log2_cost(offset_from[0])))
    else:
        off = float(np.mean(y - h))
    out = {"n": int(len(y)), "rmse_heuristic": float(np.sqrt(np.mean((h + off - y) ** 2)))}
    true = np.exp2(y)
    ideal = true.sum() / workers
    out["imbalance_none"] = list_schedule_makespan(true, np.arange(len(y)), workers) / ideal - 1
    out["imbalance_heuristic"] = list_schedule_makespan(true, np.argsort(-h, kind="stable"), workers) / ideal - 1
    out["imbalance_oracle_order"] = list_schedule_makespan(true, np.argsort(-y, kind="stable"), workers) / ideal - 1
    if pred is None and trained():
        pred = predict(X)
    if pred is not None:
        pred = np.asarray(pred, np.float64)
        out["rmse_learned"] = float(np.sqrt(np.mean((pred - y) ** 2)))
        out["imbalance_learned"] = (list_schedule_makespan(true, np.argsort(-pred, kind="stable"), workers)
                                    / ideal - 1)
    nch = np.maximum(1, np.round(np.exp2(X[:, 2]))).astype(np.int64)
    out["items_imbalance_none"] = item_makespan(true, np.arange(len(y)), nch, workers) / ideal - 1
    ests = {"heuristic": h, "oracle_order": y}
    if pred is not None:
        ests["learned"] = pred
    for name, est in ests.items():
        order, nitems = engine_items(X, est, workers)
        out["items_imbalance_" + name] = item_makespan(true, order, nitems, workers) / ideal - 1
    return out
