"""Writes ref_cases.json: golden outputs of the reference theory pair counters,
computed by the reference C library compiled from its own sources
(oracle/build_ref.sh -> oracle/_ref/libcorrfunc_ref.so) on small seeded inputs.

Inputs are not stored: each case records the generator (tests/helpers.py
`uniform` / `clustered`, numpy default_rng) and its seed. float32 inputs are
up-cast to float64 before the reference call (the FP64 contract of the engine).
Outputs are stored for the slots the reference's callers read; floating-point
values as float.hex strings.

    python tests/golden/make_golden.py
"""
import ctypes as C
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.dirname(HERE))

from corrfunc_amd import _abi  # noqa: E402
from helpers import clustered, uniform, weights  # noqa: E402

MODES3 = ("countpairs", "countpairs_rp_pi", "countpairs_s_mu")


def cases():
    out = []
    logb = [0.5, 0.8, 1.3, 2.0, 3.2, 5.0, 8.0, 12.6, 20.0]
    zero_b = [0.0, 1.0, 2.5, 5.0, 7.5, 10.0, 15.0, 22.0]
    for mode in MODES3 + ("countpairs_xi", "countpairs_wp"):
        for isa in ("fallback", "avx512f", "avx"):
            out.append(dict(mode=mode, gen="uniform", n=2000, L=100.0, seed=101, dtype="f64", bins=logb,
                            periodic=True, weights=True, ravg=True, isa=isa, refine=[2, 2, 1], autocorr=1,
                            pimax=20.0, mu_max=0.9, nmu_bins=9))
        out.append(dict(mode=mode, gen="clustered", n=3000, L=150.0, seed=102, dtype="f32", bins=zero_b,
                        periodic=True, weights=True, ravg=True, isa="fallback", refine=[2, 2, 1], autocorr=1,
                        pimax=25.0, mu_max=1.0, nmu_bins=12, nparent=20, sigma=3.0))
        out.append(dict(mode=mode, gen="uniform", n=2500, L=80.0, seed=103, dtype="f64", bins=logb,
                        periodic=True, weights=False, ravg=False, isa="fallback", refine=[3, 1, 2], autocorr=1,
                        pimax=15.5, mu_max=0.5, nmu_bins=5))
    for mode in MODES3:
        out.append(dict(mode=mode, gen="uniform", n=1500, n2=2500, L=100.0, seed=104, dtype="f64", bins=logb,
                        periodic=False, weights=True, ravg=True, isa="fallback", refine=[1, 1, 1], autocorr=0,
                        pimax=18.0, mu_max=1.0, nmu_bins=7))
        out.append(dict(mode=mode, gen="clustered", n=2000, n2=1800, L=120.0, seed=105, dtype="f32", bins=zero_b,
                        periodic=True, weights=True, ravg=True, isa="avx512f", refine=[2, 2, 1], autocorr=0,
                        pimax=22.0, mu_max=0.7, nmu_bins=10, nparent=15, sigma=2.0))
        out.append(dict(mode=mode, gen="uniform", n=1200, L=[51.0, 75.0, 123.0], seed=106, dtype="f64",
                        bins=list(np.linspace(0.01, 0.49 * 51.0, 12)), periodic=True, weights=False, ravg=True,
                        isa="fallback", refine=[2, 2, 2], autocorr=1, pimax=24.0, mu_max=0.5, nmu_bins=10,
                        max_cells=3))
    return out


def make_inputs(c):
    dt = np.float64 if c["dtype"] == "f64" else np.float32
    L = c["L"]
    Lmin = min(L) if isinstance(L, list) else L

    def gen(n, seed):
        if c["gen"] == "uniform":
            if isinstance(L, list):
                rng = np.random.default_rng(seed)
                return [rng.uniform(0, L[d], n).astype(dt) for d in range(3)]
            return uniform(n, L, seed, dt)
        return clustered(n, L, seed, nparent=c.get("nparent", 20), sigma=c.get("sigma", 2.0), dtype=dt)

    X = gen(c["n"], c["seed"])
    W = weights(c["n"], c["seed"] + 1000, dt) if c["weights"] else None
    X2 = W2 = None
    if not c["autocorr"]:
        X2 = gen(c["n2"], c["seed"] + 1)
        W2 = weights(c["n2"], c["seed"] + 2000, dt) if c["weights"] else None
    return X, W, X2, W2, Lmin


def call(lib, c, prefix="", upcast=False):
    X, W, X2, W2, Lmin = make_inputs(c)
    if upcast:
        f = lambda a: None if a is None else (a.astype(np.float64) if hasattr(a, "dtype") else [v.astype(np.float64) for v in a])
        X, W, X2, W2 = f(X), f(W), f(X2), f(W2)
    mode = c["mode"]
    kw = dict(nthreads=2, binfile=np.array(c["bins"]), X1=X[0], Y1=X[1], Z1=X[2])
    if W is not None:
        kw.update(weights1=W, weight_type="pair_product")
    if mode in ("countpairs_xi", "countpairs_wp"):
        kw.update(boxsize=Lmin)
    else:
        kw.update(autocorr=c["autocorr"])
        if not c["autocorr"]:
            kw.update(X2=X2[0], Y2=X2[1], Z2=X2[2])
            if W2 is not None:
                kw.update(weights2=W2)
    if mode in ("countpairs_rp_pi", "countpairs_wp"):
        kw.update(pimax=c["pimax"])
    if mode == "countpairs_s_mu":
        kw.update(mu_max=c["mu_max"], nmu_bins=c["nmu_bins"])
    o = _abi.make_options(periodic=c["periodic"], boxsize=c["L"] if c["periodic"] else None,
                          need_avg_sep=c["ravg"], isa=c["isa"], refine=tuple(c["refine"]),
                          max_cells_per_dim=c.get("max_cells", 100))
    return _abi.call_mode(lib, mode, prefix, options=o, **kw)


def main():
    ref = _abi.bind(C.CDLL(os.path.join(ROOT, "oracle", "_ref", "libcorrfunc_ref.so")))
    out = []
    for c in cases():
        r = call(ref, c, upcast=True)
        assert r is not None, c
        s = _abi.reported_slots(c["mode"], r)
        rec = dict(case=c, slots=[int(v) for v in s], npairs=[int(v) for v in r["npairs"][s]],
                   ravg=[float(v).hex() for v in r["ravg"][s]], weightavg=[float(v).hex() for v in r["weightavg"][s]])
        for k in ("xi", "wp"):
            if k in r:
                rec[k] = [float(v).hex() for v in r[k][1:]]
        out.append(rec)
    with open(os.path.join(HERE, "ref_cases.json"), "w") as f:
        json.dump(out, f, separators=(",", ":"))
    print(len(out), "cases; total pairs", sum(sum(r["npairs"]) for r in out))


if __name__ == "__main__":
    main()
