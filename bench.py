#!/usr/bin/env python
"""Pair-counting throughput on one MI355X: pairs tested per second.

Workload (one "step"): a DDrppi call followed by a wp call on a clustered
catalog -- N = 4,194,304 float32 points in a periodic box L = 1000 made of
Neyman-Scott clumps (2000 uniform parents, multinomial child counts, Gaussian
offsets of sigma = 2 per axis), 14 log-spaced rp bins 0.1-30, pimax = 40,
pair_product weights U(0.5, 1.5), averages of rp and of the weights requested.
Inputs are generated with numpy (seed 1002) and copied to HBM before timing;
the timed region is the two C-ABI calls on HBM-resident inputs (gridding, work
list, pair kernel, histogram reduction, finalisation).

"Pairs tested" is the reference's unit of work: the sum over its neighbour-cell
pair list (default options) of N_i * N_j (N(N-1)/2 for a cell with itself),
reported by the engine for each call.

The tuned form (FP32 filter + FP64 replay, persistent waves) is the reported
value; the plain FP64 GPU form runs the same steps for comparison. The CPU
baseline is the reference library compiled from its own sources
(oracle/_ref/libcorrfunc_ref.so), run on the host cores on the same catalog
(FP64 path: float32 inputs up-cast), whose counts are also compared with the
GPU's.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP32_VECTOR_TFLOPS = 157.3   # MI355X published FP32 vector rate
FLOPS_PER_PAIR_RP = 6             # dx, dy, dz subtractions, 2 squares, 1 add


def clustered_catalog(seed=1002, n=4194304, L=1000.0, nparent=2000, sigma=2.0):
    rng = np.random.default_rng(seed)
    parents = rng.uniform(0, L, (nparent, 3))
    counts = rng.multinomial(n, np.full(nparent, 1.0 / nparent))
    k = np.repeat(np.arange(nparent), counts)
    pos = (parents[k] + rng.normal(0, sigma, (n, 3))) % L
    pos = pos.astype(np.float32)
    pos[pos >= np.float32(L)] = 0
    w = rng.uniform(0.5, 1.5, n).astype(np.float32)
    return [np.ascontiguousarray(pos[:, d]) for d in range(3)], w


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--n", type=int, default=4194304)
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; the multi-GPU design is described in DESIGN.md "
              "(no figure at %d GPUs is produced)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    from corrfunc_amd import _abi, _lib

    L, pimax = 1000.0, 40.0
    X, W = clustered_catalog(n=args.n)
    bins = np.logspace(np.log10(0.1), np.log10(30.0), 15)
    binpath, _ = _abi.write_bin_file(bins)
    lib = _lib.lib()
    dev = [_lib.DeviceArray(a) for a in X + [W]]
    ptr = [C.c_void_p(d.ptr) for d in dev]
    n = len(X[0])

    def extra():
        e = _abi.get_extra_options(_abi.WEIGHT_PAIR_PRODUCT)
        e.weights0.weights[0] = ptr[3].value
        return e

    def step():
        o = _abi.make_options(periodic=True, boxsize=L, need_avg_sep=True, isa="fastest", float_type=4)
        r1 = _abi.results_countpairs_rp_pi()
        e = extra()
        st = lib.countpairs_rp_pi(n, ptr[0], ptr[1], ptr[2], 0, None, None, None, 16, 1, binpath.encode(), pimax,
                                  C.byref(r1), C.byref(o), C.byref(e))
        assert st == 0, "countpairs_rp_pi failed"
        s1 = _lib.last_stats()
        out1 = _abi._collect("countpairs_rp_pi", r1)
        lib.free_results_rp_pi(C.byref(r1))
        o = _abi.make_options(periodic=True, boxsize=L, need_avg_sep=True, isa="fastest", float_type=4)
        r2 = _abi.results_countpairs_wp()
        e = extra()
        st = lib.countpairs_wp(n, ptr[0], ptr[1], ptr[2], L, 16, binpath.encode(), pimax, C.byref(r2), C.byref(o),
                               C.byref(e))
        assert st == 0, "countpairs_wp failed"
        s2 = _lib.last_stats()
        out2 = _abi._collect("countpairs_wp", r2)
        lib.free_results_wp(C.byref(r2))
        return out1, out2, s1, s2

    def timed(kind, steps, warmup):
        _lib.set_engine(kind)
        for _ in range(warmup):
            step()
        t0 = time.perf_counter()
        kms, outs = [], None
        for _ in range(steps):
            outs = step()
            kms.append((outs[2]["kernel_ms"], outs[3]["kernel_ms"]))
        # every call returns after its results are back on the host, so the
        # device is idle here
        return time.perf_counter() - t0, outs, np.array(kms)

    t_tuned, out_t, kms_t = timed(_lib.ENGINE_TUNED, args.steps, args.warmup)
    tested = out_t[2]["tested_pairs"] + out_t[3]["tested_pairs"]
    value = tested * args.steps / t_tuned
    useful = (out_t[0]["npairs"].sum() + out_t[1]["npairs"].sum()) / 2 * args.steps / t_tuned

    plain = None
    parity_plain = None
    if not args.no_plain:
        t_plain, out_p, kms_p = timed(_lib.ENGINE_PLAIN, max(1, args.steps // 2), 1)
        plain = tested * max(1, args.steps // 2) / t_plain
        parity_plain = bool(np.array_equal(out_p[0]["npairs"], out_t[0]["npairs"]) and
                            np.array_equal(out_p[1]["npairs"], out_t[1]["npairs"]))
        _lib.set_engine(_lib.ENGINE_TUNED)

    # roofline of the dominant kernel (the DDrppi pair kernel of the tuned form)
    k_ms = float(kms_t[:, 0].mean())
    achieved = out_t[2]["tested_pairs"] * FLOPS_PER_PAIR_RP / (k_ms * 1e-3) / 1e12
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_k_pairs_tuned_rppi_clustered4M.txt")
    if os.path.isfile(tpath):
        for line in open(tpath):
            if line.startswith("traffic_bytes_per_launch"):
                traffic = float(line.split()[1])

    cpu = None
    parity_cpu = None
    if not args.no_cpu:
        ref_so = os.path.join(ROOT, "oracle", "_ref", "libcorrfunc_ref.so")
        if os.path.isfile(ref_so):
            ref = _abi.bind(C.CDLL(ref_so))
            cores = min(os.cpu_count() or 1, 16)
            X64 = [a.astype(np.float64) for a in X]
            W64 = W.astype(np.float64)
            t0 = time.perf_counter()
            r1 = _abi.call_mode(ref, "countpairs_rp_pi", options=_abi.make_options(periodic=True, boxsize=L,
                                need_avg_sep=True, isa="fastest"), autocorr=1, nthreads=cores, binfile=bins,
                                X1=X64[0], Y1=X64[1], Z1=X64[2], weights1=W64, weight_type="pair_product",
                                pimax=pimax)
            r2 = _abi.call_mode(ref, "countpairs_wp", options=_abi.make_options(need_avg_sep=True, isa="fastest"),
                                nthreads=cores, binfile=bins, X1=X64[0], Y1=X64[1], Z1=X64[2], weights1=W64,
                                weight_type="pair_product", pimax=pimax, boxsize=L)
            t_cpu = time.perf_counter() - t0
            cpu = {"value": tested / t_cpu, "unit": "pairs tested/s", "cores": cores, "kind": "reference",
                   "sample": "the full catalog, one DDrppi + one wp call (one step), FP64 path on up-cast "
                             "inputs, isa=fastest, %.1f s" % t_cpu}
            s = _abi.reported_slots("countpairs_rp_pi", r1)
            parity_cpu = bool(np.array_equal(r1["npairs"][s], out_t[0]["npairs"][s]) and
                              np.array_equal(r2["npairs"][1:], out_t[1]["npairs"][1:]))
    os.remove(binpath)

    line = {
        "metric": "pairs tested/s on 1 MI355X: tuned vs plain GPU form, npairs bit-exact vs FP64",
        "value": value,
        "unit": "pairs tested/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_tuned / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 filter + f64 decisions and sums",
        "data": "synthetic (numpy default_rng seed 1002)",
        "config": {"workload": "DDrppi + wp, Neyman-Scott clustered catalog N=%d float32, periodic L=1000, "
                               "14 log rp bins 0.1-30, pimax=40, pair_product weights U(0.5,1.5)" % n,
                   "tested_pairs_per_step": tested, "parallelism": "single GPU"},
        "plain_value": plain,
        "tuned_over_plain": (value / plain) if plain else None,
        "useful_pairs_per_s": useful,
        "parity": {"tuned_vs_plain_npairs_equal": parity_plain, "gpu_vs_cpu_reference_npairs_equal": parity_cpu},
        "roofline": {"bound": "valu", "kernel": "k_pairs_tuned<rp_pi>", "achieved": achieved,
                     "peak": PEAK_FP32_VECTOR_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_VECTOR_TFLOPS,
                     "traffic": traffic, "kernel_ms": k_ms,
                     "algorithmic": "%d FP32 flops per tested pair x %.4g tested pairs per launch"
                                    % (FLOPS_PER_PAIR_RP, out_t[2]["tested_pairs"])},
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
