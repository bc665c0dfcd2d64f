#!/usr/bin/env python
"""Pair-counting throughput on one MI355X: pairs tested per second.

Five workloads (--workload, default 5: the largest single-GPU workload by pairs
tested per step); one "step" is one pass of the hot path
over the workload's synthetic input, which is generated with numpy and copied
to HBM before timing (the timed region is the C-ABI calls on HBM-resident
inputs: gridding, work list, pair kernel, histogram reduction, finalisation):

  1  DD autocorrelation: N = 1,048,576 uniform float64 points (RandomState(42)
     x then y then z) in a periodic box L = 420, 20 log r bins 0.1-25, ravg.
  2  DDrppi + wp on a clustered catalog: N = 4,194,304 float32 points, periodic
     L = 1000, Neyman-Scott clumps (2000 uniform parents, multinomial child
     counts, Gaussian offsets sigma = 2 per axis), 14 log rp bins 0.1-30,
     pimax = 40, pair_product weights U(0.5, 1.5), rp and weight averages.
  3  Batched DD autocorrelations: 256 uniform float32 catalogs in a periodic
     box L = 500 with sizes round(exp(U(ln 5e4, ln 2e6))), 18 log r bins
     0.5-50, one batched call (one persistent pair-kernel launch).
  4  xi on a highly clustered catalog: N = 16,777,216 float32 points, 60 % in
     20,000 NFW halos (log-normal occupancy, concentration U(5, 15)), 40 %
     uniform, periodic L = 1000, 25 log r bins 0.1-40; the work list is
     ordered by the learned cost model, trained (untimed) on timings the engine
     records on a second catalog of the same recipe.
  5  DDsmu cross-correlation: 2,000,000 clustered points (the workload-2
     recipe) x 20,000,000 uniform randoms, float32, periodic L = 1000, 30
     linear s bins 0-150, 40 mu bins to mu_max = 1, pair_product weights;
     the work list is ordered by the learned cost model, trained (untimed) on
     timings the engine records on a second clustered catalog of the recipe.

"Pairs tested" is the reference's unit of work: the sum over its neighbour-cell
pair list (default options) of N_i * N_j (N(N-1)/2 for a cell with itself),
reported by the engine for each call.

The tuned form (integer filter on quantized positions + FP64 replay,
persistent waves, cost-ordered work list) is the reported value; the plain FP64 GPU form runs the same steps
for comparison (workload 3: one plain call per catalog). The CPU baseline is
the reference library compiled from its own sources (oracle/_ref), run on the
host cores on the workload (or a stated sample of it) in the FP64 path, whose
counts are also compared with the GPU's.
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

# Roofline of the pair kernels: algorithmic FP32 work of the distance test per
# tested pair (3-D: 3 sub + 3 mul + 2 add for r^2; projected: 2 sub + 2 mul +
# 1 add for rp^2 and 1 sub for dz), against MI355X's FP32 vector peak
# (256 CUs x 128 lanes x 2 (FMA) x 2.4 GHz = 157.3 TFLOP/s). The plain FP64
# form is also priced against the FP64 vector peak, 78.6 TFLOP/s.
PEAK_FP32_TFLOPS = 157.3
PEAK_FP64_TFLOPS = 78.6
FLOPS_PER_PAIR = {"countpairs": 8, "countpairs_xi": 8, "countpairs_s_mu": 8,
                  "countpairs_rp_pi": 6, "countpairs_wp": 6, "countpairs_mocks": 8,
                  "countpairs_mocks_s_mu": 8}
SCHED_NAMES = {0: "none", 1: "heuristic", 2: "learned"}
METRIC = "pairs tested/s on 1 MI355X: tuned vs plain GPU form, npairs bit-exact vs FP64"


def _opts(**kw):
    from corrfunc_amd import _abi
    return _abi.make_options(**kw)


class Call:
    """One C-ABI call on device-resident inputs."""

    def __init__(self, mode, X1, W1=None, X2=None, W2=None, autocorr=1, bins=None, L=None, pimax=None,
                 mu_max=None, nmu_bins=None, float_type=4, need_avg=True):
        from corrfunc_amd import _abi, _lib
        self.mode, self.autocorr, self.L = mode, autocorr, L
        self.pimax, self.mu_max, self.nmu_bins, self.float_type, self.need_avg = pimax, mu_max, nmu_bins, \
            float_type, need_avg
        self.dev = [_lib.DeviceArray(a) for a in X1]
        self.n1 = len(X1[0])
        self.w1 = _lib.DeviceArray(W1) if W1 is not None else None
        self.dev2 = [_lib.DeviceArray(a) for a in X2] if X2 is not None else None
        self.n2 = len(X2[0]) if X2 is not None else 0
        self.w2 = _lib.DeviceArray(W2) if W2 is not None else None
        self.binpath, _ = _abi.write_bin_file(bins)

    def __call__(self):
        from corrfunc_amd import _abi, _lib
        lib = _lib.lib()
        o = _opts(periodic=True, boxsize=self.L, need_avg_sep=self.need_avg, isa="fastest",
                  float_type=self.float_type)
        e = _abi.get_extra_options(_abi.WEIGHT_PAIR_PRODUCT if self.w1 else _abi.WEIGHT_NONE)
        if self.w1:
            e.weights0.weights[0] = self.w1.ptr
            if self.w2:
                e.weights1.weights[0] = self.w2.ptr
        p = [C.c_void_p(d.ptr) for d in self.dev]
        q = [C.c_void_p(d.ptr) for d in self.dev2] if self.dev2 else [None, None, None]
        res = _abi._RES[self.mode]()
        b = self.binpath.encode()
        m = self.mode
        if m == "countpairs":
            st = lib.countpairs(self.n1, *p, self.n2, *q, 16, self.autocorr, b, C.byref(res), C.byref(o), C.byref(e))
        elif m == "countpairs_rp_pi":
            st = lib.countpairs_rp_pi(self.n1, *p, self.n2, *q, 16, self.autocorr, b, self.pimax, C.byref(res),
                                      C.byref(o), C.byref(e))
        elif m == "countpairs_s_mu":
            st = lib.countpairs_s_mu(self.n1, *p, self.n2, *q, 16, self.autocorr, b, self.mu_max, self.nmu_bins,
                                     C.byref(res), C.byref(o), C.byref(e))
        elif m == "countpairs_xi":
            st = lib.countpairs_xi(self.n1, *p, self.L, 16, b, C.byref(res), C.byref(o), C.byref(e))
        else:
            st = lib.countpairs_wp(self.n1, *p, self.L, 16, b, self.pimax, C.byref(res), C.byref(o), C.byref(e))
        if st != 0:
            raise RuntimeError("%s failed" % m)
        out = _abi._collect(m, res)
        getattr(lib, _abi._FREE[m])(C.byref(res))
        return out, _lib.last_stats()


def cpu_reference():
    from corrfunc_amd import _abi
    so = os.path.join(ROOT, "oracle", "_ref", "libcorrfunc_ref.so")
    if not os.path.isfile(so):
        return None
    return _abi.bind(C.CDLL(so))


def median_time(fn, runs):
    """Runs fn() `runs` times; returns (median seconds, every time, last result)."""
    ts, r = [], None
    for _ in range(max(1, runs)):
        t0 = time.perf_counter()
        r = fn()
        ts.append(time.perf_counter() - t0)
    return float(np.median(ts)), ts, r


def cpu_model():
    """The host CPU's model name (lscpu / /proc/cpuinfo)."""
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def npairs_equal(mode, a, b):
    from corrfunc_amd import _abi
    s = _abi.reported_slots(mode, a)
    return bool(np.array_equal(a["npairs"][s], b["npairs"][s]))


# ----------------------------------------------------------------- workloads
class Workload:
    name = ""
    description = ""
    dtype = "i16/i32 filter + f64 decisions and sums"
    data = ""
    dominant = ""        # mode of the dominant pair kernel

    def step(self):
        """Runs one step; returns (list of outputs, list of per-call stats)."""
        return [], []

    def plain_step(self):
        return self.step()

    plain_steps = 0     # default plain steps (0: max(1, steps // 2))

    def cpu(self, cores, gpu_outs, tested, runs):
        """Times the reference on the host `runs` times; returns (median seconds,
        pairs tested in the timed sample, sample description, counts equal to the
        GPU's[, kind]) or None."""
        return None


class W1(Workload):
    name = "dd_uniform_1M"
    dominant = "countpairs"
    data = "synthetic (numpy RandomState(42), float64)"

    def __init__(self, args):
        from corrfunc_amd import catalogs
        n = args.n or 1048576
        self.X = catalogs.uniform_box(n, 420.0, 42)
        self.bins = np.logspace(np.log10(0.1), np.log10(25.0), 21)
        self.call = Call("countpairs", self.X, bins=self.bins, L=420.0, float_type=8)
        self.description = ("DD autocorrelation, N=%d uniform float64, periodic L=420, 20 log r bins 0.1-25, "
                            "ravg" % n)

    def step(self):
        out, st = self.call()
        return [out], [st]

    def cpu(self, cores, gpu_outs, tested, runs):
        from corrfunc_amd import _abi
        ref = cpu_reference()
        if ref is None:
            return None
        dt, _, r = median_time(lambda: _abi.call_mode(
            ref, "countpairs", options=_opts(periodic=True, boxsize=420.0, need_avg_sep=True, isa="fastest"),
            autocorr=1, nthreads=cores, binfile=self.bins, X1=self.X[0], Y1=self.X[1], Z1=self.X[2]), runs)
        return dt, tested, "the full workload (one call), isa=fastest", npairs_equal("countpairs", r, gpu_outs[0])


class W2(Workload):
    name = "rppi_wp_clustered_4M"
    dominant = "countpairs_rp_pi"
    data = "synthetic (numpy default_rng seed 1002)"

    def __init__(self, args):
        from corrfunc_amd import catalogs
        n = args.n or 4194304
        self.X, self.W = catalogs.neyman_scott(n=n, L=1000.0, nparent=2000, sigma=2.0, seed=1002)
        self.bins = np.logspace(np.log10(0.1), np.log10(30.0), 15)
        self.rp = Call("countpairs_rp_pi", self.X, self.W, bins=self.bins, L=1000.0, pimax=40.0)
        self.wp = Call("countpairs_wp", self.X, self.W, bins=self.bins, L=1000.0, pimax=40.0)
        self.description = ("DDrppi + wp, Neyman-Scott clustered catalog N=%d float32, periodic L=1000, 14 log rp "
                            "bins 0.1-30, pimax=40, pair_product weights U(0.5,1.5)" % n)

    def step(self):
        a, sa = self.rp()
        b, sb = self.wp()
        return [a, b], [sa, sb]

    def cpu(self, cores, gpu_outs, tested, runs):
        from corrfunc_amd import _abi
        ref = cpu_reference()
        if ref is None:
            return None
        X64 = [a.astype(np.float64) for a in self.X]
        W64 = self.W.astype(np.float64)

        def one():
            r1 = _abi.call_mode(ref, "countpairs_rp_pi", options=_opts(periodic=True, boxsize=1000.0,
                                                                       need_avg_sep=True, isa="fastest"),
                                autocorr=1, nthreads=cores, binfile=self.bins, X1=X64[0], Y1=X64[1], Z1=X64[2],
                                weights1=W64, weight_type="pair_product", pimax=40.0)
            r2 = _abi.call_mode(ref, "countpairs_wp", options=_opts(need_avg_sep=True, isa="fastest"),
                                nthreads=cores, binfile=self.bins, X1=X64[0], Y1=X64[1], Z1=X64[2], weights1=W64,
                                weight_type="pair_product", pimax=40.0, boxsize=1000.0)
            return r1, r2
        dt, _, (r1, r2) = median_time(one, runs)
        ok = npairs_equal("countpairs_rp_pi", r1, gpu_outs[0]) and npairs_equal("countpairs_wp", r2, gpu_outs[1])
        return dt, tested, "the full workload, one DDrppi + one wp call (one step), FP64 path on up-cast inputs, " \
                        "isa=fastest", ok


class W3(Workload):
    name = "batch_dd_256"
    dominant = "countpairs"
    data = "synthetic (numpy default_rng seed 1003)"

    def __init__(self, args):
        from corrfunc_amd import _lib, catalogs
        scale = (args.n / 1.35e8) if args.n else 1.0
        self.cats = catalogs.batch_catalogs(256, 50000, 2000000, 500.0, 1003, scale=scale)
        self.bins = np.logspace(np.log10(0.5), np.log10(50.0), 19)
        self.dev = [[_lib.DeviceArray(a) for a in c] for c in self.cats]
        from corrfunc_amd import _abi
        self.binpath, _ = _abi.write_bin_file(self.bins)
        n = sum(len(c[0]) for c in self.cats)
        self.graph = None
        if args.graph:
            # captured once (untimed), replayed each step on the same device arrays
            self.graph = _lib.BatchGraph([tuple(d) for d in self.dev], self.bins,
                                         options=_opts(periodic=True, boxsize=500.0, need_avg_sep=True,
                                                       isa="fastest"))
        self.description = ("batched DD autocorrelation, 256 uniform float32 catalogs (%d points, sizes "
                            "log-uniform 50k-2M), periodic L=500, 18 log r bins 0.5-50, one batched call%s"
                            % (n, ", replayed from a captured HIP graph" if args.graph else ""))

    def step(self):
        from corrfunc_amd import _lib
        if self.graph is not None:
            return self.graph.launch(), [_lib.last_stats()]
        outs = _lib.countpairs_batch(self.dev, self.bins, options=_opts(periodic=True, boxsize=500.0,
                                                                        need_avg_sep=True, isa="fastest"))
        return outs, [_lib.last_stats()]

    def plain_step(self):
        # the plain form has no batched launch: one call per catalog
        from corrfunc_amd import _abi, _lib
        outs, stats = [], []
        for d in self.dev:
            o = _opts(periodic=True, boxsize=500.0, need_avg_sep=True, isa="fastest", float_type=4)
            res = _abi.results_countpairs()
            st = _lib.lib().countpairs(d[0].shape[0], C.c_void_p(d[0].ptr), C.c_void_p(d[1].ptr),
                                       C.c_void_p(d[2].ptr), 0, None, None, None, 16, 1, self.binpath.encode(),
                                       C.byref(res), C.byref(o), C.byref(_abi.get_extra_options(_abi.WEIGHT_NONE)))
            if st != 0:
                raise RuntimeError("countpairs failed")
            outs.append(_abi._collect("countpairs", res))
            _lib.lib().free_results(C.byref(res))
            stats.append(_lib.last_stats())
        return outs, stats

    plain_steps = 2

    def cpu(self, cores, gpu_outs, tested, runs):
        # bounded sample: the smallest catalogs, in order, until ~2e10 pairs tested
        from corrfunc_amd import _abi
        ref = cpu_reference()
        if ref is None:
            return None
        order = np.argsort([len(c[0]) for c in self.cats])
        per = self._tested_per_catalog()
        tested_total = sum(per)
        sample, frac_pairs = [], 0.0
        for i in order:
            sample.append(i)
            frac_pairs += per[i] / tested_total
            if frac_pairs * tested_total >= 2e10:
                break
        cats64 = {i: [a.astype(np.float64) for a in self.cats[i]] for i in sample}

        def one():
            return [_abi.call_mode(ref, "countpairs", options=_opts(periodic=True, boxsize=500.0,
                                                                    need_avg_sep=True, isa="fastest"),
                                   autocorr=1, nthreads=cores, binfile=self.bins, X1=cats64[i][0],
                                   Y1=cats64[i][1], Z1=cats64[i][2]) for i in sample]
        t_sum, _, rs = median_time(one, runs)
        ok = all(npairs_equal("countpairs", r, gpu_outs[i]) for r, i in zip(rs, sample))
        k = len(sample) - 1
        return t_sum, frac_pairs * tested_total, "the %d smallest catalogs (%.2f%% of the pairs tested), sequential reference " \
                                  "calls, isa=fastest" % (k + 1, 100 * frac_pairs), ok

    def _tested_per_catalog(self):
        if getattr(self, "_per", None) is None:
            from corrfunc_amd import _lib
            self._per = []
            for d in self.dev:
                _lib.countpairs_batch([d], self.bins, options=_opts(periodic=True, boxsize=500.0, isa="fastest"))
                self._per.append(_lib.last_stats()["tested_pairs"])
        return self._per


class W4(Workload):
    name = "xi_nfw_16M"
    dominant = "countpairs_xi"
    data = "synthetic (numpy default_rng seeds 1004 (measured) and 1014 (cost-model training))"

    def __init__(self, args):
        from corrfunc_amd import catalogs, scheduler
        n = args.n or 16777216
        self.bins = np.logspace(np.log10(0.1), np.log10(40.0), 26)
        # cost model: trained on timings recorded on a second catalog of the same recipe
        if args.sched == "learned":
            train_cat = catalogs.nfw_halos(n=n, L=1000.0, nhalo=20000, seed=1014)
            tc = Call("countpairs_xi", train_cat, bins=self.bins, L=1000.0)
            scheduler.clear()
            scheduler.record(True)
            tc()
            scheduler.record(False)
            Xs, ys = scheduler.samples()
            t0 = time.perf_counter()
            self.train_loss = scheduler.train(Xs, ys, epochs=8, batch=1024, lr=3e-3, seed=0)
            self.train_s = time.perf_counter() - t0
            self.train_n = len(ys)
            del tc, train_cat
        self.X = catalogs.nfw_halos(n=n, L=1000.0, nhalo=20000, seed=1004)
        self.call = Call("countpairs_xi", self.X, bins=self.bins, L=1000.0)
        self.description = ("xi, N=%d float32, 60%% in 20,000 NFW halos + 40%% uniform, periodic L=1000, 25 log r "
                            "bins 0.1-40, learned cost-model scheduling" % n)

    def step(self):
        out, st = self.call()
        return [out], [st]

    def cpu(self, cores, gpu_outs, tested, runs):
        from corrfunc_amd import _abi
        ref = cpu_reference()
        if ref is None:
            return None
        X64 = [a.astype(np.float64) for a in self.X]
        dt, _, r = median_time(lambda: _abi.call_mode(
            ref, "countpairs_xi", options=_opts(need_avg_sep=True, isa="fastest"), nthreads=cores,
            binfile=self.bins, X1=X64[0], Y1=X64[1], Z1=X64[2], boxsize=1000.0), runs)
        return dt, tested, "the full workload (one call), isa=fastest", npairs_equal("countpairs_xi", r, gpu_outs[0])


class W5(Workload):
    name = "smu_cross_2M_20M"
    dominant = "countpairs_s_mu"
    data = "synthetic (numpy default_rng seeds 1005/10050 (measured) and 1015 (cost-model training))"
    plain_steps = 2

    def __init__(self, args):
        from corrfunc_amd import catalogs, scheduler
        nd = args.n or 2000000
        nr = 10 * nd
        rng = np.random.default_rng(10050)
        self.R, self.WR = catalogs.uniform32(nr, 1000.0, rng, with_weights=True)
        self.bins = np.linspace(0.0, 150.0, 31)
        if args.sched == "learned":
            # cost model: trained on timings recorded on a second clustered
            # catalog of the recipe against the same randoms (untimed)
            D2, W2 = catalogs.neyman_scott(n=nd, L=1000.0, nparent=2000, sigma=2.0, seed=1015)
            tc = Call("countpairs_s_mu", D2, W2, self.R, self.WR, autocorr=0, bins=self.bins, L=1000.0,
                      mu_max=1.0, nmu_bins=40)
            scheduler.clear()
            scheduler.record(True)
            tc()
            scheduler.record(False)
            Xs, ys = scheduler.samples()
            t0 = time.perf_counter()
            self.train_loss = scheduler.train(Xs, ys, epochs=20, batch=512, lr=3e-3, seed=0)
            self.train_s = time.perf_counter() - t0
            self.train_n = len(ys)
            del tc, D2, W2
        self.D, self.WD = catalogs.neyman_scott(n=nd, L=1000.0, nparent=2000, sigma=2.0, seed=1005)
        self.call = Call("countpairs_s_mu", self.D, self.WD, self.R, self.WR, autocorr=0, bins=self.bins, L=1000.0,
                         mu_max=1.0, nmu_bins=40)
        self.description = ("DDsmu cross-correlation, %d clustered (Neyman-Scott) x %d uniform float32, periodic "
                            "L=1000, 30 linear s bins 0-150, 40 mu bins, mu_max=1, pair_product weights" % (nd, nr))

    def step(self):
        out, st = self.call()
        return [out], [st]

    def cpu(self, cores, gpu_outs, tested, runs):
        # bounded sample: every 64th data point against all randoms, on the CPU
        # and (untimed) on the GPU, whose counts must agree
        from corrfunc_amd import _abi
        ref = cpu_reference()
        if ref is None:
            return None
        sub = [a[::64].copy() for a in self.D]
        wsub = self.WD[::64].copy()
        kw = dict(autocorr=0, binfile=self.bins, X1=sub[0], Y1=sub[1], Z1=sub[2], weights1=wsub,
                  X2=self.R[0], Y2=self.R[1], Z2=self.R[2], weights2=self.WR, weight_type="pair_product",
                  mu_max=1.0, nmu_bins=40)
        from corrfunc_amd import _lib
        g = _abi.call_mode(_lib.lib(), "countpairs_s_mu", options=_opts(periodic=True, boxsize=1000.0,
                                                                        need_avg_sep=True, isa="fastest"),
                           nthreads=16, **kw)
        pairs = _lib.last_stats()["tested_pairs"]
        kw64 = dict(kw)
        for k in ("X1", "Y1", "Z1", "X2", "Y2", "Z2", "weights1", "weights2"):
            kw64[k] = kw[k].astype(np.float64)
        dt, _, r = median_time(lambda: _abi.call_mode(
            ref, "countpairs_s_mu", options=_opts(periodic=True, boxsize=1000.0, need_avg_sep=True, isa="fastest"),
            nthreads=cores, **kw64), runs)
        return dt, pairs, "every 64th data point (%d) x all %d randoms, isa=fastest" % (len(sub[0]), len(self.R[0])), \
            npairs_equal("countpairs_s_mu", r, g)


class W6(Workload):
    name = "rppi_mocks_sky_1M"
    dominant = "countpairs_mocks"
    data = "synthetic on-sky catalog (numpy default_rng seed 1006)"

    def __init__(self, args):
        from corrfunc_amd import _abi, catalogs
        n = args.n or 1000000
        self.A, self.W = catalogs.sky_mock(n)
        self.bins = np.logspace(np.log10(0.1), np.log10(25.0), 15)
        self.binpath, _ = _abi.write_bin_file(self.bins)
        self.description = ("DDrppi_mocks autocorrelation, N=%d on-sky float64 (RA 120-240, DEC 0-60, cz 3000-30000 "
                            "km/s, half in 3000 clumps), cosmology 1 (cz -> distance on the host), 14 log rp bins "
                            "0.1-25, pimax=40, pair_product weights, rpavg" % n)

    def _call(self, lib, prefix=""):
        from corrfunc_amd import _abi
        o = _opts(periodic=False, need_avg_sep=True, isa="fallback")
        return _abi.call_mocks(lib, "countpairs_mocks", prefix, autocorr=1, nthreads=16, binfile=self.binpath,
                               RA1=self.A[0], DEC1=self.A[1], CZ1=self.A[2], weights1=self.W,
                               weight_type="pair_product", pimax=40.0, cosmology=1, options=o)

    def step(self):
        from corrfunc_amd import _lib
        out = self._call(_lib.lib())
        if out is None:
            raise RuntimeError("countpairs_mocks failed")
        return [out], [_lib.last_stats()]

    def cpu(self, cores, gpu_outs, tested, runs):
        # the reference's mocks engines need GSL, absent here: the oracle (a C
        # restatement with OpenMP) is the CPU baseline for this workload
        from corrfunc_amd import _abi
        so = os.path.join(ROOT, "oracle", "liboracle.so")
        if not os.path.isfile(so):
            return None
        ora = _abi.bind(C.CDLL(so), "oracle_")
        dt, _, r = median_time(lambda: self._call(ora, "oracle_"), runs)
        return dt, tested, "the full workload (one call), the oracle's C restatement, fallback order", \
            npairs_equal("countpairs_mocks", r, gpu_outs[0]), "port"


WORKLOADS = {1: W1, 2: W2, 3: W3, 4: W4, 5: W5, 6: W6}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", type=int, default=5, choices=sorted(WORKLOADS))
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--plain-steps", type=int, default=0,
                    help="default: 2 for workloads 3 and 5 (seconds per plain step), else max(1, steps // 2)")
    ap.add_argument("--cpu-runs", type=int, default=5, help="timed CPU-baseline runs (the median is reported)")
    ap.add_argument("--n", type=int, default=0, help="override the workload's point count (smaller test runs)")
    ap.add_argument("--sched", default="learned", choices=["none", "heuristic", "learned"])
    ap.add_argument("--sched-ab", action="store_true", help="also time the pair kernel under every scheduler")
    ap.add_argument("--graph", action="store_true",
                    help="workload 3: capture the batch once into a HIP graph and time its replays")
    ap.add_argument("--autotune", action="store_true",
                    help="tuned form with the engine's autotuner (refine factors, LDS replication); the pairs "
                         "tested stay those of the default options, counted by an untimed default step")
    ap.add_argument("--sweep", default="", metavar="VAR=v1,v2,...",
                    help="also time the tuned step with the engine variable VAR (read on every call, e.g. "
                         "CORRFUNC_AMD_DENSE_FRAC) set to each value")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="first copy MIB MiB device-to-device 3 times with the engine's copy kernel "
                         "(the known-size kernel that scales the HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU; several GPUs would split the cell-pair work list and "
              "sum the histograms with one all-reduce (no figure at %d GPUs is produced)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    from corrfunc_amd import _lib, scheduler

    if args.copy_calibration > 0:
        nb = args.copy_calibration << 20
        a = _lib.DeviceArray(np.ones(nb, np.uint8))
        b = _lib.DeviceArray(np.zeros(nb, np.uint8))
        for _ in range(3):
            b.copy_from(a)
        a.free()
        b.free()
        print("copy calibration: 3 x %d bytes device-to-device" % nb, file=sys.stderr, flush=True)

    scheduler.set_scheduler(args.sched)
    t_setup = time.perf_counter()
    wl = WORKLOADS[args.workload](args)
    print("setup %.1f s: %s" % (time.perf_counter() - t_setup, wl.description), file=sys.stderr, flush=True)

    def timed(kind, steps, warmup, fn):
        _lib.set_engine(kind)
        for _ in range(warmup):
            fn()
        t0 = time.perf_counter()
        outs, stats = None, []
        for _ in range(steps):
            outs, st = fn()
            stats.append(st)
        # every call returns after its results are back on the host, so the
        # device is idle here
        return time.perf_counter() - t0, outs, stats

    canonical = None
    if args.autotune:
        # the unit of work is the default options' tested pairs: count them first
        _lib.set_engine(_lib.ENGINE_TUNED)
        _lib.set_autotune(False)
        _, st0 = wl.step()
        canonical = sum(s["tested_pairs"] for s in st0)
        _lib.set_autotune(True)
    t_tuned, out_t, st_t = timed(_lib.ENGINE_TUNED, args.steps, args.warmup, wl.step)
    _lib.set_autotune(False)
    tested = canonical if canonical is not None else sum(s["tested_pairs"] for s in st_t[-1])
    engine_tested = sum(s["tested_pairs"] for s in st_t[-1])
    value = tested * args.steps / t_tuned
    useful = sum(float(o["npairs"].sum()) for o in out_t) / 2 * args.steps / t_tuned
    print("tuned: %.4g pairs tested/s (%.1f ms/step)" % (value, t_tuned / args.steps * 1e3), file=sys.stderr,
          flush=True)

    plain = parity_plain = None
    if not args.no_plain:
        ps = args.plain_steps or wl.plain_steps or max(1, args.steps // 2)
        t_plain, out_p, st_p = timed(_lib.ENGINE_PLAIN, ps, 0 if args.workload in (3, 5) else 1, wl.plain_step)
        plain = tested * ps / t_plain
        parity_plain = all(np.array_equal(a["npairs"], b["npairs"]) for a, b in zip(out_p, out_t))
        _lib.set_engine(_lib.ENGINE_TUNED)
        print("plain: %.4g pairs tested/s" % plain, file=sys.stderr, flush=True)

    # roofline of the dominant kernel: the tuned pair kernel of the workload's
    # main call; its duration comes from HIP events the engine records around
    # the launch on its own stream
    dom = [s for s in st_t[0]][0]
    k_ms = float(np.mean([s[0]["kernel_ms"] for s in st_t]))
    dom_pairs = dom["tested_pairs"]
    flops = FLOPS_PER_PAIR[wl.dominant]
    achieved = dom_pairs * flops / (k_ms * 1e-3) / 1e12
    plain_roof = None
    if plain is not None:
        pk_ms = float(np.mean([s[0]["kernel_ms"] for s in st_p]))
        pa = dom_pairs * flops / (pk_ms * 1e-3) / 1e12
        plain_roof = {"bound": "valu", "kernel": "k_pairs_plain", "achieved": pa, "peak": PEAK_FP64_TFLOPS,
                      "unit": "TFLOP/s", "frac": pa / PEAK_FP64_TFLOPS, "kernel_ms": pk_ms,
                      "algorithmic": "%d FP64 flops per tested pair x %.4g tested pairs per launch"
                                     % (flops, dom_pairs)}
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic_k_pairs_tuned_%s.txt" % wl.name)
    if os.path.isfile(tpath):
        for line in open(tpath):
            if line.startswith("traffic_bytes_per_launch"):
                traffic = float(line.split()[1])

    ab = None
    if args.sched_ab:
        ab = {}
        cur = scheduler.get_scheduler()
        for name in ("none", "heuristic", "learned"):
            scheduler.set_scheduler(name)
            ms = []
            for _ in range(max(1, args.steps)):
                _, st = wl.step()
                ms.append(st[0]["kernel_ms"])
            ab[name] = {"kernel_ms": float(np.median(ms)), "sched_used": int(st[0]["sched"])}
        scheduler.set_scheduler(cur)

    sweep = None
    if args.sweep:
        var, vals = args.sweep.split("=", 1)
        old = os.environ.get(var)
        sweep = {}
        for v in vals.split(","):
            os.environ[var] = v
            wl.step()
            ms, ks = [], []
            for _ in range(max(1, args.steps)):
                t0 = time.perf_counter()
                outs, st = wl.step()
                ms.append((time.perf_counter() - t0) * 1e3)
                ks.append(st[0]["kernel_ms"])
            same = all(np.array_equal(a["npairs"], b["npairs"]) for a, b in zip(outs, out_t))
            sweep[v] = {"ms_per_step": float(np.median(ms)), "kernel_ms": float(np.median(ks)), "npairs_equal": same,
                        "hist_copies": st[0]["hist_copies"]}
            print("sweep %s=%s: %.2f ms/step, kernel %.3f ms" % (var, v, sweep[v]["ms_per_step"],
                                                                sweep[v]["kernel_ms"]), file=sys.stderr, flush=True)
        if old is None:
            del os.environ[var]
        else:
            os.environ[var] = old

    cpu = parity_cpu = None
    if not args.no_cpu:
        # every host core this process may run on (OpenMP threads of the reference)
        cores = len(os.sched_getaffinity(0))
        r = wl.cpu(cores, out_t, tested, args.cpu_runs)
        if r is not None:
            dt, pairs, sample, parity_cpu = r[:4]
            kind = r[4] if len(r) > 4 else "reference"
            cpu = {"value": pairs / dt, "unit": "pairs tested/s", "cores": cores, "kind": kind,
                   "cpu_model": cpu_model(),
                   "sample": "%s; median of %d runs, %.2f s" % (sample, max(1, args.cpu_runs), dt)}

    line = {
        "metric": METRIC,
        "value": value,
        "unit": "pairs tested/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_tuned / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": wl.dtype,
        "data": wl.data,
        "config": {"workload": wl.description, "tested_pairs_per_step": tested, "parallelism": "single GPU",
                   # the scheduler that ordered the dominant call's work list (the
                   # engine's own statistic), and the one asked for
                   "scheduler": SCHED_NAMES[st_t[-1][0]["sched"]], "scheduler_requested": args.sched,
                   "autotune": bool(args.autotune),
                   "engine_tested_pairs_per_step": engine_tested,
                   "refine": list(st_t[-1][0]["refine"]), "autotuned": st_t[-1][0]["autotuned"],
                   "hist_copies": st_t[-1][0]["hist_copies"], "queue_fetch": st_t[-1][0]["queue_fetch"],
                   "drain_pairs": st_t[-1][0]["drain_pairs"],
                   "workgroup_threads": st_t[-1][0]["workgroup_threads"],
                   "global_hist": st_t[-1][0]["global_hist"]},
        "plain_value": plain,
        "tuned_over_plain": (value / plain) if plain else None,
        "useful_pairs_per_s": useful,
        "parity": {"tuned_vs_plain_npairs_equal": parity_plain, "gpu_vs_cpu_reference_npairs_equal": parity_cpu},
        "roofline": {"bound": "valu", "kernel": "k_pairs_tuned<%s>" % wl.dominant, "achieved": achieved,
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": traffic, "kernel_ms": k_ms,
                     "algorithmic": "%d FP32 flops per tested pair x %.4g tested pairs per launch"
                                    % (flops, dom_pairs)},
        "plain_roofline": plain_roof,
        "cpu_baseline": cpu,
    }
    if ab is not None:
        line["scheduler_ab"] = ab
    if sweep is not None:
        line["sweep"] = {args.sweep.split("=", 1)[0]: sweep}
    if hasattr(wl, "train_loss"):
        line["cost_model"] = {"train_samples": wl.train_n, "train_loss": wl.train_loss, "train_s": wl.train_s}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
