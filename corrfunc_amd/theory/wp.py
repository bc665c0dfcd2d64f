"""Projected correlation w_p(rp) with analytic randoms; mirrors Corrfunc.theory.wp
(reference Corrfunc/theory/wp.py:17-552): wp() and find_fastest_wp_bin_refs()."""
import itertools
import sys
import time

import numpy as np

from . import _common

__all__ = ("wp", "find_fastest_wp_bin_refs")


def find_fastest_wp_bin_refs(boxsize, pimax, nthreads, binfile, X, Y, Z,
                             verbose=False, output_rpavg=False,
                             max_cells_per_dim=100,
                             isa=r'fastest',
                             maxbinref=3, nrepeats=3,
                             return_runtimes=False):
    """Times wp() for every bin refine factor triple (nx, ny, nz) in
    [1, maxbinref]^3, `nrepeats` calls each, and returns the fastest triple;
    with return_runtimes also a structured array (nx, ny, nz, avg_time,
    sigma_time) sorted by (avg_time, sigma_time), as the reference
    (Corrfunc/theory/wp.py:17-241). A triple with a failed call keeps an
    infinite time (the reference averages the calls that succeeded) and a
    message is printed. (The engine's own autotuner,
    cf_amd_set_autotune, chooses refine factors per workload signature
    without this sweep.)"""
    bin_refs = np.arange(1, maxbinref + 1)
    dtype = np.dtype([("nx", np.int64), ("ny", np.int64), ("nz", np.int64), ("avg_time", np.float64),
                      ("sigma_time", np.float64)])
    all_runtimes = np.zeros(maxbinref ** 3, dtype=dtype)
    all_runtimes["avg_time"] = np.inf
    all_runtimes["sigma_time"] = np.inf
    for ii, (nx, ny, nz) in enumerate(itertools.product(bin_refs, bin_refs, bin_refs)):
        total = total_sqr = 0.0
        failed = False
        for _ in range(nrepeats):
            t0 = time.time()
            try:
                wp(boxsize, pimax, nthreads, binfile, X, Y, Z, verbose=verbose, output_rpavg=output_rpavg,
                   xbin_refine_factor=int(nx), ybin_refine_factor=int(ny), zbin_refine_factor=int(nz),
                   max_cells_per_dim=max_cells_per_dim, isa=isa)
            except RuntimeError:
                print("RuntimeError occurred with perms = ({0}, {1}, {2})".format(nx, ny, nz))
                print("Continuing...")
                sys.stdout.flush()
                failed = True
                continue
            dt = time.time() - t0
            total += dt
            total_sqr += dt * dt
        all_runtimes[ii]["nx"] = nx
        all_runtimes[ii]["ny"] = ny
        all_runtimes[ii]["nz"] = nz
        if not failed:
            avg = total / nrepeats
            all_runtimes[ii]["avg_time"] = avg
            all_runtimes[ii]["sigma_time"] = np.sqrt(max(total_sqr / nrepeats - avg * avg, 0.0))
    all_runtimes.sort(order=("avg_time", "sigma_time"))
    best = (int(all_runtimes[0]["nx"]), int(all_runtimes[0]["ny"]), int(all_runtimes[0]["nz"]))
    if not return_runtimes:
        return best
    return best, all_runtimes


def wp(boxsize, pimax, nthreads, binfile, X, Y, Z,
       weights=None, weight_type=None, verbose=False, output_rpavg=False,
       xbin_refine_factor=2, ybin_refine_factor=2,
       zbin_refine_factor=1, max_cells_per_dim=100,
       copy_particles=True, enable_min_sep_opt=True,
       c_api_timer=False, c_cell_timer=False, isa='fastest'):
    """Fields (rmin, rmax, rpavg, wp, npairs, weightavg)."""
    weights, _ = _common.process_weights(weights, None, X, None, weight_type, autocorr=True)
    out = _common.run("countpairs_wp", autocorr=1, nthreads=nthreads, binfile=binfile, X1=X, Y1=Y, Z1=Z,
                      weights1=weights, weight_type=weight_type, boxsize=boxsize, verbose=verbose,
                      need_avg=output_rpavg, refine=(xbin_refine_factor, ybin_refine_factor, zbin_refine_factor),
                      max_cells_per_dim=max_cells_per_dim, copy_particles=copy_particles,
                      enable_min_sep_opt=enable_min_sep_opt, c_api_timer=c_api_timer, isa=isa, pimax=pimax,
                      c_cell_timer=c_cell_timer)
    n = out["nbin"]
    dt = np.dtype([("rmin", np.float64), ("rmax", np.float64), ("rpavg", np.float64), ("wp", np.float64),
                   ("npairs", np.uint64), ("weightavg", np.float64)])
    res = np.zeros(n - 1, dtype=dt)
    res["rmin"] = out["rupp"][:-1]
    res["rmax"] = out["rupp"][1:]
    res["rpavg"] = out["ravg"][1:]
    res["wp"] = out["wp"][1:]
    res["npairs"] = out["npairs"][1:]
    res["weightavg"] = out["weightavg"][1:]
    ret = (res,)
    if c_api_timer:
        ret += (out["api_time"],)
    if c_cell_timer:
        # one row per cell pair: N1, N2, time_in_ns (device wall clock of the
        # waves on the pair), cellidx1, cellidx2, tid (compute unit)
        ret += (out["cell_timings"],)
    return ret[0] if len(ret) == 1 else ret
