"""Projected correlation w_p(rp) with analytic randoms; mirrors Corrfunc.theory.wp
(reference Corrfunc/theory/wp.py:283-552)."""
import numpy as np

from . import _common

__all__ = ("wp",)


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
