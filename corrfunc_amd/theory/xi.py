"""Periodic auto-correlation xi(r) with analytic randoms; mirrors Corrfunc.theory.xi
(reference Corrfunc/theory/xi.py:18-250)."""
import numpy as np

from . import _common

__all__ = ("xi",)


def xi(boxsize, nthreads, binfile, X, Y, Z,
       weights=None, weight_type=None, verbose=False, output_ravg=False,
       xbin_refine_factor=2, ybin_refine_factor=2,
       zbin_refine_factor=1, max_cells_per_dim=100,
       copy_particles=True, enable_min_sep_opt=True,
       c_api_timer=False, isa=r'fastest'):
    """Fields (rmin, rmax, ravg, xi, npairs, weightavg)."""
    weights, _ = _common.process_weights(weights, None, X, None, weight_type, autocorr=True)
    out = _common.run("countpairs_xi", autocorr=1, nthreads=nthreads, binfile=binfile, X1=X, Y1=Y, Z1=Z,
                      weights1=weights, weight_type=weight_type, boxsize=boxsize, verbose=verbose,
                      need_avg=output_ravg, refine=(xbin_refine_factor, ybin_refine_factor, zbin_refine_factor),
                      max_cells_per_dim=max_cells_per_dim, copy_particles=copy_particles,
                      enable_min_sep_opt=enable_min_sep_opt, c_api_timer=c_api_timer, isa=isa)
    n = out["nbin"]
    dt = np.dtype([("rmin", np.float64), ("rmax", np.float64), ("ravg", np.float64), ("xi", np.float64),
                   ("npairs", np.uint64), ("weightavg", np.float64)])
    res = np.zeros(n - 1, dtype=dt)
    res["rmin"] = out["rupp"][:-1]
    res["rmax"] = out["rupp"][1:]
    res["ravg"] = out["ravg"][1:]
    res["xi"] = out["xi"][1:]
    res["npairs"] = out["npairs"][1:]
    res["weightavg"] = out["weightavg"][1:]
    if not c_api_timer:
        return res
    return res, out["api_time"]
