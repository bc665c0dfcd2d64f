"""2-D pair counts DD(rp, pi); mirrors Corrfunc.theory.DDrppi (reference Corrfunc/theory/DDrppi.py:16-334)."""
import numpy as np

from . import _common

__all__ = ("DDrppi",)


def DDrppi(autocorr, nthreads, pimax, binfile, X1, Y1, Z1, weights1=None,
           periodic=True, boxsize=None,
           X2=None, Y2=None, Z2=None, weights2=None,
           verbose=False, output_rpavg=False,
           xbin_refine_factor=2, ybin_refine_factor=2,
           zbin_refine_factor=1, max_cells_per_dim=100,
           copy_particles=True, enable_min_sep_opt=True,
           c_api_timer=False, isa=r'fastest', weight_type=None):
    """Pair counts in bins of (rp, pi), pi bins of width pimax/int(pimax).
    Fields (rmin, rmax, rpavg, pimax, npairs, weightavg); `pimax` holds the
    upper edge of each pi bin."""
    if not autocorr:
        if X2 is None or Y2 is None or Z2 is None:
            raise ValueError("Must pass valid arrays for X2/Y2/Z2 for computing cross-correlation")
    if periodic and boxsize is None:
        raise ValueError("Must specify a boxsize if periodic=True")
    weights1, weights2 = _common.process_weights(weights1, weights2, X1, X2, weight_type, autocorr)
    out = _common.run("countpairs_rp_pi", autocorr=autocorr, nthreads=nthreads, binfile=binfile, X1=X1, Y1=Y1,
                      Z1=Z1, weights1=weights1, X2=X2, Y2=Y2, Z2=Z2, weights2=weights2, weight_type=weight_type,
                      periodic=periodic, boxsize=_common.boxsize_tuple(boxsize), verbose=verbose,
                      need_avg=output_rpavg, refine=(xbin_refine_factor, ybin_refine_factor, zbin_refine_factor),
                      max_cells_per_dim=max_cells_per_dim, copy_particles=copy_particles,
                      enable_min_sep_opt=enable_min_sep_opt, c_api_timer=c_api_timer, isa=isa, pimax=pimax)
    n, m = out["nbin"], out["npibin"]
    dpi = out["pimax"] / m
    rows = []
    for i in range(1, n):
        for j in range(m):
            idx = i * (m + 1) + j
            rows.append((out["rupp"][i - 1], out["rupp"][i], out["ravg"][idx], (j + 1) * dpi, out["npairs"][idx],
                         out["weightavg"][idx]))
    dt = np.dtype([("rmin", np.float64), ("rmax", np.float64), ("rpavg", np.float64), ("pimax", np.float64),
                   ("npairs", np.uint64), ("weightavg", np.float64)])
    res = np.array(rows, dtype=dt)
    if not c_api_timer:
        return res
    return res, out["api_time"]
