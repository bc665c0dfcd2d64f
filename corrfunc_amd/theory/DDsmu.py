"""2-D pair counts DD(s, mu); mirrors Corrfunc.theory.DDsmu (reference Corrfunc/theory/DDsmu.py:16-360)."""
import numpy as np

from . import _common

__all__ = ("DDsmu",)


def DDsmu(autocorr, nthreads, binfile, mu_max, nmu_bins,
          X1, Y1, Z1, weights1=None, periodic=True, boxsize=None,
          X2=None, Y2=None, Z2=None, weights2=None,
          verbose=False, output_savg=False,
          fast_divide_and_NR_steps=0,
          xbin_refine_factor=2, ybin_refine_factor=2,
          zbin_refine_factor=1, max_cells_per_dim=100,
          copy_particles=True, enable_min_sep_opt=True,
          c_api_timer=False, isa=r'fastest', weight_type=None):
    """Pair counts in bins of (s, mu = |dz|/s). Fields (smin, smax, savg,
    mu_max, npairs, weightavg); `mu_max` holds the upper edge of each mu bin."""
    if not autocorr:
        if X2 is None or Y2 is None or Z2 is None:
            raise ValueError("Must pass valid arrays for X2/Y2/Z2 for computing cross-correlation")
    if periodic and boxsize is None:
        raise ValueError("Must specify a boxsize if periodic=True")
    weights1, weights2 = _common.process_weights(weights1, weights2, X1, X2, weight_type, autocorr)
    out = _common.run("countpairs_s_mu", autocorr=autocorr, nthreads=nthreads, binfile=binfile, X1=X1, Y1=Y1,
                      Z1=Z1, weights1=weights1, X2=X2, Y2=Y2, Z2=Z2, weights2=weights2, weight_type=weight_type,
                      periodic=periodic, boxsize=_common.boxsize_tuple(boxsize), verbose=verbose,
                      need_avg=output_savg, refine=(xbin_refine_factor, ybin_refine_factor, zbin_refine_factor),
                      max_cells_per_dim=max_cells_per_dim, copy_particles=copy_particles,
                      enable_min_sep_opt=enable_min_sep_opt, c_api_timer=c_api_timer, isa=isa, mu_max=mu_max,
                      nmu_bins=nmu_bins, fast_divide_and_NR_steps=fast_divide_and_NR_steps)
    n, m = out["nbin"], out["nmu_bins"]
    dmu = out["mu_max"] / m
    rows = []
    for i in range(1, n):
        for j in range(m):
            idx = i * (m + 1) + j
            rows.append((out["rupp"][i - 1], out["rupp"][i], out["ravg"][idx], (j + 1) * dmu, out["npairs"][idx],
                         out["weightavg"][idx]))
    dt = np.dtype([("smin", np.float64), ("smax", np.float64), ("savg", np.float64), ("mu_max", np.float64),
                   ("npairs", np.uint64), ("weightavg", np.float64)])
    res = np.array(rows, dtype=dt)
    if not c_api_timer:
        return res
    return res, out["api_time"]
