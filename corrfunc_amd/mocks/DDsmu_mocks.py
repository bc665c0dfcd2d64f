"""DD(s, mu) of on-sky catalogs; mirrors Corrfunc.mocks.DDsmu_mocks
(reference Corrfunc/mocks/DDsmu_mocks.py:17-352)."""
import numpy as np

from . import _common

__all__ = ("DDsmu_mocks",)


def DDsmu_mocks(autocorr, cosmology, nthreads, mu_max, nmu_bins, binfile,
                RA1, DEC1, CZ1, weights1=None,
                RA2=None, DEC2=None, CZ2=None, weights2=None,
                is_comoving_dist=False,
                verbose=False, output_savg=False,
                fast_divide_and_NR_steps=0,
                xbin_refine_factor=2, ybin_refine_factor=2,
                zbin_refine_factor=1, max_cells_per_dim=100,
                copy_particles=True, enable_min_sep_opt=True,
                c_api_timer=False, isa='fastest', weight_type=None):
    """Pair counts in (s, mu) of on-sky catalogs, mu = cosine of the angle between
    the separation and the pair's mean line of sight, nmu_bins linear bins on
    [0, mu_max). Fields (smin, smax, savg, mumax, npairs, weightavg); `mumax`
    holds the upper edge of each mu bin. The separations and mu are computed in
    FP64 whatever the input precision."""
    if not np.isscalar(mu_max):
        raise TypeError("The parameter `mu_max` = {0}, has size = {1}. The code is expecting a scalar quantity "
                        "(and not not a list, array)".format(mu_max, np.size(mu_max)))
    if mu_max <= 0.0 or mu_max > 1.0:
        raise ValueError("The parameter `mu_max` = {0}, is the max. of cosine of an angle and should be within "
                         "(0.0, 1.0]".format(mu_max))
    out = _common.run("countpairs_mocks_s_mu", autocorr=autocorr, cosmology=cosmology, nthreads=nthreads,
                      binfile=binfile, RA1=RA1, DEC1=DEC1, CZ1=CZ1, weights1=weights1, RA2=RA2, DEC2=DEC2, CZ2=CZ2,
                      weights2=weights2, is_comoving_dist=is_comoving_dist, verbose=verbose, need_avg=output_savg,
                      fast_divide_and_NR_steps=fast_divide_and_NR_steps,
                      refine=(xbin_refine_factor, ybin_refine_factor, zbin_refine_factor),
                      max_cells_per_dim=max_cells_per_dim, copy_particles=copy_particles,
                      enable_min_sep_opt=enable_min_sep_opt, c_api_timer=c_api_timer, isa=isa,
                      weight_type=weight_type, mu_max=mu_max, nmu_bins=nmu_bins)
    res = _common.table(out, mu_max / float(out["nmu_bins"]), ("smin", "smax", "savg", "mumax"))
    if not c_api_timer:
        return res
    return res, out["api_time"]
