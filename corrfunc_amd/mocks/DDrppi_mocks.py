"""DD(rp, pi) of on-sky catalogs; mirrors Corrfunc.mocks.DDrppi_mocks
(reference Corrfunc/mocks/DDrppi_mocks.py:17-402)."""
from . import _common

__all__ = ("DDrppi_mocks",)


def DDrppi_mocks(autocorr, cosmology, nthreads, pimax, binfile,
                 RA1, DEC1, CZ1, weights1=None,
                 RA2=None, DEC2=None, CZ2=None, weights2=None,
                 is_comoving_dist=False,
                 verbose=False, output_rpavg=False,
                 fast_divide_and_NR_steps=0,
                 xbin_refine_factor=2, ybin_refine_factor=2,
                 zbin_refine_factor=1, max_cells_per_dim=100,
                 copy_particles=True, enable_min_sep_opt=True,
                 c_api_timer=False, isa=r'fastest', weight_type=None):
    """Pair counts in (rp, pi) of catalogs given as RA, DEC (degrees) and cz
    (km/s; with is_comoving_dist=True, comoving distance in Mpc/h), converted
    with cosmology 1 (LasDamas) or 2 (Planck). pi is the separation along the
    pair's mean line of sight, in int(pimax) bins of width pimax/int(pimax).
    Fields (rmin, rmax, rpavg, pimax, npairs, weightavg); `pimax` holds the
    upper edge of each pi bin. RA / DEC arrays are wrapped in place as the
    reference does (negative RA + 180, DEC above 90 - 90)."""
    out = _common.run("countpairs_mocks", autocorr=autocorr, cosmology=cosmology, nthreads=nthreads,
                      binfile=binfile, RA1=RA1, DEC1=DEC1, CZ1=CZ1, weights1=weights1, RA2=RA2, DEC2=DEC2, CZ2=CZ2,
                      weights2=weights2, is_comoving_dist=is_comoving_dist, verbose=verbose, need_avg=output_rpavg,
                      fast_divide_and_NR_steps=fast_divide_and_NR_steps,
                      refine=(xbin_refine_factor, ybin_refine_factor, zbin_refine_factor),
                      max_cells_per_dim=max_cells_per_dim, copy_particles=copy_particles,
                      enable_min_sep_opt=enable_min_sep_opt, c_api_timer=c_api_timer, isa=isa,
                      weight_type=weight_type, pimax=pimax)
    res = _common.table(out, out["pimax"] / out["npibin"], ("rmin", "rmax", "rpavg", "pimax"))
    if not c_api_timer:
        return res
    return res, out["api_time"]
