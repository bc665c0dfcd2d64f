"""Shared argument handling of the corrfunc_amd.mocks functions: the reference
wrappers' checks (Corrfunc/mocks/DDrppi_mocks.py:328-402,
DDsmu_mocks.py:250-352) and the option filling of its CPython extension
(mocks/python_bindings/_countpairs_mocks.c:1100-1210)."""
import numpy as np

from .. import _abi
from .._lib import lib
from ..theory import _common as _tc
from ..utils import fix_ra_dec


def run(mode, *, autocorr, cosmology, nthreads, binfile, RA1, DEC1, CZ1, weights1, RA2, DEC2, CZ2, weights2,
        is_comoving_dist, verbose, need_avg, fast_divide_and_NR_steps, refine, max_cells_per_dim, copy_particles,
        enable_min_sep_opt, c_api_timer, isa, weight_type, pimax=None, mu_max=None, nmu_bins=None):
    if not autocorr:
        if RA2 is None or DEC2 is None or CZ2 is None:
            raise ValueError("Must pass valid arrays for RA2/DEC2/CZ2 for computing cross-correlation")
    else:
        RA2 = DEC2 = CZ2 = None
    RA1, DEC1, CZ1 = [np.asarray(a) for a in (RA1, DEC1, CZ1)]
    if RA2 is not None:
        RA2, DEC2, CZ2 = [np.asarray(a) for a in (RA2, DEC2, CZ2)]
    weights1, weights2 = _tc.process_weights(weights1, weights2, RA1, RA2, weight_type, autocorr)
    fix_ra_dec(RA1, DEC1)
    if not autocorr:
        fix_ra_dec(RA2, DEC2)
    integer_isa = _tc.translate_isa_string_to_enum(isa)
    if weight_type is not None and weight_type not in ("pair_product", "p"):
        raise ValueError('unknown weight_type "{0}"'.format(weight_type))
    opts = _abi.make_options(periodic=False, verbose=verbose, need_avg_sep=need_avg, refine=refine,
                             max_cells_per_dim=max_cells_per_dim, copy_particles=copy_particles,
                             enable_min_sep_opt=enable_min_sep_opt, c_api_timer=c_api_timer, isa=integer_isa,
                             fast_divide_and_NR_steps=fast_divide_and_NR_steps)
    opts.is_comoving_dist = int(bool(is_comoving_dist))
    out = _abi.call_mocks(lib(), mode, autocorr=autocorr, nthreads=nthreads, binfile=binfile, RA1=RA1, DEC1=DEC1,
                          CZ1=CZ1, weights1=weights1, RA2=RA2, DEC2=DEC2, CZ2=CZ2, weights2=weights2,
                          weight_type=weight_type, pimax=pimax, mu_max=mu_max, nmu_bins=nmu_bins,
                          cosmology=cosmology, options=opts)
    if out is None:
        raise RuntimeError("RuntimeError occurred")
    return out


def table(out, upper, names):
    """(lower edge, upper edge, average, column upper edge, npairs, weightavg) rows
    for every reported (bin, column) slot; `upper` = the column width."""
    n = out["nbin"]
    m = out["npibin"] if "npibin" in out else out["nmu_bins"]
    rows = []
    for i in range(1, n):
        for j in range(m):
            idx = i * (m + 1) + j
            rows.append((out["rupp"][i - 1], out["rupp"][i], out["ravg"][idx], (j + 1) * upper, out["npairs"][idx],
                         out["weightavg"][idx]))
    dt = np.dtype([(names[0], np.float64), (names[1], np.float64), (names[2], np.float64), (names[3], np.float64),
                   ("npairs", np.uint64), ("weightavg", np.float64)])
    return np.array(rows, dtype=dt)
