"""Pair counts -> correlation functions (host numpy), the immediate callers of
the pair counters in user workflows.

  convert_3d_counts_to_cf      Corrfunc/utils.py:27-164   (Landy-Szalay xi)
  convert_rp_pi_counts_to_wp   Corrfunc/utils.py:167-322  (projected w_p)

Same arguments, checks, NaN convention and errors as the reference.
"""
import numpy as np

__all__ = ("convert_3d_counts_to_cf", "convert_rp_pi_counts_to_wp", "translate_isa_string_to_enum",
           "return_file_with_rbins", "fix_ra_dec", "fix_cz", "convert_to_native_endian", "is_native_endian")


def _npairs(a):
    """The npairs column of a results array, or the array itself."""
    a = np.asarray(a)
    if a.dtype.names is not None and "npairs" in a.dtype.names:
        return a["npairs"]
    return a


def convert_3d_counts_to_cf(ND1, ND2, NR1, NR2, D1D2, D1R2, D2R1, R1R2, estimator="LS"):
    """Landy-Szalay correlation function from raw pair counts,

        xi = (fN1 fN2 D1D2 - fN1 D1R2 - fN2 D2R1 + R1R2) / R1R2,
        fN1 = NR1 / ND1,  fN2 = NR2 / ND2,

    NaN where R1R2 is 0. The counts may be arrays or the structured arrays the
    theory functions return (their npairs column is used)."""
    counts = [_npairs(a) for a in (D1D2, D1R2, D2R1, R1R2)]
    nbins = len(counts[0])
    if any(len(c) != nbins for c in counts[1:]):
        raise ValueError("Pair counts must have the same number of elements (same bins)")
    if not ("LS" in estimator or "Landy" in estimator):
        raise ValueError("Only the Landy-Szalay estimator is supported. Pass estimator='LS'. (Got estimator = {0})"
                         .format(estimator))
    dd, dr, rd, rr = counts
    f1 = np.float64(NR1) / np.float64(ND1)
    f2 = np.float64(NR2) / np.float64(ND2)
    cf = np.full(nbins, np.nan)
    ok = rr > 0
    cf[ok] = (f1 * f2 * dd[ok] - f1 * dr[ok] - f2 * rd[ok] + rr[ok]) / rr[ok]
    return cf


def convert_rp_pi_counts_to_wp(ND1, ND2, NR1, NR2, D1D2, D1R2, D2R1, R1R2, nrpbins, pimax, dpi=1.0,
                               estimator="LS"):
    """Projected correlation function w_p(rp) = 2 dpi sum_pi xi(rp, pi) from
    (rp, pi) pair counts laid out rp-major with npibins = len / nrpbins pi bins
    per rp bin; NaN propagates from any pi bin with R1R2 = 0."""
    if dpi <= 0.0:
        raise ValueError("Binsize along the line of sight (dpi) = {0}must be positive".format(dpi))
    xirppi = convert_3d_counts_to_cf(ND1, ND2, NR1, NR2, D1D2, D1R2, D2R1, R1R2, estimator=estimator)
    npibins = len(xirppi) // nrpbins
    if npibins * nrpbins != len(xirppi):
        raise ValueError("Number of pi bins could not be calculated correctly. Expected to find that the total "
                         "number of bins = {0} would be the product of the number of pi bins = {1} and the number "
                         "of rp bins = {2}".format(len(xirppi), npibins, nrpbins))
    if dpi * npibins != pimax:
        raise ValueError("Pimax = {0} should be equal to the product of npibins = {1} and dpi = {2}. Check your "
                         "binning scheme.".format(pimax, npibins, dpi))
    return 2.0 * dpi * np.array([np.sum(xirppi[i * npibins:(i + 1) * npibins]) for i in range(nrpbins)])


def fix_ra_dec(ra, dec):
    """Wraps RA into [0, 360] and DEC into [-90, 90] in place, as the reference's
    Corrfunc/utils.py:421-461 does: negative RA are shifted by +180 and DEC above
    90 by -90 (every element, once). Returns the arrays cast to RA's dtype."""
    try:
        input_dtype = ra.dtype
    except AttributeError:
        raise TypeError("Input RA array must be a numpy array")
    if ra is None or dec is None:
        raise ValueError("RA or DEC must be valid arrays")
    if np.size(ra) and np.min(ra) < 0.0:
        print("Warning: found negative RA values, wrapping into [0.0, 360.0]  range")
        ra += 180.0
    if np.size(dec) and np.max(dec) > 90.0:
        print("Warning: found DEC values more than 90.0; wrapping into [-90.0, 90.0] range")
        dec -= 90.0
    return ra.astype(input_dtype), dec.astype(input_dtype)


def fix_cz(cz):
    """Multiplies cz by the speed of light (299800 km/s) in place when its maximum
    is below 10, i.e. when redshifts were passed (Corrfunc/utils.py:384-418).
    Returns the array cast to its own dtype."""
    try:
        input_dtype = cz.dtype
    except AttributeError:
        raise TypeError("Input cz array must be a numpy array")
    if np.max(cz) < 10.0:
        cz *= 299800.0
    return cz.astype(input_dtype)


def translate_isa_string_to_enum(isa):
    """Instruction-set name -> the C API's enum value (Corrfunc/utils.py:464-518):
    FASTEST -1, FALLBACK 0, SSE42 6, AVX 7, AVX2 8, AVX512F 9. The engine uses the
    value to pick which FP64 rounding order of the reference it reproduces."""
    from .theory import _common
    return _common.translate_isa_string_to_enum(isa)


def return_file_with_rbins(rbins):
    """(binfile, delete_after_use) as Corrfunc/utils.py:324-381: an existing file
    name is returned as is (IOError if it does not exist); an array of >= 1 edges
    is sorted and written as consecutive "lo hi" lines to a temporary file that
    the caller deletes."""
    import os
    import tempfile
    if isinstance(rbins, str):
        if os.path.exists(rbins):
            return rbins, False
        raise IOError("Could not find file = `{0}` containing the bins".format(rbins))
    if len(rbins) >= 1:
        rbins = sorted(rbins)
        with tempfile.NamedTemporaryFile(delete=False, mode="w") as f:
            for i in range(len(rbins) - 1):
                f.write("{0} {1}\n".format(rbins[i], rbins[i + 1]))
            name = f.name
        return name, True
    raise TypeError("Input `binfile` was not a valid array (>= 1 element).Num elements = {0}".format(len(rbins)))


def convert_to_native_endian(array, warn=False):
    """The array in native byte order (a swapped copy if needed, with a warning
    when `warn`); None stays None (Corrfunc/utils.py:866-921)."""
    import sys
    import warnings
    if array is None:
        return array
    array = np.asanyarray(array)
    little = sys.byteorder == "little"
    if (array.dtype.byteorder == "<") != little and array.dtype.byteorder not in ("=", "|"):
        if warn:
            warnings.warn("One or more input array has non-native endianness!  A copy will be made with the "
                          "correct endianness.")
        return array.byteswap().view(array.dtype.newbyteorder())
    return array


def is_native_endian(array):
    """True for None and for arrays in native byte order (Corrfunc/utils.py:924-964)."""
    import sys
    if array is None:
        return True
    array = np.asanyarray(array)
    little = sys.byteorder == "little"
    return array.dtype.byteorder in ("=", "|") or (array.dtype.byteorder == "<") == little

