"""Catalog readers: fast-food binary files and ascii/csv column files.

Same functions, defaults and errors as the reference's Corrfunc/io.py:22-284
(read_fastfood_catalog, read_ascii_catalog, read_catalog). The C programs and
C callers use read_positions / read_columns_into_array of the engine library
(include/corrfunc_io.h) instead.
"""
import os
import struct

import numpy as np

try:
    import pandas as pd
except ImportError:  # the numpy reader below is used instead
    pd = None

__all__ = ("read_fastfood_catalog", "read_ascii_catalog", "read_catalog")


def _marker(f):
    b = f.read(4)
    if len(b) != 4:
        raise AssertionError("fast-food file seems to be truncated")
    return struct.unpack("@i", b)[0]


def read_fastfood_catalog(filename, return_dtype=None, need_weights=None):
    """X, Y, Z (and W with need_weights) of a fast-food file.

    Layout (Corrfunc/io.py:96-137): a 20-byte header record of 5 ints whose
    second is the particle count, a 36-byte and a 4-byte record that are
    skipped, then one record per column of float32 or float64 values, each
    framed by its byte count. return_dtype (float64 by default) must be
    numpy.float32 or numpy.float64 (ValueError); columns of the other precision
    are converted. A missing file raises IOError and a malformed record an
    AssertionError.
    """
    if return_dtype is None:
        return_dtype = np.float64
    if return_dtype not in [np.float32, np.float64]:
        raise ValueError("Return data-type must be set and a valid numpy float")
    if not os.path.exists(filename):
        raise IOError("Could not find file = {0}".format(filename))

    with open(filename, "rb") as f:
        skip1 = _marker(f)
        idat = struct.unpack("@iiiii", f.read(20))
        skip2 = _marker(f)
        assert skip1 == 20 and skip2 == 20, "fast-food file seems to be incorrect (reading idat)"
        ngal = idat[1]
        # the fdat (9 floats) and znow (1 float) records
        f.seek(4 + 36 + 4 + 4 + 4 + 4, 1)
        skip1 = _marker(f)
        assert skip1 == ngal * 4 or skip1 == ngal * 8, "fast-food file seems to be corrupt (padding bytes)"
        f.seek(-4, 1)
        cols = {}
        for field in "xyz" + ("w" if need_weights else ""):
            skip1 = _marker(f)
            assert skip1 == ngal * 4 or skip1 == ngal * 8, "fast-food file seems to be corrupt (padding bytes a)"
            input_dtype = np.float32 if skip1 // ngal == 4 else np.float64
            arr = np.fromfile(f, input_dtype, ngal)
            assert arr.size == ngal, "fast-food file seems to be truncated"
            skip2 = _marker(f)
            assert skip2 == skip1, "fast-food file seems to be corrupt (record markers differ)"
            cols[field] = arr if input_dtype == return_dtype else arr.astype(return_dtype)

    out = [np.array(cols[k]) for k in "xyz"]
    if need_weights:
        out.append(np.array(cols["w"]))
    return out


def read_ascii_catalog(filename, return_dtype=None):
    """The first three whitespace-separated columns of a text file as X, Y, Z
    arrays of return_dtype (float64 by default); Corrfunc/io.py:141-224.
    A missing file raises IOError."""
    if return_dtype is None:
        return_dtype = np.float64
    if not os.path.exists(filename):
        raise IOError("Could not find file = {0}".format(filename))
    if pd is not None:
        # round_trip: correctly rounded values (pandas' default parser can be 1 ulp off)
        df = pd.read_csv(filename, header=None, engine="c", sep=r"\s+", float_precision="round_trip")
        return tuple(np.asarray(df[k], dtype=return_dtype) for k in range(3))
    a = np.loadtxt(filename, dtype=return_dtype, usecols=(0, 1, 2), ndmin=2)
    return a[:, 0].copy(), a[:, 1].copy(), a[:, 2].copy()


_DEFAULT_EXTS = {".ff": read_fastfood_catalog, ".txt": read_ascii_catalog, ".dat": read_ascii_catalog,
                 ".csv": read_ascii_catalog}


def read_catalog(filebase=None, return_dtype=np.float64):
    """X, Y, Z of a catalog file: fast-food when the extension contains '.ff',
    ascii otherwise (Corrfunc/io.py:227-279). Without filebase, the package's
    test catalog `gals_Mr19` with extension .ff, .txt, .dat or .csv is looked
    for next to the package (this package ships none, so that raises IOError
    unless one is placed there)."""
    if filebase is None:
        base = os.path.join(os.path.dirname(os.path.abspath(__file__)), "../theory/tests/data/", "gals_Mr19")
        for ext, reader in _DEFAULT_EXTS.items():
            if os.path.exists(base + ext):
                x, y, z = reader(base + ext, return_dtype)[:3]
                return x, y, z
        raise IOError("Could not locate {0} with any of these extensions = {1}".format(base, _DEFAULT_EXTS.keys()))
    if os.path.exists(filebase):
        ext = os.path.splitext(filebase)[1]
        reader = read_fastfood_catalog if ".ff" in ext else read_ascii_catalog
        x, y, z = reader(filebase, return_dtype)[:3]
        return x, y, z
    raise IOError("Could not locate file {0}".format(filebase))
