"""Catalog readers: the C entry points read_positions / read_columns_into_array
of the engine library (host code, no GPU call) and the Python module
corrfunc_amd.io, against known answers of the reference (the doctest table of
Corrfunc/io.py:161-183 and the head of its Mr19 mock fast-food file) and
against files written here in the reference's layout (io/io.c:98-184)."""
import ctypes as C
import gzip
import json
import os
import struct

import numpy as np
import pytest

from corrfunc_amd import _lib
from corrfunc_amd import io as cio

from helpers import write_fastfood

HERE = os.path.dirname(os.path.abspath(__file__))
FIX = json.load(open(os.path.join(HERE, "golden", "io_mr19_mock.json")))
REF_FILE = os.path.join("/root/reference", FIX["file"])


@pytest.fixture(scope="module")
def lib():
    L = C.CDLL(_lib.LIB_PATH)
    L.read_columns_into_array.restype = C.c_int64
    L.read_columns_into_array.argtypes = [C.c_char_p, C.c_char_p, C.c_size_t, C.c_int, C.POINTER(C.c_void_p)]
    L.read_positions.restype = C.c_int64
    return L


_free = C.CDLL(None).free
_free.argtypes = [C.c_void_p]


def c_read(lib, path, fmt, size, nfields):
    ptrs = (C.c_void_p * nfields)()
    n = lib.read_columns_into_array(str(path).encode(), fmt.encode(), size, nfields, ptrs)
    if n < 0:
        return n, None
    dt = np.float32 if size == 4 else np.float64
    out = []
    for p in ptrs:
        out.append(np.ctypeslib.as_array(C.cast(p, C.POINTER(C.c_float if size == 4 else C.c_double)),
                                         shape=(max(n, 1),))[:n].astype(dt, copy=True))
        _free(p)
    return n, out


def fixture_columns():
    return [np.array([float.fromhex(h) for h in col]) for col in FIX["head"]]


def test_fixture_head_matches_doctest_table():
    cols = fixture_columns()
    rows = np.array(FIX["doctest_rows"])
    assert rows.shape == (20, 3)
    for k in range(3):
        assert np.array_equal(np.round(cols[k], 5), rows[:, k]), k


@pytest.mark.parametrize("size", [4, 8])
def test_c_reader_on_fixture_head(lib, tmp_path, size):
    cols = fixture_columns()
    p = tmp_path / "mock.ff"
    write_fastfood(p, cols, idat=[FIX["idat"][0], 20] + FIX["idat"][2:], fdat=FIX["fdat"], znow=FIX["znow"])
    n, got = c_read(lib, p, "f", size, 4)
    assert n == 20
    dt = np.float32 if size == 4 else np.float64
    for k in range(4):
        assert np.array_equal(got[k], cols[k].astype(dt)), k
    # read_positions: the same columns through its variadic outputs
    x, y, z = C.c_void_p(), C.c_void_p(), C.c_void_p()
    n2 = lib.read_positions(str(p).encode(), b"f", C.c_size_t(8), C.c_int(3), C.byref(x), C.byref(y), C.byref(z))
    assert n2 == 20
    for k, q in enumerate((x, y, z)):
        a = np.ctypeslib.as_array(C.cast(q, C.POINTER(C.c_double)), shape=(20,)).copy()
        _free(q)
        assert np.array_equal(a, cols[k])


@pytest.mark.skipif(not os.path.isfile(REF_FILE), reason="reference data file not mounted")
def test_readers_on_reference_file(lib):
    n, got = c_read(lib, REF_FILE, "f", 8, 4)
    assert n == FIX["ngal"]
    cols = fixture_columns()
    for k in range(4):
        assert np.array_equal(got[k][:20], cols[k]), k
    x, y, z, w = cio.read_fastfood_catalog(REF_FILE, need_weights=True)
    for k, a in enumerate((x, y, z, w)):
        assert a.dtype == np.float64 and np.array_equal(a, got[k]), k
    x32, _, _ = cio.read_fastfood_catalog(REF_FILE, return_dtype=np.float32)
    assert x32.dtype == np.float32 and np.array_equal(x32, got[0].astype(np.float32))
    rx, ry, rz = cio.read_catalog(REF_FILE)
    assert np.array_equal(rz, got[2])
    rows = np.array(FIX["doctest_rows"])
    assert np.array_equal(np.round(got[2][:20], 5), rows[:, 2])


@pytest.mark.parametrize("fsize", [4, 8])
@pytest.mark.parametrize("size", [4, 8])
def test_fastfood_precision_conversion(lib, tmp_path, fsize, size):
    rng = np.random.default_rng(5)
    dt = np.float32 if fsize == 4 else np.float64
    cols = [rng.uniform(-1e3, 1e3, 1001).astype(dt) for _ in range(3)]
    p = tmp_path / "c.ff"
    write_fastfood(p, cols)
    n, got = c_read(lib, p, "fast-food", size, 3)
    assert n == 1001
    want_dt = np.float32 if size == 4 else np.float64
    for k in range(3):
        assert got[k].dtype == want_dt and np.array_equal(got[k], cols[k].astype(want_dt))
    py = cio.read_fastfood_catalog(str(p), return_dtype=want_dt)
    for k in range(3):
        assert py[k].dtype == want_dt and np.array_equal(py[k], cols[k].astype(want_dt))


def test_fastfood_errors(lib, tmp_path, capfd):
    cols = [np.arange(10, dtype=np.float64) for _ in range(3)]
    bad = tmp_path / "bad.ff"
    write_fastfood(bad, cols, markers={1: 79})  # column 1's leading byte count is wrong
    assert c_read(lib, bad, "f", 8, 3)[0] == -1
    with pytest.raises(AssertionError):
        cio.read_fastfood_catalog(str(bad))
    trunc = tmp_path / "trunc.ff"
    write_fastfood(trunc, cols)
    trunc.write_bytes(trunc.read_bytes()[:-40])
    assert c_read(lib, trunc, "f", 8, 3)[0] == -1
    with pytest.raises(AssertionError):
        cio.read_fastfood_catalog(str(trunc))
    hdr = tmp_path / "hdr.ff"
    with open(hdr, "wb") as f:  # header record of 4 ints instead of 5
        f.write(struct.pack("@i4ii", 16, 0, 10, 0, 0, 16))
    assert c_read(lib, hdr, "f", 8, 3)[0] == -1
    with pytest.raises(AssertionError):
        cio.read_fastfood_catalog(str(hdr))
    # a precision that is neither 4 nor 8 bytes per value
    odd = tmp_path / "odd.ff"
    write_fastfood(odd, [np.zeros(10, np.int16) for _ in range(3)])
    assert c_read(lib, odd, "f", 8, 3)[0] == -1
    good = tmp_path / "good.ff"
    write_fastfood(good, cols)
    assert c_read(lib, good, "f", 2, 3)[0] == -1  # size must be 4 or 8
    assert c_read(lib, good, "f", 8, 0)[0] == -1  # at least one field
    assert c_read(lib, good, "x", 8, 3)[0] == -1  # unknown format
    assert c_read(lib, tmp_path / "missing.ff", "f", 8, 3)[0] == -1
    with pytest.raises(IOError):
        cio.read_fastfood_catalog(str(tmp_path / "missing.ff"))
    with pytest.raises(ValueError):
        cio.read_fastfood_catalog(str(good), return_dtype=np.int32)
    capfd.readouterr()


def test_fastfood_empty_catalog(lib, tmp_path):
    p = tmp_path / "empty.ff"
    write_fastfood(p, [np.zeros(0, np.float64) for _ in range(3)])
    n, got = c_read(lib, p, "f", 8, 3)
    assert n == 0 and all(g.size == 0 for g in got)


@pytest.mark.parametrize("fmt", ["a", "c", "ascii", "csv"])
def test_text_reader(lib, tmp_path, fmt, capfd):
    p = tmp_path / "cat.txt"
    p.write_text("1.5 2.5 3.5 9\n"
                 "4,5,6\n"
                 "7\t8.25\t-9e-3 extra text\n"
                 "# a comment line\n"
                 "\n"
                 "10 11\n"
                 "  12   1e2\t,  0x1p-2\n")
    for size, dt in ((8, np.float64), (4, np.float32)):
        n, got = c_read(lib, p, fmt, size, 3)
        assert n == 4
        assert np.array_equal(got[0], np.array([1.5, 4, 7, 12], dt))
        assert np.array_equal(got[1], np.array([2.5, 5, 8.25, 100], dt))
        assert np.array_equal(got[2], np.array([3.5, 6, -9e-3, 0.25], dt))
    err = capfd.readouterr().err
    assert err.count("Skipping that line") == 6  # 3 bad lines, two reads


def test_text_reader_large_and_python_twin(lib, tmp_path):
    rng = np.random.default_rng(9)
    a = rng.uniform(0, 420, (5000, 4))
    p = tmp_path / "big.dat"
    np.savetxt(p, a, fmt="%.17g")
    n, got = c_read(lib, p, "a", 8, 4)
    assert n == 5000
    for k in range(4):
        assert np.array_equal(got[k], a[:, k])
    x, y, z = cio.read_ascii_catalog(str(p))
    assert np.array_equal(x, a[:, 0]) and np.array_equal(y, a[:, 1]) and np.array_equal(z, a[:, 2])
    x32, _, _ = cio.read_ascii_catalog(str(p), return_dtype=np.float32)
    assert x32.dtype == np.float32 and np.array_equal(x32, a[:, 0].astype(np.float32))
    rx, ry, rz = cio.read_catalog(str(p))
    assert np.array_equal(rz, a[:, 2])
    with pytest.raises(IOError):
        cio.read_catalog(str(tmp_path / "nope.dat"))
    with pytest.raises(IOError):
        cio.read_ascii_catalog(str(tmp_path / "nope.dat"))


def test_gzip_fallback(lib, tmp_path):
    cols = [np.linspace(0, 1, 77) + k for k in range(3)]
    plain = tmp_path / "z.ff"
    write_fastfood(plain, cols)
    gz = tmp_path / "z.ff.gz"
    gz.write_bytes(gzip.compress(plain.read_bytes()))
    plain.unlink()
    n, got = c_read(lib, plain, "f", 8, 3)
    assert n == 77 and plain.exists()
    for k in range(3):
        assert np.array_equal(got[k], cols[k])
