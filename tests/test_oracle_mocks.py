"""The oracle's on-sky (mocks) pair counters against the reference's known
answers and against closed forms.

Pins: the reference's own test output mocks/tests/Mr19_mock.DD (DD(rp, pi) of
the Mr19 mock with the cz -> distance conversion of cosmology 1; the input
catalog is read from the reference tree where it is mounted), the doctest table
of Corrfunc/mocks/DDrppi_mocks.py (inputs regenerated with numpy's legacy
RandomState(42)), and pairs built with a known (rp, pi) / (s, mu)."""
import json
import math
import os

import numpy as np
import pytest

from corrfunc_amd import _abi
from corrfunc_amd import io as cio

HERE = os.path.dirname(os.path.abspath(__file__))
F = json.load(open(os.path.join(HERE, "golden", "mocks_tables.json")))
MR19 = os.path.join("/root/reference", F["mr19_dd"]["input"])


def bins_file(tmp_path, edges):
    p = tmp_path / "bins"
    p.write_text("".join("%r %r\n" % (float(a), float(b)) for a, b in edges))
    return str(p)


def opts(comoving, need_avg=True):
    o = _abi.make_options(periodic=False, need_avg_sep=need_avg, isa="fallback")
    o.is_comoving_dist = int(comoving)
    return o


@pytest.mark.skipif(not os.path.isfile(MR19), reason="reference mock catalog not mounted")
def test_mr19_mock_ddrppi_golden(oracle, tmp_path):
    ra, dec, cz, w = cio.read_fastfood_catalog(MR19, need_weights=True)
    r = _abi.call_mocks(oracle, "countpairs_mocks", "oracle_", autocorr=1, nthreads=8,
                        binfile=bins_file(tmp_path, F["bins"]), RA1=ra, DEC1=dec, CZ1=cz, weights1=w,
                        weight_type="pair_product", pimax=40.0, cosmology=1, options=opts(False))
    s = _abi.reported_slots("countpairs_mocks", r)
    g = np.array(F["mr19_dd"]["rows"])
    assert len(s) == len(g) == 560
    assert np.array_equal(r["npairs"][s], g[:, 0].astype(np.uint64))
    # the golden file prints 8 decimals
    assert np.all(np.abs(r["ravg"][s] - g[:, 1]) <= 1e-8 + 1e-7 * np.abs(g[:, 1]))
    assert np.all(np.abs(r["weightavg"][s] - g[:, 4]) <= 1e-8 + 1e-7 * np.abs(g[:, 4]))
    assert np.allclose(np.repeat(np.log10(np.array(F["bins"])[:, 1]), 40), g[:, 2], atol=1e-8)


def doctest_inputs():
    """Corrfunc/mocks/DDrppi_mocks.py:251-266."""
    N, boxsize = 100000, 420.0
    np.random.seed(42)
    X = np.random.uniform(-0.5 * boxsize, 0.5 * boxsize, N)
    Y = np.random.uniform(-0.5 * boxsize, 0.5 * boxsize, N)
    Z = np.random.uniform(-0.5 * boxsize, 0.5 * boxsize, N)
    CZ = np.sqrt(X * X + Y * Y + Z * Z)
    inv = 1.0 / CZ
    X *= inv
    Y *= inv
    Z *= inv
    DEC = 90.0 - np.arccos(Z) * 180.0 / math.pi
    RA = (np.arctan2(Y, X) * 180.0 / math.pi) + 180.0
    return RA, DEC, CZ, np.ones_like(X)


def test_ddrppi_mocks_doctest_table(oracle, tmp_path):
    RA, DEC, CZ, w = doctest_inputs()
    r = _abi.call_mocks(oracle, "countpairs_mocks", "oracle_", autocorr=1, nthreads=8,
                        binfile=bins_file(tmp_path, F["bins"]), RA1=RA, DEC1=DEC, CZ1=CZ, weights1=w,
                        weight_type="pair_product", pimax=40.0, cosmology=1, options=opts(True))
    s = _abi.reported_slots("countpairs_mocks", r)[519:]
    rows = np.array(F["ddrppi_doctest"]["rows"])
    assert np.array_equal(r["npairs"][s], rows[:, 4].astype(np.uint64))
    assert np.array_equal(np.round(r["ravg"][s], 6), rows[:, 2])
    assert np.all(r["weightavg"][s] == 1.0)


def _sky(pos):
    """Cartesian -> (RA, DEC, distance) in degrees, as the doctest does."""
    x, y, z = pos[:, 0], pos[:, 1], pos[:, 2]
    d = np.sqrt(x * x + y * y + z * z)
    dec = 90.0 - np.degrees(np.arccos(z / d))
    ra = np.degrees(np.arctan2(y, x)) % 360.0
    return ra, dec, d


def constructed_pairs(rng, sep, cosang, n=300, far=2000.0):
    """n pairs far apart from each other; pair k has separation sep[k] and
    cos(angle between s and the mean line of sight) cosang[k]."""
    out = []
    for k in range(n):
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        v = rng.normal(size=3)
        v -= v.dot(u) * u
        v /= np.linalg.norm(v)
        l = u * (far + 400.0 * k)  # pairs 400 Mpc/h apart along the sky
        sv = sep[k] * (cosang[k] * u + math.sqrt(1 - cosang[k] ** 2) * v)
        out += [l - sv / 2, l + sv / 2]
    return np.array(out)


def test_los_split_closed_form_rppi(oracle, tmp_path):
    rng = np.random.default_rng(3)
    n = 300
    edges = np.array([[0.5, 2.0], [2.0, 5.0], [5.0, 10.0]])
    rp_mid = np.array([1.25, 3.5, 7.5])
    kb = rng.integers(0, 3, n)
    pb = rng.integers(0, 10, n)  # pi bins of width 1 up to pimax 10
    rp = rp_mid[kb]
    pi = pb + 0.5
    sep = np.sqrt(rp ** 2 + pi ** 2)
    pos = constructed_pairs(rng, sep, pi / sep, n)
    ra, dec, d = _sky(pos)
    r = _abi.call_mocks(oracle, "countpairs_mocks", "oracle_", autocorr=1, nthreads=4,
                        binfile=bins_file(tmp_path, edges), RA1=ra, DEC1=dec, CZ1=d, pimax=10.0, cosmology=1,
                        options=opts(True))
    want = np.zeros((4, 11), np.uint64)
    for k, p in zip(kb, pb):
        want[k + 1, p] += 2  # autocorrelation counts each pair twice
    got = r["npairs"].reshape(5, 11)[:4]
    assert np.array_equal(got[1:, :10], want[1:, :10])
    assert np.allclose(r["ravg"].reshape(5, 11)[1:4, :10][want[1:, :10] > 0],
                       np.repeat(rp_mid, 10).reshape(3, 10)[want[1:, :10] > 0], rtol=1e-9)


def test_los_split_closed_form_smu(oracle, tmp_path):
    rng = np.random.default_rng(4)
    n = 300
    edges = np.array([[1.0, 4.0], [4.0, 9.0], [9.0, 16.0]])
    s_mid = np.array([2.5, 6.5, 12.5])
    kb = rng.integers(0, 3, n)
    mb = rng.integers(0, 8, n)  # 8 mu bins up to mu_max 0.8
    mu = (mb + 0.5) * 0.1
    pos = constructed_pairs(rng, s_mid[kb], mu, n)
    ra, dec, d = _sky(pos)
    r = _abi.call_mocks(oracle, "countpairs_mocks_s_mu", "oracle_", autocorr=1, nthreads=4,
                        binfile=bins_file(tmp_path, edges), RA1=ra, DEC1=dec, CZ1=d, mu_max=0.8, nmu_bins=8,
                        cosmology=2, options=opts(True))
    want = np.zeros((4, 9), np.uint64)
    for k, m in zip(kb, mb):
        want[k + 1, m] += 2
    got = r["npairs"].reshape(5, 9)[:4]
    assert np.array_equal(got[1:, :8], want[1:, :8])
    # mu beyond mu_max is not counted in any reported column
    pos2 = constructed_pairs(rng, np.full(50, 6.5), np.full(50, 0.95), 50)
    ra2, dec2, d2 = _sky(pos2)
    r2 = _abi.call_mocks(oracle, "countpairs_mocks_s_mu", "oracle_", autocorr=1, nthreads=4,
                         binfile=bins_file(tmp_path, edges), RA1=ra2, DEC1=dec2, CZ1=d2, mu_max=0.8, nmu_bins=8,
                         cosmology=1, options=opts(True))
    assert r2["npairs"][_abi.reported_slots("countpairs_mocks_s_mu", r2)].sum() == 0


def test_cross_of_identical_sets_is_autocorrelation(oracle, tmp_path):
    rng = np.random.default_rng(8)
    n = 3000
    ra = rng.uniform(150, 200, n)
    dec = rng.uniform(10, 40, n)
    cz = rng.uniform(6000, 15000, n)
    edges = np.array([[0.5, 2.0], [2.0, 6.0], [6.0, 15.0]])
    kw = dict(nthreads=4, binfile=bins_file(tmp_path, edges), pimax=20.0, cosmology=1)
    a = _abi.call_mocks(oracle, "countpairs_mocks", "oracle_", autocorr=1, RA1=ra, DEC1=dec, CZ1=cz,
                        options=opts(False), **kw)
    c = _abi.call_mocks(oracle, "countpairs_mocks", "oracle_", autocorr=0, RA1=ra, DEC1=dec, CZ1=cz, RA2=ra,
                        DEC2=dec, CZ2=cz, options=opts(False), **kw)
    s = _abi.reported_slots("countpairs_mocks", a)
    assert a["npairs"][s].sum() > 1000
    assert np.array_equal(a["npairs"][s], c["npairs"][s])  # rpmin > 0: no self pairs in either


def test_input_fixes_and_cosmologies(oracle, tmp_path):
    rng = np.random.default_rng(9)
    n = 2000
    ra = rng.uniform(0, 60, n)
    dec = rng.uniform(-20, 20, n)
    z = rng.uniform(0.02, 0.05, n)
    edges = np.array([[1.0, 5.0], [5.0, 20.0]])
    kw = dict(autocorr=1, nthreads=4, binfile=bins_file(tmp_path, edges), pimax=20.0)
    base = _abi.call_mocks(oracle, "countpairs_mocks", "oracle_", RA1=ra, DEC1=dec, CZ1=z * 299800.0, cosmology=1,
                           options=opts(False), **kw)
    # redshifts (max < 10) are taken for z and multiplied by c; RA in [-180, 180)
    # is shifted by 180 and DEC in [0, 180] by -90: shift the inputs the other way
    shifted = [ra - 180.0, dec + 90.0, z.copy()]
    keep = [a.copy() for a in shifted]
    fixed = _abi.call_mocks(oracle, "countpairs_mocks", "oracle_", RA1=shifted[0], DEC1=shifted[1], CZ1=shifted[2],
                            cosmology=1, options=opts(False), **kw)
    for a, k in zip(shifted, keep):  # the library fixes a copy, not the caller's arrays
        assert np.array_equal(a, k)
    s = _abi.reported_slots("countpairs_mocks", base)
    assert base["npairs"][s].sum() > 0
    tot_b, tot_f = int(base["npairs"][s].sum()), int(fixed["npairs"][s].sum())
    assert abs(tot_b - tot_f) <= 2 * 1e-3 * tot_b  # same catalog up to the rounding of the shifts
    planck = _abi.call_mocks(oracle, "countpairs_mocks", "oracle_", RA1=ra, DEC1=dec, CZ1=z * 299800.0, cosmology=2,
                             options=opts(False), **kw)
    assert not np.array_equal(planck["npairs"][s], base["npairs"][s])


def test_errors(oracle, tmp_path, capfd):
    rng = np.random.default_rng(2)
    ra, dec, cz = rng.uniform(0, 30, 100), rng.uniform(0, 30, 100), rng.uniform(5000, 6000, 100)
    good = bins_file(tmp_path, [[1.0, 5.0], [5.0, 10.0]])
    kw = dict(autocorr=1, nthreads=1, RA1=ra, DEC1=dec, CZ1=cz)
    assert _abi.call_mocks(oracle, "countpairs_mocks", "oracle_", binfile=good, pimax=10.0, cosmology=3,
                           options=opts(False), **kw) is None
    zero = bins_file(tmp_path, [[0.0, 5.0], [5.0, 10.0]])
    assert _abi.call_mocks(oracle, "countpairs_mocks", "oracle_", binfile=zero, pimax=10.0, cosmology=1,
                           options=opts(False), **kw) is None
    for mu_max, nmu in ((1.5, 5), (0.0, 5), (0.5, 0)):
        assert _abi.call_mocks(oracle, "countpairs_mocks_s_mu", "oracle_", binfile=good, mu_max=mu_max,
                               nmu_bins=nmu, cosmology=1, options=opts(False), **kw) is None
    dec_bad = dec.copy()
    dec_bad[3] = 200.0
    assert _abi.call_mocks(oracle, "countpairs_mocks", "oracle_", binfile=good, pimax=10.0, cosmology=1,
                           options=opts(False), autocorr=1, nthreads=1, RA1=ra, DEC1=dec_bad, CZ1=cz) is None
    capfd.readouterr()
