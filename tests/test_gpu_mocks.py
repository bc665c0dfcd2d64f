"""On-sky (mocks) pair counters on the GPU engine vs the FP64 oracle, through
the C ABI: npairs bit-exact, rpavg / savg / weightavg within 1e-10 relative;
and the DDrppi_mocks doctest table of the reference (inputs regenerated here)."""
import json
import math
import os

import numpy as np
import pytest

from corrfunc_amd import _abi, _lib
from helpers import rel_err

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
F = json.load(open(os.path.join(HERE, "golden", "mocks_tables.json")))
MODES = ("countpairs_mocks", "countpairs_mocks_s_mu")
ENGINES = {"tuned": (_lib.ENGINE_TUNED, "0"), "dense": (_lib.ENGINE_TUNED, "1"),
           "sparse": (_lib.ENGINE_TUNED, "2"), "plain": (_lib.ENGINE_PLAIN, "0")}


def _select(eng):
    kind, dense = ENGINES[eng]
    os.environ["CORRFUNC_AMD_DENSE"] = dense
    _lib.set_engine(kind)


def bins_file(tmp_path, edges):
    p = tmp_path / "bins"
    p.write_text("".join("%r %r\n" % (float(a), float(b)) for a, b in edges))
    return str(p)


def opts(comoving, need_avg=True, refine=None):
    o = _abi.make_options(periodic=False, need_avg_sep=need_avg, isa="fallback",
                          **({} if refine is None else dict(refine=refine)))
    o.is_comoving_dist = int(comoving)
    return o


def sky_catalog(n, seed, comoving, dtype, clustered=True):
    """A patch of sky with clumps: RA 150-210, DEC 5-45, cz 6000-24000 km/s (or
    comoving distance 60-240 Mpc/h)."""
    rng = np.random.default_rng(seed)
    ra = rng.uniform(150, 210, n)
    dec = rng.uniform(5, 45, n)
    r = rng.uniform(6000, 24000, n) if not comoving else rng.uniform(60, 240, n)
    if clustered:
        k = n // 3
        c = rng.integers(0, 30, k)
        cra, cdec = rng.uniform(155, 205, 30), rng.uniform(8, 42, 30)
        cr = rng.uniform(8000, 22000, 30) if not comoving else rng.uniform(80, 220, 30)
        ra[:k] = cra[c] + rng.normal(0, 0.3, k)
        dec[:k] = cdec[c] + rng.normal(0, 0.3, k)
        r[:k] = cr[c] + rng.normal(0, 150.0 if not comoving else 1.5, k)
    return [a.astype(dtype) for a in (ra, dec, r)]


def assert_same(mode, got, want, what=""):
    assert got is not None and want is not None, what
    s = _abi.reported_slots(mode, want)
    gn, wn = got["npairs"][s], want["npairs"][s]
    if not np.array_equal(gn, wn):
        bad = np.nonzero(gn != wn)[0][:8]
        raise AssertionError("%s %s npairs differ at slots %s: got %s want %s" % (what, mode, s[bad], gn[bad], wn[bad]))
    assert rel_err(got["ravg"][s], want["ravg"][s]) <= 1e-10, (what, "avg")
    assert rel_err(got["weightavg"][s], want["weightavg"][s]) <= 1e-10, (what, "weightavg")


def _kw(mode, binfile, A, W=None, B=None, W2=None):
    kw = dict(autocorr=1 if B is None else 0, nthreads=4, binfile=binfile, RA1=A[0], DEC1=A[1], CZ1=A[2])
    if B is not None:
        kw.update(RA2=B[0], DEC2=B[1], CZ2=B[2])
    if W is not None:
        kw.update(weights1=W, weight_type="pair_product")
        if B is not None:
            kw["weights2"] = W2
    if mode == "countpairs_mocks":
        kw["pimax"] = 30.0
    else:
        kw.update(mu_max=0.9, nmu_bins=9)
    return kw


@pytest.mark.parametrize("eng", list(ENGINES))
@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("comoving", [False, True], ids=["cz", "dist"])
@pytest.mark.parametrize("dtype", [np.float64, np.float32], ids=["f64", "f32"])
def test_engine_matches_oracle(engine, oracle, tmp_path, eng, mode, comoving, dtype):
    _select(eng)
    A = sky_catalog(12000, 5, comoving, dtype)
    B = sky_catalog(9000, 6, comoving, dtype)
    rng = np.random.default_rng(7)
    W, W2 = rng.uniform(0.5, 1.5, 12000).astype(dtype), rng.uniform(0.5, 1.5, 9000).astype(dtype)
    edges = np.array([[0.3, 1.0], [1.0, 3.0], [3.0, 8.0], [8.0, 20.0]])
    binfile = bins_file(tmp_path, edges)
    for cosmology in (1, 2):
        for cross in (False, True):
            kw = _kw(mode, binfile, A, W, B if cross else None, W2)
            got = _abi.call_mocks(engine, mode, options=opts(comoving), cosmology=cosmology, **kw)
            want = _abi.call_mocks(oracle, mode, "oracle_", options=opts(comoving), cosmology=cosmology, **kw)
            assert_same(mode, got, want, "cosmology=%d cross=%d" % (cosmology, cross))
            assert want["npairs"].sum() > 0


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("refine", [(1, 1, 1), (2, 2, 1), (3, 3, 2)])
def test_options(engine, oracle, tmp_path, mode, refine):
    _select("tuned")
    A = sky_catalog(15000, 9, False, np.float64)
    edges = np.array([[0.5, 2.0], [2.0, 5.0], [5.0, 12.0]])
    kw = _kw(mode, bins_file(tmp_path, edges), A)
    for need_avg in (True, False):
        got = _abi.call_mocks(engine, mode, options=opts(False, need_avg, refine), cosmology=1, **kw)
        want = _abi.call_mocks(oracle, mode, "oracle_", options=opts(False, need_avg, refine), cosmology=1, **kw)
        assert_same(mode, got, want, "refine=%s need_avg=%d" % (refine, need_avg))


@pytest.mark.parametrize("mode", MODES)
def test_call_sequence(engine, oracle, tmp_path, mode):
    """Back-to-back calls that grow, shrink and drop weights or the second set:
    the engine keeps its converted positions between calls, and no call may see
    an earlier call's points or weights."""
    _select("tuned")
    binfile = bins_file(tmp_path, np.array([[0.4, 1.5], [1.5, 4.0], [4.0, 10.0]]))
    rng = np.random.default_rng(17)
    seq = [(12000, 9000, True, np.float64), (5000, None, False, np.float64), (7000, None, True, np.float32),
           (3000, 2000, False, np.float64), (20000, 4000, True, np.float64), (4000, None, True, np.float64)]
    for k, (n1, n2, weighted, dtype) in enumerate(seq):
        A = sky_catalog(n1, 100 + k, False, dtype)
        B = sky_catalog(n2, 200 + k, False, dtype) if n2 else None
        W = rng.uniform(0.5, 1.5, n1).astype(dtype) if weighted else None
        W2 = rng.uniform(0.5, 1.5, n2).astype(dtype) if weighted and n2 else None
        kw = _kw(mode, binfile, A, W, B, W2)
        got = _abi.call_mocks(engine, mode, options=opts(False), cosmology=1 + k % 2, **kw)
        want = _abi.call_mocks(oracle, mode, "oracle_", options=opts(False), cosmology=1 + k % 2, **kw)
        assert_same(mode, got, want, "call %d" % k)
        assert want["npairs"].sum() > 0


def test_ddrppi_mocks_doctest_table_gpu(engine, tmp_path):
    """Corrfunc/mocks/DDrppi_mocks.py:251-317 (uniform points in a 420 box around
    the observer, distances given)."""
    _select("tuned")
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
    r = _abi.call_mocks(engine, "countpairs_mocks", autocorr=1, nthreads=2, binfile=bins_file(tmp_path, F["bins"]),
                        RA1=RA, DEC1=DEC, CZ1=CZ, weights1=np.ones_like(X), weight_type="pair_product", pimax=40.0,
                        cosmology=1, options=opts(True))
    s = _abi.reported_slots("countpairs_mocks", r)[519:]
    rows = np.array(F["ddrppi_doctest"]["rows"])
    assert np.array_equal(r["npairs"][s], rows[:, 4].astype(np.uint64))
    assert np.array_equal(np.round(r["ravg"][s], 6), rows[:, 2])
    assert np.all(np.abs(r["weightavg"][s] - 1.0) <= 1e-12)


def test_empty_and_fixed_inputs(engine, oracle, tmp_path):
    _select("tuned")
    edges = np.array([[1.0, 4.0], [4.0, 10.0]])
    binfile = bins_file(tmp_path, edges)
    empty = [np.zeros(0), np.zeros(0), np.zeros(0)]
    r = _abi.call_mocks(engine, "countpairs_mocks", autocorr=1, nthreads=1, binfile=binfile, RA1=empty[0],
                        DEC1=empty[1], CZ1=empty[2], pimax=10.0, cosmology=1, options=opts(False))
    assert r is not None and r["npairs"].sum() == 0 and r["nbin"] == 3
    # RA in [-180, 180], DEC in [0, 180] and redshifts instead of cz are fixed as the reference does
    rng = np.random.default_rng(12)
    ra, dec, z = rng.uniform(-170, -120, 8000), rng.uniform(100, 130, 8000), rng.uniform(0.02, 0.06, 8000)
    for mode in MODES:
        kw = _kw(mode, binfile, [ra, dec, z])
        got = _abi.call_mocks(engine, mode, options=opts(False), cosmology=1, **kw)
        want = _abi.call_mocks(oracle, mode, "oracle_", options=opts(False), cosmology=1, **kw)
        assert_same(mode, got, want, "fixed inputs")
        assert want["npairs"].sum() > 0


def test_errors(engine, tmp_path, capfd):
    _select("tuned")
    A = sky_catalog(500, 3, False, np.float64)
    good = bins_file(tmp_path, [[1.0, 5.0], [5.0, 10.0]])
    kw = dict(autocorr=1, nthreads=1, RA1=A[0], DEC1=A[1], CZ1=A[2])
    assert _abi.call_mocks(engine, "countpairs_mocks", binfile=good, pimax=10.0, cosmology=7,
                           options=opts(False), **kw) is None
    zero = bins_file(tmp_path, [[0.0, 5.0], [5.0, 10.0]])
    assert _abi.call_mocks(engine, "countpairs_mocks", binfile=zero, pimax=10.0, cosmology=1,
                           options=opts(False), **kw) is None
    assert _abi.call_mocks(engine, "countpairs_mocks_s_mu", binfile=good, mu_max=1.5, nmu_bins=5, cosmology=1,
                           options=opts(False), **kw) is None
    capfd.readouterr()


def test_python_wrappers(tmp_path):
    """corrfunc_amd.mocks.DDrppi_mocks / DDsmu_mocks: the doctest call of the
    reference wrapper (Corrfunc/mocks/DDrppi_mocks.py:241-317), and DDsmu_mocks
    against the C entry point."""
    from corrfunc_amd.mocks import DDrppi_mocks, DDsmu_mocks
    _select("tuned")
    N, boxsize = 100000, 420.0
    np.random.seed(42)
    X = np.random.uniform(-0.5 * boxsize, 0.5 * boxsize, N)
    Y = np.random.uniform(-0.5 * boxsize, 0.5 * boxsize, N)
    Z = np.random.uniform(-0.5 * boxsize, 0.5 * boxsize, N)
    weights = np.ones_like(X)
    CZ = np.sqrt(X * X + Y * Y + Z * Z)
    inv_cz = 1.0 / CZ
    X *= inv_cz
    Y *= inv_cz
    Z *= inv_cz
    DEC = 90.0 - np.arccos(Z) * 180.0 / math.pi
    RA = (np.arctan2(Y, X) * 180.0 / math.pi) + 180.0
    binfile = bins_file(tmp_path, F["bins"])
    results = DDrppi_mocks(1, 1, 2, 40.0, binfile, RA, DEC, CZ, weights1=weights, weight_type='pair_product',
                           output_rpavg=True, is_comoving_dist=True)
    assert results.dtype.names == ("rmin", "rmax", "rpavg", "pimax", "npairs", "weightavg")
    rows = np.array(F["ddrppi_doctest"]["rows"])
    got = results[519:]
    assert np.array_equal(got["npairs"], rows[:, 4].astype(np.uint64))
    for k, name in enumerate(("rmin", "rmax", "rpavg", "pimax")):
        assert np.array_equal(np.round(got[name], 6 if k < 3 else 1), rows[:, k]), name
    res2, t = DDsmu_mocks(1, 2, 2, 0.5, 5, binfile, RA, DEC, CZ, is_comoving_dist=True, output_savg=True,
                          c_api_timer=True)
    assert res2.dtype.names == ("smin", "smax", "savg", "mumax", "npairs", "weightavg") and t > 0
    raw = _abi.call_mocks(_lib.lib(), "countpairs_mocks_s_mu", autocorr=1, nthreads=2, binfile=binfile, RA1=RA,
                          DEC1=DEC, CZ1=CZ, mu_max=0.5, nmu_bins=5, cosmology=2, options=opts(True))
    s = _abi.reported_slots("countpairs_mocks_s_mu", raw)
    assert np.array_equal(res2["npairs"], raw["npairs"][s])
    assert np.allclose(res2["mumax"][:5], [0.1, 0.2, 0.3, 0.4, 0.5])
