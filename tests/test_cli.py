"""Command-line programs DD, DDrppi, DDsmu, wp, xi, DDrppi_mocks and DDsmu_mocks
(corrfunc_amd/bin).

CPU tests: argument checking and read errors, which end before any pair
counting (the reference programs' rules, theory/DD/DD.c:85-131 and siblings).
GPU tests: each program on catalogs written here (fast-food positions, ascii
weights) against the oracle on the same float32 inputs: npairs exact, the
printed columns to the 7 significant digits of "%e"."""
import os
import subprocess

import numpy as np
import pytest

from corrfunc_amd import _abi

from helpers import weights, write_fastfood

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "corrfunc_amd", "bin")
PROGS = ("DD", "DDrppi", "DDsmu", "wp", "xi", "DDrppi_mocks", "DDsmu_mocks")
NARGS = {"DD": 7, "DDrppi": 8, "DDsmu": 9, "wp": 6, "xi": 5, "DDrppi_mocks": 8, "DDsmu_mocks": 9}


def prog(name):
    p = os.path.join(BIN, name)
    if not os.path.isfile(p):
        pytest.fail("command-line program %s not built (make -C corrfunc_amd/cli)" % p)
    return p


def run(name, *args):
    return subprocess.run([prog(name)] + [str(a) for a in args], capture_output=True, text=True, timeout=120)


@pytest.mark.parametrize("name", PROGS)
def test_missing_arguments(name):
    r = run(name)
    assert r.returncode == 1 and "Missing required parameters" in r.stderr


@pytest.mark.parametrize("name", PROGS)
def test_bad_optional_argument_count(name):
    r = run(name, *(["x"] * (NARGS[name] + 2)))
    assert r.returncode == 1 and "optional arguments" in r.stderr
    if name in ("wp", "xi"):  # one weight set only
        r = run(name, *(["x"] * (NARGS[name] + 5)))
        assert r.returncode == 1 and "optional arguments" in r.stderr


def _args(name, f1, f2, fmt, binfile, L, nthreads=2):
    if name == "DD":
        return [f1, fmt, f2, fmt, binfile, L, nthreads]
    if name == "DDrppi":
        return [f1, fmt, f2, fmt, binfile, 20.0, L, nthreads]
    if name == "DDsmu":
        return [f1, fmt, f2, fmt, binfile, 1.0, 5, L, nthreads]
    if name == "wp":
        return [L, f1, fmt, binfile, 20.0, nthreads]
    if name == "DDrppi_mocks":
        return [f1, fmt, f2, fmt, binfile, 20.0, 1, nthreads]
    if name == "DDsmu_mocks":
        return [f1, fmt, f2, fmt, binfile, 0.9, 6, 2, nthreads]
    return [L, f1, fmt, binfile, nthreads]


@pytest.mark.parametrize("name", PROGS)
def test_argument_errors_before_counting(name, tmp_path):
    missing = str(tmp_path / "none.ff")
    binfile = str(tmp_path / "bins")
    r = run(name, *_args(name, missing, missing, "f", binfile, 100.0, nthreads=0))
    assert r.returncode == 1 and "Nthreads" in r.stderr
    r = run(name, *(_args(name, missing, missing, "f", binfile, 100.0) + ["bogus", missing, "a"]))
    assert r.returncode == 1 and "Unknown weight method" in r.stderr
    r = run(name, *_args(name, missing, missing, "f", binfile, 100.0))
    assert r.returncode == 1 and "Could not find file" in r.stderr


def _parse(text):
    rows = [line.split() for line in text.strip().splitlines()]
    return np.array([[float(v) for v in row] for row in rows])


def _close(printed, exact):
    exact = np.asarray(exact, dtype=np.float64)
    return np.all(np.abs(printed - exact) <= 5e-6 * np.maximum(np.abs(exact), 1e-300) + 1e-300)


MODE = {"DD": "countpairs", "DDrppi": "countpairs_rp_pi", "DDsmu": "countpairs_s_mu", "wp": "countpairs_wp",
        "xi": "countpairs_xi"}


@pytest.mark.gpu
@pytest.mark.parametrize("cross", [False, True])
@pytest.mark.parametrize("name", PROGS[:5])
def test_program_matches_oracle(name, cross, tmp_path, oracle):
    if cross and name in ("wp", "xi"):
        pytest.skip("wp and xi are autocorrelations")
    L, n = 100.0, 4000
    rng = np.random.default_rng(77 + len(name))
    X1 = [rng.uniform(0, L, n).astype(np.float32) for _ in range(3)]
    X2 = [rng.uniform(0, L, n + 500).astype(np.float32) for _ in range(3)]
    W1, W2 = weights(n, 3, np.float32), weights(n + 500, 4, np.float32)
    f1, f2 = str(tmp_path / "d1.ff"), str(tmp_path / "d2.ff")
    write_fastfood(f1, X1)
    write_fastfood(f2, X2)
    w1, w2 = str(tmp_path / "w1.txt"), str(tmp_path / "w2.txt")
    np.savetxt(w1, W1.astype(np.float64), fmt="%.9g")  # 9 digits: the float32 values exactly
    np.savetxt(w2, W2.astype(np.float64), fmt="%.9g")
    edges = np.logspace(np.log10(0.5), np.log10(12.0), 9) if name != "DDsmu" else np.linspace(0, 12.0, 7)
    binfile = str(tmp_path / "bins")
    with open(binfile, "w") as f:
        for lo, hi in zip(edges[:-1], edges[1:]):
            f.write("%r %r\n" % (float(lo), float(hi)))
    args = _args(name, f1, f2 if cross else f1, "f", binfile, L)
    args += ["pair_product", w1, "a"] + ([w2, "a"] if cross else [])
    r = run(name, *args)
    assert r.returncode == 0, r.stderr[-2000:]
    got = _parse(r.stdout)

    kw = dict(autocorr=0 if cross else 1, nthreads=2, binfile=binfile, X1=X1[0], Y1=X1[1], Z1=X1[2], weights1=W1,
              weight_type="pair_product")
    if cross:
        kw.update(X2=X2[0], Y2=X2[1], Z2=X2[2], weights2=W2)
    if name in ("DDrppi", "wp"):
        kw["pimax"] = 20.0
    if name == "DDsmu":
        kw.update(mu_max=1.0, nmu_bins=5)
    if name in ("wp", "xi"):
        kw["boxsize"] = L
        for k in ("autocorr",):
            kw.pop(k)
    opts = _abi.make_options(periodic=True, boxsize=L, need_avg_sep=False, isa="fallback", float_type=4)
    want = _abi.call_mode(oracle, MODE[name], "oracle_", options=opts, **kw)
    s = _abi.reported_slots(MODE[name], want)
    assert got.shape[0] == len(s)
    npairs_col = 3
    assert np.array_equal(got[:, npairs_col].astype(np.uint64), want["npairs"][s])
    wcol = 4 if name in ("DD", "DDrppi", "DDsmu") else 5
    assert _close(got[:, wcol], want["weightavg"][s])
    assert np.all(got[:, 2] == 0.0) or name == "DDsmu"  # averages are printed only with OUTPUT_RPAVG
    if name == "wp":
        assert _close(got[:, 4], want["wp"][s])
    if name == "xi":
        assert _close(got[:, 4], want["xi"][s])
    if name in ("DD", "wp", "xi"):
        assert _close(got[:, 1], edges[1:])
    if name == "DDrppi":
        assert _close(got[:, 0], np.repeat(np.log10(edges[1:]), 20))
        assert _close(got[:, 1], np.tile(np.arange(1, 21, dtype=np.float64), len(edges) - 1))
    if name == "DDsmu":
        assert _close(got[:, 2], np.tile(np.arange(1, 6) / 5.0, len(edges) - 1))


@pytest.mark.gpu
@pytest.mark.parametrize("cross", [False, True])
@pytest.mark.parametrize("name", ["DDrppi_mocks", "DDsmu_mocks"])
def test_mocks_program_matches_oracle(name, cross, tmp_path, oracle):
    """On-sky programs (double positions, averages printed): fast-food RA / DEC /
    cz catalogs, ascii weights; columns npairs, average, log10(upper edge),
    pi or mu upper edge, weightavg."""
    rng = np.random.default_rng(91 + len(name))

    def cat(n):
        return [rng.uniform(150, 200, n), rng.uniform(10, 40, n), rng.uniform(6000, 12000, n)]
    A, B = cat(5000), cat(4000)
    W1, W2 = rng.uniform(0.5, 1.5, 5000), rng.uniform(0.5, 1.5, 4000)
    f1, f2 = str(tmp_path / "s1.ff"), str(tmp_path / "s2.ff")
    write_fastfood(f1, A)
    write_fastfood(f2, B)
    w1, w2 = str(tmp_path / "w1.txt"), str(tmp_path / "w2.txt")
    np.savetxt(w1, W1, fmt="%.17g")
    np.savetxt(w2, W2, fmt="%.17g")
    edges = np.logspace(np.log10(0.5), np.log10(15.0), 7)
    binfile = str(tmp_path / "bins")
    with open(binfile, "w") as f:
        for lo, hi in zip(edges[:-1], edges[1:]):
            f.write("%r %r\n" % (float(lo), float(hi)))
    args = _args(name, f1, f2 if cross else f1, "f", binfile, 0.0)
    args += ["pair_product", w1, "a"] + ([w2, "a"] if cross else [])
    r = run(name, *args)
    assert r.returncode == 0, r.stderr[-2000:]
    got = _parse(r.stdout)
    mode = "countpairs_mocks" if name == "DDrppi_mocks" else "countpairs_mocks_s_mu"
    kw = dict(autocorr=0 if cross else 1, nthreads=2, binfile=binfile, RA1=A[0], DEC1=A[1], CZ1=A[2], weights1=W1,
              weight_type="pair_product")
    if cross:
        kw.update(RA2=B[0], DEC2=B[1], CZ2=B[2], weights2=W2)
    if mode == "countpairs_mocks":
        kw.update(pimax=20.0, cosmology=1)
    else:
        kw.update(mu_max=0.9, nmu_bins=6, cosmology=2)
    o = _abi.make_options(periodic=False, need_avg_sep=True, isa="fallback")
    want = _abi.call_mocks(oracle, mode, "oracle_", options=o, **kw)
    s = _abi.reported_slots(mode, want)
    assert got.shape[0] == len(s)
    assert np.array_equal(got[:, 0].astype(np.uint64), want["npairs"][s])
    assert want["npairs"][s].sum() > 0
    assert np.all(np.abs(got[:, 1] - want["ravg"][s]) <= 1e-8 + 1e-9 * np.abs(want["ravg"][s]))
    assert np.all(np.abs(got[:, 4] - want["weightavg"][s]) <= 1e-8 + 1e-9 * np.abs(want["weightavg"][s]))
    m = 20 if mode == "countpairs_mocks" else 6
    assert np.allclose(got[:, 2], np.repeat(np.log10(edges[1:]), m), atol=1e-8)
    step = 1.0 if mode == "countpairs_mocks" else 0.15
    assert np.allclose(got[:, 3], np.tile(np.arange(1, m + 1) * step, len(edges) - 1), atol=1e-8)
