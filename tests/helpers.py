"""Input generators and comparison helpers shared by the tests."""
import struct

import numpy as np

from corrfunc_amd import _abi

MODES = ("countpairs", "countpairs_rp_pi", "countpairs_s_mu", "countpairs_xi", "countpairs_wp")


def uniform(n, L, seed, dtype=np.float64):
    rng = np.random.default_rng(seed)
    return [rng.uniform(0, L, n).astype(dtype) for _ in range(3)]


def clustered(n, L, seed, nparent=50, sigma=2.0, dtype=np.float32):
    """Neyman-Scott clumps: uniform parents, Gaussian children, wrapped into [0, L)."""
    rng = np.random.default_rng(seed)
    parents = rng.uniform(0, L, (nparent, 3))
    k = rng.integers(0, nparent, n)
    pos = (parents[k] + rng.normal(0, sigma, (n, 3))) % L
    pos = pos.astype(dtype)
    pos[pos >= dtype(L)] = 0  # float rounding may map L-eps to L
    return [pos[:, 0].copy(), pos[:, 1].copy(), pos[:, 2].copy()]


def weights(n, seed, dtype):
    return np.random.default_rng(seed).uniform(0.5, 1.5, n).astype(dtype)


def rel_err(a, b):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    d = np.abs(a - b)
    m = np.maximum(np.abs(a), np.abs(b))
    m[m == 0] = 1.0
    return float((d / m).max()) if d.size else 0.0


def run(lib, mode, prefix="", options=None, **kw):
    return _abi.call_mode(lib, mode, prefix, options=options if options is not None else _abi.make_options(), **kw)


def assert_same(mode, got, want, tol=1e-10, what=""):
    """npairs bit-exact on the reported slots; averages within `tol` relative."""
    assert got is not None and want is not None, what
    s = _abi.reported_slots(mode, want)
    gn, wn = got["npairs"][s], want["npairs"][s]
    if not np.array_equal(gn, wn):
        bad = np.nonzero(gn != wn)[0][:8]
        raise AssertionError("%s %s npairs differ at slots %s: got %s want %s" % (what, mode, s[bad], gn[bad], wn[bad]))
    assert rel_err(got["ravg"][s], want["ravg"][s]) <= tol, (what, mode, "ravg")
    assert rel_err(got["weightavg"][s], want["weightavg"][s]) <= tol, (what, mode, "weightavg")
    # xi = ratio - 1 and wp = (ratio - 1) * 2 pimax are compared through the ratio
    # DD / RR, whose relative error is what the sums bound.
    # (an empty catalog gives NaN in both: compared as equal)
    if "xi" in want:
        a, b = got["xi"][1:], want["xi"][1:]
        assert np.array_equal(np.isnan(a), np.isnan(b)), (what, mode, "xi NaN")
        a, b = a[~np.isnan(b)], b[~np.isnan(b)]
        assert np.all(np.abs(a - b) <= tol * (1.0 + np.abs(b))), (what, mode, "xi")
    if "wp" in want:
        t = 2.0 * want["pimax"]
        a, b = got["wp"][1:], want["wp"][1:]
        assert np.array_equal(np.isnan(a), np.isnan(b)), (what, mode, "wp NaN")
        a, b = a[~np.isnan(b)], b[~np.isnan(b)]
        assert np.all(np.abs(a - b) <= tol * (t + np.abs(b))), (what, mode, "wp")


def mode_kwargs(mode, X, W=None, X2=None, W2=None, autocorr=1, bins=None, L=None, pimax=None, mu_max=1.0,
                nmu_bins=10, nthreads=2):
    kw = dict(nthreads=nthreads, binfile=bins, X1=X[0], Y1=X[1], Z1=X[2])
    if W is not None:
        kw.update(weights1=W, weight_type="pair_product")
    if mode in ("countpairs_xi", "countpairs_wp"):
        kw.update(boxsize=L)
    else:
        kw.update(autocorr=autocorr)
        if not autocorr:
            kw.update(X2=X2[0], Y2=X2[1], Z2=X2[2])
            if W is not None:
                kw.update(weights2=W2)
    if mode in ("countpairs_rp_pi", "countpairs_wp"):
        kw.update(pimax=pimax)
    if mode == "countpairs_s_mu":
        kw.update(mu_max=mu_max, nmu_bins=nmu_bins)
    return kw


def write_fastfood(path, cols, idat=None, fdat=None, znow=0.0, markers=None):
    """A fast-food file: 5-int header record (second int = rows), a 9-float and
    a 1-float record, one record per column; markers[k] overrides column k's
    leading byte count (to plant corruption)."""
    n = len(cols[0])
    idat = list(idat) if idat is not None else [0, n, 0, 0, 0]
    fdat = list(fdat) if fdat is not None else [0.0] * 9

    def rec(f, payload, lead=None):
        f.write(struct.pack("@i", len(payload) if lead is None else lead))
        f.write(payload)
        f.write(struct.pack("@i", len(payload)))

    with open(path, "wb") as f:
        rec(f, struct.pack("@5i", *idat))
        rec(f, np.asarray(fdat, np.float32).tobytes())
        rec(f, np.asarray([znow], np.float32).tobytes())
        for k, c in enumerate(cols):
            rec(f, np.ascontiguousarray(c).tobytes(), None if markers is None else markers.get(k))
