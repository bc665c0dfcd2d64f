"""Synthetic catalogs of the five benchmark workloads (numpy, seeded).

Every generator returns float32 (or float64 where stated) coordinate arrays
in [0, L) with the value L mapped to 0 after the float32 cast, so periodic
gridding sees every particle inside the box.

  uniform_box        DD autocorrelation input: N uniform points, float64,
                     RandomState(42).uniform for X then Y then Z (the recipe
                     of the reference's doctests, Corrfunc/theory/DD.py:189-202).
  neyman_scott       clustered catalog: uniform parents, multinomial child
                     counts, Gaussian offsets per axis, wrapped periodically.
  batch_catalogs     many uniform catalogs with log-uniform sizes.
  nfw_halos          a fraction of the points in NFW-profile halos with
                     log-normal occupancy, the rest uniform.
"""
import numpy as np


def _wrap32(pos, L):
    pos = np.mod(pos, L).astype(np.float32)
    pos[pos >= np.float32(L)] = 0
    return [np.ascontiguousarray(pos[:, d]) for d in range(3)]


def uniform_box(n=1048576, L=420.0, seed=42):
    """float64 uniform points (x, y, z), drawn axis by axis from RandomState(seed)."""
    rs = np.random.RandomState(seed)
    return [rs.uniform(0, L, n) for _ in range(3)]


def neyman_scott(n=4194304, L=1000.0, nparent=2000, sigma=2.0, seed=1002, with_weights=True):
    """Clustered float32 catalog: `nparent` uniform parents, child counts
    multinomial(n, 1/nparent), N(0, sigma^2) offsets per axis, wrapped into
    [0, L). Weights U(0.5, 1.5) float32. Returns (X, Y, Z), W (or None)."""
    rng = np.random.default_rng(seed)
    parents = rng.uniform(0, L, (nparent, 3))
    counts = rng.multinomial(n, np.full(nparent, 1.0 / nparent))
    k = np.repeat(np.arange(nparent), counts)
    pos = parents[k] + rng.normal(0, sigma, (n, 3))
    X = _wrap32(pos, L)
    W = rng.uniform(0.5, 1.5, n).astype(np.float32) if with_weights else None
    return X, W


def uniform32(n, L, rng, with_weights=False):
    """float32 uniform points from a numpy Generator (+ U(0.5,1.5) weights)."""
    X = _wrap32(rng.uniform(0, L, (n, 3)), L)
    W = rng.uniform(0.5, 1.5, n).astype(np.float32) if with_weights else None
    return X, W


def batch_sizes(ncat=256, nmin=50000, nmax=2000000, seed=1003):
    """Catalog sizes round(exp(U(ln nmin, ln nmax)))."""
    rng = np.random.default_rng(seed)
    return np.rint(np.exp(rng.uniform(np.log(nmin), np.log(nmax), ncat))).astype(np.int64), rng


def batch_catalogs(ncat=256, nmin=50000, nmax=2000000, L=500.0, seed=1003, scale=1.0):
    """`ncat` uniform float32 catalogs in [0, L)^3 with log-uniform sizes
    (multiplied by `scale`, for reduced test versions)."""
    sizes, rng = batch_sizes(ncat, nmin, nmax, seed)
    sizes = np.maximum(1, np.rint(sizes * scale).astype(np.int64))
    return [uniform32(int(s), L, rng)[0] for s in sizes]


def _nfw_m(x):
    return np.log1p(x) - x / (1.0 + x)


def _nfw_radii(u, c):
    """x in [0, c] with m(x) = u m(c), m(x) = ln(1+x) - x/(1+x) (bisection, 60 steps)."""
    target = u * _nfw_m(c)
    lo = np.zeros_like(c)
    hi = c.copy()
    for _ in range(60):
        mid = 0.5 * (lo + hi)
        below = _nfw_m(mid) < target
        lo = np.where(below, mid, lo)
        hi = np.where(below, hi, mid)
    return 0.5 * (lo + hi)


def nfw_halos(n=16777216, L=1000.0, nhalo=20000, frac_halo=0.6, sigma_occ=1.0, r_scale=1.0, seed=1004):
    """float32 catalog: round(frac_halo*n) points in `nhalo` NFW halos, the rest
    uniform. Occupancies: log-normal(0, sigma_occ) weights, multinomial draw of
    the halo points. Concentration c ~ U(5, 15); halo radius
    R_h = r_scale * (occ / mean occ)^(1/3); radii by the inverse of the NFW
    enclosed-mass profile on [0, c] (in units of R_h / c); isotropic directions."""
    rng = np.random.default_rng(seed)
    nh = int(round(frac_halo * n))
    centres = rng.uniform(0, L, (nhalo, 3))
    occw = rng.lognormal(0.0, sigma_occ, nhalo)
    occ = rng.multinomial(nh, occw / occw.sum())
    conc = rng.uniform(5.0, 15.0, nhalo)
    rh = r_scale * np.cbrt(np.maximum(occ, 1) / (nh / nhalo))
    k = np.repeat(np.arange(nhalo), occ)
    x = _nfw_radii(rng.uniform(0, 1, nh), conc[k])
    r = x * (rh[k] / conc[k])
    v = rng.normal(0, 1, (nh, 3))
    v /= np.linalg.norm(v, axis=1)[:, None]
    pos = np.concatenate([centres[k] + v * r[:, None], rng.uniform(0, L, (n - nh, 3))])
    return _wrap32(pos, L)


def sky_mock(n=1000000, ra_range=(120.0, 240.0), dec_range=(0.0, 60.0), cz_range=(3000.0, 30000.0),
             frac_clump=0.5, nclump=3000, sigma_deg=0.15, sigma_cz=200.0, seed=1006):
    """float64 on-sky catalog (RA, DEC in degrees, cz in km/s) and pair_product
    weights U(0.5, 1.5): (1 - frac_clump) of the points uniform in solid angle
    and in volume (cz^3 uniform) over the window, the rest in `nclump` Gaussian
    clumps (sigma_deg on the sky, sigma_cz along the line of sight), clipped to
    the window."""
    rng = np.random.default_rng(seed)
    s0, s1 = np.sin(np.radians(dec_range[0])), np.sin(np.radians(dec_range[1]))

    def uniform_sky(m):
        ra = rng.uniform(ra_range[0], ra_range[1], m)
        dec = np.degrees(np.arcsin(rng.uniform(s0, s1, m)))
        cz = np.cbrt(rng.uniform(cz_range[0] ** 3, cz_range[1] ** 3, m))
        return ra, dec, cz

    nc = int(round(frac_clump * n))
    cra, cdec, ccz = uniform_sky(nclump)
    k = rng.integers(0, nclump, nc)
    ra_c = np.clip(cra[k] + rng.normal(0, sigma_deg, nc), *ra_range)
    dec_c = np.clip(cdec[k] + rng.normal(0, sigma_deg, nc), *dec_range)
    cz_c = np.clip(ccz[k] + rng.normal(0, sigma_cz, nc), *cz_range)
    ra_u, dec_u, cz_u = uniform_sky(n - nc)
    ra = np.concatenate([ra_c, ra_u])
    dec = np.concatenate([dec_c, dec_u])
    cz = np.concatenate([cz_c, cz_u])
    w = rng.uniform(0.5, 1.5, n)
    return [ra, dec, cz], w
