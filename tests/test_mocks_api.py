"""Argument checks of corrfunc_amd.mocks (raised before any pair counting, as
in the reference wrappers Corrfunc/mocks/DDrppi_mocks.py:333-345 and
DDsmu_mocks.py:256-277) and corrfunc_amd.utils.fix_ra_dec."""
import numpy as np
import pytest

from corrfunc_amd.mocks import DDrppi_mocks, DDsmu_mocks
from corrfunc_amd.utils import fix_ra_dec


def _cat(n=10):
    rng = np.random.default_rng(0)
    return rng.uniform(0, 10, n), rng.uniform(0, 10, n), rng.uniform(5000, 6000, n)


def test_cross_needs_second_set():
    ra, dec, cz = _cat()
    with pytest.raises(ValueError):
        DDrppi_mocks(0, 1, 1, 40.0, np.array([1.0, 2.0]), ra, dec, cz)
    with pytest.raises(ValueError):
        DDsmu_mocks(0, 1, 1, 1.0, 5, np.array([1.0, 2.0]), ra, dec, cz, RA2=ra)


@pytest.mark.parametrize("mu_max,exc", [(1.5, ValueError), (0.0, ValueError), (-0.1, ValueError),
                                        ([0.5, 0.6], TypeError)])
def test_mu_max_checks(mu_max, exc):
    ra, dec, cz = _cat()
    with pytest.raises(exc):
        DDsmu_mocks(1, 1, 1, mu_max, 5, np.array([1.0, 2.0]), ra, dec, cz)


def test_bad_isa_and_weight_type():
    ra, dec, cz = _cat()
    with pytest.raises(ValueError):
        DDrppi_mocks(1, 1, 1, 40.0, np.array([1.0, 2.0]), ra, dec, cz, isa="sse5")
    with pytest.raises(ValueError):
        DDrppi_mocks(1, 1, 1, 40.0, np.array([1.0, 2.0]), ra, dec, cz, weights1=np.ones(10), weight_type="bogus")


def test_fix_ra_dec():
    ra = np.array([-170.0, 10.0, 150.0])
    dec = np.array([100.0, 95.0, 170.0])
    r, d = fix_ra_dec(ra, dec)
    assert np.array_equal(ra, [10.0, 190.0, 330.0]) and np.array_equal(dec, [10.0, 5.0, 80.0])
    assert r.dtype == ra.dtype
    ra2, dec2 = np.array([10.0]), np.array([20.0])
    fix_ra_dec(ra2, dec2)
    assert ra2[0] == 10.0 and dec2[0] == 20.0
    with pytest.raises(TypeError):
        fix_ra_dec([1.0], [2.0])
