"""Counts -> statistics converters (host numpy), pinned by hand-computed
values, the estimator's algebra and a full pipeline on the oracle's counts."""
import numpy as np
import pytest

from corrfunc_amd.utils import convert_3d_counts_to_cf, convert_rp_pi_counts_to_wp
from helpers import run, uniform


def test_landy_szalay_hand_values():
    dd = np.array([10, 40, 0, 7])
    dr = np.array([30, 60, 5, 0])
    rd = np.array([20, 80, 5, 0])
    rr = np.array([100, 400, 0, 50])
    # ND1 = ND2 = 10, NR1 = NR2 = 20 -> fN = 2
    cf = convert_3d_counts_to_cf(10, 10, 20, 20, dd, dr, rd, rr)
    assert cf[0] == pytest.approx((4 * 10 - 2 * 30 - 2 * 20 + 100) / 100)   # 0.4
    assert cf[1] == pytest.approx((4 * 40 - 2 * 60 - 2 * 80 + 400) / 400)   # 0.7
    assert np.isnan(cf[2])                                                  # RR = 0
    assert cf[3] == pytest.approx((4 * 7 + 50) / 50)
    # asymmetric sample sizes use fN1 for D1R2 and fN2 for D2R1
    cf2 = convert_3d_counts_to_cf(10, 5, 30, 10, dd[:1], dr[:1], rd[:1], rr[:1])
    assert cf2[0] == pytest.approx((3 * 2 * 10 - 3 * 30 - 2 * 20 + 100) / 100)


def test_unclustered_counts_give_zero():
    rr = np.array([1000.0, 5000.0])
    # Poisson-free expectation: DD = RR (ND/NR)^2, DR = RD = RR (ND/NR)
    cf = convert_3d_counts_to_cf(100, 100, 400, 400, rr / 16, rr / 4, rr / 4, rr)
    assert np.allclose(cf, 0.0, atol=1e-15)


def test_structured_results_and_errors():
    dt = np.dtype([("rmin", float), ("npairs", np.uint64)])
    a = np.zeros(3, dt)
    a["npairs"] = [4, 5, 6]
    cf = convert_3d_counts_to_cf(1, 1, 1, 1, a, a, a, a)
    assert np.allclose(cf, 0.0)
    with pytest.raises(ValueError):
        convert_3d_counts_to_cf(1, 1, 1, 1, a, a, a, a[:2])
    with pytest.raises(ValueError):
        convert_3d_counts_to_cf(1, 1, 1, 1, a, a, a, a, estimator="natural")


def test_wp_sums_pi_bins():
    nrp, npi = 3, 4
    rr = np.full(nrp * npi, 100)
    dd = np.arange(nrp * npi) + 100           # xi = dd/rr - 1 with all f = 1 and DR = RD = RR
    wp = convert_rp_pi_counts_to_wp(1, 1, 1, 1, dd, rr, rr, rr, nrp, 4.0, dpi=1.0)
    xi = dd / 100.0 - 1
    assert wp == pytest.approx([2 * xi[i * npi:(i + 1) * npi].sum() for i in range(nrp)])
    with pytest.raises(ValueError):
        convert_rp_pi_counts_to_wp(1, 1, 1, 1, dd, rr, rr, rr, nrp, 5.0, dpi=1.0)   # pimax != dpi * npibins
    with pytest.raises(ValueError):
        convert_rp_pi_counts_to_wp(1, 1, 1, 1, dd, rr, rr, rr, 5, 4.0)               # 12 not divisible by 5
    with pytest.raises(ValueError):
        convert_rp_pi_counts_to_wp(1, 1, 1, 1, dd, rr, rr, rr, nrp, 4.0, dpi=0.0)
    rr0 = rr.copy()
    rr0[5] = 0
    assert np.isnan(convert_rp_pi_counts_to_wp(1, 1, 1, 1, dd, rr0, rr0, rr0, nrp, 4.0)[1])


def test_pipeline_on_uniform_points(oracle):
    """D and R drawn from the same uniform distribution: xi ~ 0 within Poisson noise."""
    L = 100.0
    D = uniform(4000, L, 5)
    R = uniform(12000, L, 6)
    bins = np.linspace(2.0, 12.0, 6)
    kw = dict(nthreads=2, binfile=bins)
    from corrfunc_amd import _abi
    o = lambda: _abi.make_options(periodic=True, boxsize=L)  # noqa: E731
    dd = run(oracle, "countpairs", "oracle_", options=o(), autocorr=1, X1=D[0], Y1=D[1], Z1=D[2], **kw)
    dr = run(oracle, "countpairs", "oracle_", options=o(), autocorr=0, X1=D[0], Y1=D[1], Z1=D[2],
             X2=R[0], Y2=R[1], Z2=R[2], **kw)
    rr = run(oracle, "countpairs", "oracle_", options=o(), autocorr=1, X1=R[0], Y1=R[1], Z1=R[2], **kw)
    cf = convert_3d_counts_to_cf(4000, 4000, 12000, 12000, dd["npairs"][1:], dr["npairs"][1:], dr["npairs"][1:],
                                 rr["npairs"][1:])
    assert np.all(np.abs(cf) < 0.05)


def test_small_helpers(tmp_path):
    """translate_isa_string_to_enum, return_file_with_rbins, fix_cz and the
    endianness helpers behave as the reference's (Corrfunc/utils.py:324-964)."""
    import os
    import sys
    from corrfunc_amd import utils as U
    assert [U.translate_isa_string_to_enum(k) for k in ("fastest", "FALLBACK", "sse42", "AVX", "avx2", "AVX512F")] \
        == [-1, 0, 6, 7, 8, 9]
    with pytest.raises(ValueError):
        U.translate_isa_string_to_enum("sse5")
    with pytest.raises(TypeError):
        U.translate_isa_string_to_enum(3)
    name, delete = U.return_file_with_rbins([3.0, 1.0, 2.0])
    assert delete and open(name).read() == "1.0 2.0\n2.0 3.0\n"
    os.remove(name)
    p = tmp_path / "b"
    p.write_text("1 2\n")
    assert U.return_file_with_rbins(str(p)) == (str(p), False)
    with pytest.raises(IOError):
        U.return_file_with_rbins(str(tmp_path / "missing"))
    with pytest.raises(TypeError):
        U.return_file_with_rbins([])
    z = np.array([0.01, 0.02])
    assert np.allclose(U.fix_cz(z), [2998.0, 5996.0]) and z[0] == 2998.0
    cz = np.array([3000.0, 5000.0])
    assert np.array_equal(U.fix_cz(cz), [3000.0, 5000.0])
    swapped = np.arange(5, dtype=(">" if sys.byteorder == "little" else "<") + "f8")
    assert not U.is_native_endian(swapped)
    nat = U.convert_to_native_endian(swapped)
    assert U.is_native_endian(nat) and np.array_equal(nat, np.arange(5.0))
    a = np.arange(3.0)
    assert U.convert_to_native_endian(a) is a and U.convert_to_native_endian(None) is None
    assert U.is_native_endian(None)
