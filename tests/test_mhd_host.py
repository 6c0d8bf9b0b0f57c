"""Host-only pieces of the MHD step: the product's Coriolis coupling table
(closed forms of the Gaunt and Elsasser integrals) against the oracle's
restatement of the reference's exact-arithmetic algorithm
(cal_gaunt_itgs.f90, interact_coriolis_rlm.f90), and the grid helpers."""
import numpy as np

from calypso_amd import mhd
from oracle import mhd_step


def test_coriolis_table_matches_reference_algorithm():
    L = 31
    cor, idx = mhd.coriolis_table(L)
    c = mhd_step.CoriolisRlm(L)
    want = np.array([c.sw11, c.sw21, c.sw12, c.sw22, c.sw13, c.tw13, c.tw23, c.tw14,
                     c.tw24, c.tw11, c.tw12, c.g3rlm])
    np.testing.assert_allclose(cor, want, rtol=1e-13, atol=1e-15)
    assert np.array_equal(idx[0], c.jk1)
    assert np.array_equal(idx[1], c.jk2)
    assert np.array_equal(idx[2], c.jl1)


def test_dynamobench_parameters():
    p, r = mhd.dynamobench_case1()
    assert r.size == 49 and abs(r[0] - 0.35 / 0.65) < 1e-15 and abs(r[-1] - 1 / 0.65) < 1e-15
    np.testing.assert_array_equal(r, mhd_step.chebyshev_radius(48, (1 / 0.65) * 0.35, 1 / 0.65))
    assert abs(p.coef_buo - 65000.0) < 1e-9 and p.coef_cor == 2000.0
    assert abs(p.coef_lor - 200.0) < 1e-12 and p.coef_d_magne == 0.2
