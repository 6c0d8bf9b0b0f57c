"""Host-only pieces of the MHD step: the product's Coriolis coupling table
(closed forms of the Gaunt and Elsasser integrals) against the oracle's
restatement of the reference's exact-arithmetic algorithm
(cal_gaunt_itgs.f90, interact_coriolis_rlm.f90), and the grid helpers."""
import ctypes

import numpy as np
import pytest

import calypso_amd
from calypso_amd import _lib, mhd
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
    r_icb = 1.0 * 0.35 / (1.0 - 0.35)        # set_ctl_gen_shell_grids.f90:179-185
    np.testing.assert_array_equal(r, mhd_step.chebyshev_radius(48, r_icb, r_icb + 1.0))
    assert abs(p.coef_buo - 65000.0) < 1e-9 and p.coef_cor == 2000.0
    assert abs(p.coef_lor - 200.0) < 1e-12 and p.coef_d_magne == 0.2


def test_velocity_wall_kinds():
    """bc_velocity names map to the C ABI's kinds; other names, and other
    kinds at the C ABI, are rejected before any device work."""
    p, _ = mhd.dynamobench_case1()
    assert (p.bc_velo_icb, p.bc_velo_cmb) == (mhd.VELO_RIGID, mhd.VELO_RIGID)
    p, _ = mhd.dynamobench_case1(bc_velo=("free_slip_sph", "non_slip_sph"))
    assert (p.bc_velo_icb, p.bc_velo_cmb) == (mhd.VELO_FREE_SLIP, mhd.VELO_RIGID)
    with pytest.raises(_lib.CalypsoError, match="not supported"):
        mhd.dynamobench_case1(bc_velo=("rotatable_inner_core", "non_slip_sph"))
    lib = calypso_amd.load()
    p, r = mhd.dynamobench_case1()
    p.bc_velo_cmb = 2
    h = ctypes.c_void_p()
    rc = lib.calypso_mhd_create(ctypes.byref(p), r.ctypes.data_as(ctypes.c_void_p), None,
                                ctypes.byref(h))
    assert rc != 0 and "velocity boundary kind 2" in lib.calypso_last_error().decode()
    assert not h.value


def test_inner_core_model_parameters_and_rejections():
    """dynamobench_case2 gives the reference's case 2 grid (the radii its
    gen_sph_grids writes) and boundary kinds; combinations outside the two
    supported models are rejected at the C ABI before any device work."""
    import json
    import os
    ref = json.load(open(os.path.join(os.path.dirname(__file__), "golden",
                                      "dynamobench_case2_steps.json")))
    p, r = mhd.dynamobench_case2()
    assert (p.nr, p.kr_icb, p.kr_cmb) == (96, ref["icb_id"], ref["cmb_id"])
    np.testing.assert_allclose(r, np.array(ref["radius"]), rtol=1e-15, atol=0)
    assert (p.bc_velo_icb, p.bc_magne_icb) == (mhd.VELO_ROT_INNER_CORE, mhd.MAGNE_TO_CENTER)
    assert abs(p.coef_buo - 0.65 * 110.0 / 1e-3) < 1e-9
    lib = calypso_amd.load()

    def create(q, rad):
        h = ctypes.c_void_p()
        rc = lib.calypso_mhd_create(ctypes.byref(q), rad.ctypes.data_as(ctypes.c_void_p), None,
                                    ctypes.byref(h))
        assert not h.value
        return rc, lib.calypso_last_error().decode()

    q, rad = mhd.dynamobench_case2()
    q.bc_magne_icb = mhd.MAGNE_INSULATOR          # inner core without the field to the centre
    rc, msg = create(q, rad)
    assert rc != 0 and "points below the ICB" in msg
    q, rad = mhd.dynamobench_case2()
    q.bc_velo_icb = mhd.VELO_RIGID                 # grid below the ICB with a rigid wall
    rc, msg = create(q, rad)
    assert rc != 0 and "points below the ICB" in msg
    q, rad = mhd.dynamobench_case1()
    q.bc_velo_icb = mhd.VELO_ROT_INNER_CORE        # rotatable core on a grid without one
    rc, msg = create(q, rad)
    assert rc != 0 and "points below the ICB" in msg
    q, rad = mhd.dynamobench_case2()
    q.bc_magne_icb = 7
    rc, msg = create(q, rad)
    assert rc != 0 and "magnetic boundary kind 7" in msg
