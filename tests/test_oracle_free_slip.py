"""The oracle's free-slip velocity walls (oracle/mhd_step.py) pinned three
ways: the wall tables reproduce every quadratic that satisfies the wall
condition (the mathematics), the tables and each wall routine agree with the
reference's own routines compiled from its sources (oracle/_ref), and the
whole step keeps a solid-body rotation exactly, as free-slip walls must and
rigid walls must not."""
import ctypes
import json
import os

import numpy as np
import pytest

import oracle
from oracle import mhd_step as ms

_P, _I, _D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref not built")


def _radius(nri=12):
    r_cmb = 1.0 / (1.0 - 0.35)
    return ms.chebyshev_radius(nri, 0.35 * r_cmb, r_cmb)


def _ptr(a):
    return a.ctypes.data_as(_P)


def test_wall_tables_reproduce_the_wall_quadratics():
    """A quadratic f about the wall point r0 with f(r0) = 0 and
    f - r f' + r^2/2 f'' = 0 (the poloidal free-slip condition), or with
    2 f - r f' = 0 (the toroidal one): the tables give f'(r0) and f''(r0)
    exactly from f's values at the wall and at its neighbour."""
    r = _radius()
    bc = ms.BoundaryFdm(r, 1, r.size)
    for r0, rn, fvp, fvt, nb in ((r[0], r[1], bc.free_vp_icb, bc.free_vt_icb, 2),
                                 (r[-1], r[-2], bc.free_vp_cmb, bc.free_vt_cmb, 0)):
        h = rn - r0
        # poloidal: a = 0, b = r0 c / 2
        c = 1.7
        b = 0.5 * r0 * c
        f0, fn = 0.0, b * h + 0.5 * c * h * h
        assert fvp[1][1] * f0 + fvp[nb][1] * fn == pytest.approx(b, rel=1e-13)
        assert fvp[1][2] * f0 + fvp[nb][2] * fn == pytest.approx(c, rel=1e-13)
        # toroidal: b = 2 a / r0
        a, c = 0.9, -0.4
        b = 2.0 * a / r0
        f0, fn = a, a + b * h + 0.5 * c * h * h
        assert fvt[1][1] * f0 == pytest.approx(b, rel=1e-13)
        assert fvt[1][2] * f0 + fvt[nb][2] * fn == pytest.approx(c, rel=1e-13)


@needs_ref
def test_wall_tables_match_the_reference():
    r = _radius()
    bc = ms.BoundaryFdm(r, 1, r.size)
    coef = np.zeros((4, 3, 3))
    oracle.ref().ref_free_slip_coefs(_I(r.size), _I(1), _I(r.size), _ptr(r), _ptr(coef))
    for t, mine in enumerate((bc.free_vp_icb, bc.free_vt_icb, bc.free_vp_cmb, bc.free_vt_cmb)):
        np.testing.assert_allclose(np.array(mine), coef[t], rtol=1e-14, atol=1e-14 * np.abs(coef[t]).max())


@needs_ref
@pytest.mark.parametrize("which", [0, 1, 2, 3])
@pytest.mark.parametrize("wall", [0, 1])
def test_wall_routines_match_the_reference(which, wall):
    """0 v_and_w, 1 vpol2, 2 rot2, 3 w_diffuse2 on seeded fields, at the ICB
    (wall 0) and the CMB (wall 1)."""
    r = _radius()
    L = 3
    J = (L + 1) ** 2
    nri = r.size
    g3 = np.array([float(int(np.sqrt(j)) * (int(np.sqrt(j)) + 1)) for j in range(J)])
    bc = ms.BoundaryFdm(r, 1, nri)
    rng = np.random.default_rng(5 + which + 10 * wall)
    d = rng.standard_normal((6, nri, J))
    fix = np.zeros((2, 3, 3))                 # Fortran fix_fld(0:2, 3, 2)
    fix[0] = np.array(bc.fix_fld_icb).T
    fix[1] = np.array(bc.fix_fld_cmb).T
    coef_d = 0.37
    want = d.copy()
    oracle.ref().ref_free_slip_wall(_I(which), _I(wall), _I(nri), _I(J), _I(1), _I(nri),
                                    _ptr(r), _ptr(g3), _D(coef_d), _ptr(fix), _ptr(want))
    got = d.copy()
    f, o = got[:3], got[3:]
    k = 0 if wall == 0 else nri - 1
    fvp = bc.free_vp_icb if wall == 0 else bc.free_vp_cmb
    fvt = bc.free_vt_icb if wall == 0 else bc.free_vt_cmb
    r3 = bc.r_icb if wall == 0 else bc.r_cmb
    ffl = bc.fix_fld_icb if wall == 0 else bc.fix_fld_cmb
    if which == 0:
        (ms.icb_free_v_and_w if wall == 0 else ms.cmb_free_v_and_w)(fvp, fvt, f, o, k)
    elif which == 1:
        (ms.icb_free_vpol2 if wall == 0 else ms.cmb_free_vpol2)(fvp, f, k)
    elif which == 2:
        (ms.icb_free_rot2 if wall == 0 else ms.cmb_free_rot2)(fvp, fvt, r3, g3, f, o, k)
    else:
        (ms.icb_free_w_diffuse2 if wall == 0 else ms.cmb_free_w_diffuse2)(
            ffl, fvt, r3, g3, coef_d, f, o, k)
    assert not np.array_equal(want, d), "the reference routine changed nothing"
    np.testing.assert_allclose(got, want, rtol=1e-13, atol=1e-13 * np.abs(want).max())


@needs_ref
@pytest.mark.parametrize("wall", [0, 1])
def test_matrix_rows_match_the_reference(wall):
    r = _radius()
    nri, J = r.size, 9
    g3 = np.array([float(int(np.sqrt(j)) * (int(np.sqrt(j)) + 1)) for j in range(J)])
    bc = ms.BoundaryFdm(r, 1, nri)
    rng = np.random.default_rng(3 + wall)
    mat = rng.standard_normal((J, nri, 3))
    for fdm in ((bc.free_vp_icb, bc.free_vt_icb) if wall == 0
                else (bc.free_vp_cmb, bc.free_vt_cmb)):
        want = mat.copy()
        k = 1 if wall == 0 else nri
        r3 = np.array(bc.r_icb if wall == 0 else bc.r_cmb)
        tab = np.ascontiguousarray(np.array(fdm))          # Fortran fdm(3, -1:1)
        oracle.ref().ref_add_fix_flux_rows(_I(wall), _I(nri), _I(J), _I(k), _ptr(r3),
                                           _ptr(tab), _ptr(g3), _D(0.21), _ptr(want))
        got = mat.copy()
        if wall == 0:
            ms.add_fix_flux_icb_rows(got, 0, bc.r_icb, fdm, g3, 0.21)
        else:
            ms.add_fix_flux_cmb_rows(got, nri - 1, bc.r_cmb, fdm, g3, 0.21)
        np.testing.assert_allclose(got, want, rtol=1e-14, atol=1e-14)


def _solid_body_change(free_icb, free_cmb, steps=20, bc_hook=None):
    case = ms.shell_case(L=7, nth=12, nphi=24, nri=16, dt=1e-4,
                         free_icb=free_icb, free_cmb=free_cmb)
    m = ms.MhdStep(case)
    if bc_hook is not None:
        bc_hook(m)
        m._matrices()
    m.initialize(seeded=(1, 0.0, 0.0, 0.0))
    m.trans_per_temp_to_temp()
    m.f["velo"][2][:, 2] = 0.1 * m.radius ** 2          # T_10 = r^2: solid rotation
    m._start()
    t0 = m.f["velo"][2][:, 2].copy()
    for _ in range(steps):
        m.step()
    return np.abs(m.f["velo"][2][:, 2] - t0).max() / np.abs(t0).max()


def test_free_slip_keeps_solid_body_rotation():
    assert _solid_body_change(True, True) < 1e-12
    # a rigid wall on either side brakes it
    assert _solid_body_change(False, False) > 1e-2
    assert _solid_body_change(True, False) > 1e-2
    assert _solid_body_change(False, True) > 1e-2


def test_planted_toroidal_wall_mistake_breaks_solid_rotation():
    """The toroidal wall condition written f - r f' = 0 (a dropped factor 2)
    must spoil the invariant."""
    def wrong(m):
        r, bc = m.radius, m.bc
        r0, h = r[0], r[1] - r[0]
        mvt = ms.inverse33([[1.0, 0.0, 0.0], [1.0, -r0, 0.0], [1.0, h, 0.5 * h * h]])
        bc.free_vt_icb = [[0.0, 0.0, 0.0], [1.0, mvt[1][0], mvt[2][0]], [0.0, 0.0, mvt[2][2]]]
    assert _solid_body_change(True, True, bc_hook=wrong) > 1e-6


FREE_STEPS = os.path.join(os.path.dirname(__file__), "golden", "dynamobench_free_slip_steps.json")


@pytest.mark.parametrize("run,walls", [("free_free", (True, True)), ("free_rigid", (True, False))])
def test_free_slip_energies_match_reference_program(run, walls):
    """The oracle's Dynamobench case 1 with free-slip velocity walls against
    the energies the reference's own MHD program (built from its sources)
    writes after every one of 20 steps of the same run
    (tests/golden/dynamobench_free_slip_steps.json): within 1e-12 (reached:
    2e-13)."""
    ref = json.load(open(FREE_STEPS))["runs"][run]
    s = ms.MhdStep(ms.shell_case(free_icb=walls[0], free_cmb=walls[1]))
    s.initialize()
    for _ in range(20):
        s.step()
        np.testing.assert_allclose(s.energies(), np.array(ref["steps"][str(s.i_step)]),
                                   rtol=1e-12, atol=0, err_msg=f"{run} step {s.i_step}")
