"""CPU pins of the oracle's explicit time integration and volume mean squares.

explicit update (cal_vorticity_terms_adams.f90:31-51,
cal_diff_adv_src_explicit.f90:129-154): Adams-Bashforth 2 integrates a
forcing linear in time exactly, so a run started with the Euler step that
seeds the history must land on the closed-form integral; a wrong history
sign or weight breaks that.

mean squares (cal_rms_by_sph_spectr.f90:63-154, radial_int_for_sph_spec.f90:
66-94): the per-mode spectral formulas are checked against a quadrature of
|u|^2 / f^2 on the physical grid (oracle Legendre + longitude transforms of
the same spectrum), and the radial average against a closed form.
"""
import numpy as np
import pytest

import oracle


def _linear_forcing_run(kind, dt, nsteps, a, b, c, d0):
    """x' = c*D + s*(a + b t) with D fixed; returns x after nsteps from t=0."""
    n = a.size
    x = np.zeros(n)
    pre = np.zeros(n)
    sign = 1.0 if kind == 0 else -1.0
    for i in range(nsteps):
        t = i * dt
        force = sign * (a + b * t)   # kind 1 takes the advection term A with x' = -A
        scheme = 2 if i == 0 else 0
        x, pre = oracle.explicit_update(kind, scheme, dt, c, x, d0, force, pre)
    return x


@pytest.mark.parametrize("kind", [0, 1])
def test_adams_bashforth_exact_for_linear_forcing(kind):
    rng = np.random.default_rng(5)
    n, dt, nsteps = 17, 1e-2, 40
    a, b, d0 = rng.standard_normal(n), rng.standard_normal(n), rng.standard_normal(n)
    c = 0.4
    x = _linear_forcing_run(kind, dt, nsteps, a, b, c, d0)
    # Euler first step: x1 = dt (c d0 + a); AB2 after that is exact for a + b t
    T = nsteps * dt
    x1 = dt * (c * d0 + a)
    exact = x1 + c * d0 * (T - dt) + a * (T - dt) + 0.5 * b * (T * T - dt * dt)
    assert np.max(np.abs(x - exact)) < 1e-12


def test_euler_scheme_keeps_history_untouched():
    n = 5
    x0 = np.arange(n, dtype=float)
    pre = np.full(n, 7.0)
    x, p = oracle.explicit_update(0, 1, 0.5, 2.0, x0, np.ones(n), np.ones(n), pre)
    assert np.array_equal(p, pre)
    assert np.array_equal(x, x0 + 0.5 * (2.0 + 1.0))
    x, p = oracle.explicit_update(1, 2, 0.5, 2.0, x0, np.ones(n), np.ones(n), pre)
    assert np.array_equal(p, -np.ones(n))
    assert np.array_equal(x, x0 + 0.5 * (2.0 - 1.0))


def _physical_mean_square(L, nth, radius, sp, ncomp, nvec, nscl):
    """r^2 <|f|^2>_sphere at each radius by Gauss-Legendre x uniform-phi
    quadrature of the synthesized field (exact for degree <= 2L)."""
    grid = oracle.Grid(L, nth, radius)
    nphi = 4 * (L + 1)
    vr = oracle.leg_bwd(grid, ncomp, nvec, nscl, sp)
    f = oracle.fft_bwd(L, nth, grid.nr, nphi, ncomp, vr).reshape(ncomp, nphi, nth, grid.nr)
    sq = (f ** 2).sum(axis=0)                       # (phi, t, k)
    mean = np.einsum("ptk,t->k", sq, grid.weight) / (2.0 * nphi)
    return mean * radius ** 2


def _trapezoid(radius, y, kr_st, kr_ed):
    s = 0.0
    for kr in range(kr_st, kr_ed):
        s += 0.5 * (radius[kr] - radius[kr - 1]) * (y[kr - 1] + y[kr])
    return 3.0 * s / (radius[kr_ed - 1] ** 3 - radius[kr_st - 1] ** 3)


@pytest.mark.parametrize("part", ["pol", "tor"])
def test_vector_mean_square_matches_grid_quadrature(part):
    L, nth, nr = 9, 12, 7
    radius = np.linspace(0.6, 1.6, nr)
    J = (L + 1) ** 2
    l_of_j = oracle.rlm_index(L)[0][:, 1]
    rng = np.random.default_rng(11)
    rj = np.zeros((3, nr, J))
    if part == "pol":
        rj[0] = rng.standard_normal((nr, J))
        rj[1] = rng.standard_normal((nr, J))
    else:
        rj[2] = rng.standard_normal((nr, J))
    sp = np.ascontiguousarray(rj.transpose(1, 2, 0)).ravel()   # sp[nd + 3 (j + J k)]
    phys = _physical_mean_square(L, nth, radius, sp, 3, 1, 0)
    want = _trapezoid(radius, phys, 2, nr)
    got = oracle.sph_mean_square(0, radius, l_of_j, 2, nr, rj)
    idx = 0 if part == "pol" else 1
    assert abs(got[idx] - want) < 1e-10 * abs(want)
    assert abs(got[1 - idx]) == 0.0
    assert got[2] == got[0] + got[1]
    half = oracle.sph_mean_square(1, radius, l_of_j, 2, nr, rj)
    assert np.array_equal(half, 0.5 * got)


def test_scalar_mean_square_matches_grid_quadrature():
    L, nth, nr = 8, 11, 6
    radius = np.linspace(0.5, 1.5, nr)
    J = (L + 1) ** 2
    l_of_j = oracle.rlm_index(L)[0][:, 1]
    rng = np.random.default_rng(3)
    rj = rng.standard_normal((1, nr, J))
    sp = np.ascontiguousarray(rj.transpose(1, 2, 0)).ravel()
    phys = _physical_mean_square(L, nth, radius, sp, 1, 0, 1)
    want = _trapezoid(radius, phys, 1, nr)
    got = oracle.sph_mean_square(2, radius, l_of_j, 1, nr, rj)
    assert abs(got[0] - want) < 1e-10 * abs(want)


def test_volume_average_closed_form():
    # a uniform scalar (l = 0 only) of value d: <f^2>_vol -> d^2 as dr -> 0
    nr = 2001
    radius = np.linspace(0.35, 1.35, nr)
    l_of_j = np.array([0, 1, 1, 1], dtype=np.int32)
    rj = np.zeros((1, nr, 4))
    rj[0, :, 0] = 1.7
    got = oracle.sph_mean_square(2, radius, l_of_j, 1, nr, rj)[0]
    assert abs(got - 1.7 ** 2) < 1e-6
