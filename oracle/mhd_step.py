"""Test-only CPU restatement of the reference's full MHD time step
(SPH_analyze_MHD, programs/SPH_MHD/SPH_analyzer_MHD.f90:115-202) for a
rotating Boussinesq shell with rigid walls, insulating magnetic boundaries,
fixed-temperature boundaries and the spherical-shell reference temperature --
the model of the reference's Dynamobench_case1 test -- and the same shell with
free-slip velocity walls at either boundary (bc_velocity free_slip_sph).

It is the checker for the GPU step, never the product.  Each routine cites the
reference routine it restates; numpy arrays are indexed [k, j] with k the
radial point (0-based) and j the rj mode, j = l(l+1) + m.

With DynamoBenchmarkCase2 it restates the model of the reference's
Dynamobench_case2 test: the radial grid continues through a conducting inner
core to the centre and beyond the core-mantle boundary
(count_chebyshev_ext_layers / set_chebyshev_distance_shell), the magnetic
field is solved from the first point above the centre (sph_to_center:
set_sph_exp_nod_center.f90, center_sph_matrices.f90,
m_coef_fdm_to_center.f90) to the insulating CMB and continued above it as a
potential field (ext_outside_potential_with_j, extend_potential_field.f90), and
the inner core turns as a solid body about any axis under the viscous,
Lorentz and Coriolis torques at the ICB (bc_velocity rot_inner_core:
cal_inner_core_rotation.f90, cal_sph_nod_icb_rotate_velo2,
inner_core_rot_z_coriolis_rlm).  The rj centre node that the reference adds
for degree-0 scalars is not kept: with a fixed ICB temperature its l = 0
solve is the same as the shell's (const_radial_mat_scalar00_sph with
add_scalar_poisson_mat_no_fld leaves the centre row decoupled).

Composition and the diagnostic-only fields (pressure, grad T, viscous
diffusion of u) are not part of this model and are not restated.
"""
from __future__ import annotations

import math
from fractions import Fraction

import numpy as np

from . import (Grid, fft_bwd, fft_fwd, leg_bwd, leg_fwd, ludcmp, lubksb,
               products_mhd, rlm_index, sph_mean_square, sph_normalizations)


# --------------------------------------------------------------- helpers --

def _det22(a11, a12, a21, a22):
    """cal_det_22_matrix (cal_inverse_small_matrix.f90): a11 a22 - a21 a12."""
    return a11 * a22 - a21 * a12


def inverse33(a):
    """cal_inverse_33_matrix (SERIAL_src/BASE/cal_inverse_small_matrix.f90):
    cofactor expansion along the first row, adjugate / det.  a is a 3x3 list
    of rows (a[i][j] = Fortran a(i+1, j+1))."""
    det = 0.0
    for k in range(3):
        cols = [c for c in range(3) if c != k]
        d = _det22(a[1][cols[0]], a[1][cols[1]], a[2][cols[0]], a[2][cols[1]])
        det = det + (-1) ** k * d * a[0][k]
    a_det = 1.0 / det
    inv = [[0.0] * 3 for _ in range(3)]
    for kj in range(3):
        for ki in range(3):
            rows = [r for r in range(3) if r != ki]
            cols = [c for c in range(3) if c != kj]
            d = _det22(a[rows[0]][cols[0]], a[rows[0]][cols[1]],
                       a[rows[1]][cols[0]], a[rows[1]][cols[1]])
            inv[kj][ki] = (-1) ** (ki + kj) * d * a_det
    return inv


def chebyshev_radius(nri, r_icb, r_cmb):
    """set_chebyshev_distance_shell with no points outside the shell
    (SERIAL_src/SPH_SPECTR_src/set_radial_grid_sph_shell.f90:30-70)."""
    pi = 4.0 * math.atan(1.0)
    shell = r_cmb - r_icb
    return np.array([r_icb + 0.5 * shell * (1.0 - math.cos(pi * float(k) / float(nri)))
                     for k in range(nri + 1)])


def chebyshev_radius_ext(nri, r_icb, r_cmb, r_min, r_max):
    """count_chebyshev_ext_layers then set_chebyshev_distance_shell
    (set_radial_grid_sph_shell.f90:60-109, :171-229): Chebyshev points in the
    shell, mirrored Chebyshev spacing below the ICB down to r_min (then evenly
    spaced) and above the CMB up to r_max.  Returns (radius, kr_icb, kr_cmb)
    with 1-based boundary indices."""
    pi = 4.0 * math.atan(1.0)
    shell = r_cmb - r_icb
    half_n = nri // 2
    dr = 0.5 * shell * (1.0 - math.cos(pi / float(nri)))
    r, k = r_icb, 0
    while not (r <= r_min or k >= half_n):
        k += 1
        r = r_icb - 0.5 * shell * (1.0 - math.cos(pi * float(k) / float(nri)))
        dr = 0.5 * shell * (math.cos(pi * float(k - 1) / float(nri))
                            - math.cos(pi * float(k) / float(nri)))
    if k == half_n:
        k = int((r - r_min) / dr) + half_n
    n_icore = max(k - 1, 0)
    r, k = r_cmb, 0
    while not (r >= r_max or k >= half_n):
        k += 1
        r = r_cmb + 0.5 * shell * (1.0 - math.cos(pi * float(k) / float(nri)))
        dr = 0.5 * shell * (-math.cos(pi * float(k) / float(nri))
                            + math.cos(pi * float(k - 1) / float(nri)))
    while r < r_max:
        k += 1
        r = r + dr
    n_ext = 0 if k <= 1 else k
    k_icb = n_icore + 1
    k_cmb = k_icb + nri
    n = k_cmb + n_ext
    rg = np.zeros(n)
    for kk in range(1, k_icb - half_n):
        rg[kk - 1] = (r_icb - 0.5 * shell) * float(kk) / float(k_icb - half_n)
    for kk in range(max(1, k_icb - half_n), k_icb):
        rg[kk - 1] = r_icb - 0.5 * shell * (1.0 - math.cos(pi * float(kk - k_icb) / float(nri)))
    for kk in range(k_icb, k_cmb + 1):
        rg[kk - 1] = r_icb + 0.5 * shell * (1.0 - math.cos(pi * float(kk - k_icb) / float(nri)))
    ked = min(n, k_cmb + half_n)
    for kk in range(k_cmb + 1, ked + 1):
        rg[kk - 1] = r_cmb + 0.5 * shell * (1.0 - math.cos(pi * float(kk - k_cmb) / float(nri)))
    if ked > k_cmb:
        dr = rg[ked - 1] - rg[ked - 2]
    for kk in range(k_cmb + half_n + 1, n + 1):
        rg[kk - 1] = rg[kk - 2] + dr
    return rg, k_icb, k_cmb


def center_fdm(r1, r2):
    """cal_2nd_to_center_fixed_fdm (m_coef_fdm_to_center.f90:73-109): the
    three-point Taylor inverse at the first point with the centre as the
    point below (value there fixed to zero).  Returns [pt + 1][c] for pt
    -1 (centre), 0 (first point), +1 (second point), c = 0, 1, 2."""
    dn, dp = r1, r2 - r1
    m = inverse33([[1.0, 0.0, 0.0], [1.0, -dn, dn * dn / 2.0], [1.0, dp, dp * dp / 2.0]])
    return [[m[c][1] for c in range(3)], [m[c][0] for c in range(3)],
            [m[c][2] for c in range(3)]]


def fdm_coefs(radius, below_icb=False):
    """d1nod_mat_fdm_2 / d2nod_mat_fdm_2: 3-point Taylor inverse at every
    point (set_dr_for_nonequi, nod_r_2nd_fdm_coefs_nonequi,
    SERIAL_src/SPH_SPECTR_src/set_radius_func_noequi.f90:173-197; copy order
    copy_fdm_nod_coefs_from_mat, m_fdm_coefs.f90).  When the grid has points
    below the ICB (below_icb), the first point's point below is the centre
    (distance radius[0]).  Returns (d1, d2), each (nr, 3) for offsets
    -1, 0, +1."""
    nr = radius.size
    dr = np.zeros(nr)
    for kr in range(1, nr - 1):
        dr[kr] = radius[kr + 1] - radius[kr]
    dr[nr - 1] = radius[nr - 1] - radius[nr - 2]
    dr[0] = radius[1] - radius[0]
    d1 = np.zeros((nr, 3))
    d2 = np.zeros((nr, 3))
    for k in range(nr):
        if k == 0:
            dp, dn = dr[0], (radius[0] if below_icb else dr[0])
        elif k == nr - 1:
            dp, dn = dr[k], dr[k]
        else:
            dp, dn = dr[k], dr[k - 1]
        m = inverse33([[1.0, 0.0, 0.0],
                       [1.0, dp, 0.5 * dp * dp],
                       [1.0, -dn, 0.5 * dn * dn]])
        d1[k] = (m[1][2], m[1][0], m[1][1])
        d2[k] = (m[2][2], m[2][0], m[2][1])
    return d1, d2


class BoundaryFdm:
    """cal_fdm_coefs_4_BCs (MHD_src/sph_MHD/t_boundary_params_sph_MHD.f90) with
    the one-sided formulas of cal_fdm_coefs_4_boundaries.f90."""

    def __init__(self, radius, kr_in, kr_out):
        r = radius
        i0 = kr_in - 1
        dp1, dp2 = r[i0 + 1] - r[i0], r[i0 + 2] - r[i0]
        m = inverse33([[1.0, 0.0, 0.0], [1.0, dp1, dp1 * dp1 / 2.0],
                       [1.0, dp2, dp2 * dp2 / 2.0]])
        # fix_fld_icb[i][c]: point kr_in+i, c = 0 value, 1 d/dr, 2 d2/dr2
        self.fix_fld_icb = [[m[c][i] for c in range(3)] for i in range(3)]
        m = inverse33([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [dp1, 1.0, dp1 * dp1 / 2.0]])
        # fix_dr_icb[i+1][c]: i = -1 the given d/dr, 0 point kr_in, 1 kr_in+1
        self.fix_dr_icb = [[m[c][i] for c in range(3)] for i in range(3)]
        o = kr_out - 1
        dn1, dn2 = r[o] - r[o - 1], r[o] - r[o - 2]
        m = inverse33([[1.0, 0.0, 0.0], [1.0, -dn1, dn1 * dn1 / 2.0],
                       [1.0, -dn2, dn2 * dn2 / 2.0]])
        # fix_fld_cmb[i][c]: point kr_out-i
        self.fix_fld_cmb = [[m[c][i] for c in range(3)] for i in range(3)]
        m = inverse33([[1.0, 0.0, 0.0], [0.0, 1.0, 0.0], [1.0, -dn1, dn1 * dn1 / 2.0]])
        # fix_dr_cmb[i+1][c]: i = -1 point kr_out-1, 0 point kr_out, 1 given d/dr
        self.fix_dr_cmb = [[m[c][2] for c in range(3)], [m[c][0] for c in range(3)],
                           [m[c][1] for c in range(3)]]
        self.r_icb = (r[i0], 1.0 / r[i0], (1.0 / r[i0]) ** 2)
        self.r_cmb = (r[o], 1.0 / r[o], (1.0 / r[o]) ** 2)
        # free-slip walls (cal_fdm2_ICB_free_vp / _vt, m_coef_fdm_free_ICB.f90:
        # 85-173; cal_fdm2_CMB_free_vp / _vt, m_coef_fdm_free_CMB.f90): the
        # second Taylor row is the wall condition, f - r f' + r^2/2 f'' = 0 for
        # the poloidal scalar (with f = 0) and 2 f - r f' = 0 for the toroidal
        # one.  Stored as [pt + 1][c] like fix_dr_*: pt -1 / 0 / +1 the points
        # below / at / above the wall, c = 0 value, 1 d/dr, 2 d2/dr2.
        r0 = r[i0]
        mvp = inverse33([[1.0, 0.0, 0.0], [1.0, -r0, 0.5 * r0 * r0],
                         [1.0, dp1, 0.5 * dp1 * dp1]])
        mvt = inverse33([[1.0, 0.0, 0.0], [2.0, -r0, 0.0], [1.0, dp1, 0.5 * dp1 * dp1]])
        self.free_vp_icb = [[0.0, 0.0, 0.0], [1.0, mvp[1][0], mvp[2][0]],
                            [0.0, mvp[1][2], mvp[2][2]]]
        self.free_vt_icb = [[0.0, 0.0, 0.0], [1.0, mvt[1][0], mvt[2][0]],
                            [0.0, 0.0, mvt[2][2]]]
        r0 = r[o]
        mvp = inverse33([[1.0, 0.0, 0.0], [1.0, -r0, 0.5 * r0 * r0],
                         [1.0, -dn1, 0.5 * dn1 * dn1]])
        mvt = inverse33([[1.0, 0.0, 0.0], [2.0, -r0, 0.0], [1.0, -dn1, 0.5 * dn1 * dn1]])
        self.free_vp_cmb = [[0.0, mvp[1][2], mvp[2][2]], [1.0, mvp[1][0], mvp[2][0]],
                            [0.0, 0.0, 0.0]]
        self.free_vt_cmb = [[0.0, 0.0, mvt[2][2]], [1.0, mvt[1][0], mvt[2][0]],
                            [1.0, 1.0, 1.0]]


# ----------------------------------------------- free-slip wall formulas --
# Each takes the wall coefficients of BoundaryFdm ([pt + 1][c]) and arrays
# indexed [component, k, j] (matrices [j, k, band]) and updates the wall row
# in place.

def add_fix_flux_icb_rows(m, ki, r_icb, fdm, g3, coef):
    """add_fix_flux_icb_poisson_mat (set_sph_scalar_mat_bc.f90:78-101)."""
    m[:, ki, 1] = m[:, ki, 1] - coef * (fdm[1][2] - g3 * r_icb[2])
    m[:, ki + 1, 0] = m[:, ki + 1, 0] - coef * fdm[2][2]


def add_fix_flux_cmb_rows(m, ko, r_cmb, fdm, g3, coef):
    """add_fix_flux_cmb_poisson_mat (set_sph_scalar_mat_bc.f90:152-175)."""
    m[:, ko - 1, 2] = m[:, ko - 1, 2] - coef * fdm[0][2]
    m[:, ko, 1] = m[:, ko, 1] - coef * (fdm[1][2] - g3 * r_cmb[2])


def icb_free_v_and_w(fvp, fvt, u, w, ki):
    """cal_sph_nod_icb_free_v_and_w (set_sph_exp_free_ICB.f90:60-90)."""
    d1s = fvp[2][1] * u[0, ki + 1]
    d2s = fvp[2][2] * u[0, ki + 1]
    d1t = fvt[1][1] * u[2, ki]
    u[0, ki] = 0.0
    u[1, ki] = d1s
    w[0, ki] = u[2, ki]
    w[1, ki] = d1t
    w[2, ki] = -d2s


def cmb_free_v_and_w(fvp, fvt, u, w, ko):
    """cal_sph_nod_cmb_free_v_and_w (set_sph_exp_free_CMB.f90:58-90)."""
    d1s = fvp[0][1] * u[0, ko - 1]
    d2s = fvp[0][2] * u[0, ko - 1]
    d1t = fvt[0][1] * u[2, ko - 1] + fvt[1][1] * u[2, ko]
    u[0, ko] = 0.0
    u[1, ko] = d1s
    w[0, ko] = u[2, ko]
    w[1, ko] = d1t
    w[2, ko] = -d2s


def icb_free_vpol2(fvp, u, ki):
    """cal_sph_nod_icb_free_vpol2 (set_sph_exp_free_ICB.f90:94-119)."""
    d1s = fvp[1][1] * u[0, ki] + fvp[2][1] * u[0, ki + 1]
    u[0, ki] = 0.0
    u[1, ki] = d1s


def cmb_free_vpol2(fvp, u, ko):
    """cal_sph_nod_cmb_free_vpol2 (set_sph_exp_free_CMB.f90:94-118)."""
    d1s = fvp[0][1] * u[0, ko - 1] + fvp[1][1] * u[0, ko]
    u[0, ko] = 0.0
    u[1, ko] = d1s


def icb_free_rot2(fvp, fvt, r_icb, g3, src, rot, ki):
    """cal_sph_nod_icb_free_rot2 (set_sph_exp_free_ICB.f90:123-154)."""
    d2s = fvp[1][2] * src[0, ki] + fvp[2][2] * src[0, ki + 1]
    d1t = fvt[1][1] * src[2, ki]
    rot[0, ki] = src[2, ki]
    rot[1, ki] = d1t
    rot[2, ki] = -d2s + g3 * r_icb[2] * src[0, ki]


def cmb_free_rot2(fvp, fvt, r_cmb, g3, src, rot, ko):
    """cal_sph_nod_cmb_free_rot2 (set_sph_exp_free_CMB.f90:122-154)."""
    d2s = fvp[0][2] * src[0, ko - 1] + fvp[1][2] * src[0, ko]
    d1t = fvt[0][1] * src[2, ko - 1] + fvt[1][1] * src[2, ko]
    rot[0, ko] = src[2, ko]
    rot[1, ko] = d1t
    rot[2, ko] = -d2s + g3 * r_cmb[2] * src[0, ko]


def icb_free_w_diffuse2(ffl, fvt, r_icb, g3, coef, w, out, ki):
    """cal_sph_nod_icb_free_w_diffuse2 (set_sph_exp_free_ICB.f90:194-228):
    component 0 (the vorticity's poloidal part, the toroidal velocity's
    image) through the free-slip formula, component 2 through the three-point
    fixed-value one."""
    d2s = fvt[1][2] * w[0, ki] + fvt[2][2] * w[0, ki + 1]
    d2t = ffl[0][2] * w[2, ki] + ffl[1][2] * w[2, ki + 1] + ffl[2][2] * w[2, ki + 2]
    out[0, ki] = coef * (d2s - g3 * r_icb[2] * w[0, ki])
    out[2, ki] = coef * (d2t - g3 * r_icb[2] * w[2, ki])


def cmb_free_w_diffuse2(ffl, fvt, r_cmb, g3, coef, w, out, ko):
    """cal_sph_nod_cmb_free_w_diffuse2 (set_sph_exp_free_CMB.f90:196-230)."""
    d2s = fvt[0][2] * w[0, ko - 1] + fvt[1][2] * w[0, ko]
    d2t = ffl[2][2] * w[2, ko - 2] + ffl[1][2] * w[2, ko - 1] + ffl[0][2] * w[2, ko]
    out[0, ko] = coef * (d2s - g3 * r_cmb[2] * w[0, ko])
    out[2, ko] = coef * (d2t - g3 * r_cmb[2] * w[2, ko])


# ------------------------------------------------------------ Gaunt ints --

def _fact(n):
    return math.factorial(n)


def gaunt_gs(m1, l1, m2, l2, m3, l3):
    """Adams-Gaunt integral of three Schmidt harmonics, exact rational form of
    gs_mul_fact / gs_prime (SERIAL_src/spherical_harmonics/normgi_by_fact.f90,
    normgi_by_primes.f90): the same closed form, products of factorials."""
    s = (l1 + l2 + l3) // 2
    itmax = min(l1 - m1, l3 - m3, l2 + l3 - m1)
    itmin = -min(l2 - l3 + m1, l1 + m1, 0)
    pre = Fraction(_fact(2 * s - 2 * l3) * _fact(s),
                   _fact(s - l1) * _fact(s - l2) * _fact(s - l3) * _fact(2 * s + 1))
    sq = Fraction(_fact(l1 - m1) * _fact(l2 + m2) * _fact(l3 + m3) * _fact(l3 - m3),
                  _fact(l1 + m1) * _fact(l2 - m2))
    tsum = Fraction(0)
    for it in range(itmin, itmax + 1):
        tsum += Fraction((-1) ** it * _fact(l1 + m1 + it) * _fact(l2 + l3 - m1 - it),
                         _fact(l1 - m1 - it) * _fact(l2 - l3 + m1 + it)
                         * _fact(l3 - m3 - it) * _fact(it))
    # the square of the value is a moderate rational even where its factors
    # overflow a double (high degrees): one rounding, then the square root
    x = (-1) ** (s - l2 - m3) * 2 * pre * tsum
    val = math.copysign(math.sqrt(float(x * x * sq)), float(x > 0) - 0.5) if x != 0 else 0.0
    for m in (m1, m2, m3):
        if m != 0:
            val *= math.sqrt(2.0)
    return val


def _normki(gi, m1, m2, m3):
    if m1 * m2 * m3 > 0:
        return -gi * 0.5 if (m2 < 0 and m3 < 0) else gi * 0.5
    if m2 == 0 and m1 * m3 > 0:
        return gi
    if m3 == 0 and m2 * m1 > 0:
        return gi
    if m3 == 0 and m2 == 0 and m1 == 0:
        return 2.0 * gi
    return 0.0


def leadki(l1, m1, l2, m2, l3, m3):
    """leadki (SERIAL_src/spherical_harmonics/cal_gaunt_itgs.f90)."""
    if not (l1 + l2 >= l3 and l2 + l3 >= l1 and l3 + l1 >= l2):
        return 0.0
    if (l1 + l2 + l3) % 2 != 0:
        return 0.0
    if not (m1 + m2 + m3 == 0 or m1 + m2 - m3 == 0 or m1 - m2 + m3 == 0 or -m1 + m2 + m3 == 0):
        return 0.0
    a1, a2, a3 = abs(m1), abs(m2), abs(m3)
    if a1 == a2 + a3:
        return _normki(gaunt_gs(a1, l1, a2, l2, a3, l3), m1, m2, m3)
    if a2 == a1 + a3:
        return _normki(gaunt_gs(a2, l2, a1, l1, a3, l3), m2, m1, m3)
    if a3 == a2 + a1:
        return _normki(gaunt_gs(a3, l3, a2, l2, a1, l1), m3, m2, m1)
    return 0.0


def _leades(m2, l2, m3, l3, gg):
    a1 = float((l2 + m2) * (l3 - m3))
    a2 = float((l3 + m3) * (l2 - m2))
    a3 = float((l2 + m2 - 1) * (l3 + m3 + 1))
    a4 = float((l3 + m3 - 1) * (l2 + m2 + 1))
    a5 = float((l2 - m2) * (l3 - m3))
    if m2 - 1 == 0:
        a3 *= 2.0
    if m3 == 0:
        a3 *= 0.5
        a4 *= 0.5
    if m2 == 0:
        a3 *= 0.5
        a4 *= 0.5
    if m3 - 1 == 0:
        a4 *= 2.0
    return -(math.sqrt(a1) * (math.sqrt(a3) * gg[0] + math.sqrt(a5) * gg[1])
             - math.sqrt(a2) * (math.sqrt(a4) * gg[2] + math.sqrt(a5) * gg[3])) * 0.5


def _setei(l1, m1, l2, m2, l3, m3):
    """setei (SERIAL_src/spherical_harmonics/normei.f90)."""
    if l2 == l3 and m2 == m3:
        return 0.0
    gg = [0.0] * 4
    if m3 + 1 > l3 or abs(m2 - 1) > l2 - 1:
        gg[0] = 0.0
    elif m2 == 0:
        gg[0] = -gaunt_gs(m3 + 1, l3, 1, l2 - 1, m1, l1)
    else:
        gg[0] = gaunt_gs(m1, l1, m2 - 1, l2 - 1, m3 + 1, l3)
    gg[1] = 0.0 if m2 > l2 - 1 else gaunt_gs(m1, l1, m2, l2 - 1, m3, l3)
    if m2 + 1 > l2 or abs(m3 - 1) > l3 - 1:
        gg[2] = 0.0
    elif m3 == 0:
        gg[2] = -gaunt_gs(m2 + 1, l2, 1, l3 - 1, m1, l1)
    else:
        gg[2] = gaunt_gs(m1, l1, m2 + 1, l2, m3 - 1, l3 - 1)
    gg[3] = 0.0 if m3 > l3 - 1 else gaunt_gs(m1, l1, m2, l2, m3, l3 - 1)
    return _leades(m2, l2, m3, l3, gg)


def _normli(ei, m1, m2, m3, mm1):
    if m1 * m2 * m3 < 0:
        return ei * 0.5 if (m3 > 0 and m2 > 0) else -ei * 0.5
    if m2 == 0 and m1 * m3 < 0:
        return -float(int(m1 / mm1)) * ei
    if m3 == 0 and m2 * m1 < 0:
        return -float(int(m1 / mm1)) * ei
    return 0.0


def leadli(l1, m1, l2, m2, l3, m3):
    """leadli, the Elsasser integral (cal_gaunt_itgs.f90)."""
    if not (l1 + l2 >= l3 and l2 + l3 >= l1 and l3 + l1 >= l2):
        return 0.0
    if (l1 + l2 + l3) % 2 != 1:
        return 0.0
    if not (m1 + m2 + m3 == 0 or m1 + m2 - m3 == 0 or m1 - m2 + m3 == 0 or -m1 + m2 + m3 == 0):
        return 0.0
    a1, a2, a3 = abs(m1), abs(m2), abs(m3)
    if a1 == a2 + a3:
        return _normli(_setei(l1, a1, l2, a2, l3, a3), m1, m2, m3, a1)
    if a2 == a1 + a3:
        return -_normli(_setei(l2, a2, l1, a1, l3, a3), m2, m1, m3, a2)
    if a3 == a2 + a1:
        return -_normli(_setei(l3, a3, l2, a2, l1, a1), m3, m2, m1, a3)
    return 0.0


class CoriolisRlm:
    """Coupling coefficients of the curl of the Coriolis force:
    cal_gaunt_coriolis_rlm / interact_rot_coriolis_rlm
    (MHD_src/sph_MHD/interact_coriolis_rlm.f90), indexed by rj mode
    (the coupling only pairs (l+-1, m) and (l, -m) of one rj mode)."""

    def __init__(self, L):
        J = (L + 1) ** 2
        self.jk1 = np.zeros(J, dtype=np.int64)   # (l-1, m)
        self.jk2 = np.zeros(J, dtype=np.int64)   # (l+1, m)
        self.jl1 = np.zeros(J, dtype=np.int64)   # (l, -m)
        gi = np.zeros((J, 2))
        ei = np.zeros(J)
        for j in range(J):
            l3 = int(math.isqrt(j))
            m3 = j - l3 * (l3 + 1)
            l2 = l3 - 1
            if 1 <= l2 <= L and abs(m3) <= l2:
                gi[j, 0] = leadki(1, 0, l2, m3, l3, m3)
                k1 = (l2, m3)
            else:
                k1 = (l3, m3)
            l2 = l3 + 1
            if 1 <= l2 <= L and abs(m3) <= l2:
                gi[j, 1] = leadki(1, 0, l2, m3, l3, m3)
                k2 = (l2, m3)
            else:
                k2 = (l3, m3)
            e1 = (l3, m3)
            if l3 >= 1:
                ei[j] = leadli(1, 0, l3, -m3, l3, m3)
                e1 = (l3, -m3)
            if j == 0:
                k1 = k2 = e1 = (0, 0)
                gi[j, 1] = 0.0
                ei[j] = 0.0
            self.jk1[j] = k1[0] * (k1[0] + 1) + k1[1]
            self.jk2[j] = k2[0] * (k2[0] + 1) + k2[1]
            self.jl1[j] = e1[0] * (e1[0] + 1) + e1[1]
        g = np.array([sph_normalizations(int(math.isqrt(j)), j - int(math.isqrt(j)) *
                                         (int(math.isqrt(j)) + 1)) for j in range(J)])
        g[0, 2] = 0.5                                   # g_sph_rlm(3) at l = m = 0
        g3, g4, g5, g17 = g[:, 2], g[:, 3], g[:, 4], g[:, 16]
        self.sw11 = (2.0 - g4 - g3) * gi[:, 0] * g17
        self.sw21 = (2.0 - g5 - g3) * gi[:, 1] * g17
        self.sw12 = g4 * (2.0 - g4 + g3) * gi[:, 0] * g17 * 0.5
        self.sw22 = g5 * (2.0 - g5 + g3) * gi[:, 1] * g17 * 0.5
        self.sw13 = 2.0 * ei * g17
        self.tw13 = -(g3 * (2.0 + g4 - g3) + 2.0 * (-2.0 + g4 + g3)) * gi[:, 0] * g17 * 0.5
        self.tw23 = -(g3 * (2.0 + g5 - g3) + 2.0 * (-2.0 + g5 + g3)) * gi[:, 1] * g17 * 0.5
        self.tw14 = -(-2.0 + g4 + g3) * gi[:, 0] * g17
        self.tw24 = -(-2.0 + g5 + g3) * gi[:, 1] * g17
        self.tw11 = -2.0 * ei * g17
        self.tw12 = -g3 * ei * g17
        self.g3rlm = g3
        self.gi, self.ei = gi, ei


# ------------------------------------------------------------- the model --

class DynamoBenchmarkCase1:
    """Parameters of the reference's tests/Dynamobench_case1 (control_MHD,
    control_sph_shell): L = 31, 48 x 96 grid, 48 Chebyshev intervals,
    r_i/r_o = 0.35 with r_o - r_i = 1, E = 1e-3, modified Ra = 100, Pr = 1,
    Pm = 5, Crank-Nicolson weight 0.6, dt = 5e-5."""
    L, nth, nphi, nri = 31, 48, 96, 48
    ratio, shell = 0.35, 1.0
    dt = 5.0e-5
    coef_imp = 0.6
    Pr, Ra_mod, E, Pm = 1.0, 1.0e2, 1.0e-3, 5.0
    depth_high_t, depth_low_t = 0.5384615384615384, 1.5384615384615384
    high_temp, low_temp = 1.0, 0.0
    isig = 400
    # velocity walls (bc_velocity): False non_slip_sph, True free_slip_sph
    free_icb, free_cmb = False, False


class DynamoBenchmarkCase2(DynamoBenchmarkCase1):
    """Parameters of the reference's tests/Dynamobench_case2 (control_MHD,
    control_sph_shell): case 1's shell and numbers, Ra_mod = 110, the radial
    grid extended to Min_radius 0 and Max_radius 2 (24 inner-core points, 49
    in the shell, 23 above the CMB: 96 in all), the magnetic field solved to
    the centre (bc_magnetic_field sph_to_center), a rotatable inner core
    (bc_velocity rot_inner_core) and the dynamo_benchmark_2 initial field."""
    Ra_mod = 1.1e2
    inner_core = True
    r_min, r_max = 0.0, 2.0


def shell_case(L=31, nth=48, nphi=96, nri=48, dt=5.0e-5, coef_imp=0.6, Ra_mod=1.0e2,
               free_icb=False, free_cmb=False):
    """The Dynamobench case 1 physics (rotating Boussinesq shell, rigid walls,
    insulating magnetic boundaries, fixed temperature) on another grid, time
    step, Crank-Nicolson weight or modified Rayleigh number, or with free-slip
    velocity walls."""
    return type("ShellCase", (DynamoBenchmarkCase1,),
                dict(L=L, nth=nth, nphi=nphi, nri=nri, dt=dt, coef_imp=coef_imp, Ra_mod=Ra_mod,
                     free_icb=free_icb, free_cmb=free_cmb))


class MhdStep:
    """State d_rj and the reference's step sequence (see module docstring)."""

    def __init__(self, case=DynamoBenchmarkCase1):
        c = case
        self.c = c
        L = self.L = c.L
        self.J = J = (L + 1) ** 2
        # set_ctl_gen_shell_grids.f90:179-185: r_ICB from the shell size and
        # radius ratio, r_CMB = r_ICB + size
        r_icb = c.shell * c.ratio / (1.0 - c.ratio)
        r_cmb = r_icb + c.shell
        # inner_core: rotatable inner core with the magnetic field solved to
        # the centre and a potential field above the CMB (Dynamobench_case2)
        self.inner_core = bool(getattr(c, "inner_core", False))
        if self.inner_core:
            self.radius, self.kr_in, self.kr_out = chebyshev_radius_ext(
                c.nri, r_icb, r_cmb, c.r_min, c.r_max)
        else:
            self.radius = chebyshev_radius(c.nri, r_icb, r_cmb)
            self.kr_in, self.kr_out = 1, self.radius.size
        # set_radius_dat_4_sph_dynamo: r_ICB, r_CMB are the grid points of
        # the ICB and CMB groups
        self.r_icb = float(self.radius[self.kr_in - 1])
        self.r_cmb = float(self.radius[self.kr_out - 1])
        nr = self.nr = self.radius.size
        # first point of the magnetic field's range (1 = the point above the
        # centre when the field is solved to the centre)
        self.kb_in = 1 if self.inner_core else self.kr_in
        self.ar1 = 1.0 / self.radius
        self.ar2 = self.ar1 ** 2
        self.d1, self.d2 = fdm_coefs(self.radius, self.kr_in > 1)
        self.bc = BoundaryFdm(self.radius, self.kr_in, self.kr_out)
        if self.inner_core:
            self.ctr1 = center_fdm(float(self.radius[0]), float(self.radius[1]))
            self.r_ctr1 = (float(self.radius[0]), 1.0 / float(self.radius[0]),
                           (1.0 / float(self.radius[0])) ** 2)
            # cal_fdm1_coef_fix_fld_ICB (cal_fdm_coefs_4_boundaries.f90:135-151)
            dr1 = float(self.radius[self.kr_in]) - self.r_icb
            self.fdm1_icb = (-1.0 / dr1, 1.0 / dr1)
        self.free_icb, self.free_cmb = bool(c.free_icb), bool(c.free_cmb)
        # rj mode degrees and normalisations (copy_sph_normalization_2_rj: no
        # l = 0 override in rj)
        self.l_of_j = np.array([int(math.isqrt(j)) for j in range(J)])
        self.m_of_j = np.array([j - l * (l + 1) for j, l in enumerate(self.l_of_j)])
        self.g1 = self.l_of_j.astype(np.float64)
        self.g3 = (self.l_of_j * (self.l_of_j + 1)).astype(np.float64)
        # coefficients (material_property.f90, construct_MHD_coefficient.f90)
        self.coef_temp = 1.0
        self.coef_d_temp = 1.0 * c.Pr ** -1.0
        self.coef_velo = 1.0
        self.coef_d_velo = 1.0
        self.coef_buo = 1.0 * (1.0 - 0.35) * c.Ra_mod ** 1.0 * c.E ** -1.0
        self.coef_cor = 1.0 * 2.0 ** 1.0 * c.E ** -1.0
        self.coef_lor = 1.0 * c.Pm ** -1.0 * c.E ** -1.0
        self.coef_magne = 1.0
        self.coef_d_magne = 1.0 * c.Pm ** -1.0
        self.coef_induct = 1.0
        self.coef_press = 1.0 * c.E ** -1.0
        self.coef_imp_v = self.coef_imp_t = self.coef_imp_b = c.coef_imp
        self.coef_exp_v = 1.0 - c.coef_imp
        self.coef_exp_t = 1.0 - c.coef_imp
        self.coef_exp_b = 1.0 - c.coef_imp
        self.dt = c.dt
        # reference temperature (set_ref_temp_sph_mhd, set_reference_sph_mhd.f90)
        dh, dl, hi, lo = c.depth_high_t, c.depth_low_t, c.high_temp, c.low_temp
        self.reft0 = (dh * dl * self.ar1 * (hi - lo) - dh * hi + dl * lo) / (dl - dh)
        self.reft1 = -dh * dl * self.ar2 * (hi - lo) / (dl - dh)
        # below the ICB the high temperature, above the CMB the low one
        self.reft0[:self.kr_in - 1] = hi
        self.reft1[:self.kr_in - 1] = 0.0
        self.reft0[self.kr_out:] = lo
        self.reft1[self.kr_out:] = 0.0
        # temperature boundary values after adjust_sph_temp_bc_by_reftemp
        self.T_icb = np.zeros(J)
        self.T_cmb = np.zeros(J)
        self.T_icb[0] = hi - self.reft0[self.kr_in - 1]
        self.T_cmb[0] = lo - self.reft0[self.kr_out - 1]
        # transform geometry and Coriolis couplings
        self.grid = Grid(L, c.nth, self.radius)
        jlm, _ = rlm_index(L)
        self.rj_of_rlm = jlm[:, 0].astype(np.int64)      # rlm j -> rj j
        self.cor = CoriolisRlm(L)
        self.f = {}
        self._matrices()

    # ---------------------------------------------------------- fields --
    def zeros(self, n=1):
        return np.zeros((n, self.nr, self.J)) if n > 1 else np.zeros((self.nr, self.J))

    # --------------------------------------------------------- matrices --
    def _matrices(self):
        """const_radial_mat_vort_2step, const_radial_mat_4_temp_sph,
        const_radial_mat_4_magne_sph (const_r_mat_4_vector_sph.f90,
        const_r_mat_4_scalar_sph.f90) for rigid walls, insulating and
        fixed-temperature boundaries; band storage mat3[j, k, 0..2] =
        Fortran mat3(1..3, k, j)."""
        nr, J = self.nr, self.J
        ki, ko = self.kr_in - 1, self.kr_out - 1
        bc = self.bc
        d1, d2, ar1, ar2, g3 = self.d1, self.d2, self.ar1, self.ar2, self.g3

        def unit_evo():
            m = np.zeros((J, nr, 3))
            m[:, :, 1] = 1.0
            return m

        def unit_poisson():
            m = np.zeros((J, nr, 3))
            m[:, :ki, 1] = 1.0
            m[:, ko + 1:, 1] = 1.0
            return m

        def add_vector_poisson(m, coef, k_lo=None):
            for k in range((ki if k_lo is None else k_lo) + 1, ko):
                m[:, k - 1, 2] = m[:, k - 1, 2] - coef * d2[k, 0]
                m[:, k, 1] = m[:, k, 1] - coef * (d2[k, 1] - g3 * ar2[k])
                m[:, k + 1, 0] = m[:, k + 1, 0] - coef * d2[k, 2]

        def add_scalar_poisson(m, coef):
            for k in range(ki + 1, ko):
                m[:, k - 1, 2] = m[:, k - 1, 2] - coef * (d2[k, 0] + 2.0 * ar1[k] * d1[k, 0])
                m[:, k, 1] = m[:, k, 1] - coef * (d2[k, 1] + 2.0 * ar1[k] * d1[k, 1]
                                                  - g3 * ar2[k])
                m[:, k + 1, 0] = m[:, k + 1, 0] - coef * (d2[k, 2] + 2.0 * ar1[k] * d1[k, 2])

        def add_fix_flux_icb(m, coef, fdm=None):
            fdm = bc.fix_dr_icb if fdm is None else fdm
            add_fix_flux_icb_rows(m, ki, bc.r_icb, fdm, g3, coef)

        def add_fix_flux_cmb(m, coef, fdm=None):
            fdm = bc.fix_dr_cmb if fdm is None else fdm
            add_fix_flux_cmb_rows(m, ko, bc.r_cmb, fdm, g3, coef)

        def set_fix_fld_icb(m):
            m[:, ki, 1] = 1.0
            m[:, ki + 1, 0] = 0.0

        def set_fix_fld_cmb(m):
            m[:, ko - 1, 2] = 0.0
            m[:, ko, 1] = 1.0

        # velocity (vorticity two-step)
        coef_dvt = self.coef_imp_v * self.coef_d_velo * self.dt
        vt, wt = unit_evo(), unit_evo()
        vs = unit_poisson()
        add_vector_poisson(vt, coef_dvt)
        add_vector_poisson(wt, coef_dvt)
        add_vector_poisson(vs, 1.0)
        # walls (const_radial_mat_vort_2step, const_r_mat_4_vector_sph.f90:
        # 85-135): free slip puts the free-slip formulas in the toroidal and
        # poloidal rows, rigid walls fix the toroidal value
        add_fix_flux_icb(wt, coef_dvt)
        if self.free_icb:
            add_fix_flux_icb(vt, coef_dvt, bc.free_vt_icb)
            add_fix_flux_icb(vs, 1.0, bc.free_vp_icb)
        else:
            set_fix_fld_icb(vt)
            add_fix_flux_icb(vs, 1.0)
        add_fix_flux_cmb(wt, coef_dvt)
        if self.free_cmb:
            add_fix_flux_cmb(vt, coef_dvt, bc.free_vt_cmb)
            add_fix_flux_cmb(vs, 1.0, bc.free_vp_cmb)
        else:
            set_fix_fld_cmb(vt)
            add_fix_flux_cmb(vs, 1.0)
        if self.inner_core:
            # set_icore_viscous_matrix / set_rotate_icb_vt_sph_mat
            # (cal_inner_core_rotation.f90:55-71, :185-208): the l = 1
            # toroidal rows at the ICB carry the viscous torque on the core
            ar = 1.0 / self.r_icb
            for j in (1, 2, 3):
                vt[j, ki, 1] = (1.0 - self.coef_imp_v * self.dt * self.coef_d_velo * 5.0
                                * (self.fdm1_icb[0] - 2.0 * self.ar1[ki]) * ar)
                vt[j, ki + 1, 0] = -(self.coef_imp_v * self.dt * self.coef_d_velo * 5.0
                                     * ar * self.fdm1_icb[1])
        vp = self._mat_product_3band(wt, vs)
        # temperature
        coef = self.coef_imp_t * self.coef_d_temp * self.dt
        tm = unit_evo()
        add_scalar_poisson(tm, coef)
        set_fix_fld_icb(tm)
        set_fix_fld_cmb(tm)
        # magnetic field
        coef_dbt = self.coef_imp_b * self.coef_d_magne * self.dt
        bs, bt = unit_evo(), unit_evo()
        kb = self.kb_in - 1
        add_vector_poisson(bs, coef_dbt, kb)
        add_vector_poisson(bt, coef_dbt, kb)
        g1 = self.g1
        if self.inner_core:
            # add_vector_poisson_mat_center (center_sph_matrices.f90:160-180)
            ctr, rc = self.ctr1, self.r_ctr1
            for m in (bs, bt):
                m[:, 0, 1] = m[:, 0, 1] - coef_dbt * (ctr[1][2] - g3 * rc[2])
                m[:, 1, 0] = m[:, 1, 0] - coef_dbt * ctr[2][2]
        else:
            bs[:, ki, 1] = bs[:, ki, 1] - coef_dbt * (bc.fix_dr_icb[1][2] - g3 * bc.r_icb[2]
                                                      + (g1 + 1.0) * bc.r_icb[1]
                                                      * bc.fix_dr_icb[0][2])
            bs[:, ki + 1, 0] = bs[:, ki + 1, 0] - coef_dbt * bc.fix_dr_icb[2][2]
            set_fix_fld_icb(bt)
        bs[:, ko - 1, 2] = bs[:, ko - 1, 2] - coef_dbt * bc.fix_dr_cmb[0][2]
        bs[:, ko, 1] = bs[:, ko, 1] - coef_dbt * (bc.fix_dr_cmb[1][2] - g3 * bc.r_cmb[2]
                                                  - g1 * bc.r_cmb[1] * bc.fix_dr_cmb[2][2])
        set_fix_fld_cmb(bt)
        self.mats = {"vp": vp, "vt": vt, "temp": tm, "bs": bs, "bt": bt}
        self.lu = {}
        for name, m in self.mats.items():
            nb = 5 if name == "vp" else 3
            self.lu[name] = ludcmp(nb, m)[:2]

    def _mat_product_3band(self, a_l, a_r):
        """cal_mat_product_3band_mul (SOLVER_src/.../mat_product_3band_mul.f90)
        into 5-band storage p[j, k, 0..4] = Fortran a_prod(1..5, k, j)."""
        nr, J = self.nr, self.J
        st, ed = self.kr_in - 1, self.kr_out - 1
        L_, R_ = a_l, a_r
        p = np.zeros((J, nr, 5))
        p[:, :, 2] = 1.0          # identity outside kr_in..kr_out (m_radial_matrices_sph)

        def rows(k, skip5=False, skip1=False):
            if not skip5:
                p[:, k - 2, 4] = L_[:, k - 1, 2] * R_[:, k - 2, 2]
            p[:, k - 1, 3] = L_[:, k - 1, 2] * R_[:, k - 1, 1] + L_[:, k, 1] * R_[:, k - 1, 2]
            p[:, k, 2] = (L_[:, k - 1, 2] * R_[:, k, 0] + L_[:, k, 1] * R_[:, k, 1]
                          + L_[:, k + 1, 0] * R_[:, k, 2])
            p[:, k + 1, 1] = L_[:, k, 1] * R_[:, k + 1, 0] + L_[:, k + 1, 0] * R_[:, k + 1, 1]
            if not skip1:
                p[:, k + 2, 0] = L_[:, k + 1, 0] * R_[:, k + 2, 0]

        for k in range(st + 2, ed - 1):
            rows(k)
        p[:, st, 2] = 1.0
        p[:, st + 1, 1] = 0.0
        p[:, st + 2, 0] = 0.0
        rows(st + 1, skip5=True)
        rows(ed - 1, skip1=True)
        p[:, ed - 2, 4] = 0.0
        p[:, ed - 1, 3] = 0.0
        p[:, ed, 2] = 1.0
        return p

    def _solve(self, name, x):
        nb = 5 if name == "vp" else 3
        lu, piv = self.lu[name]
        return lubksb(nb, lu, piv, x)

    # ------------------------------------------------- radial operators --
    def _interior(self):
        return range(self.kr_in, self.kr_out - 1)   # 0-based k = kr_in+1..kr_out-1

    def _interior_b(self):
        return range(self.kb_in, self.kr_out - 1)   # the magnetic field's range

    def _center_rot2(self, src, rot):
        """cal_sph_nod_center_rot2 (set_sph_exp_nod_center.f90:96-124): the
        first point with the centre value taken as zero (only the offsets 0,
        +1 of the three-point formulas)."""
        d1, d2, g3 = self.d1, self.d2, self.g3
        d2s = d2[0, 1] * src[0, 0] + d2[0, 2] * src[0, 1]
        d1t = d1[0, 1] * src[2, 0] + d1[0, 2] * src[2, 1]
        rot[0, 0] = src[2, 0]
        rot[1, 0] = d1t
        rot[2, 0] = -(d2s - g3 * self.ar2[0] * src[0, 0])

    def set_inner_core_rotation(self):
        """set_inner_core_rotation / set_inner_core_rot_l1
        (cal_inner_core_rotation.f90:42-52, :212-245): the l = 1 toroidal
        velocity and vorticity of the solid-body turning core below the ICB."""
        if not self.inner_core:
            return
        u, w = self.f["velo"], self.f["vort"]
        ki = self.kr_in - 1
        r = self.radius
        ar1, ar2 = self.ar1[ki], self.ar2[ki]
        for j in (1, 2, 3):
            for k in range(ki):
                ratio = r[k] * r[k] * ar2
                u[2, k, j] = ratio * u[2, ki, j]
                w[0, k, j] = ratio * w[0, ki, j]
                w[1, k, j] = 2.0 * w[0, ki, j] * r[k] * ar2
            w[1, ki, j] = 2.0 * w[0, ki, j] * ar1

    def _rotate_velo2(self, u):
        """cal_sph_nod_icb_rotate_velo2 (set_sph_exp_rigid_ICB.f90:84-146)
        with no imposed ICB rotation: zero velocity up to the ICB except the
        l = 1 toroidal modes, which keep their ICB value and turn the core
        below it as a solid body."""
        ki = self.kr_in - 1
        u[0, :ki + 1] = 0.0
        u[1, :ki + 1] = 0.0
        keep = u[2, ki, 1:4].copy()
        u[2, :ki + 1] = 0.0
        u[2, ki, 1:4] = keep
        ar2 = self.bc.r_icb[2]
        for k in range(ki):
            u[2, k, 1:4] = keep * ar2 * self.radius[k] * self.radius[k]

    def _dr(self, f, k, which=1):
        d = self.d1 if which == 1 else self.d2
        return d[k, 0] * f[k - 1] + d[k, 1] * f[k] + d[k, 2] * f[k + 1]

    def grad_vp_and_vorticity(self):
        """const_grad_vp_and_vorticity: at rigid walls cal_sph_nod_icb_rigid_velo2,
        cal_sph_nod_icb_rigid_rot2, cal_sph_nod_cmb_rigid_v_and_w
        (set_sph_exp_rigid_ICB.f90 / _CMB.f90), at free-slip walls
        cal_sph_nod_*_free_v_and_w, inside cal_sph_diff_pol_and_rot2
        (cal_sph_exp_rotation.f90)."""
        u, w = self.f["velo"], self.f["vort"]
        bc, g3 = self.bc, self.g3
        ki, ko = self.kr_in - 1, self.kr_out - 1
        if self.free_icb:
            icb_free_v_and_w(bc.free_vp_icb, bc.free_vt_icb, u, w, ki)
        else:
            if self.inner_core:
                self._rotate_velo2(u)
            else:
                u[0, :ki + 1] = 0.0
                u[1, :ki + 1] = 0.0
                u[2, :ki + 1] = 0.0          # vt_ICB_bc = 0
            fdr, ffl = bc.fix_dr_icb, bc.fix_fld_icb
            d2s = fdr[1][2] * u[0, ki] + fdr[2][2] * u[0, ki + 1]
            d1t = ffl[0][1] * u[2, ki] + ffl[1][1] * u[2, ki + 1] + ffl[2][1] * u[2, ki + 2]
            w[0, ki] = u[2, ki]
            w[1, ki] = d1t
            w[2, ki] = -(d2s - g3 * bc.r_icb[2] * u[0, ki])
        if self.free_cmb:
            cmb_free_v_and_w(bc.free_vp_cmb, bc.free_vt_cmb, u, w, ko)
        else:
            u[0, ko] = 0.0
            u[1, ko] = 0.0
            u[2, ko] = 0.0
            fdr, ffl = bc.fix_dr_cmb, bc.fix_fld_cmb
            d2s = fdr[0][2] * u[0, ko - 1]
            d1t = ffl[2][1] * u[2, ko - 2] + ffl[1][1] * u[2, ko - 1] + ffl[0][1] * u[2, ko]
            w[0, ko] = u[2, ko]
            w[1, ko] = d1t
            w[2, ko] = -d2s
        self._diff_pol_and_rot(u, w)

    def _diff_pol_and_rot(self, f, rot, interior=None):
        g3 = self.g3
        for k in (self._interior() if interior is None else interior):
            d1s = self._dr(f[0], k, 1)
            d2s = self._dr(f[0], k, 2)
            d1t = self._dr(f[2], k, 1)
            f[1, k] = d1s
            rot[0, k] = f[2, k]
            rot[1, k] = d1t
            rot[2, k] = -d2s + g3 * self.ar2[k] * f[0, k]

    def grad_bp_and_current(self):
        """const_grad_bp_and_current with insulating boundaries:
        cal_sph_nod_icb_ins_b_and_j / cal_sph_nod_cmb_ins_b_and_j
        (cal_sph_exp_nod_icb_ins.f90, _cmb_ins.f90) then
        cal_sph_diff_pol_and_rot2 (no points outside the shell to extend)."""
        b, jc = self.f["magne"], self.f["current"]
        bc, g1, g3 = self.bc, self.g1, self.g3
        ki, ko = self.kr_in - 1, self.kr_out - 1
        if self.inner_core:
            # sph_center_fld_and_curl (set_sph_exp_nod_center.f90:42-71)
            d1 = self.d1
            b[1, 0] = d1[0, 1] * b[0, 0] + d1[0, 2] * b[0, 1]
            self._center_rot2(b, jc)
        else:
            fdr, ffl = bc.fix_dr_icb, bc.fix_fld_icb
            d1s = (g1 + 1.0) * bc.r_icb[1] * b[0, ki]
            d2s = (((g1 + 1.0) * bc.r_icb[1] * fdr[0][2] + fdr[1][2]) * b[0, ki]
                   + fdr[2][2] * b[0, ki + 1])
            d1t = ffl[1][1] * b[2, ki + 1] + ffl[2][1] * b[2, ki + 2]
            b[1, ki] = d1s
            b[2, ki] = 0.0
            jc[0, ki] = 0.0
            jc[1, ki] = d1t
            jc[2, ki] = -(d2s - g3 * bc.r_icb[2] * b[0, ki])
        fdr, ffl = bc.fix_dr_cmb, bc.fix_fld_cmb
        d1s = -g1 * bc.r_cmb[1] * b[0, ko]
        d2s = (fdr[0][2] * b[0, ko - 1]
               + (fdr[1][2] - g1 * bc.r_cmb[1] * fdr[2][2]) * b[0, ko])
        d1t = ffl[2][1] * b[2, ko - 2] + ffl[1][1] * b[2, ko - 1]
        b[1, ko] = d1s
        b[2, ko] = 0.0
        jc[0, ko] = 0.0
        jc[1, ko] = d1t
        jc[2, ko] = -(d2s - g3 * bc.r_cmb[2] * b[0, ko])
        self._diff_pol_and_rot(b, jc, self._interior_b())
        # ext_outside_potential_with_j (extend_potential_field.f90:169-201)
        for k in range(ko + 1, self.nr):
            ratio = self.r_cmb * self.ar1[k]
            pw = np.array([ratio ** int(l) for l in self.l_of_j])
            b[0, k] = b[0, ko] * pw
            b[1, k] = -b[0, ko] * pw * g1 * self.ar1[k]
            b[2, k] = 0.0
            jc[:, k] = 0.0

    def vorticity_diffusion(self):
        """const_sph_vorticirty_diffusion (const_sph_diffusion.f90) with
        cal_sph_nod_icb_rgd_w_diffuse2 / cal_sph_nod_cmb_rgd_w_diffuse2; the
        d/dr component is not used by the step and is not formed."""
        w, out = self.f["vort"], self.f["w_diffuse"]
        coef, g3, bc = self.coef_d_velo, self.g3, self.bc
        for k in self._interior():
            d2s = self._dr(w[0], k, 2)
            d2t = self._dr(w[2], k, 2)
            out[0, k] = coef * (d2s - g3 * self.ar2[k] * w[0, k])
            out[2, k] = coef * (d2t - g3 * self.ar2[k] * w[2, k])
        ki, ko = self.kr_in - 1, self.kr_out - 1
        ffl = bc.fix_fld_icb
        if self.free_icb:
            icb_free_w_diffuse2(ffl, bc.free_vt_icb, bc.r_icb, g3, coef, w, out, ki)
        else:
            for c in (0, 2):
                d2 = ffl[0][2] * w[c, ki] + ffl[1][2] * w[c, ki + 1] + ffl[2][2] * w[c, ki + 2]
                out[c, ki] = coef * (d2 - g3 * bc.r_icb[2] * w[c, ki])
        if self.inner_core:
            # cal_icore_viscous_drag_l1 (cal_inner_core_rotation.f90:249-276):
            # the viscous torque on the core replaces the l = 1 ICB row
            mat_0 = self.fdm1_icb[0] - 2.0 * self.ar1[ki]
            mat_1 = self.fdm1_icb[1]
            ar = 1.0 / self.r_icb
            for j in (1, 2, 3):
                out[0, ki, j] = 5.0 * coef * ar * (mat_0 * w[0, ki, j] + mat_1 * w[0, ki + 1, j])
        ffl = bc.fix_fld_cmb
        if self.free_cmb:
            cmb_free_w_diffuse2(ffl, bc.free_vt_cmb, bc.r_cmb, g3, coef, w, out, ko)
        else:
            for c in (0, 2):
                d2 = ffl[2][2] * w[c, ko - 2] + ffl[1][2] * w[c, ko - 1] + ffl[0][2] * w[c, ko]
                out[c, ko] = coef * (d2 - g3 * bc.r_cmb[2] * w[c, ko])

    def magnetic_diffusion(self):
        """const_sph_mag_diffuse_by_j (const_sph_rotation.f90):
        cal_sph_nod_diffuse_by_rot2 on the current density, then
        cal_sph_nod_icb_ins_diffuse2 / cal_sph_nod_cmb_ins_diffuse2 on B."""
        b, jc, out = self.f["magne"], self.f["current"], self.f["b_diffuse"]
        coef, g1, g3, bc = self.coef_d_magne, self.g1, self.g3, self.bc
        for k in self._interior_b():
            d2s = self._dr(jc[0], k, 2)
            d1t = self._dr(jc[2], k, 1)
            out[0, k] = -coef * jc[2, k]
            out[1, k] = -coef * d1t
            out[2, k] = coef * (d2s - g3 * self.ar2[k] * jc[0, k])
        ki, ko = self.kr_in - 1, self.kr_out - 1
        if self.inner_core:
            # cal_sph_nod_center_diffuse2 (set_sph_exp_nod_center.f90:128-156)
            d2 = self.d2
            d2s = d2[0, 1] * b[0, 0] + d2[0, 2] * b[0, 1]
            d2t = d2[0, 1] * b[2, 0] + d2[0, 2] * b[2, 1]
            out[0, 0] = coef * (d2s - g3 * self.ar2[0] * b[0, 0])
            out[2, 0] = coef * (d2t - g3 * self.ar2[0] * b[2, 0])
        else:
            fdr, ffl = bc.fix_dr_icb, bc.fix_fld_icb
            d2s = (((g1 + 1.0) * bc.r_icb[1] * fdr[0][2] + fdr[1][2]) * b[0, ki]
                   + fdr[2][2] * b[0, ki + 1])
            d2t = ffl[0][2] * b[2, ki] + ffl[1][2] * b[2, ki + 1] + ffl[2][2] * b[2, ki + 2]
            out[0, ki] = coef * (d2s - g3 * bc.r_icb[2] * b[0, ki])
            out[2, ki] = coef * (d2t - g3 * bc.r_icb[2] * b[2, ki])
        fdr, ffl = bc.fix_dr_cmb, bc.fix_fld_cmb
        d2s = (fdr[0][2] * b[0, ko - 1]
               + (fdr[1][2] - g1 * bc.r_cmb[1] * fdr[2][2]) * b[0, ko])
        d2t = ffl[2][2] * b[2, ko - 2] + ffl[1][2] * b[2, ko - 1] + ffl[0][2] * b[2, ko]
        out[0, ko] = coef * (d2s - g3 * bc.r_cmb[2] * b[0, ko])
        out[2, ko] = coef * (d2t - g3 * bc.r_cmb[2] * b[2, ko])

    def scalar_diffusion(self):
        """const_sph_scalar_diffusion: cal_sph_nod_scalar_diffuse2
        (cal_sph_exp_diffusion.f90), cal_sph_fix_scalar_in_diffuse2 /
        cal_sph_out_fix_scalar_diffuse2 (cal_sph_exp_fixed_scalar.f90)."""
        t, out = self.f["temp"], self.f["t_diffuse"]
        coef, g3, bc = self.coef_d_temp, self.g3, self.bc
        for k in self._interior():
            d1s = self._dr(t, k, 1)
            d2s = self._dr(t, k, 2)
            out[k] = coef * (d2s + 2.0 * self.ar1[k] * d1s - g3 * self.ar2[k] * t[k])
        ki, ko = self.kr_in - 1, self.kr_out - 1
        ffl = bc.fix_fld_icb
        d1t = ffl[0][1] * self.T_icb + ffl[1][1] * t[ki + 1] + ffl[2][1] * t[ki + 2]
        d2t = ffl[0][2] * self.T_icb + ffl[1][2] * t[ki + 1] + ffl[2][2] * t[ki + 2]
        t[ki] = self.T_icb
        out[ki] = coef * (d2t + 2.0 * bc.r_icb[1] * d1t - g3 * bc.r_icb[2] * t[ki])
        ffl = bc.fix_fld_cmb
        d1t = ffl[2][1] * t[ko - 2] + ffl[1][1] * t[ko - 1] + ffl[0][1] * self.T_cmb
        d2t = ffl[2][2] * t[ko - 2] + ffl[1][2] * t[ko - 1] + ffl[0][2] * self.T_cmb
        t[ko] = self.T_cmb
        out[ko] = coef * (d2t + 2.0 * bc.r_cmb[1] * d1t - g3 * bc.r_cmb[2] * t[ko])

    def rot_buoyancy(self):
        """cal_rot_buoyancy_sph_MHD (cal_rot_buoyancies_sph_MHD.f90)."""
        ki, ko = self.kr_in - 1, self.kr_out - 1
        rb = self.f["rot_buoyancy"]
        rb[ki:ko + 1] = self.coef_buo * self.f["temp"][ki:ko + 1] * self.radius[ki:ko + 1, None]

    def _force_rot(self, src, rot, rigid):
        """const_sph_force_rot2 / const_sph_rotation_uxb: boundary rows
        (cal_sph_nod_icb_rigid_rot2 + cal_sph_nod_cmb_rigid_rot2 for the
        momentum terms, cal_sph_nod_icb_ins_vp_rot2 +
        cal_sph_nod_cmb_ins_vp_rot2 for the induction) and
        cal_sph_nod_vect_w_div_rot2 inside."""
        bc, g1, g3 = self.bc, self.g1, self.g3
        ki, ko = self.kr_in - 1, self.kr_out - 1
        if rigid:
            # sel_bc_sph_vorticity (select_exp_velocity_bc.f90:158-190): the
            # velocity walls' formulas, cal_sph_nod_icb_free_rot2 /
            # cal_sph_nod_cmb_free_rot2 at a free-slip wall
            if self.free_icb:
                icb_free_rot2(bc.free_vp_icb, bc.free_vt_icb, bc.r_icb, g3, src, rot, ki)
            else:
                fdr, ffl = bc.fix_dr_icb, bc.fix_fld_icb
                d2s = fdr[1][2] * src[0, ki] + fdr[2][2] * src[0, ki + 1]
                d1t = ffl[0][1] * src[2, ki] + ffl[1][1] * src[2, ki + 1] + ffl[2][1] * src[2, ki + 2]
                rot[0, ki] = src[2, ki]
                rot[1, ki] = d1t
                rot[2, ki] = -(d2s - g3 * bc.r_icb[2] * src[0, ki])
            if self.free_cmb:
                cmb_free_rot2(bc.free_vp_cmb, bc.free_vt_cmb, bc.r_cmb, g3, src, rot, ko)
            else:
                fdr, ffl = bc.fix_dr_cmb, bc.fix_fld_cmb
                d2s = fdr[0][2] * src[0, ko - 1] + fdr[1][2] * src[0, ko]
                d1t = ffl[2][1] * src[2, ko - 2] + ffl[1][1] * src[2, ko - 1] + ffl[0][1] * src[2, ko]
                rot[0, ko] = src[2, ko]
                rot[1, ko] = d1t
                rot[2, ko] = -(d2s - g3 * bc.r_cmb[2] * src[0, ko])
        else:
            if self.inner_core:
                self._center_rot2(src, rot)      # sel_bc_sph_rotation_uxb
            else:
                rot[0, ki] = src[2, ki]
                rot[1, ki] = (g1 + 1.0) * bc.r_icb[1] * src[2, ki]
                rot[2, ki] = 0.0
            rot[0, ko] = src[2, ko]
            rot[1, ko] = -g1 * bc.r_cmb[1] * src[2, ko]
            rot[2, ko] = 0.0
        for k in (self._interior() if rigid else self._interior_b()):
            d1d = self._dr(src[1], k, 1)
            d1t = self._dr(src[2], k, 1)
            rot[0, k] = src[2, k]
            rot[1, k] = d1t
            rot[2, k] = -d1d + g3 * self.ar2[k] * src[0, k]

    def scalar_advect(self):
        """const_sph_scalar_advect: cal_sph_nod_vect_div2 inside,
        cal_sph_div_flux_4_fix_in / _out at fixed-value boundaries (the
        boundary value of the scalar stands in for the flux there, as the
        reference does)."""
        flux, out = self.f["h_flux"], self.f["h_advect"]
        bc, g3 = self.bc, self.g3
        gg = np.maximum(g3, 0.5)
        for k in self._interior():
            d1s = self._dr(flux[0], k, 1)
            out[k] = (d1s - flux[1, k]) * gg * self.ar2[k]
        ki, ko = self.kr_in - 1, self.kr_out - 1
        ffl = bc.fix_fld_icb
        d1s = ffl[0][1] * self.T_icb + ffl[1][1] * flux[0, ki + 1] + ffl[2][1] * flux[0, ki + 2]
        out[ki] = (d1s - flux[1, ki]) * gg * bc.r_icb[2]
        ffl = bc.fix_fld_cmb
        d1s = ffl[2][1] * flux[0, ko - 2] + ffl[1][1] * flux[0, ko - 1] + ffl[0][1] * self.T_cmb
        out[ko] = (d1s - flux[1, ko]) * gg * bc.r_cmb[2]

    # --------------------------------------------------------- nonlinear --
    def coriolis(self):
        """sum_rot_coriolis_rlm_10 (sum_coriolis_terms_rlm.f90:88-165) on the
        rj fields: omega_rlm = (r^2/2, r, 1) (set_rotatio_spectr)."""
        u, w, c = self.f["velo"], self.f["vort"], self.cor
        r = self.radius[:, None]
        ar = 1.0 / self.radius[:, None]
        om0, om1, om2 = 0.5 * r ** 2, r, 1.0
        k1, k2, l1 = c.jk1, c.jk2, c.jl1
        sp_dvp_k1, sp_dvp_k2 = u[1][:, k1], u[1][:, k2]
        sp_wp_l1 = w[0][:, l1]
        sp_vp_k1, sp_vp_k2 = u[0][:, k1], u[0][:, k2]
        sp_d2vp_l1 = ar * ar * c.g3rlm[None, :] * u[0][:, l1] - w[2][:, l1]
        sp_vp_l1 = u[0][:, l1]
        sp_wp_k1, sp_wp_k2 = w[0][:, k1], w[0][:, k2]
        sp_dwp_k1 = w[1][:, k1] - 2.0 * ar * w[0][:, k1]
        sp_dwp_k2 = w[1][:, k2] - 2.0 * ar * w[0][:, k2]
        pol = (c.sw11 * om0 * sp_dvp_k1 + c.sw21 * om0 * sp_dvp_k2 + c.sw13 * om0 * sp_wp_l1
               + c.sw12 * om1 * sp_vp_k1 + c.sw22 * om1 * sp_vp_k2)
        tor = (c.tw11 * om0 * sp_d2vp_l1 - c.tw12 * om2 * sp_vp_l1 + c.tw13 * om1 * sp_wp_k1
               + c.tw23 * om1 * sp_wp_k2 + c.tw14 * om0 * sp_dwp_k1 + c.tw24 * om0 * sp_dwp_k2)
        pol = -self.coef_cor * pol * ar * ar
        tor = -self.coef_cor * tor * ar * ar
        return pol, tor

    def nonlinear(self):
        """nonlinear (cal_nonlinear.f90): backward transform of u, w, B, J, T,
        grid products, forward transform, Coriolis in spectral space, curl /
        divergence of the forces, reference-temperature advection, force sum."""
        f, J, nr = self.f, self.J, self.nr
        perm = self.rj_of_rlm
        sp = np.zeros((nr, J, 13))
        for c0, name in ((0, "velo"), (3, "vort"), (6, "magne"), (9, "current")):
            for c in range(3):
                sp[:, :, c0 + c] = f[name][c][:, perm]
        sp[:, :, 12] = f["temp"][:, perm]
        vr = leg_bwd(self.grid, 13, 4, 1, sp.ravel())
        nphi = self.c.nphi
        fld = fft_bwd(self.L, self.c.nth, nr, nphi, 13, vr)
        nnod = nr * self.c.nth * nphi
        frc = products_mhd(nnod, fld, self.coef_velo, self.coef_lor, self.coef_induct,
                           self.coef_temp)
        vr2 = fft_fwd(self.L, self.c.nth, nr, nphi, 19, frc)
        sp2 = leg_fwd(self.grid, 19, 6, 1, vr2).reshape(nr, J, 19)
        inv = np.empty(J, dtype=np.int64)
        inv[perm] = np.arange(J)

        def take(c0):
            out = np.zeros((3, nr, J))
            for c in range(3):
                out[c] = sp2[:, inv, c0 + c]
            return out

        f["m_advect"] = take(0)
        rc = take(6)
        rc[0], rc[2] = self.coriolis()
        ki = self.kr_in - 1
        if self.inner_core:
            # inner_core_rot_z_coriolis_rlm (sum_coriolis_terms_rlm.f90:
            # 316-351): on the core's l = 1 ICB rows only the z rotation's
            # coupling of the x and y turns remains
            w0 = f["vort"][0]
            om0 = 0.5 * self.r_icb ** 2
            rc[0, ki, 2] = 0.0
            rc[0, ki, 1] = -(2.0 * self.coef_cor * self.r_icb * om0 * w0[ki, 3])
            rc[0, ki, 3] = 2.0 * self.coef_cor * self.r_icb * om0 * w0[ki, 1]
        f["rot_Coriolis"] = rc
        f["lorentz"] = take(9)
        f["vp_induct"] = take(12)
        f["h_flux"] = take(15)
        self._force_rot(f["m_advect"], f["rot_inertia"], rigid=True)
        self._force_rot(f["lorentz"], f["rot_Lorentz"], rigid=True)
        if self.inner_core:
            self.int_icore_toroidal_lorentz()
        self._force_rot(f["vp_induct"], f["induction"], rigid=False)
        self.scalar_advect()
        ki, ko = self.kr_in - 1, self.kr_out - 1
        ha = f["h_advect"]
        for k in range(ki, ko + 1):
            ha[k] = ha[k] + self.coef_temp * self.g3 * self.ar2[k] * self.reft1[k] * f["velo"][0, k]
        fo = f["forces"]
        fo[0] = -f["rot_inertia"][0] + f["rot_Coriolis"][0] + f["rot_Lorentz"][0]
        fo[2] = (-f["rot_inertia"][2] + f["rot_Coriolis"][2] + f["rot_Lorentz"][2]
                 + f["rot_buoyancy"])

    def int_icore_toroidal_lorentz(self):
        """int_icore_tor_lorentz_l1 (cal_inner_core_rotation.f90:280-314):
        the Lorentz torque on the core, 5/2 r_i^-3 times the trapezoid
        integral of the l = 1 toroidal Lorentz force times r^2 from the centre
        (first interval f(r_1) r_1^3) to the ICB, in the ICB rows of the
        toroidal force and of the poloidal curl."""
        lor, rl = self.f["lorentz"], self.f["rot_Lorentz"]
        ki = self.kr_in - 1
        r = self.radius
        ar = 1.0 / self.r_icb
        for j in (1, 2, 3):
            sk = lor[2, 0, j] * (r[0] * r[0] * r[0])
            for k in range(ki):
                sk = sk + ((lor[2, k, j] * (r[k] * r[k]) + lor[2, k + 1, j] * (r[k + 1] * r[k + 1]))
                           * (r[k + 1] - r[k]))
            lor[2, ki, j] = 0.5 * 5.0 * sk * (ar * ar * ar)
            rl[0, ki, j] = lor[2, ki, j]

    # ------------------------------------------------------ initialisation --
    def initialize(self, seeded=None):
        """SPH_initialize_MHD: dynamo_benchmark_1 initial fields
        (set_initial_sph_dynamo.f90, initial_magne_dynamobench.f90), the
        perturbation temperature, the fields derived from them
        (set_sph_field_to_start) and the first nonlinear terms.
        seeded=(seed, amp_u, amp_b, amp_t) starts instead from the seeded
        random fields: poloidal / toroidal u and B (l >= 1) and the
        temperature perturbation (l >= 0), amplitude (l+1)^-2 times the wall
        envelope, field ids 0, 2, 6, 8, 12 (oracle.seeded_spectrum), with the
        reference temperature added to the l = 0 mode."""
        nr, J = self.nr, self.J
        f = self.f
        for name in ("velo", "vort", "magne", "current", "w_diffuse", "b_diffuse",
                     "forces", "pre_mom", "rot_inertia", "rot_Coriolis", "rot_Lorentz",
                     "induction", "pre_uxb", "m_advect", "lorentz", "vp_induct", "h_flux"):
            f[name] = np.zeros((3, nr, J))
        for name in ("temp", "t_diffuse", "rot_buoyancy", "h_advect", "pre_heat"):
            f[name] = np.zeros((nr, J))
        if seeded is not None:
            from oracle import seeded_spectrum
            seed, amp_u, amp_b, amp_t = seeded
            r, L = self.radius, self.L
            f["velo"][0] = seeded_spectrum(seed, 0, L, r, amp_u, 2.0, True, 1)
            f["velo"][2] = seeded_spectrum(seed, 2, L, r, amp_u, 2.0, True, 1)
            f["magne"][0] = seeded_spectrum(seed, 6, L, r, amp_b, 2.0, True, 1)
            f["magne"][2] = seeded_spectrum(seed, 8, L, r, amp_b, 2.0, True, 1)
            f["temp"] = seeded_spectrum(seed, 12, L, r, amp_t, 2.0, True, 0)
            f["temp"][:, 0] = self.reft0 + f["temp"][:, 0]
            self._start()
            return
        pi = 4.0 * math.atan(1.0)
        r = self.radius
        shell = self.r_cmb - self.r_icb
        f["temp"][:, 0] = self.reft0
        m = int((self.c.isig % 1000) / 100)
        jj = m * (m + 1) + m
        # scalar libm arithmetic per radius (pow, sin, cos as the Fortran has them)
        rc, ri = self.r_cmb, self.r_icb
        ki, ko = self.kr_in - 1, self.kr_out - 1
        for k, x in enumerate(float(v) for v in r):
            if k < ki or k > ko:
                continue
            xr = 2.0 * x - 1.0 * (rc + ri) / shell
            f["temp"][k, jj] = ((1.0 - 3.0 * (xr * xr) + 3.0 * math.pow(xr, 4) - math.pow(xr, 6))
                                * 0.1 * 3.0 / math.sqrt(2.0 * pi))
        js, jt = 2, 6                       # (l, m) = (1, 0), (2, 0)
        if self.inner_core:
            # initial_b_dynamobench_2 (initial_magne_dynamobench.f90:103-147),
            # from the first point to the CMB
            for k in range(ko + 1):
                x = float(r[k])
                f["magne"][0][k, js] = ((5.0 / 2.0) * (x * x) * (4.0 * rc - 3.0 * x)
                                        / (rc + 3.0))
                f["magne"][1][k, js] = (5.0 / 2.0) * x * (8.0 * rc - 9.0 * x) / (rc + 3.0)
                f["current"][2][k, js] = 5.0 * 6.0 * x / (3.0 + rc)
                s_ = math.sin(pi * x / rc)
                c_ = math.cos(pi * x / rc)
                f["magne"][2][k, jt] = (10.0 / 3.0) * x * s_
                f["current"][0][k, jt] = f["magne"][2][k, jt]
                f["current"][1][k, jt] = (10.0 / 3.0) * (s_ + (pi / rc) * x * c_)
            self._start()
            return
        for k, x in enumerate(float(v) for v in r):
            f["magne"][0][k, js] = (5.0 / 8.0) * (-3.0 * math.pow(x, 3) + (4.0 * rc) * (x * x)
                                                  - math.pow(ri, 4) / x)
            f["magne"][1][k, js] = (5.0 / 8.0) * (-9.0 * (x * x) + 8.0 * rc * x
                                                  + math.pow(ri, 4) / (x * x))
        f["current"][2][:, js] = (5.0 * 3.0 / 2.0) * r
        sin_ = np.array([math.sin(pi * (x - self.r_icb)) for x in r])   # libm, as Fortran
        cos_ = np.array([math.cos(pi * (x - self.r_icb)) for x in r])
        f["magne"][2][:, jt] = (10.0 / 3.0) * r * sin_
        f["current"][0][:, jt] = f["magne"][2][:, jt]
        f["current"][1][:, jt] = (10.0 / 3.0) * (sin_ + pi * r * cos_)
        self._start()

    def _start(self):
        self.sync_temp_by_per_temp()
        # set_sph_field_to_start
        self.grad_vp_and_vorticity()
        self.set_inner_core_rotation()
        self.vorticity_diffusion()
        self.rot_buoyancy()
        self.scalar_diffusion()
        self.grad_bp_and_current()
        self.magnetic_diffusion()
        self.nonlinear()
        self.i_step = 0

    def sync_temp_by_per_temp(self):
        self.f["temp"][:, 0] = self.f["temp"][:, 0] - self.reft0

    def trans_per_temp_to_temp(self):
        self.f["temp"][:, 0] = self.f["temp"][:, 0] + self.reft0

    # --------------------------------------------------------------- step --
    def explicit(self, euler, first):
        """cal_expricit_sph_euler / cal_expricit_sph_adams
        (cal_momentum_eq_explicit.f90:54-128)."""
        f, dt = self.f, self.dt
        a0, a1 = 3.0 / 2.0, -1.0 / 2.0
        ki, ko = self.kr_in - 1, self.kr_out - 1
        sl = slice(ki, ko + 1)
        w, wd, fo, pm = f["vort"], f["w_diffuse"], f["forces"], f["pre_mom"]
        b, bd, ind, pu = f["magne"], f["b_diffuse"], f["induction"], f["pre_uxb"]
        t, td, ha, ph = f["temp"], f["t_diffuse"], f["h_advect"], f["pre_heat"]
        for c in (0, 2):
            if euler:
                w[c, sl] = w[c, sl] + dt * (self.coef_exp_v * wd[c, sl] + fo[c, sl])
                b[c] = b[c] + dt * (self.coef_exp_b * bd[c] + ind[c])
            else:
                w[c, sl] = w[c, sl] + dt * (self.coef_exp_v * wd[c, sl] + a0 * fo[c, sl]
                                            + a1 * pm[c, sl])
                pm[c, sl] = fo[c, sl]
                b[c] = b[c] + dt * (self.coef_exp_b * bd[c] + a0 * ind[c] + a1 * pu[c])
                pu[c] = ind[c]
        if euler:
            t[sl] = t[sl] + dt * (self.coef_exp_t * td[sl] - ha[sl])
        else:
            t[sl] = t[sl] + dt * (self.coef_exp_t * td[sl] - a0 * ha[sl] + a1 * ph[sl])
            ph[sl] = -ha[sl]
        if first:
            for c in (0, 2):
                pm[c] = fo[c]
                pu[c] = ind[c]
            ph[sl] = -ha[sl]

    def crank_nicolson(self):
        """s_cal_sol_sph_MHD_crank (cal_sol_sph_MHD_crank.f90:44-101) with
        cal_sol_velo_by_vort_sph_crank, cal_sol_temperature_sph_crank,
        cal_sol_magne_sph_crank (cal_sol_sph_fluid_crank.f90)."""
        f = self.f
        u, w = f["velo"], f["vort"]
        ki, ko = self.kr_in - 1, self.kr_out - 1
        u[0] = w[2].copy()
        u[2] = w[0].copy()
        u[0][:, 0] = 0.0
        u[2][:, 0] = 0.0
        bc = self.bc
        if self.free_icb:
            icb_free_vpol2(bc.free_vp_icb, u, ki)
        elif self.inner_core:
            self._rotate_velo2(u)
        else:
            u[0, :ki + 1] = 0.0
            u[1, :ki + 1] = 0.0
            u[2, :ki + 1] = 0.0
        if self.free_cmb:
            cmb_free_vpol2(bc.free_vp_cmb, u, ko)
        else:
            u[0, ko] = 0.0
            u[1, ko] = 0.0
            u[2, ko] = 0.0
        u[0] = self._solve("vp", u[0])
        u[2] = self._solve("vt", u[2])
        self.grad_vp_and_vorticity()
        # temperature
        t = f["temp"]
        t[ko:] = self.T_cmb
        t[:ki + 1] = self.T_icb
        f["temp"] = t = self._solve("temp", t)
        # magnetic field
        b = f["magne"]
        b[0][:, 0] = 0.0
        b[2][:, 0] = 0.0
        if not self.inner_core:
            b[1, ki] = (self.g1 + 1.0) * self.bc.r_icb[1] * b[0, ki]
            b[2, ki] = 0.0
        b[1, ko] = -self.g1 * self.bc.r_cmb[1] * b[0, ko]
        b[2, ko] = 0.0
        b[0] = self._solve("bs", b[0])
        b[2] = self._solve("bt", b[2])
        self.grad_bp_and_current()
        # update_after_*
        self.set_inner_core_rotation()
        self.vorticity_diffusion()
        self.rot_buoyancy()
        self.scalar_diffusion()
        self.magnetic_diffusion()

    def step(self):
        """SPH_analyze_MHD for one step; returns nothing, advances i_step."""
        self.i_step += 1
        first = self.i_step == 1
        self.explicit(euler=first, first=first)
        self.crank_nicolson()
        self.nonlinear()
        self.trans_per_temp_to_temp()
        self._last_energy = None
        self.sync_temp_by_per_temp()

    def pressure(self):
        """pressure_4_sph_mhd (lead_fields_4_sph_mhd.f90:70-95) at an output
        step, from the fields after the step and the forces of its nonlinear
        stage: the divergences of the advection and Lorentz force
        (const_sph_div_force, const_sph_divergence.f90:56-78, with
        cal_sph_nod_vect_div2 cal_sph_exp_rotation.f90:210-239 inside and
        cal_sph_nod_nobc_in/out_div2 cal_sph_exp_nod_none_bc.f90:133-160,
        269-296 on the walls), of the buoyancy (cal_div_buoyancy_sph_MHD
        cal_div_buoyancies_sph_MHD.f90:124-145 over the perturbation
        temperature and its gradient: cal_sph_nod_gradient_2
        cal_sph_exp_1st_diff.f90:40-75, dsdr_sph_fix_scalar_in/out_2
        cal_sph_exp_fixed_scalar.f90:76-106, 208-238,
        normalize_sph_average_grad :79-97); the wall rows replaced by radial
        forces (s_const_radial_forces_on_bc const_radial_forces_on_bc.f90:
        26-60: the advection's then the Lorentz force's radial part both
        into the inertia term, so the Lorentz force's stays; the buoyancy's
        coef T r, cal_r_buoyancy_on_sph cal_r_buoyancies_on_sph.f90:97-113);
        the sum set_MHD_terms_to_div_force (cal_div_of_forces.f90:81-95: the
        Coriolis vector is not formed on the grid or in rlm, so its term is
        zero); the Poisson solve with the matrix of
        const_radial_mat_4_press_sph (const_r_mat_4_scalar_sph.f90:73-129:
        set_unit_mat_4_poisson, add_scalar_poisson_mat_sph
        set_radial_mat_sph.f90:141-204, add_icb/cmb_scalar_poisson_mat
        set_sph_scalar_mat_bc.f90:105-128, 179-202, ludcmp_3band), and
        adjust_by_ave_pressure_on_CMB (set_reference_sph_mhd.f90:53-73).
        Returns d_rj of the pressure, (nr, J)."""
        f, nr, J = self.f, self.nr, self.J
        ki, ko = self.kr_in - 1, self.kr_out - 1
        bc, g3, r = self.bc, self.g3, self.radius
        gh = np.maximum(g3, 0.5)
        g13 = np.where(self.l_of_j == 0, 1.0, 1.0 / np.where(g3 == 0.0, 1.0, g3))
        t = f["temp"]

        def div(v):
            out = np.zeros((nr, J))
            for k in range(ki + 1, ko):
                out[k] = (self._dr(v[0], k, 1) - v[1, k]) * gh * self.ar2[k]
            ffl = bc.fix_fld_icb
            d1 = ffl[0][1] * v[0, ki] + ffl[1][1] * v[0, ki + 1] + ffl[2][1] * v[0, ki + 2]
            out[ki] = (d1 - v[1, ki]) * gh * bc.r_icb[2]
            ffl = bc.fix_fld_cmb
            d1 = ffl[2][1] * v[0, ko - 2] + ffl[1][1] * v[0, ko - 1] + ffl[0][1] * v[0, ko]
            out[ko] = (d1 - v[1, ko]) * gh * bc.r_cmb[2]
            return out

        # gradient of the perturbation temperature (poloidal part)
        gt = np.zeros((nr, J))
        for k in range(ki + 1, ko):
            gt[k] = self._dr(t, k, 1) * g13 * r[k] ** 2
        ffl = bc.fix_fld_icb
        d1 = ffl[0][1] * self.T_icb + ffl[1][1] * t[ki + 1] + ffl[2][1] * t[ki + 2]
        gt[ki] = d1 * g13 * bc.r_icb[0] ** 2
        ffl = bc.fix_fld_cmb
        d1 = ffl[2][1] * t[ko - 2] + ffl[1][1] * t[ko - 1] + ffl[0][1] * self.T_cmb
        gt[ko] = d1 * g13 * bc.r_cmb[0] ** 2
        gt[:, 0] = 2.0 * gt[:, 0]

        div_in = div(f["m_advect"])
        div_lor = div(f["lorentz"])
        div_buo = np.zeros((nr, J))
        for k in range(ki, ko + 1):
            div_buo[k] = self.coef_buo * (3.0 * t[k] + gt[k] * r[k] * g3 * self.ar1[k])
        for k in (ki, ko):
            div_buo[k] = self.coef_buo * t[k] * r[k]
            div_in[k] = f["lorentz"][0, k] * gh * self.ar2[k]
        rhs = -div_in + div_lor + div_buo
        self.press_terms = {"div_inertia": div_in, "div_Lorentz_force": div_lor,
                            "div_buoyancy": div_buo, "grad_temp_pol": gt}

        coef_p = -self.coef_press
        d1c, d2c, ar1 = self.d1, self.d2, self.ar1
        m = np.zeros((J, nr, 3))
        m[:, :ki, 1] = 1.0
        m[:, ko + 1:, 1] = 1.0
        for k in range(ki + 1, ko):
            m[:, k - 1, 2] = m[:, k - 1, 2] - coef_p * (d2c[k, 0] + 2.0 * ar1[k] * d1c[k, 0])
            m[:, k, 1] = m[:, k, 1] - coef_p * (d2c[k, 1] + 2.0 * ar1[k] * d1c[k, 1]
                                                - g3 * self.ar2[k])
            m[:, k + 1, 0] = m[:, k + 1, 0] - coef_p * (d2c[k, 2] + 2.0 * ar1[k] * d1c[k, 2])
        fdr = bc.fix_dr_icb
        m[:, ki, 1] = m[:, ki, 1] - coef_p * (fdr[1][2] + 2.0 * bc.r_icb[1] * fdr[1][1]
                                              - g3 * bc.r_icb[2])
        m[:, ki + 1, 0] = m[:, ki + 1, 0] - coef_p * (fdr[2][2] + 2.0 * bc.r_icb[1] * fdr[2][1])
        fdr = bc.fix_dr_cmb
        m[:, ko - 1, 2] = m[:, ko - 1, 2] - coef_p * (fdr[0][2] + 2.0 * bc.r_cmb[1] * fdr[0][1])
        m[:, ko, 1] = m[:, ko, 1] - coef_p * (fdr[1][2] + 2.0 * bc.r_cmb[1] * fdr[1][1]
                                              - g3 * bc.r_cmb[2])
        lu, piv = ludcmp(3, m)[:2]
        p = lubksb(3, lu, piv, rhs)
        p[ki:ko + 1, 0] = p[ki:ko + 1, 0] - p[ko, 0]
        return p

    def energy_fluxes(self):
        """enegy_fluxes_4_sph_mhd (lead_fields_4_sph_mhd.f90:98-121) at an
        output step: the snapshot's backward transform of u, B, J, the
        magnetic induction (the curl of u x B of the step's nonlinear stage)
        and the temperature perturbation (sph_back_trans_snapshot_MHD), the
        scalar fluxes on the grid (s_cal_energy_flux_rtp,
        cal_energy_flux_rtp.f90:101-150: the Lorentz force coef_lor J x B as
        cal_cross_prod_w_coef makes it, dotted with u; its negative; B dotted
        with the induction; cal_buoyancy_flux_rtp_smp :222-252 coef_buo T u_r r
        over the perturbation temperature, the spherical-shell reference's
        branch), and their forward transform (sph_forward_trans_snapshot_MHD).
        Returns (4, nr, J): Lorentz_work, work_against_Lorentz,
        magnetic_ene_generation, buoyancy_flux."""
        f, J, nr = self.f, self.J, self.nr
        perm = self.rj_of_rlm
        sp = np.zeros((nr, J, 13))
        for c0, name in ((0, "velo"), (3, "magne"), (6, "current"), (9, "induction")):
            for c in range(3):
                sp[:, :, c0 + c] = f[name][c][:, perm]
        sp[:, :, 12] = f["temp"][:, perm]
        nth, nphi = self.c.nth, self.c.nphi
        vr = leg_bwd(self.grid, 13, 4, 1, sp.ravel())
        g = fft_bwd(self.L, nth, nr, nphi, 13, vr).reshape(13, -1)
        u, b, cj, ind, t = g[0:3], g[3:6], g[6:9], g[9:12], g[12]
        fl = [(cj[1] * b[2] - cj[2] * b[1]) * self.coef_lor,
              (cj[2] * b[0] - cj[0] * b[2]) * self.coef_lor,
              (cj[0] * b[1] - cj[1] * b[0]) * self.coef_lor]
        ujb = fl[0] * u[0] + fl[1] * u[1] + fl[2] * u[2]
        r = np.tile(self.radius, t.size // nr)
        flux = np.stack([ujb, -1.0 * ujb, ind[0] * b[0] + ind[1] * b[1] + ind[2] * b[2],
                         self.coef_buo * t * u[0] * r])
        vr2 = fft_fwd(self.L, nth, nr, nphi, 4, flux)
        sp2 = leg_fwd(self.grid, 4, 0, 4, vr2).reshape(nr, J, 4)
        inv = np.empty(J, dtype=np.int64)
        inv[perm] = np.arange(J)
        return np.stack([sp2[:, inv, q] for q in range(4)])

    def energies(self):
        """The sph_pwr_volume_s.dat columns K_ene_pol, K_ene_tor, K_ene,
        temperature, M_ene_pol, M_ene_tor, M_ene over the whole shell, with the
        full (reference + perturbation) temperature."""
        f = self.f
        t = f["temp"].copy()
        t[:, 0] = t[:, 0] + self.reft0
        lj = self.l_of_j.astype(np.int32)
        ke = sph_mean_square(1, self.radius, lj, self.kr_in, self.kr_out, f["velo"])
        ts = sph_mean_square(2, self.radius, lj, self.kr_in, self.kr_out, t[None])
        me = sph_mean_square(1, self.radius, lj, self.kr_in, self.kr_out, f["magne"])
        return np.array([ke[0], ke[1], ke[2], ts[0], me[0], me[1], me[2]])
