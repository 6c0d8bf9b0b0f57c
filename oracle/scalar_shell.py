"""TEST INFRASTRUCTURE, NOT PRODUCT CODE: numpy restatement of the
reference's step for scalars diffusing in a shell at rest (no velocity or
magnetic field evolved) with boundary values and fluxes given per spherical
harmonic mode -- the physics of the reference's tests/heterogineous_temp
(temperature and composition, bc_spectr.btx).

Followed routines (MHD_src/sph_MHD unless noted):
  * initial fields: set_noize_scalar_sph (set_initial_sph_dynamo.f90:390-441)
  * boundary data: read_boundary_spectr_file, set_fixed_scalar_bc_by_file,
    set_fixed_gradient_bc_by_file (MHD_src/IO/m_sph_boundary_input_data.f90:
    147-300), the inner flux negated (inner_sph_bc_scalar_sph,
    set_bc_sph_scalars.f90)
  * matrices: const_radial_mat_4_scalar_sph (const_r_mat_4_scalar_sph.f90),
    add_scalar_poisson_mat_sph (set_radial_mat_sph.f90:175-204),
    set_fix_fld_*_poisson_mat / add_fix_flux_*_poisson_mat
    (set_sph_scalar_mat_bc.f90), LU ludcmp_3band
  * diffusion: cal_sph_nod_scalar_diffuse2 (cal_sph_exp_diffusion.f90),
    cal_sph_fix_scalar_in/out_diffuse2 (cal_sph_exp_fixed_scalar.f90),
    cal_sph_in/out_fix_flux_diffuse2 (cal_sph_exp_fixed_flux.f90)
  * advection: const_sph_scalar_advect (zero in the shell at rest except
    at fixed-flux walls: cal_div_sph_in/out_fix_flux_2,
    cal_sph_exp_fixed_flux.f90)
  * explicit step: scalar_diff_advect_euler / _adams
    (cal_diff_adv_src_explicit.f90), Adams history from the first step
    (set_ini_adams_scalar)
  * Crank-Nicolson: cal_sol_scalar_sph_crank (cal_sol_sph_fluid_crank.f90:
    309-391), set_fixed_scalar_sph, adjust_in/out_fixed_flux_sph
    (set_scalar_boundary_sph.f90), lubksb_3band_mul
  * energies: volume mean square of each scalar (cal_rms_by_sph_spectr.f90)
Pinned by the reference's own tests/heterogineous_temp/reference/
sph_pwr_volume_s.dat (steps 50 and 100, tests/test_oracle_scalar_shell.py).
"""
from __future__ import annotations

import math

import numpy as np

from . import ludcmp, lubksb, sph_mean_square
from .mhd_step import BoundaryFdm, chebyshev_radius, fdm_coefs

FIXED, FLUX = 0, 1


def read_bc_spectr(path):
    """The reference's spectral boundary file: {(field, group): [(l, m, v)]}."""
    lines = [l.strip() for l in open(path) if l.strip() and not l.lstrip().startswith("#")]
    n = int(lines[0].split()[0])
    out, i = {}, 1
    for _ in range(n):
        field, group, nm = lines[i].split()[0], lines[i + 1].split()[0], int(lines[i + 2].split()[0])
        rows = []
        for q in range(nm):
            t = lines[i + 3 + q].split()
            rows.append((int(t[0]), int(t[1]), float(t[2])))
        out[(field.lower(), group.upper())] = rows
        i += 3 + nm
    return out


class ScalarShell:
    """Scalars (each with a diffusivity, a Crank-Nicolson weight and one
    boundary condition per wall) in a shell at rest."""

    def __init__(self, L, nri, r_icb, r_cmb, dt, scalars):
        """scalars: list of dicts with coef_d, coef_imp and
        icb=(kind, [(l, m, v)]), cmb=(kind, [(l, m, v)]) (kind FIXED or FLUX)."""
        self.L = L
        self.J = J = (L + 1) ** 2
        self.radius = chebyshev_radius(nri, r_icb, r_cmb)
        self.nr = nr = self.radius.size
        self.dt = dt
        self.ar1 = 1.0 / self.radius
        self.ar2 = self.ar1 ** 2
        self.d1, self.d2 = fdm_coefs(self.radius)
        self.bc = BoundaryFdm(self.radius, 1, nr)
        self.l_of_j = np.array([int(math.isqrt(j)) for j in range(J)])
        self.m_of_j = np.array([j - l * (l + 1) for j, l in enumerate(self.l_of_j)])
        self.g3 = (self.l_of_j * (self.l_of_j + 1)).astype(np.float64)
        self.sc = []
        for s in scalars:
            d = dict(coef_d=s["coef_d"], coef_imp=s["coef_imp"],
                     coef_exp=1.0 - s["coef_imp"])
            for wall in ("icb", "cmb"):
                kind, rows = s[wall]
                v = np.zeros(J)
                for l, m, x in rows:
                    if l <= L and abs(m) <= l:
                        v[l * (l + 1) + m] = x
                if wall == "icb" and kind == FLUX:
                    v = -v          # outward normal at the inner wall
                d[wall] = (kind, v)
            d["lu"] = ludcmp(3, self._matrix(d))[:2]
            self.sc.append(d)
        self.x = [np.zeros((nr, J)) for _ in self.sc]
        self.diff = [np.zeros((nr, J)) for _ in self.sc]
        self.adv = [self._advection(q) for q in range(len(self.sc))]
        self.pre = [np.zeros((nr, J)) for _ in self.sc]
        self.i_step = 0

    def _advection(self, q):
        """Divergence of the advective flux (const_sph_scalar_advect,
        const_sph_divergence.f90): zero in a shell at rest, except in the
        rows of a fixed-flux wall, where the wall formula
        (cal_div_sph_in/out_fix_flux_2, cal_sph_exp_fixed_flux.f90) takes
        the prescribed flux as the flux there: (q_j - 0) max(l(l+1), 1/2)
        / r^2.  (At a fixed-value wall the row is overwritten by the wall
        value before it is used.)"""
        s, bc = self.sc[q], self.bc
        a = np.zeros((self.nr, self.J))
        gm = np.maximum(self.g3, 0.5)
        if s["icb"][0] == FLUX:
            a[0] = (s["icb"][1] - 0.0) * gm * bc.r_icb[2]
        if s["cmb"][0] == FLUX:
            a[-1] = (s["cmb"][1] - 0.0) * gm * bc.r_cmb[2]
        return a

    # ----------------------------------------------------------- matrix --
    def _matrix(self, s):
        nr, J = self.nr, self.J
        ki, ko = 0, nr - 1
        m = np.zeros((J, nr, 3))
        m[:, :, 1] = 1.0
        coef = s["coef_imp"] * s["coef_d"] * self.dt
        d1, d2, ar1, ar2, g3 = self.d1, self.d2, self.ar1, self.ar2, self.g3
        for k in range(ki + 1, ko):
            m[:, k - 1, 2] = m[:, k - 1, 2] - coef * (d2[k, 0] + 2.0 * ar1[k] * d1[k, 0])
            m[:, k, 1] = m[:, k, 1] - coef * (d2[k, 1] + 2.0 * ar1[k] * d1[k, 1] - g3 * ar2[k])
            m[:, k + 1, 0] = m[:, k + 1, 0] - coef * (d2[k, 2] + 2.0 * ar1[k] * d1[k, 2])
        bc = self.bc
        if s["icb"][0] == FLUX:
            fdr = bc.fix_dr_icb
            m[:, ki, 1] = m[:, ki, 1] - coef * (fdr[1][2] - g3 * bc.r_icb[2])
            m[:, ki + 1, 0] = m[:, ki + 1, 0] - coef * fdr[2][2]
        else:
            m[:, ki, 1] = 1.0
            m[:, ki + 1, 0] = 0.0
        if s["cmb"][0] == FLUX:
            fdr = bc.fix_dr_cmb
            m[:, ko - 1, 2] = m[:, ko - 1, 2] - coef * fdr[0][2]
            m[:, ko, 1] = m[:, ko, 1] - coef * (fdr[1][2] - g3 * bc.r_cmb[2])
        else:
            m[:, ko - 1, 2] = 0.0
            m[:, ko, 1] = 1.0
        return m

    # ------------------------------------------------------------ fields --
    def set_noise(self):
        """set_noize_scalar_sph: zero, then for the modes after (l, m) = (0, 0)
        in the reference's local order -- orders m >= 0, degrees l >= 1 --
        and radial points ICB+2..CMB-2 the profile
        (1 - 3x^2 + 3x^4 - x^6) 1e-4 6 / sqrt(pi)."""
        pi = 4.0 * math.atan(1.0)
        r = self.radius
        ki, ko = 0, self.nr - 1
        shell = r[ko] - r[ki]
        modes = (self.m_of_j >= 0) & (self.l_of_j >= 1)
        for x in self.x:
            x[:] = 0.0
            for k in range(ki + 2, ko - 1):
                xr = 2.0 * float(r[k]) - (float(r[ki + 2]) + float(r[ko - 2])) / shell
                v = (1.0 - 3.0 * xr ** 2 + 3.0 * math.pow(xr, 4) - math.pow(xr, 6)) \
                    * 1.0e-4 * 6.0 / math.sqrt(pi)
                x[k, modes] = v

    def _diffusion(self, q):
        """update_after_heat_sph / update_after_composit_sph
        (cal_sol_sph_MHD_crank.f90:191-236): the radial gradient first
        (const_radial_grad_scalar; its fixed-value wall rows,
        dsdr_sph_fix_scalar_in_2 / _out_2, put the wall values into the
        field), then the diffusion (const_sph_scalar_diffusion: interior,
        then the wall rows)."""
        s, x, out = self.sc[q], self.x[q], self.diff[q]
        bc, g3, coef = self.bc, self.g3, s["coef_d"]
        ki, ko = 0, self.nr - 1
        d1, d2 = self.d1, self.d2
        if s["icb"][0] == FIXED:
            x[ki] = s["icb"][1]
        if s["cmb"][0] == FIXED:
            x[ko] = s["cmb"][1]
        for k in range(ki + 1, ko):
            d1s = d1[k, 0] * x[k - 1] + d1[k, 1] * x[k] + d1[k, 2] * x[k + 1]
            d2s = d2[k, 0] * x[k - 1] + d2[k, 1] * x[k] + d2[k, 2] * x[k + 1]
            out[k] = coef * (d2s + 2.0 * self.ar1[k] * d1s - g3 * self.ar2[k] * x[k])
        kind, v = s["icb"]
        if kind == FIXED:
            f = bc.fix_fld_icb
            d1t = f[0][1] * v + f[1][1] * x[ki + 1] + f[2][1] * x[ki + 2]
            d2t = f[0][2] * v + f[1][2] * x[ki + 1] + f[2][2] * x[ki + 2]
            x[ki] = v
            out[ki] = coef * (d2t + 2.0 * bc.r_icb[1] * d1t - g3 * bc.r_icb[2] * x[ki])
        else:
            f = bc.fix_dr_icb
            d2t = f[0][2] * v + f[1][2] * x[ki] + f[2][2] * x[ki + 1]
            out[ki] = coef * (d2t + 2.0 * bc.r_icb[1] * v - g3 * bc.r_icb[2] * x[ki])
        kind, v = s["cmb"]
        if kind == FIXED:
            f = bc.fix_fld_cmb
            d1t = f[2][1] * x[ko - 2] + f[1][1] * x[ko - 1] + f[0][1] * v
            d2t = f[2][2] * x[ko - 2] + f[1][2] * x[ko - 1] + f[0][2] * v
            x[ko] = v
            out[ko] = coef * (d2t + 2.0 * bc.r_cmb[1] * d1t - g3 * bc.r_cmb[2] * x[ko])
        else:
            f = bc.fix_dr_cmb
            d2t = f[0][2] * x[ko - 1] + f[1][2] * x[ko] + f[2][2] * v
            out[ko] = coef * (d2t + 2.0 * bc.r_cmb[1] * v - g3 * bc.r_cmb[2] * x[ko])

    def initialize(self):
        """SPH_initialize_MHD after the initial fields: the diffusion terms
        (set_sph_field_to_start; fixed walls take their values)."""
        for q in range(len(self.sc)):
            self._diffusion(q)
        self.i_step = 0

    def step(self):
        self.i_step += 1
        dt = self.dt
        bc = self.bc
        ki, ko = 0, self.nr - 1
        first = self.i_step == 1
        for q, s in enumerate(self.sc):
            x, adv, pre = self.x[q], self.adv[q], self.pre[q]
            # scalar_diff_advect_euler on the first step (and the Adams
            # history started: set_ini_adams_scalar), then
            # scalar_diff_advect_adams (cal_diff_adv_src_explicit.f90)
            if first:
                x[:] = x + dt * (s["coef_exp"] * self.diff[q] - adv)
                pre[:] = -adv
            else:
                x[:] = x + dt * (s["coef_exp"] * self.diff[q] - 1.5 * adv + (-0.5) * pre)
                pre[:] = -adv
            coef = dt * s["coef_imp"] * s["coef_d"]
            kind, v = s["cmb"]
            if kind == FIXED:
                x[ko] = v
            else:
                x[ko] = x[ko] + coef * (bc.fix_dr_cmb[2][2] + 2.0 * bc.r_cmb[1]) * v
            kind, v = s["icb"]
            if kind == FIXED:
                x[ki] = v
            else:
                x[ki] = x[ki] + coef * (bc.fix_dr_icb[0][2] + 2.0 * bc.r_icb[1]) * v
            lu, piv = s["lu"]
            self.x[q] = lubksb(3, lu, piv, x)
            self._diffusion(q)

    def mean_squares(self):
        lj = self.l_of_j.astype(np.int32)
        return np.array([sph_mean_square(2, self.radius, lj, 1, self.nr, x[None])[0]
                         for x in self.x])


def heterogeneous_temp_case(bc_path):
    """tests/heterogineous_temp: L=4, 64 Chebyshev intervals, r_i/r_o=0.35,
    dt=1e-4, Pr=1 (temperature diffusivity 1), composite Prandtl 2
    (composition diffusivity 1/2); Crank-Nicolson weight 0.6 for the
    temperature and the default 0.5 for the composition (the test's control
    file names `coef_imp_d_ctl`, which the reader does not know: it reads
    `coef_imp_c_ctl`, m_ctl_data_mhd_evo_scheme.f90:183, and keeps 0.5,
    set_control_4_model.f90:287-291); temperature fixed at
    the ICB and flux at the CMB, composition flux at the ICB and fixed at the
    CMB, all from the spectral file."""
    bc = read_bc_spectr(bc_path)
    r_cmb = 1.0 / (1.0 - 0.35)
    r_icb = r_cmb * 0.35
    scalars = [
        dict(coef_d=1.0 * 1.0 ** -1.0, coef_imp=0.6,
             icb=(FIXED, bc[("temperature", "ICB")]), cmb=(FLUX, bc[("heat_flux", "CMB")])),
        dict(coef_d=1.0 * 2.0 ** -1.0, coef_imp=0.5,
             icb=(FLUX, bc[("composite_flux", "ICB")]), cmb=(FIXED, bc[("composition", "CMB")])),
    ]
    return ScalarShell(4, 64, r_icb, r_cmb, 1.0e-4, scalars)
