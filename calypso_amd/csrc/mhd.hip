// mhd.hip -- the spectral MHD time step on the GPU (gfx950), FP64.
//
// One step follows SPH_analyze_MHD (programs/SPH_MHD/SPH_analyzer_MHD.f90:
// 115-202) for a rotating Boussinesq shell with rigid or free-slip walls,
// insulating magnetic boundaries and fixed-temperature boundaries around the
// spherical-shell reference temperature, or the same shell around a
// rotatable conducting inner core (points ki > 0 below the ICB) with the
// magnetic field solved from the first point above the centre (kb = 0) and
// continued above the CMB as a potential field:
//
//   k_explicit      cal_expricit_sph_euler / _adams
//                   (cal_momentum_eq_explicit.f90:54-128)
//   k_cn_rhs        right-hand sides and boundary rows of the Crank-Nicolson
//                   solves (cal_sol_sph_fluid_crank.f90:53-391)
//   lubksb x 5      banded solves, poloidal velocity 5-band, the rest 3-band
//   k_bc_rows       wall rows of grad / curl (set_sph_exp_rigid_*.f90,
//                   set_sph_exp_free_*.f90,
//   k_grad_interior cal_sph_exp_nod_*_ins.f90, cal_sph_exp_rotation.f90)
//   k_diffusion     diffusion terms and buoyancy (const_sph_diffusion.f90,
//                   const_sph_rotation.f90, cal_rot_buoyancies_sph_MHD.f90)
//   k_pack          rj -> rlm, 13 components (u, w, B, J, T)
//   Legendre bwd -> FFT, grid products, FFT (one pencil kernel) -> Legendre fwd
//   k_unpack        rlm -> rj of the forward components; k_coriolis the curl of the
//                   Coriolis force in spectral space
//                   (sum_coriolis_terms_rlm.f90:88-165)
//   k_icore_torque  the Lorentz torque on the inner core (cal_inner_core_rotation.f90)
//   k_rot_forces    curl of the forces, curl of u x B, divergence of u T with
//                   the reference-temperature advection, the force sum
//                   (cal_sph_field_by_rotation.f90, cal_nonlinear.f90)
//
// The whole state stays in HBM as d_rj (kMhdNumComp components of nr*J
// doubles, mode fastest, like the reference's d_rj); one thread per (k, j)
// node, so every access to a component is coalesced over j.  Pointwise and
// stencil kernels are HBM-bound; FMA contraction is off so their arithmetic
// follows the reference's operation order.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "hip_check.hpp"
#include "mhd_internal.hpp"
#include "seeded.hpp"
#include "radial_internal.hpp"

namespace calypso {


#pragma clang fp contract(off)

#define FLD(c) (p.d + (size_t)(c) * p.nnod)

__device__ __forceinline__ double dr1(const MhdDev &p, const double *f, int k, size_t i)
{
    return p.d1[3 * k + 0] * f[i - p.J] + p.d1[3 * k + 1] * f[i] + p.d1[3 * k + 2] * f[i + p.J];
}

__device__ __forceinline__ double dr2(const MhdDev &p, const double *f, int k, size_t i)
{
    return p.d2[3 * k + 0] * f[i - p.J] + p.d2[3 * k + 1] * f[i] + p.d2[3 * k + 2] * f[i + p.J];
}

// x^n for n >= 0 by binary powering (the integer power of Fortran's x**n)
__device__ inline double powi(double x, int n)
{
    double y = (n & 1) ? x : 1.0;
    while (n >>= 1) {
        x = x * x;
        if (n & 1) y = y * x;
    }
    return y;
}

// ---------------------------------------------------------- explicit ----
// node i = k J + j -> (k, j) without a 64-bit division sequence: the host
// keeps nnod below 2^31, so a double quotient and one correction are exact
struct NodeKJ {
    int k, j;
};
__device__ inline NodeKJ node_kj(const MhdDev &p, size_t i)
{
    const int ii = (int)i;
    int k = (int)((double)ii * p.rJ);
    int j = ii - k * p.J;
    if (j < 0) {
        --k;
        j += p.J;
    } else if (j >= p.J) {
        ++k;
        j -= p.J;
    }
    return {k, j};
}

__global__ void k_explicit(MhdDev p, int euler, int first)
{
    const double a0 = 3.0 / 2.0, a1 = -1.0 / 2.0, dt = p.dt;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < p.nnod;
         i += (size_t)gridDim.x * blockDim.x) {
        const int k = node_kj(p, i).k;
        const bool in_shell = (k >= p.ki && k <= p.ko);
        for (int c = 0; c <= 2; c += 2) {
            double *w = FLD(kVort + c), *pm = FLD(kPreMom + c);
            const double wd = FLD(kWDiffuse + c)[i], fo = FLD(kForces + c)[i];
            if (in_shell) {
                if (euler) {
                    w[i] = w[i] + dt * (p.coef_exp_v * wd + fo);
                } else {
                    w[i] = w[i] + dt * (p.coef_exp_v * wd + a0 * fo + a1 * pm[i]);
                    pm[i] = fo;
                }
            }
            double *b = FLD(kMagne + c), *pu = FLD(kPreUxB + c);
            const double bd = FLD(kBDiffuse + c)[i], ind = FLD(kInduction + c)[i];
            if (euler) {
                b[i] = b[i] + dt * (p.coef_exp_b * bd + ind);
            } else {
                b[i] = b[i] + dt * (p.coef_exp_b * bd + a0 * ind + a1 * pu[i]);
                pu[i] = ind;
            }
            if (first) {
                pm[i] = fo;
                pu[i] = ind;
            }
        }
        if (in_shell) {
            double *t = FLD(kTemp), *ph = FLD(kPreHeat);
            const double td = FLD(kTDiffuse)[i], ha = FLD(kHAdvect)[i];
            if (euler) {
                t[i] = t[i] + dt * (p.coef_exp_t * td - ha);
            } else {
                t[i] = t[i] + dt * (p.coef_exp_t * td - a0 * ha + a1 * ph[i]);
                ph[i] = -ha;
            }
            if (first) ph[i] = -ha;
        }
    }
}

// ------------------------------------------- Crank-Nicolson right sides --
// cal_sol_velo_by_vort_sph_crank: u_pol <- w_tor, u_tor <- w_pol,
// delete_zero_degree_comp, cal_sph_nod_icb/cmb_rigid_velo2 or
// cal_sph_nod_icb/cmb_free_vpol2;
// cal_sol_scalar_sph_crank: set_fixed_scalar_sph at both walls;
// cal_sol_magne_sph_crank: delete_zero_degree_comp,
// cal_sph_nod_icb/cmb_ins_mag2.
__global__ void k_cn_rhs(MhdDev p)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < p.nnod;
         i += (size_t)gridDim.x * blockDim.x) {
        const NodeKJ kj = node_kj(p, i);
        const int k = kj.k, j = kj.j;
        const double *wtor = FLD(kVort + 2);
        double up = wtor[i], ut = FLD(kVort)[i];
        if (j == 0) {
            up = 0.0;
            ut = 0.0;
        }
        if (k == p.ki && p.free_i) {
            // the poloidal wall value after delete_zero_degree_comp is the
            // neighbour's w_tor (zero for l = 0)
            const double up1 = j == 0 ? 0.0 : wtor[i + p.J];
            FLD(kVelo + 1)[i] = p.fvp_i[1][1] * up + p.fvp_i[2][1] * up1;
            up = 0.0;
        } else if (k == p.ko && p.free_c) {
            const double upn = j == 0 ? 0.0 : wtor[i - p.J];
            FLD(kVelo + 1)[i] = p.fvp_c[0][1] * upn + p.fvp_c[1][1] * up;
            up = 0.0;
        } else if (k <= p.ki || k == p.ko) {
            up = 0.0;
            ut = 0.0;
            FLD(kVelo + 1)[i] = 0.0;
            if (p.icore && k <= p.ki && j >= 1 && j <= 3) {
                // cal_sph_nod_icb_rotate_velo2: the l = 1 toroidal velocity
                // keeps its ICB value and turns the core below as a solid body
                const double keep = FLD(kVort)[j + (size_t)p.J * p.ki];
                ut = k == p.ki ? keep : keep * p.ri[2] * p.r[k] * p.r[k];
            }
        }
        FLD(kVelo)[i] = up;
        FLD(kVelo + 2)[i] = ut;
        if (k >= p.ko) FLD(kTemp)[i] = p.t_cmb[j];
        if (k <= p.ki) FLD(kTemp)[i] = p.t_icb[j];
        double *bp = FLD(kMagne), *bdp = FLD(kMagne + 1), *bt = FLD(kMagne + 2);
        if (j == 0) {
            bp[i] = 0.0;
            bt[i] = 0.0;
        }
        if (k == p.ki && !p.icore) {
            bdp[i] = (p.g1[j] + 1.0) * p.ri[1] * bp[i];
            bt[i] = 0.0;
        }
        if (k == p.ko) {
            bdp[i] = -p.g1[j] * p.rc[1] * bp[i];
            bt[i] = 0.0;
        }
    }
}

// ------------------------------------------------ wall rows of grad/curl --
// cal_sph_nod_icb_rigid_velo2 + cal_sph_nod_icb_rigid_rot2,
// cal_sph_nod_cmb_rigid_v_and_w or cal_sph_nod_icb/cmb_free_v_and_w,
// cal_sph_nod_icb_ins_b_and_j, cal_sph_nod_cmb_ins_b_and_j.  One thread per
// mode.
__global__ void k_bc_rows(MhdDev p)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= p.J) return;
    const double g1 = p.g1[j], g3 = p.g3[j];
    double *us = FLD(kVelo), *ud = FLD(kVelo + 1), *ut = FLD(kVelo + 2);
    double *ws = FLD(kVort), *wd = FLD(kVort + 1), *wt = FLD(kVort + 2);
    const size_t J = p.J;
    if (p.free_i) {   // ICB free slip
        const size_t i0 = j + J * p.ki, i1 = i0 + J;
        const double d1s = p.fvp_i[2][1] * us[i1];
        const double d2s = p.fvp_i[2][2] * us[i1];
        const double d1t = p.fvt_i[1][1] * ut[i0];
        us[i0] = 0.0;
        ud[i0] = d1s;
        ws[i0] = ut[i0];
        wd[i0] = d1t;
        wt[i0] = -d2s;
    } else {   // ICB rigid, or the rotatable core's wall
        const bool turn = p.icore && j >= 1 && j <= 3;
        const double keep = turn ? ut[j + J * p.ki] : 0.0;
        for (int k = 0; k <= p.ki; ++k) {
            const size_t i = j + J * k;
            us[i] = 0.0;
            ud[i] = 0.0;
            ut[i] = 0.0;
        }
        if (turn) {   // cal_sph_nod_icb_rotate_velo2
            ut[j + J * p.ki] = keep;
            for (int k = 0; k < p.ki; ++k) ut[j + J * k] = keep * p.ri[2] * p.r[k] * p.r[k];
        }
        const size_t i0 = j + J * p.ki, i1 = i0 + J, i2 = i1 + J;
        const double d2s = p.fdr_i[1][2] * us[i0] + p.fdr_i[2][2] * us[i1];
        const double d1t = p.ffl_i[0][1] * ut[i0] + p.ffl_i[1][1] * ut[i1] + p.ffl_i[2][1] * ut[i2];
        ws[i0] = ut[i0];
        wd[i0] = d1t;
        wt[i0] = -(d2s - g3 * p.ri[2] * us[i0]);
    }
    if (p.free_c) {   // CMB free slip
        const size_t i0 = j + J * p.ko, n1 = i0 - J;
        const double d1s = p.fvp_c[0][1] * us[n1];
        const double d2s = p.fvp_c[0][2] * us[n1];
        const double d1t = p.fvt_c[0][1] * ut[n1] + p.fvt_c[1][1] * ut[i0];
        us[i0] = 0.0;
        ud[i0] = d1s;
        ws[i0] = ut[i0];
        wd[i0] = d1t;
        wt[i0] = -d2s;
    } else {   // CMB rigid
        const size_t i0 = j + J * p.ko, n1 = i0 - J, n2 = n1 - J;
        us[i0] = 0.0;
        ud[i0] = 0.0;
        ut[i0] = 0.0;
        const double d2s = p.fdr_c[0][2] * us[n1];
        const double d1t = p.ffl_c[2][1] * ut[n2] + p.ffl_c[1][1] * ut[n1] + p.ffl_c[0][1] * ut[i0];
        ws[i0] = ut[i0];
        wd[i0] = d1t;
        wt[i0] = -d2s;
    }
    if (p.icore && j >= 1 && j <= 3) {
        // set_inner_core_rotation (cal_inner_core_rotation.f90:212-245):
        // solid-body rotation below the ICB from the ICB's l = 1 values
        const size_t i0 = j + J * p.ki;
        const double a1 = p.ar1[p.ki], a2 = p.ar2[p.ki];
        for (int k = 0; k < p.ki; ++k) {
            const size_t i = j + J * k;
            const double ratio = p.r[k] * p.r[k] * a2;
            ut[i] = ratio * ut[i0];
            ws[i] = ratio * ws[i0];
            wd[i] = 2.0 * ws[i0] * p.r[k] * a2;
        }
        wd[i0] = 2.0 * ws[i0] * a1;
    }
    double *bs = FLD(kMagne), *bd = FLD(kMagne + 1), *bt = FLD(kMagne + 2);
    double *js = FLD(kCurrent), *jd = FLD(kCurrent + 1), *jt = FLD(kCurrent + 2);
    if (p.icore) {   // sph_center_fld_and_curl: the first point above the centre
        const size_t i0 = j, i1 = i0 + J;
        const double d1s = p.d1[1] * bs[i0] + p.d1[2] * bs[i1];
        const double d2s = p.d2[1] * bs[i0] + p.d2[2] * bs[i1];
        const double d1t = p.d1[1] * bt[i0] + p.d1[2] * bt[i1];
        bd[i0] = d1s;
        js[i0] = bt[i0];
        jd[i0] = d1t;
        jt[i0] = -(d2s - g3 * p.ar2[0] * bs[i0]);
    } else {   // ICB insulator
        const size_t i0 = j + J * p.ki, i1 = i0 + J, i2 = i1 + J;
        const double d1s = (g1 + 1.0) * p.ri[1] * bs[i0];
        const double d2s = ((g1 + 1.0) * p.ri[1] * p.fdr_i[0][2] + p.fdr_i[1][2]) * bs[i0]
                         + p.fdr_i[2][2] * bs[i1];
        const double d1t = p.ffl_i[1][1] * bt[i1] + p.ffl_i[2][1] * bt[i2];
        bd[i0] = d1s;
        bt[i0] = 0.0;
        js[i0] = 0.0;
        jd[i0] = d1t;
        jt[i0] = -(d2s - g3 * p.ri[2] * bs[i0]);
    }
    {   // CMB insulator
        const size_t i0 = j + J * p.ko, n1 = i0 - J, n2 = n1 - J;
        const double d1s = -g1 * p.rc[1] * bs[i0];
        const double d2s = p.fdr_c[0][2] * bs[n1]
                         + (p.fdr_c[1][2] - g1 * p.rc[1] * p.fdr_c[2][2]) * bs[i0];
        const double d1t = p.ffl_c[2][1] * bt[n2] + p.ffl_c[1][1] * bt[n1];
        bd[i0] = d1s;
        bt[i0] = 0.0;
        js[i0] = 0.0;
        jd[i0] = d1t;
        jt[i0] = -(d2s - g3 * p.rc[2] * bs[i0]);
    }
    // ext_outside_potential_with_j (extend_potential_field.f90:169-201): a
    // potential field above the CMB
    const int l = (int)g1;
    for (int k = p.ko + 1; k < p.nr; ++k) {
        const size_t i = j + J * k, ic = j + J * p.ko;
        const double pw = powi(p.rc[0] * p.ar1[k], l);
        bs[i] = bs[ic] * pw;
        bd[i] = -bs[ic] * pw * g1 * p.ar1[k];
        bt[i] = 0.0;
        js[i] = 0.0;
        jd[i] = 0.0;
        jt[i] = 0.0;
    }
}

// cal_sph_diff_pol_and_rot2 for (u, w) and (B, J) inside the shell.
__global__ void k_grad_interior(MhdDev p)
{
    const size_t ist = (size_t)p.J * (p.kb + 1), ied = (size_t)p.J * p.ko;
    for (size_t i = ist + blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < ied;
         i += (size_t)gridDim.x * blockDim.x) {
        const NodeKJ kj = node_kj(p, i);
        const int k = kj.k, j = kj.j;
        const double g3a = p.g3[j] * p.ar2[k];
        // (u, w) inside the shell, (B, J) from the first point above the centre
#pragma unroll
        for (int f = k > p.ki ? 0 : 1; f < 2; ++f) {
            const int src = f == 0 ? kVelo : kMagne, dst = f == 0 ? kVort : kCurrent;
            const double *s = FLD(src), *t = FLD(src + 2);
            const double d1s = dr1(p, s, k, i), d2s = dr2(p, s, k, i), d1t = dr1(p, t, k, i);
            FLD(src + 1)[i] = d1s;
            FLD(dst)[i] = t[i];
            FLD(dst + 1)[i] = d1t;
            FLD(dst + 2)[i] = -d2s + g3a * s[i];
        }
    }
}

// ------------------------------------------------ diffusion, buoyancy ----
// const_sph_vorticirty_diffusion (cal_sph_nod_vect_diffuse2 +
// cal_sph_nod_icb/cmb_rgd_w_diffuse2 or cal_sph_nod_icb/cmb_free_w_diffuse2), const_sph_mag_diffuse_by_j
// (cal_sph_nod_diffuse_by_rot2 + cal_sph_nod_icb/cmb_ins_diffuse2),
// const_sph_scalar_diffusion (cal_sph_nod_scalar_diffuse2 +
// cal_sph_fix_scalar_in/out_diffuse2), cal_rot_buoyancy_sph_MHD.
__global__ void k_diffusion(MhdDev p)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < p.nnod;
         i += (size_t)gridDim.x * blockDim.x) {
        const NodeKJ kj = node_kj(p, i);
        const int k = kj.k, j = kj.j;
        const size_t J = p.J;
        const double g1 = p.g1[j], g3 = p.g3[j];
        const double *w0 = FLD(kVort), *w2 = FLD(kVort + 2);
        const double *b0 = FLD(kMagne), *b2 = FLD(kMagne + 2);
        const double *c0 = FLD(kCurrent), *c2 = FLD(kCurrent + 2);
        const double *t = FLD(kTemp);
        double *wd0 = FLD(kWDiffuse), *wd2 = FLD(kWDiffuse + 2);
        double *bd0 = FLD(kBDiffuse), *bd1 = FLD(kBDiffuse + 1), *bd2 = FLD(kBDiffuse + 2);
        double *td = FLD(kTDiffuse);
        const double g3a = g3 * p.ar2[k];
        // vorticity and temperature: the shell
        if (k > p.ki && k < p.ko) {
            wd0[i] = p.coef_d_velo * (dr2(p, w0, k, i) - g3a * w0[i]);
            wd2[i] = p.coef_d_velo * (dr2(p, w2, k, i) - g3a * w2[i]);
            const double d1s = dr1(p, t, k, i), d2t = dr2(p, t, k, i);
            td[i] = p.coef_d_temp * (d2t + 2.0 * p.ar1[k] * d1s - g3a * t[i]);
        } else if (k == p.ki) {
            const size_t i1 = i + J, i2 = i1 + J;
            // cal_sph_nod_icb_free_w_diffuse2 at a free-slip wall
            const double dw0 = p.free_i
                ? p.fvt_i[1][2] * w0[i] + p.fvt_i[2][2] * w0[i1]
                : p.ffl_i[0][2] * w0[i] + p.ffl_i[1][2] * w0[i1] + p.ffl_i[2][2] * w0[i2];
            const double dw2 = p.ffl_i[0][2] * w2[i] + p.ffl_i[1][2] * w2[i1] + p.ffl_i[2][2] * w2[i2];
            wd0[i] = p.coef_d_velo * (dw0 - g3 * p.ri[2] * w0[i]);
            wd2[i] = p.coef_d_velo * (dw2 - g3 * p.ri[2] * w2[i]);
            if (p.icore && j >= 1 && j <= 3) {
                // cal_icore_viscous_drag_l1: the viscous torque on the core
                const double mat0 = p.fdm1_i[0] - 2.0 * p.ar1[k], mat1 = p.fdm1_i[1];
                wd0[i] = 5.0 * p.coef_d_velo * (1.0 / p.r[k]) * (mat0 * w0[i] + mat1 * w0[i1]);
            }
            const double fix = p.t_icb[j];
            const double d1t = p.ffl_i[0][1] * fix + p.ffl_i[1][1] * t[i1] + p.ffl_i[2][1] * t[i2];
            const double d2tt = p.ffl_i[0][2] * fix + p.ffl_i[1][2] * t[i1] + p.ffl_i[2][2] * t[i2];
            td[i] = p.coef_d_temp * (d2tt + 2.0 * p.ri[1] * d1t - g3 * p.ri[2] * fix);
        } else if (k == p.ko) {
            const size_t n1 = i - J, n2 = n1 - J;
            const double dw0 = p.free_c
                ? p.fvt_c[0][2] * w0[n1] + p.fvt_c[1][2] * w0[i]
                : p.ffl_c[2][2] * w0[n2] + p.ffl_c[1][2] * w0[n1] + p.ffl_c[0][2] * w0[i];
            const double dw2 = p.ffl_c[2][2] * w2[n2] + p.ffl_c[1][2] * w2[n1] + p.ffl_c[0][2] * w2[i];
            wd0[i] = p.coef_d_velo * (dw0 - g3 * p.rc[2] * w0[i]);
            wd2[i] = p.coef_d_velo * (dw2 - g3 * p.rc[2] * w2[i]);
            const double fix = p.t_cmb[j];
            const double d1t = p.ffl_c[2][1] * t[n2] + p.ffl_c[1][1] * t[n1] + p.ffl_c[0][1] * fix;
            const double d2tt = p.ffl_c[2][2] * t[n2] + p.ffl_c[1][2] * t[n1] + p.ffl_c[0][2] * fix;
            td[i] = p.coef_d_temp * (d2tt + 2.0 * p.rc[1] * d1t - g3 * p.rc[2] * fix);
        }
        // magnetic field: from its first point (the centre's neighbour or the ICB)
        if (k > p.kb && k < p.ko) {
            const double d2s = dr2(p, c0, k, i), d1t = dr1(p, c2, k, i);
            bd0[i] = -p.coef_d_magne * c2[i];
            bd1[i] = -p.coef_d_magne * d1t;
            bd2[i] = p.coef_d_magne * (d2s - g3a * c0[i]);
        } else if (k == p.kb && p.icore) {
            // cal_sph_nod_center_diffuse2
            const double d2s = p.d2[1] * b0[i] + p.d2[2] * b0[i + J];
            const double d2t = p.d2[1] * b2[i] + p.d2[2] * b2[i + J];
            bd0[i] = p.coef_d_magne * (d2s - g3 * p.ar2[0] * b0[i]);
            bd2[i] = p.coef_d_magne * (d2t - g3 * p.ar2[0] * b2[i]);
        } else if (k == p.kb) {
            const size_t i1 = i + J, i2 = i1 + J;
            const double d2s = ((g1 + 1.0) * p.ri[1] * p.fdr_i[0][2] + p.fdr_i[1][2]) * b0[i]
                             + p.fdr_i[2][2] * b0[i1];
            const double d2t = p.ffl_i[0][2] * b2[i] + p.ffl_i[1][2] * b2[i1] + p.ffl_i[2][2] * b2[i2];
            bd0[i] = p.coef_d_magne * (d2s - g3 * p.ri[2] * b0[i]);
            bd2[i] = p.coef_d_magne * (d2t - g3 * p.ri[2] * b2[i]);
        } else if (k == p.ko) {
            const size_t n1 = i - J, n2 = n1 - J;
            const double d2s = p.fdr_c[0][2] * b0[n1]
                             + (p.fdr_c[1][2] - g1 * p.rc[1] * p.fdr_c[2][2]) * b0[i];
            const double d2t = p.ffl_c[2][2] * b2[n2] + p.ffl_c[1][2] * b2[n1] + p.ffl_c[0][2] * b2[i];
            bd0[i] = p.coef_d_magne * (d2s - g3 * p.rc[2] * b0[i]);
            bd2[i] = p.coef_d_magne * (d2t - g3 * p.rc[2] * b2[i]);
        }
        if (k >= p.ki && k <= p.ko) FLD(kRotBuoyancy)[i] = p.coef_buo * t[i] * p.r[k];
    }
}

// cal_sph_fix_scalar_in/out_diffuse2 leave the wall values in place.
__global__ void k_temp_bc(MhdDev p)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= p.J) return;
    FLD(kTemp)[j + (size_t)p.J * p.ki] = p.t_icb[j];
    FLD(kTemp)[j + (size_t)p.J * p.ko] = p.t_cmb[j];
}

// ------------------------------------------------------- rj <-> rlm -----
// rj <-> rlm transposes, tiled through LDS: a work-group takes 16 degrees x
// 16 orders of one radial level.  In rj (j = l(l+1)+m) the orders of one
// degree are contiguous; in rlm the degrees of one order are contiguous with
// the node's components contiguous, so both sides move in contiguous runs.
constexpr int kTile = 16;
// components of the step's forward transform: the twelve grid products
constexpr int kNFwd = 12;

// tile origin (l0, m0) of block (blockIdx.x, blockIdx.y); false if the tile
// holds no mode (l < |m| everywhere)
__device__ inline bool tile_origin(int L, int *l0, int *m0)
{
    *m0 = -L + (int)blockIdx.x * kTile;
    *l0 = (int)blockIdx.y * kTile;
    const int mlo = *m0, mhi = min(*m0 + kTile - 1, L);
    const int minabs = (mlo <= 0 && mhi >= 0) ? 0 : min(abs(mlo), abs(mhi));
    return minabs <= min(L, *l0 + kTile - 1);
}

// A work-group walks kz radial levels of its tile: the tile's rlm indices
// are read once into LDS, and the next level's values are loaded into
// registers while the current level is stored.
__global__ __launch_bounds__(256) void k_pack(MhdDev p, int kz)
{
    __shared__ double tile[13][kTile][kTile + 1];
    __shared__ int jls[kTile][kTile];   // rlm index of (degree, order), -1 outside l >= |m|
    int l0, m0;
    if (!tile_origin(p.L, &l0, &m0)) return;
    const int k0 = blockIdx.z * kz, k1 = min(k0 + kz, p.nr);
    const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
    const int l = l0 + ty, m = m0 + tx;
    const bool ok = l <= p.L && m <= p.L && abs(m) <= l;
    const int jrj = ok ? l * (l + 1) + m : 0;
    jls[ty][tx] = ok ? p.rlm_of_rj[jrj] : -1;
    double v[13];
    auto load = [&](int k) {
        const size_t i = (size_t)p.J * k + jrj;
#pragma unroll
        for (int c = 0; c < 12; ++c) v[c] = ok ? FLD(kVelo + c)[i] : 0.0;
        v[12] = ok ? FLD(kTemp)[i] : 0.0;
    };
    load(k0);
    for (int k = k0; k < k1; ++k) {
#pragma unroll
        for (int c = 0; c < 13; ++c) tile[c][ty][tx] = v[c];
        __syncthreads();
        if (k + 1 < k1) load(k + 1);
        for (int e = threadIdx.x; e < kTile * kTile * 13; e += 256) {
            const int col = e / (kTile * 13), r = e % (kTile * 13);
            const int node = r / 13, c = r % 13;
            const int jl = jls[node][col];
            if (jl < 0) continue;
            p.sp_b[13 * ((size_t)jl + (size_t)p.J * k) + c] = tile[c][node][col];
        }
        __syncthreads();
    }
}

// Unpack the forward transform into the rj fields: the step transforms the
// twelve products only (advection 0-2, Lorentz 3-5, induction 6-8, heat
// flux 9-11; the reference's other forward slots are zero, so the poloidal
// Coriolis slot it unpacks is written as zero) through an LDS tile; kz radial
// levels per work-group as k_pack.
__global__ __launch_bounds__(256) void k_unpack(MhdDev p, int kz)
{
    constexpr int NU = kNFwd;
    constexpr int PER = (kTile * kTile * NU + 255) / 256;   // gathered values per thread
    __shared__ double tile[NU][kTile][kTile + 1];
    __shared__ int jls[kTile][kTile];
    int l0, m0;
    if (!tile_origin(p.L, &l0, &m0)) return;
    const int k0 = blockIdx.z * kz, k1 = min(k0 + kz, p.nr);
    const int tx = threadIdx.x % kTile, ty = threadIdx.x / kTile;
    {
        const int l = l0 + ty, m = m0 + tx;
        const bool ok = l <= p.L && m <= p.L && abs(m) <= l;
        jls[ty][tx] = ok ? p.rlm_of_rj[l * (l + 1) + m] : -1;
    }
    __syncthreads();
    // this thread's gathered values: rlm offset of (node, component), or -1
    int off[PER];
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int e = threadIdx.x + q * 256;
        off[q] = -1;
        if (e < kTile * kTile * NU) {
            const int col = e / (kTile * NU), r = e % (kTile * NU);
            const int node = r / NU, u = r % NU;
            const int jl = jls[node][col];
            if (jl >= 0) off[q] = NU * jl + u;
        }
    }
    double g[PER];
    auto load = [&](int k) {
        const size_t base = (size_t)NU * p.J * k;
#pragma unroll
        for (int q = 0; q < PER; ++q) g[q] = off[q] >= 0 ? p.sp_f[base + off[q]] : 0.0;
    };
    const int l = l0 + ty, m = m0 + tx;
    const bool ok = l <= p.L && m <= p.L && abs(m) <= l;
    load(k0);
    for (int k = k0; k < k1; ++k) {
#pragma unroll
        for (int q = 0; q < PER; ++q) {
            const int e = threadIdx.x + q * 256;
            if (e < kTile * kTile * NU) {
                const int col = e / (kTile * NU), r = e % (kTile * NU);
                tile[r % NU][r / NU][col] = g[q];
            }
        }
        __syncthreads();
        if (k + 1 < k1) load(k + 1);
        if (ok) {
            const size_t i = (size_t)p.J * k + l * (l + 1) + m;
#pragma unroll
            for (int c = 0; c < 3; ++c) {
                FLD(kMAdvect + c)[i] = tile[c][ty][tx];
                FLD(kLorentz + c)[i] = tile[3 + c][ty][tx];
                FLD(kVpInduct + c)[i] = tile[6 + c][ty][tx];
                FLD(kHFlux + c)[i] = tile[9 + c][ty][tx];
            }
            FLD(kRotCoriolis + 1)[i] = 0.0;
        }
        __syncthreads();
    }
}

// Curl of the Coriolis force from u and w (sum_rot_coriolis_rlm_10;
// omega = (r^2/2, r, 1)), one thread per rj node.
__global__ void k_coriolis(MhdDev p)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < p.nnod;
         i += (size_t)gridDim.x * blockDim.x) {
        const NodeKJ kj = node_kj(p, i);
        const int k = kj.k, j = kj.j;
        const size_t base = (size_t)p.J * k;
        const size_t k1 = base + p.jk1[j], k2 = base + p.jk2[j], l1 = base + p.jl1[j];
        const double *us = FLD(kVelo), *ud = FLD(kVelo + 1);
        const double *ws = FLD(kVort), *wd = FLD(kVort + 1), *wt = FLD(kVort + 2);
        const double ar = 1.0 / p.r[k], rr = p.r[k];
        const double om0 = 0.5 * rr * rr, om1 = rr, om2 = 1.0;
        const double *C = p.cor;
        const size_t J = p.J;
        const double sw11 = C[0 * J + j], sw21 = C[1 * J + j], sw12 = C[2 * J + j];
        const double sw22 = C[3 * J + j], sw13 = C[4 * J + j], tw13 = C[5 * J + j];
        const double tw23 = C[6 * J + j], tw14 = C[7 * J + j], tw24 = C[8 * J + j];
        const double tw11 = C[9 * J + j], tw12 = C[10 * J + j], g3r = C[11 * J + j];
        const double sp_d2vp_l1 = ar * ar * g3r * us[l1] - wt[l1];
        const double sp_dwp_k1 = wd[k1] - 2.0 * ar * ws[k1];
        const double sp_dwp_k2 = wd[k2] - 2.0 * ar * ws[k2];
        double pol = sw11 * om0 * ud[k1] + sw21 * om0 * ud[k2] + sw13 * om0 * ws[l1]
                   + sw12 * om1 * us[k1] + sw22 * om1 * us[k2];
        double tor = tw11 * om0 * sp_d2vp_l1 - tw12 * om2 * us[l1] + tw13 * om1 * ws[k1]
                   + tw23 * om1 * ws[k2] + tw14 * om0 * sp_dwp_k1 + tw24 * om0 * sp_dwp_k2;
        pol = -p.coef_cor * pol * ar * ar;
        tor = -p.coef_cor * tor * ar * ar;
        if (p.icore && k == p.ki && j >= 1 && j <= 3) {
            // inner_core_rot_z_coriolis_rlm (sum_coriolis_terms_rlm.f90:
            // 316-351): on the core's l = 1 ICB rows only the z rotation's
            // coupling of the x and y turns remains
            const double w11 = j == 1 ? ws[base + 3] : ws[base + 1];
            pol = j == 2 ? 0.0 : 2.0 * p.coef_cor * rr * om0 * w11;
            if (j == 1) pol = -pol;
        }
        FLD(kRotCoriolis)[i] = pol;
        FLD(kRotCoriolis + 2)[i] = tor;
    }
}

// Curl of a momentum force with the velocity walls' rows
// (const_sph_force_rot2, sel_bc_sph_vorticity: cal_sph_nod_icb/cmb_rigid_rot2
// or cal_sph_nod_icb/cmb_free_rot2) -> (pol, tor).
__device__ __forceinline__ void rot_momentum(const MhdDev &p, int src, int k, int j, size_t i,
                                             double *pol, double *tor)
{
    const double *s0 = FLD(src), *s1 = FLD(src + 1), *s2 = FLD(src + 2);
    const double g3 = p.g3[j];
    const size_t J = p.J;
    *pol = s2[i];
    if (k == p.ki && p.free_i) {
        const double d2s = p.fvp_i[1][2] * s0[i] + p.fvp_i[2][2] * s0[i + J];
        *tor = -d2s + g3 * p.ri[2] * s0[i];
    } else if (k == p.ko && p.free_c) {
        const double d2s = p.fvp_c[0][2] * s0[i - J] + p.fvp_c[1][2] * s0[i];
        *tor = -d2s + g3 * p.rc[2] * s0[i];
    } else if (k == p.ki) {
        const double d2s = p.fdr_i[1][2] * s0[i] + p.fdr_i[2][2] * s0[i + J];
        *tor = -(d2s - g3 * p.ri[2] * s0[i]);
    } else if (k == p.ko) {
        const double d2s = p.fdr_c[0][2] * s0[i - J] + p.fdr_c[1][2] * s0[i];
        *tor = -(d2s - g3 * p.rc[2] * s0[i]);
    } else {
        *tor = -dr1(p, s1, k, i) + g3 * p.ar2[k] * s0[i];
    }
}

// int_icore_tor_lorentz_l1 (cal_inner_core_rotation.f90:280-314): the
// Lorentz torque on the core, 5/2 r_i^-3 times the trapezoid integral of the
// l = 1 toroidal Lorentz force times r^2 from the centre (first interval
// f(r_1) r_1^3) to the ICB, into the ICB row of the toroidal force (whose
// wall curl then carries it).  One thread per l = 1 mode, the radial sum in
// order.
__global__ void k_icore_torque(MhdDev p)
{
    const int j = threadIdx.x + 1;
    if (j > 3) return;
    double *lor = FLD(kLorentz + 2);
    const size_t J = p.J;
    const double *r = p.r;
    double sk = lor[j] * (r[0] * r[0] * r[0]);
    for (int k = 0; k < p.ki; ++k)
        sk = sk + (lor[j + J * k] * (r[k] * r[k]) + lor[j + J * (k + 1)] * (r[k + 1] * r[k + 1]))
                * (r[k + 1] - r[k]);
    const double ar = 1.0 / r[p.ki];
    lor[j + J * p.ki] = 0.5 * 5.0 * sk * (ar * ar * ar);
}

__global__ void k_rot_forces(MhdDev p, int temp_roundtrip)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < p.nnod;
         i += (size_t)gridDim.x * blockDim.x) {
        const NodeKJ kj = node_kj(p, i);
        const int k = kj.k, j = kj.j;
        const size_t J = p.J;
        const double g1 = p.g1[j], g3 = p.g3[j];
        // curl of u x B with its rows at the field's ends (const_sph_rotation_uxb:
        // insulating walls, or cal_sph_nod_center_rot2 above the centre)
        if (k >= p.kb && k <= p.ko) {
            const double *v0 = FLD(kVpInduct), *v1 = FLD(kVpInduct + 1), *v2 = FLD(kVpInduct + 2);
            double in0 = v2[i], in1, in2;
            if (k == p.kb && p.icore) {
                in1 = p.d1[1] * v2[i] + p.d1[2] * v2[i + J];
                const double d2s = p.d2[1] * v0[i] + p.d2[2] * v0[i + J];
                in2 = -(d2s - g3 * p.ar2[0] * v0[i]);
            } else if (k == p.kb) {
                in1 = (g1 + 1.0) * p.ri[1] * v2[i];
                in2 = 0.0;
            } else if (k == p.ko) {
                in1 = -g1 * p.rc[1] * v2[i];
                in2 = 0.0;
            } else {
                in1 = dr1(p, v2, k, i);
                in2 = -dr1(p, v1, k, i) + g3 * p.ar2[k] * v0[i];
            }
            FLD(kInduction)[i] = in0;
            FLD(kInduction + 1)[i] = in1;
            FLD(kInduction + 2)[i] = in2;
        }
        // trans_per_temp_to_temp_sph then sync_temp_by_per_temp_sph
        if (temp_roundtrip && j == 0) {
            double *t = FLD(kTemp);
            t[i] = t[i] + p.reft0[k];
            t[i] = t[i] - p.reft0[k];
        }
        if (k < p.ki || k > p.ko) {
            // set_MHD_terms_to_force over every node: outside the shell only
            // the Coriolis term is nonzero
            FLD(kForces)[i] = -FLD(kRotInertia)[i] + FLD(kRotCoriolis)[i] + FLD(kRotLorentz)[i];
            FLD(kForces + 2)[i] = -FLD(kRotInertia + 2)[i] + FLD(kRotCoriolis + 2)[i]
                                + FLD(kRotLorentz + 2)[i] + FLD(kRotBuoyancy)[i];
            continue;
        }
        double ri_p, ri_t, rl_p, rl_t;
        rot_momentum(p, kMAdvect, k, j, i, &ri_p, &ri_t);
        rot_momentum(p, kLorentz, k, j, i, &rl_p, &rl_t);
        FLD(kRotInertia)[i] = ri_p;
        FLD(kRotInertia + 2)[i] = ri_t;
        FLD(kRotLorentz)[i] = rl_p;
        FLD(kRotLorentz + 2)[i] = rl_t;
        // divergence of the heat flux (const_sph_scalar_advect)
        const double *h0 = FLD(kHFlux), *h1 = FLD(kHFlux + 1);
        const double gg = fmax(g3, 0.5);
        double ha;
        if (k == p.ki) {
            const double d1s = p.ffl_i[0][1] * p.t_icb[j] + p.ffl_i[1][1] * h0[i + J]
                             + p.ffl_i[2][1] * h0[i + 2 * J];
            ha = (d1s - h1[i]) * gg * p.ri[2];
        } else if (k == p.ko) {
            const double d1s = p.ffl_c[2][1] * h0[i - 2 * J] + p.ffl_c[1][1] * h0[i - J]
                             + p.ffl_c[0][1] * p.t_cmb[j];
            ha = (d1s - h1[i]) * gg * p.rc[2];
        } else {
            ha = (dr1(p, h0, k, i) - h1[i]) * gg * p.ar2[k];
        }
        // add_reftemp_advect_sph_MHD
        ha = ha + p.coef_temp * g3 * p.ar2[k] * p.reft1[k] * FLD(kVelo)[i];
        FLD(kHAdvect)[i] = ha;
        // set_MHD_terms_to_force
        FLD(kForces)[i] = -ri_p + FLD(kRotCoriolis)[i] + rl_p;
        FLD(kForces + 2)[i] = -ri_t + FLD(kRotCoriolis + 2)[i] + rl_t + FLD(kRotBuoyancy)[i];
    }
}

__global__ void k_temp_full(MhdDev p, double *out)
{
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < p.nnod;
         i += (size_t)gridDim.x * blockDim.x) {
        const NodeKJ kj = node_kj(p, i);
        const int k = kj.k, j = kj.j;
        double t = FLD(kTemp)[i];
        if (j == 0) t = t + p.reft0[k];
        out[i] = t;
    }
}

#undef FLD

// ------------------------------------------------------------ host ------

namespace {

double det22(double a11, double a12, double a21, double a22) { return a11 * a22 - a21 * a12; }

// cal_inverse_33_matrix: first-row cofactor determinant, adjugate / det.
void inverse33(const double a[3][3], double inv[3][3])
{
    double det = 0.0;
    for (int k = 0; k < 3; ++k) {
        int c0 = k == 0 ? 1 : 0, c1 = k == 2 ? 1 : 2;
        const double d = det22(a[1][c0], a[1][c1], a[2][c0], a[2][c1]);
        det = det + ((k % 2) ? -1.0 : 1.0) * d * a[0][k];
    }
    const double a_det = 1.0 / det;
    for (int kj = 0; kj < 3; ++kj)
        for (int ki = 0; ki < 3; ++ki) {
            int r0 = ki == 0 ? 1 : 0, r1 = ki == 2 ? 1 : 2;
            int c0 = kj == 0 ? 1 : 0, c1 = kj == 2 ? 1 : 2;
            const double d = det22(a[r0][c0], a[r0][c1], a[r1][c0], a[r1][c1]);
            inv[kj][ki] = (((ki + kj) % 2) ? -1.0 : 1.0) * d * a_det;
        }
}

int isqrt_int(int j)
{
    int l = (int)std::sqrt((double)j);
    while ((l + 1) * (l + 1) <= j) ++l;
    while (l * l > j) --l;
    return l;
}

template <typename T>
int upload(T **d, const std::vector<T> &h, std::string *err)
{
    HIP_TRY(hipMalloc(d, sizeof(T) * std::max<size_t>(h.size(), 1)));
    if (!h.empty()) HIP_TRY(hipMemcpy(*d, h.data(), sizeof(T) * h.size(), hipMemcpyHostToDevice));
    return 0;
}

int blocks_for(size_t n) { return (int)std::min<size_t>((n + 255) / 256, 8192); }

}  // namespace

double gaunt_p1(int l, int m)
{
    // leadki(1,0, l-1,m, l,m) = 4 sqrt(l^2 - m^2) / ((2l+1)(2l-1)) (Schmidt
    // harmonics; the recurrence x P_l^m = (sqrt(l^2-m^2) P_{l-1}^m +
    // sqrt((l+1)^2-m^2) P_{l+1}^m) / (2l+1) and int (P_l^m)^2 = 2(2-d_m0)/(2l+1)).
    if (l < 1 || std::abs(m) > l - 1) return 0.0;
    return 4.0 * std::sqrt((double)(l * l - m * m)) / ((double)(2 * l + 1) * (double)(2 * l - 1));
}

double elsasser_p1(int l, int m)
{
    // leadli(1,0, l,-m, l,m) = -4 m / (2l+1): the phi integral leaves
    // -m pi int cos(t) d(P^2)/dt dt = -m pi int (P_l^m)^2 dx.
    if (l < 1) return 0.0;
    return -4.0 * (double)m / (double)(2 * l + 1);
}

// Coupling table of the curl of the Coriolis force (init_sum_coriolis_rlm:
// cal_gaunt_coriolis_rlm and interact_rot_coriolis_rlm,
// interact_coriolis_rlm.f90), indexed by rj mode: cor[q*J + j] for
// q = sw11 sw21 sw12 sw22 sw13 tw13 tw23 tw14 tw24 tw11 tw12 g3; the
// neighbours (l-1,m), (l+1,m), (l,-m) as rj indices.
void coriolis_table(int L, double *cor, int *jk1, int *jk2, int *jl1)
{
    const int J = (L + 1) * (L + 1);
    for (int j = 0; j < J; ++j) {
        const int l3 = isqrt_int(j), m3 = j - l3 * (l3 + 1);
        double gi1 = 0.0, gi2 = 0.0, ei = 0.0;
        int k1l = l3, k1m = m3, k2l = l3, k2m = m3, e1l = l3, e1m = m3;
        if (l3 - 1 >= 1 && std::abs(m3) <= l3 - 1) {
            gi1 = gaunt_p1(l3, m3);
            k1l = l3 - 1;
        }
        if (l3 + 1 <= L) {
            gi2 = gaunt_p1(l3 + 1, m3);
            k2l = l3 + 1;
        }
        if (l3 >= 1) {
            ei = elsasser_p1(l3, m3);
            e1m = -m3;
        }
        if (j == 0) {
            k1l = k1m = k2l = k2m = e1l = e1m = 0;
            gi2 = 0.0;
            ei = 0.0;
        }
        jk1[j] = k1l * (k1l + 1) + k1m;
        jk2[j] = k2l * (k2l + 1) + k2m;
        jl1[j] = e1l * (e1l + 1) + e1m;
        const double G3 = (j == 0) ? 0.5 : (double)(l3 * (l3 + 1));
        const double G4 = (double)((l3 - 1) * l3), G5 = (double)((l3 + 1) * (l3 + 2));
        const double G17 = (j == 0) ? 0.0 : (double)(2 * l3 + 1) / (double)(4 * l3 * (l3 + 1));
        double *C = cor;
        C[0 * J + j] = (2.0 - G4 - G3) * gi1 * G17;
        C[1 * J + j] = (2.0 - G5 - G3) * gi2 * G17;
        C[2 * J + j] = G4 * (2.0 - G4 + G3) * gi1 * G17 * 0.5;
        C[3 * J + j] = G5 * (2.0 - G5 + G3) * gi2 * G17 * 0.5;
        C[4 * J + j] = 2.0 * ei * G17;
        C[5 * J + j] = -(G3 * (2.0 + G4 - G3) + 2.0 * (-2.0 + G4 + G3)) * gi1 * G17 * 0.5;
        C[6 * J + j] = -(G3 * (2.0 + G5 - G3) + 2.0 * (-2.0 + G5 + G3)) * gi2 * G17 * 0.5;
        C[7 * J + j] = -(-2.0 + G4 + G3) * gi1 * G17;
        C[8 * J + j] = -(-2.0 + G5 + G3) * gi2 * G17;
        C[9 * J + j] = -2.0 * ei * G17;
        C[10 * J + j] = -G3 * ei * G17;
        C[11 * J + j] = G3;
    }
}

// Radial finite differences of the grid r[0..nr-1]: the three-point
// interior weights (set_dr_for_nonequi, nod_r_2nd_fdm_coefs_nonequi) and the
// one-sided wall formulas (cal_fdm_coefs_4_boundaries.f90).
void radial_fdm(const double *r, int nr, RadialFdm *out) { radial_fdm(r, nr, 0, nr - 1, out); }

// set_dr_for_nonequi, nod_r_2nd_fdm_coefs_nonequi
// (set_radius_func_noequi.f90:173-197: with points below the ICB the first
// point's lower neighbour is the centre), the wall formulas of
// cal_fdm_coefs_4_BCs at the points ki and ko, cal_2nd_to_center_fixed_fdm
// (m_coef_fdm_to_center.f90:73-109), cal_fdm1_coef_fix_fld_ICB.
void radial_fdm(const double *r, int nr, int ki, int ko, RadialFdm *out)
{
    RadialFdm &f = *out;
    std::vector<double> dr(nr);
    f.d1.assign(3 * nr, 0.0);
    f.d2.assign(3 * nr, 0.0);
    for (int k = 1; k < nr - 1; ++k) dr[k] = r[k + 1] - r[k];
    dr[nr - 1] = r[nr - 1] - r[nr - 2];
    dr[0] = r[1] - r[0];
    for (int k = 0; k < nr; ++k) {
        double dp, dn;
        if (k == 0) { dp = dr[0]; dn = ki > 0 ? r[0] : dr[0]; }
        else if (k == nr - 1) { dp = dr[k]; dn = dr[k]; }
        else { dp = dr[k]; dn = dr[k - 1]; }
        const double a[3][3] = {{1.0, 0.0, 0.0}, {1.0, dp, 0.5 * dp * dp}, {1.0, -dn, 0.5 * dn * dn}};
        double m[3][3];
        inverse33(a, m);
        f.d1[3 * k + 0] = m[1][2]; f.d1[3 * k + 1] = m[1][0]; f.d1[3 * k + 2] = m[1][1];
        f.d2[3 * k + 0] = m[2][2]; f.d2[3 * k + 1] = m[2][0]; f.d2[3 * k + 2] = m[2][1];
    }
    const double dp1 = r[ki + 1] - r[ki], dp2 = r[ki + 2] - r[ki];
    const double a1[3][3] = {{1, 0, 0}, {1, dp1, dp1 * dp1 / 2.0}, {1, dp2, dp2 * dp2 / 2.0}};
    double m[3][3];
    inverse33(a1, m);
    for (int i = 0; i < 3; ++i)
        for (int q = 0; q < 3; ++q) f.ffl_i[i][q] = m[q][i];
    const double a2[3][3] = {{1, 0, 0}, {0, 1, 0}, {dp1, 1, dp1 * dp1 / 2.0}};
    inverse33(a2, m);
    for (int i = 0; i < 3; ++i)
        for (int q = 0; q < 3; ++q) f.fdr_i[i][q] = m[q][i];
    const double dn1 = r[ko] - r[ko - 1], dn2 = r[ko] - r[ko - 2];
    const double a3[3][3] = {{1, 0, 0}, {1, -dn1, dn1 * dn1 / 2.0}, {1, -dn2, dn2 * dn2 / 2.0}};
    inverse33(a3, m);
    for (int i = 0; i < 3; ++i)
        for (int q = 0; q < 3; ++q) f.ffl_c[i][q] = m[q][i];
    const double a4[3][3] = {{1, 0, 0}, {0, 1, 0}, {1, -dn1, dn1 * dn1 / 2.0}};
    inverse33(a4, m);
    for (int q = 0; q < 3; ++q) {
        f.fdr_c[0][q] = m[q][2];
        f.fdr_c[1][q] = m[q][0];
        f.fdr_c[2][q] = m[q][1];
    }
    f.ri[0] = r[ki]; f.ri[1] = 1.0 / r[ki]; f.ri[2] = f.ri[1] * f.ri[1];
    f.rc[0] = r[ko]; f.rc[1] = 1.0 / r[ko]; f.rc[2] = f.rc[1] * f.rc[1];
    // free-slip walls (m_coef_fdm_free_ICB.f90:85-173,
    // m_coef_fdm_free_CMB.f90:85-174): the middle Taylor row is the wall
    // condition, f - r f' + r^2/2 f'' = 0 (poloidal, with f = 0 at the wall)
    // or 2 f - r f' = 0 (toroidal); the last row reaches the neighbour.
    std::memset(f.fvp_i, 0, sizeof f.fvp_i);
    std::memset(f.fvt_i, 0, sizeof f.fvt_i);
    std::memset(f.fvp_c, 0, sizeof f.fvp_c);
    std::memset(f.fvt_c, 0, sizeof f.fvt_c);
    for (int wall = 0; wall < 2; ++wall) {
        const double r0 = wall == 0 ? r[ki] : r[ko];
        const double h = wall == 0 ? dp1 : -dn1;           // neighbour offset
        const double ap[3][3] = {{1, 0, 0}, {1, -r0, 0.5 * r0 * r0}, {1, h, 0.5 * h * h}};
        const double at[3][3] = {{1, 0, 0}, {2, -r0, 0}, {1, h, 0.5 * h * h}};
        double mp[3][3], mt[3][3];
        inverse33(ap, mp);
        inverse33(at, mt);
        double (*vp)[3] = wall == 0 ? f.fvp_i : f.fvp_c;
        double (*vt)[3] = wall == 0 ? f.fvt_i : f.fvt_c;
        const int nb = wall == 0 ? 2 : 0;                  // neighbour's row
        vp[1][0] = 1.0;
        vp[1][1] = mp[1][0];
        vp[1][2] = mp[2][0];
        vp[nb][1] = mp[1][2];
        vp[nb][2] = mp[2][2];
        vt[1][0] = 1.0;
        vt[1][1] = mt[1][0];
        vt[1][2] = mt[2][0];
        vt[nb][2] = mt[2][2];
        if (wall == 1) vt[2][0] = vt[2][1] = vt[2][2] = 1.0;   // unused, as the reference fills it
    }
    {   // the centre below the first point (value there zero)
        const double dn = r[0], dp = r[1] - r[0];
        const double ac[3][3] = {{1, 0, 0}, {1, -dn, dn * dn / 2.0}, {1, dp, dp * dp / 2.0}};
        double mc[3][3];
        inverse33(ac, mc);
        for (int q = 0; q < 3; ++q) {
            f.ctr1[0][q] = mc[q][1];
            f.ctr1[1][q] = mc[q][0];
            f.ctr1[2][q] = mc[q][2];
        }
    }
    f.fdm1_i[0] = -1.0 / dp1;
    f.fdm1_i[1] = 1.0 / dp1;
}

// The boundary model a context supports (checked before any device work):
// rigid or free-slip walls with insulating magnetic walls on a grid that ends
// at the ICB and CMB, or a rotatable conducting inner core below the ICB with
// the magnetic field solved to the centre, optionally with points above the
// CMB.
int mhd_check_model(int nr, const double *radius, const MhdCoefs &co, std::string *err)
{
    if (co.bc_velo_cmb != kVeloRigid && co.bc_velo_cmb != kVeloFreeSlip) {
        *err = "mhd_create: CMB velocity boundary kind " + std::to_string(co.bc_velo_cmb)
             + " is not supported (0 rigid, 1 free slip)";
        return -1;
    }
    if (co.bc_velo_icb != kVeloRigid && co.bc_velo_icb != kVeloFreeSlip &&
        co.bc_velo_icb != kVeloRotInnerCore) {
        *err = "mhd_create: ICB velocity boundary kind " + std::to_string(co.bc_velo_icb)
             + " is not supported (0 rigid, 1 free slip, 2 rotatable inner core)";
        return -1;
    }
    if (co.bc_magne_icb != kMagneInsulator && co.bc_magne_icb != kMagneToCenter) {
        *err = "mhd_create: magnetic boundary kind " + std::to_string(co.bc_magne_icb)
             + " is not supported (0 insulating ICB, 1 field to the centre)";
        return -1;
    }
    const int ki = co.kr_icb, ko = co.kr_cmb < 0 ? nr - 1 : co.kr_cmb;
    if (ki < 0 || ko > nr - 1 || ko - ki < 4) {
        *err = "mhd_create: ICB / CMB points out of the grid or fewer than 5 points between";
        return -1;
    }
    // the rotatable inner core and the field to the centre come together, as
    // in Dynamobench case 2; points outside the shell only with them
    const bool icore = co.bc_velo_icb == kVeloRotInnerCore;
    if (icore != (co.bc_magne_icb == kMagneToCenter) || icore != (ki > 0) ||
        (!icore && ko != nr - 1)) {
        *err = "mhd_create: points below the ICB need the rotatable inner core with the "
               "magnetic field solved to the centre (and these need such points); points "
               "above the CMB are supported only with them";
        return -1;
    }
    if (icore && !(radius[0] > 0.0)) {
        *err = "mhd_create: the first point must lie above the centre";
        return -1;
    }
    return 0;
}

int mhd_create(MhdContext *c, int L, int nth, int nphi, int nr, const double *radius,
               const MhdCoefs &co, int leg_mode, hipStream_t s, std::string *err)
{
    if (L < 2 || nth < L + 1 || nphi < 2 * L + 2 || nr < 5) {
        *err = "mhd_create: need L >= 2, nth >= L+1, nphi >= 2L+2, nr >= 5";
        return -1;
    }
    if ((long long)nr * (L + 1) * (L + 1) >= (1LL << 31)) {
        *err = "mhd_create: nr (L+1)^2 must stay below 2^31 spectral nodes";
        return -1;
    }
    if (mhd_check_model(nr, radius, co, err)) return -1;
    const int ki = co.kr_icb, ko = co.kr_cmb < 0 ? nr - 1 : co.kr_cmb;
    const bool icore = co.bc_velo_icb == kVeloRotInnerCore;
    c->L = L;
    c->nth = nth;
    c->nphi = nphi;
    c->nr = nr;
    c->J = (L + 1) * (L + 1);
    c->M = 2 * L + 1;
    c->co = co;
    c->leg_mode = leg_mode;
    c->stream = s;
    c->radius.assign(radius, radius + nr);
    const int J = c->J;
    const size_t nnod = (size_t)nr * J;
    const int kb = icore ? 0 : ki;   // the magnetic field's first point
    std::vector<double> r(radius, radius + nr), ar1(nr), ar2(nr);
    for (int k = 0; k < nr; ++k) {
        ar1[k] = 1.0 / r[k];
        ar2[k] = ar1[k] * ar1[k];
    }
    RadialFdm fd;
    radial_fdm(r.data(), nr, ki, ko, &fd);
    const std::vector<double> &d1 = fd.d1, &d2 = fd.d2;
    MhdDev &p = c->p;
    p = MhdDev();
    p.nr = nr;
    p.J = J;
    p.L = L;
    p.ki = ki;
    p.ko = ko;
    p.kb = kb;
    p.icore = icore ? 1 : 0;
    p.fdm1_i[0] = fd.fdm1_i[0];
    p.fdm1_i[1] = fd.fdm1_i[1];
    p.nnod = nnod;
    p.rJ = 1.0 / (double)J;
    std::memcpy(p.ffl_i, fd.ffl_i, sizeof p.ffl_i);
    std::memcpy(p.fdr_i, fd.fdr_i, sizeof p.fdr_i);
    std::memcpy(p.ffl_c, fd.ffl_c, sizeof p.ffl_c);
    std::memcpy(p.fdr_c, fd.fdr_c, sizeof p.fdr_c);
    std::memcpy(p.fvp_i, fd.fvp_i, sizeof p.fvp_i);
    std::memcpy(p.fvt_i, fd.fvt_i, sizeof p.fvt_i);
    std::memcpy(p.fvp_c, fd.fvp_c, sizeof p.fvp_c);
    std::memcpy(p.fvt_c, fd.fvt_c, sizeof p.fvt_c);
    p.free_i = co.bc_velo_icb == kVeloFreeSlip;
    p.free_c = co.bc_velo_cmb == kVeloFreeSlip;
    std::memcpy(p.ri, fd.ri, sizeof p.ri);
    std::memcpy(p.rc, fd.rc, sizeof p.rc);
    p.dt = co.dt;
    p.coef_exp_v = 1.0 - co.coef_imp_v;
    p.coef_exp_t = 1.0 - co.coef_imp_t;
    p.coef_exp_b = 1.0 - co.coef_imp_b;
    p.coef_d_velo = co.coef_d_velo;
    p.coef_d_magne = co.coef_d_magne;
    p.coef_d_temp = co.coef_d_temp;
    p.coef_buo = co.coef_buo;
    p.coef_temp = co.coef_temp;
    p.coef_cor = co.coef_cor;
    // modes
    std::vector<int> l_of_j(J), rj_of_rlm(J), rlm_of_rj(J);
    std::vector<double> g1(J), g3(J);
    for (int j = 0; j < J; ++j) {
        const int l = isqrt_int(j);
        l_of_j[j] = l;
        g1[j] = (double)l;
        g3[j] = (double)(l * (l + 1));
    }
    {
        int jr = 0;
        for (int m = -L; m <= L; ++m)
            for (int l = std::abs(m); l <= L; ++l, ++jr) {
                rj_of_rlm[jr] = l * (l + 1) + m;
                rlm_of_rj[l * (l + 1) + m] = jr;
            }
    }
    // reference temperature and wall values (set_ref_temp_sph_mhd,
    // adjust_sph_temp_bc_by_reftemp)
    std::vector<double> reft0(nr), reft1(nr), t_icb(J, 0.0), t_cmb(J, 0.0);
    {
        const double dh = co.depth_high_t, dl = co.depth_low_t, hi = co.high_temp, lo = co.low_temp;
        for (int k = 0; k < nr; ++k) {
            if (k < ki || k > ko) {   // the high temperature below, the low one above
                reft0[k] = k < ki ? hi : lo;
                reft1[k] = 0.0;
                continue;
            }
            reft0[k] = (dh * dl * ar1[k] * (hi - lo) - dh * hi + dl * lo) / (dl - dh);
            reft1[k] = -dh * dl * ar2[k] * (hi - lo) / (dl - dh);
        }
        t_icb[0] = hi - reft0[ki];
        t_cmb[0] = lo - reft0[ko];
    }
    c->reft0 = reft0;
    std::vector<int> jk1(J), jk2(J), jl1(J);
    std::vector<double> cor(12 * (size_t)J, 0.0);
    coriolis_table(L, cor.data(), jk1.data(), jk2.data(), jl1.data());
    // constants to the device (one allocation each for doubles and ints)
    std::vector<double> dc;
    auto push = [&dc](const std::vector<double> &v) {
        size_t off = dc.size();
        dc.insert(dc.end(), v.begin(), v.end());
        return off;
    };
    const size_t o_r = push(r), o_ar1 = push(ar1), o_ar2 = push(ar2), o_d1 = push(d1),
                 o_d2 = push(d2), o_g1 = push(g1), o_g3 = push(g3), o_ti = push(t_icb),
                 o_tc = push(t_cmb), o_r0 = push(reft0), o_r1 = push(reft1), o_cor = push(cor);
    std::vector<int> ic;
    auto pushi = [&ic](const std::vector<int> &v) {
        size_t off = ic.size();
        ic.insert(ic.end(), v.begin(), v.end());
        return off;
    };
    const size_t o_k1 = pushi(jk1), o_k2 = pushi(jk2), o_l1 = pushi(jl1), o_rr = pushi(rj_of_rlm),
                 o_rj = pushi(rlm_of_rj), o_lj = pushi(l_of_j);
    if (upload(&c->d_consts, dc, err) || upload(&c->d_iconsts, ic, err)) return -1;
    if (upload(&c->d_radius, r, err)) return -1;
    p.r = c->d_consts + o_r;
    p.ar1 = c->d_consts + o_ar1;
    p.ar2 = c->d_consts + o_ar2;
    p.d1 = c->d_consts + o_d1;
    p.d2 = c->d_consts + o_d2;
    p.g1 = c->d_consts + o_g1;
    p.g3 = c->d_consts + o_g3;
    p.t_icb = c->d_consts + o_ti;
    p.t_cmb = c->d_consts + o_tc;
    p.reft0 = c->d_consts + o_r0;
    p.reft1 = c->d_consts + o_r1;
    p.cor = c->d_consts + o_cor;
    p.jk1 = c->d_iconsts + o_k1;
    p.jk2 = c->d_iconsts + o_k2;
    p.jl1 = c->d_iconsts + o_l1;
    p.rj_of_rlm = c->d_iconsts + o_rr;
    p.rlm_of_rj = c->d_iconsts + o_rj;
    c->d_l_of_j = c->d_iconsts + o_lj;
    // state and work buffers
    const size_t nrtm = (size_t)nth * nr * c->M;
    HIP_TRY(hipMalloc(&c->d_rj, sizeof(double) * kMhdNumComp * nnod));
    HIP_TRY(hipMemset(c->d_rj, 0, sizeof(double) * kMhdNumComp * nnod));
    HIP_TRY(hipMalloc(&c->d_sp_b, sizeof(double) * 13 * nnod));
    HIP_TRY(hipMalloc(&c->d_sp_f, sizeof(double) * kNFwd * nnod));
    HIP_TRY(hipMalloc(&c->d_vr_b, sizeof(double) * 13 * nrtm));
    HIP_TRY(hipMalloc(&c->d_vr_f, sizeof(double) * kNFwd * nrtm));
    HIP_TRY(hipMalloc(&c->d_scratch, sizeof(double) * (nnod + 3 * (size_t)J)));
    HIP_TRY(hipMalloc(&c->d_ms_out, sizeof(double) * 4));
    p.d = c->d_rj;
    p.sp_b = c->d_sp_b;
    p.sp_f = c->d_sp_f;
    // Crank-Nicolson matrices (const_radial_mat_vort_2step,
    // const_radial_mat_4_temp_sph, const_radial_mat_4_magne_sph), band storage
    // m[(j*nr + k)*3 + b] = Fortran mat3(b+1, k+1, j+1)
    auto M3 = [&](std::vector<double> &m, int j, int k, int b) -> double & {
        return m[((size_t)j * nr + k) * 3 + b];
    };
    std::vector<double> vt(3 * nnod, 0.0), wt(3 * nnod, 0.0), vs(3 * nnod, 0.0),
        tm(3 * nnod, 0.0), bs(3 * nnod, 0.0), bt(3 * nnod, 0.0);
    for (int j = 0; j < J; ++j)
        for (int k = 0; k < nr; ++k) {
            M3(vt, j, k, 1) = 1.0;
            M3(wt, j, k, 1) = 1.0;
            M3(tm, j, k, 1) = 1.0;
            M3(bs, j, k, 1) = 1.0;
            M3(bt, j, k, 1) = 1.0;
            M3(vs, j, k, 1) = (k < ki || k > ko) ? 1.0 : 0.0;
        }
    auto add_vector_poisson = [&](std::vector<double> &m, double coef, int k_lo) {
        for (int k = k_lo + 1; k < ko; ++k)
            for (int j = 0; j < J; ++j) {
                M3(m, j, k - 1, 2) = M3(m, j, k - 1, 2) - coef * d2[3 * k + 0];
                M3(m, j, k, 1) = M3(m, j, k, 1) - coef * (d2[3 * k + 1] - g3[j] * ar2[k]);
                M3(m, j, k + 1, 0) = M3(m, j, k + 1, 0) - coef * d2[3 * k + 2];
            }
    };
    auto add_scalar_poisson = [&](std::vector<double> &m, double coef) {
        for (int k = ki + 1; k < ko; ++k)
            for (int j = 0; j < J; ++j) {
                M3(m, j, k - 1, 2) = M3(m, j, k - 1, 2)
                                   - coef * (d2[3 * k + 0] + 2.0 * ar1[k] * d1[3 * k + 0]);
                M3(m, j, k, 1) = M3(m, j, k, 1)
                               - coef * (d2[3 * k + 1] + 2.0 * ar1[k] * d1[3 * k + 1] - g3[j] * ar2[k]);
                M3(m, j, k + 1, 0) = M3(m, j, k + 1, 0)
                                   - coef * (d2[3 * k + 2] + 2.0 * ar1[k] * d1[3 * k + 2]);
            }
    };
    // add_fix_flux_icb/cmb_poisson_mat with a wall table (fixed gradient or
    // free slip)
    auto add_fix_flux_icb = [&](std::vector<double> &m, double coef, const double (*fdm)[3]) {
        for (int j = 0; j < J; ++j) {
            M3(m, j, ki, 1) = M3(m, j, ki, 1) - coef * (fdm[1][2] - g3[j] * p.ri[2]);
            M3(m, j, ki + 1, 0) = M3(m, j, ki + 1, 0) - coef * fdm[2][2];
        }
    };
    auto add_fix_flux_cmb = [&](std::vector<double> &m, double coef, const double (*fdm)[3]) {
        for (int j = 0; j < J; ++j) {
            M3(m, j, ko - 1, 2) = M3(m, j, ko - 1, 2) - coef * fdm[0][2];
            M3(m, j, ko, 1) = M3(m, j, ko, 1) - coef * (fdm[1][2] - g3[j] * p.rc[2]);
        }
    };
    auto set_fix_fld_icb = [&](std::vector<double> &m) {
        for (int j = 0; j < J; ++j) {
            M3(m, j, ki, 1) = 1.0;
            M3(m, j, ki + 1, 0) = 0.0;
        }
    };
    auto set_fix_fld_cmb = [&](std::vector<double> &m) {
        for (int j = 0; j < J; ++j) {
            M3(m, j, ko - 1, 2) = 0.0;
            M3(m, j, ko, 1) = 1.0;
        }
    };
    const double coef_dvt = co.coef_imp_v * co.coef_d_velo * co.dt;
    add_vector_poisson(vt, coef_dvt, ki);
    add_vector_poisson(wt, coef_dvt, ki);
    add_vector_poisson(vs, 1.0, ki);
    // walls (const_radial_mat_vort_2step, const_r_mat_4_vector_sph.f90:85-135)
    add_fix_flux_icb(wt, coef_dvt, p.fdr_i);
    if (p.free_i) {
        add_fix_flux_icb(vt, coef_dvt, p.fvt_i);
        add_fix_flux_icb(vs, 1.0, p.fvp_i);
    } else {
        set_fix_fld_icb(vt);
        add_fix_flux_icb(vs, 1.0, p.fdr_i);
    }
    add_fix_flux_cmb(wt, coef_dvt, p.fdr_c);
    if (p.free_c) {
        add_fix_flux_cmb(vt, coef_dvt, p.fvt_c);
        add_fix_flux_cmb(vs, 1.0, p.fvp_c);
    } else {
        set_fix_fld_cmb(vt);
        add_fix_flux_cmb(vs, 1.0, p.fdr_c);
    }
    if (icore) {
        // set_rotate_icb_vt_sph_mat (cal_inner_core_rotation.f90:185-208): the
        // l = 1 toroidal ICB rows carry the viscous torque on the core
        const double ar = 1.0 / r[ki];
        for (int j = 1; j <= 3; ++j) {
            M3(vt, j, ki, 1) = 1.0 - co.coef_imp_v * co.dt * co.coef_d_velo * 5.0
                                   * (fd.fdm1_i[0] - 2.0 * ar1[ki]) * ar;
            M3(vt, j, ki + 1, 0) = -(co.coef_imp_v * co.dt * co.coef_d_velo * 5.0 * ar
                                     * fd.fdm1_i[1]);
        }
    }
    // poloidal velocity: 5-band product wt * vs (cal_mat_product_3band_mul),
    // the unit matrix outside the shell
    std::vector<double> vp(5 * nnod, 0.0);
    {
        auto P5 = [&](int j, int k, int b) -> double & { return vp[((size_t)j * nr + k) * 5 + b]; };
        for (int j = 0; j < J; ++j)
            for (int k = 0; k < nr; ++k) P5(j, k, 2) = 1.0;
        auto Lm = [&](int j, int k, int b) { return M3(wt, j, k, b); };
        auto Rm = [&](int j, int k, int b) { return M3(vs, j, k, b); };
        auto rows = [&](int j, int k, bool skip5, bool skip1) {
            if (!skip5) P5(j, k - 2, 4) = Lm(j, k - 1, 2) * Rm(j, k - 2, 2);
            P5(j, k - 1, 3) = Lm(j, k - 1, 2) * Rm(j, k - 1, 1) + Lm(j, k, 1) * Rm(j, k - 1, 2);
            P5(j, k, 2) = Lm(j, k - 1, 2) * Rm(j, k, 0) + Lm(j, k, 1) * Rm(j, k, 1)
                        + Lm(j, k + 1, 0) * Rm(j, k, 2);
            P5(j, k + 1, 1) = Lm(j, k, 1) * Rm(j, k + 1, 0) + Lm(j, k + 1, 0) * Rm(j, k + 1, 1);
            if (!skip1) P5(j, k + 2, 0) = Lm(j, k + 1, 0) * Rm(j, k + 2, 0);
        };
        for (int k = ki + 2; k <= ko - 2; ++k)
            for (int j = 0; j < J; ++j) rows(j, k, false, false);
        for (int j = 0; j < J; ++j) {
            P5(j, ki, 2) = 1.0;
            P5(j, ki + 1, 1) = 0.0;
            P5(j, ki + 2, 0) = 0.0;
            rows(j, ki + 1, true, false);
            rows(j, ko - 1, false, true);
            P5(j, ko - 2, 4) = 0.0;
            P5(j, ko - 1, 3) = 0.0;
            P5(j, ko, 2) = 1.0;
        }
    }
    const double coef_t = co.coef_imp_t * co.coef_d_temp * co.dt;
    add_scalar_poisson(tm, coef_t);
    set_fix_fld_icb(tm);
    set_fix_fld_cmb(tm);
    const double coef_dbt = co.coef_imp_b * co.coef_d_magne * co.dt;
    add_vector_poisson(bs, coef_dbt, kb);
    add_vector_poisson(bt, coef_dbt, kb);
    for (int j = 0; j < J; ++j) {
        if (icore) {
            // add_vector_poisson_mat_center (center_sph_matrices.f90:160-180)
            const double rc2 = ar2[0];
            for (std::vector<double> *m : {&bs, &bt}) {
                M3(*m, j, 0, 1) = M3(*m, j, 0, 1) - coef_dbt * (fd.ctr1[1][2] - g3[j] * rc2);
                M3(*m, j, 1, 0) = M3(*m, j, 1, 0) - coef_dbt * fd.ctr1[2][2];
            }
        } else {
            M3(bs, j, ki, 1) = M3(bs, j, ki, 1)
                             - coef_dbt * (p.fdr_i[1][2] - g3[j] * p.ri[2]
                                           + (g1[j] + 1.0) * p.ri[1] * p.fdr_i[0][2]);
            M3(bs, j, ki + 1, 0) = M3(bs, j, ki + 1, 0) - coef_dbt * p.fdr_i[2][2];
        }
        M3(bs, j, ko - 1, 2) = M3(bs, j, ko - 1, 2) - coef_dbt * p.fdr_c[0][2];
        M3(bs, j, ko, 1) = M3(bs, j, ko, 1)
                         - coef_dbt * (p.fdr_c[1][2] - g3[j] * p.rc[2]
                                       - g1[j] * p.rc[1] * p.fdr_c[2][2]);
    }
    if (!icore) set_fix_fld_icb(bt);
    set_fix_fld_cmb(bt);
    // LU on the GPU (ludcmp_band / ludcmp_3band)
    const std::vector<double> *mats[5] = {&vp, &vt, &tm, &bs, &bt};
    const int nbs[5] = {5, 3, 3, 3, 3};
    for (int q = 0; q < 5; ++q)
        if (factor_band_for_solve(nbs[q], nr, J, mats[q]->data(), &c->d_lu[q], &c->d_piv[q], s,
                                  err)) {
            *err = "mhd_create: " + *err;
            return -1;
        }
    // transforms
    if (leg_context_build(&c->leg, L, nth, nr, radius, nullptr, nullptr, kNFwd, 4, 0, leg_mode, err))
        return -1;
    if (fft_context_build(&c->fft, L, nth, nr, nphi, err)) return -1;
    c->fused_fft = fft_mhd_fused_fits(&c->fft);
    if (!c->fused_fft) {   // long pencils: grids in HBM between the three stages
        const size_t nrtp = (size_t)nr * nth * nphi;
        HIP_TRY(hipMalloc(&c->d_grid_b, sizeof(double) * 13 * nrtp));
        HIP_TRY(hipMalloc(&c->d_grid_f, sizeof(double) * kNFwd * nrtp));
    }
    HIP_TRY(hipDeviceSynchronize());
    return 0;
}

void mhd_free(MhdContext *c)
{
    (void)hipStreamSynchronize(c->stream);   // replays of the step graph
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
    if (c->gstream) (void)hipStreamDestroy(c->gstream);
    c->gstream = nullptr;
    leg_context_free(&c->leg);
    fft_context_free(&c->fft);
    double *bufs[] = {c->d_rj, c->d_sp_b, c->d_sp_f, c->d_vr_b, c->d_vr_f, c->d_consts,
                      c->d_radius, c->d_scratch, c->d_ms_out, c->d_grid_b, c->d_grid_f,
                      c->d_press, c->d_lu_p};
    for (double *b : bufs)
        if (b) (void)hipFree(b);
    if (c->d_iconsts) (void)hipFree(c->d_iconsts);
    if (c->d_piv_p) (void)hipFree(c->d_piv_p);
    for (int q = 0; q < 5; ++q) {
        if (c->d_lu[q]) (void)hipFree(c->d_lu[q]);
        if (c->d_piv[q]) (void)hipFree(c->d_piv[q]);
    }
}

static void drop_graph(MhdContext *c);

int mhd_set_option(MhdContext *c, const std::string &name, long long value, std::string *err)
{
    if (name == "fused_fft") {
        if (value && !fft_mhd_fused_fits(&c->fft)) {
            *err = "fused_fft: the pencil does not fit in LDS at this size";
            return -1;
        }
        if (!value && !c->d_grid_b) {
            const size_t nrtp = (size_t)c->nr * c->nth * c->nphi;
            HIP_TRY(hipMalloc(&c->d_grid_b, sizeof(double) * 13 * nrtp));
            HIP_TRY(hipMalloc(&c->d_grid_f, sizeof(double) * kNFwd * nrtp));
        }
        c->fused_fft = value != 0;
        drop_graph(c);
        return 0;
    }
    if (name == "legendre_autotune") {
        c->leg.autotune = value != 0;
        return 0;
    }
    if (name == "graph") {
        c->use_graph = value != 0;
        drop_graph(c);
        return 0;
    }
    *err = "unknown option " + name;
    return -1;
}

int mhd_field_io(MhdContext *c, int comp, double *host, bool to_host, std::string *err)
{
    if (comp < 0 || comp >= kMhdNumComp) {
        *err = "field component out of range";
        return -1;
    }
    const size_t n = c->p.nnod;
    double *d = c->d_rj + (size_t)comp * n;
    if (to_host)
        HIP_TRY(hipMemcpyAsync(host, d, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    else
        HIP_TRY(hipMemcpyAsync(d, host, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

// The caller's local mode order of its rj arrays: the reference's one-domain
// rj modes come grouped by order m (const_global_rj_modes_by_rlm), while the
// state here holds mode l(l+1)+m at j.  Host copies through
// mhd_field_io_caller permute between the two.
int mhd_set_rj_order(MhdContext *c, int J, const int *j_global, std::string *err)
{
    if (J != c->J) {
        *err = "rj mode order: " + std::to_string(J) + " modes given, the grid has " +
               std::to_string(c->J);
        return -1;
    }
    std::vector<char> seen((size_t)J, 0);
    for (int j = 0; j < J; ++j) {
        const int g = j_global[j];
        if (g < 0 || g >= J || seen[(size_t)g]) {
            *err = "rj mode order: not a permutation of 0..J-1";
            return -1;
        }
        seen[(size_t)g] = 1;
    }
    c->rj_order.assign(j_global, j_global + J);
    return 0;
}

int mhd_field_io_caller(MhdContext *c, int comp, double *host, bool to_host, std::string *err)
{
    if (c->rj_order.empty()) return mhd_field_io(c, comp, host, to_host, err);
    const int J = c->J, nr = c->nr;
    std::vector<double> own((size_t)J * nr);
    if (!to_host) {
        for (int k = 0; k < nr; ++k)
            for (int j = 0; j < J; ++j)
                own[(size_t)c->rj_order[j] + (size_t)J * k] = host[(size_t)j + (size_t)J * k];
    }
    if (mhd_field_io(c, comp, own.data(), to_host, err)) return -1;
    if (to_host) {
        for (int k = 0; k < nr; ++k)
            for (int j = 0; j < J; ++j)
                host[(size_t)j + (size_t)J * k] = own[(size_t)c->rj_order[j] + (size_t)J * k];
    }
    return 0;
}

// Gather d_rj components into an rlm spectrum, component fastest (the layout
// of the reference's send buffer after copy_snap_spectr_to_send)
__global__ void k_gather_rlm(MhdDev p, int ncomp, const int *slot, double *sp)
{
    const size_t n = (size_t)ncomp * p.J * p.nr;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        const int comp = (int)(e % ncomp);
        const size_t q = e / ncomp;
        const int jr = (int)(q % p.J), k = (int)(q / p.J);
        sp[e] = p.d[(size_t)slot[comp] * p.nnod + (size_t)k * p.J + p.rj_of_rlm[jr]];
    }
}

// Fields on the physical grid at an output step: sph_back_trans_snapshot_MHD
// (sph_transforms_4_MHD.f90:200-224) without the pole values.  The vectors
// (first component of each: pol, dpol/dr, tor) and scalars named by d_rj
// component go through the context's backward Legendre transform and FFT.
int mhd_snapshot_rtp(MhdContext *c, int nvec, const int *vec, int nscl, const int *scl,
                     double *host, std::string *err)
{
    if (nvec < 0 || nscl < 0 || nvec + nscl < 1 || 3 * nvec + nscl > 64 ||
        (nvec && !vec) || (nscl && !scl) || !host) {
        *err = "snapshot: need 1 <= 3 nvec + nscl <= 64 and the component lists";
        return -1;
    }
    std::vector<int> slot;
    for (int i = 0; i < nvec; ++i) {
        if (vec[i] < 0 || vec[i] + 3 > kMhdNumComp) {
            *err = "snapshot: vector component out of range";
            return -1;
        }
        for (int q = 0; q < 3; ++q) slot.push_back(vec[i] + q);
    }
    for (int i = 0; i < nscl; ++i) {
        if (scl[i] < 0 || scl[i] >= kMhdNumComp) {
            *err = "snapshot: scalar component out of range";
            return -1;
        }
        slot.push_back(scl[i]);
    }
    const int ncomp = (int)slot.size();
    const size_t nrlm = (size_t)ncomp * c->J * c->nr;
    const size_t nrtm = (size_t)ncomp * c->nth * c->nr * c->M;
    const size_t nrtp = (size_t)ncomp * c->nth * c->nr * c->nphi;
    hipStream_t s = c->stream;
    double *d_sp = nullptr, *d_vr = nullptr, *d_grid = nullptr;
    int *d_slot = nullptr;
    auto release = [&]() {
        if (d_sp) (void)hipFree(d_sp);
        if (d_vr) (void)hipFree(d_vr);
        if (d_grid) (void)hipFree(d_grid);
        if (d_slot) (void)hipFree(d_slot);
    };
    int rc = -1;
    do {
        hipError_t e = hipMalloc(&d_sp, nrlm * sizeof(double));
        if (e == hipSuccess) e = hipMalloc(&d_vr, nrtm * sizeof(double));
        if (e == hipSuccess) e = hipMalloc(&d_grid, nrtp * sizeof(double));
        if (e == hipSuccess) e = hipMalloc(&d_slot, ncomp * sizeof(int));
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_slot, slot.data(), ncomp * sizeof(int), hipMemcpyHostToDevice, s);
        if (e != hipSuccess) {
            *err = std::string("snapshot: ") + hipGetErrorString(e);
            break;
        }
        hipLaunchKernelGGL(k_gather_rlm, dim3(blocks_for(nrlm)), dim3(256), 0, s, c->p, ncomp,
                           d_slot, d_sp);
        e = hipGetLastError();
        if (e != hipSuccess) {
            *err = std::string("snapshot gather: ") + hipGetErrorString(e);
            break;
        }
        if (leg_backward_dev(&c->leg, ncomp, nvec, nscl, d_sp, d_vr, s, err)) break;
        if (fft_backward_dev(&c->fft, ncomp, d_vr, d_grid, s, err)) break;
        e = hipMemcpyAsync(host, d_grid, nrtp * sizeof(double), hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            *err = std::string("snapshot copy: ") + hipGetErrorString(e);
            break;
        }
        rc = 0;
    } while (false);
    release();
    return rc;
}

// Output-step energy fluxes on the grid (s_cal_energy_flux_rtp,
// cal_energy_flux_rtp.f90:101-150 and cal_buoyancy_flux_rtp_smp :222-252).
// g holds 13 grid components, node fastest: velocity, magnetic field, current
// density and magnetic induction as (r, theta, phi), then the temperature
// perturbation.  The Lorentz force is the step's coef_lor (J x B)
// (cal_cross_prod_w_coef); flux ids: 0 Lorentz_work u.F, 1
// work_against_Lorentz -u.F, 2 magnetic_ene_generation B.induction, 3
// buoyancy_flux coef_buo T u_r r.
struct FluxIds {
    int n;
    int id[kMaxEnergyFluxes];
};

__global__ void k_energy_flux(size_t nnod, int nr, const double *__restrict__ g,
                              const double *__restrict__ radius, double coef_lor,
                              double coef_buo, FluxIds ids, double *__restrict__ out)
{
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < nnod;
         i += (size_t)gridDim.x * blockDim.x) {
        const double u1 = g[i], u2 = g[nnod + i], u3 = g[2 * nnod + i];
        const double b1 = g[3 * nnod + i], b2 = g[4 * nnod + i], b3 = g[5 * nnod + i];
        const double j1 = g[6 * nnod + i], j2 = g[7 * nnod + i], j3 = g[8 * nnod + i];
        const double n1 = g[9 * nnod + i], n2 = g[10 * nnod + i], n3 = g[11 * nnod + i];
        const double t = g[12 * nnod + i];
        const double f1 = (j2 * b3 - j3 * b2) * coef_lor;
        const double f2 = (j3 * b1 - j1 * b3) * coef_lor;
        const double f3 = (j1 * b2 - j2 * b1) * coef_lor;
        for (int q = 0; q < ids.n; ++q) {
            double v;
            switch (ids.id[q]) {
            case 0: v = f1 * u1 + f2 * u2 + f3 * u3; break;
            case 1: v = -1.0 * (f1 * u1 + f2 * u2 + f3 * u3); break;
            case 2: v = n1 * b1 + n2 * b2 + n3 * b3; break;
            default: v = coef_buo * t * u1 * radius[i % nr]; break;
            }
            out[(size_t)q * nnod + i] = v;
        }
    }
}

// rlm spectrum (component fastest) back to d_rj order: out(comp, k, j)
__global__ void k_scatter_rlm(MhdDev p, int ncomp, const double *__restrict__ sp,
                              double *__restrict__ out)
{
    const size_t n = (size_t)ncomp * p.J * p.nr;
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n;
         e += (size_t)gridDim.x * blockDim.x) {
        const int comp = (int)(e % ncomp);
        const size_t q = e / ncomp;
        const int jr = (int)(q % p.J), k = (int)(q / p.J);
        out[(size_t)comp * p.nnod + (size_t)k * p.J + p.rj_of_rlm[jr]] = sp[e];
    }
}

// enegy_fluxes_4_sph_mhd (lead_fields_4_sph_mhd.f90:98-121): the snapshot's
// backward transform of u, B, J, the induction and T, the fluxes on the grid,
// and the forward transform of the flux scalars (sph_forward_trans_snapshot_MHD)
// to spectra in the caller's rj order, host(nflux, nr, J).
int mhd_energy_fluxes(MhdContext *c, int nflux, const int *ids, double *host, std::string *err)
{
    if (nflux < 1 || nflux > kMaxEnergyFluxes || !ids || !host) {
        *err = "energy fluxes: need 1 <= nflux <= 8, the flux ids and an output array";
        return -1;
    }
    FluxIds fid{};
    fid.n = nflux;
    for (int q = 0; q < nflux; ++q) {
        if (ids[q] < 0 || ids[q] > 3) {
            *err = "energy fluxes: flux id must be 0 (Lorentz_work), 1 (work_against_Lorentz), "
                   "2 (magnetic_ene_generation) or 3 (buoyancy_flux)";
            return -1;
        }
        fid.id[q] = ids[q];
    }
    if (!c->initialized) {
        *err = "energy fluxes: the state is not initialised (calypso_mhd_initialize)";
        return -1;
    }
    static const int slot[13] = {0, 1, 2, 6, 7, 8, 9, 10, 11, kInduction, kInduction + 1,
                                 kInduction + 2, 12};
    const int ncomp = 13;
    const size_t nnod_rtp = (size_t)c->nth * c->nr * c->nphi;
    const size_t nrlm = (size_t)ncomp * c->J * c->nr;
    const size_t nrtm = (size_t)ncomp * c->nth * c->nr * c->M;
    const size_t nj = (size_t)c->J * c->nr;
    hipStream_t s = c->stream;
    double *d_sp = nullptr, *d_vr = nullptr, *d_grid = nullptr, *d_flux = nullptr;
    int *d_slot = nullptr;
    auto release = [&]() {
        if (d_sp) (void)hipFree(d_sp);
        if (d_vr) (void)hipFree(d_vr);
        if (d_grid) (void)hipFree(d_grid);
        if (d_flux) (void)hipFree(d_flux);
        if (d_slot) (void)hipFree(d_slot);
    };
    int rc = -1;
    do {
        hipError_t e = hipMalloc(&d_sp, nrlm * sizeof(double));
        if (e == hipSuccess) e = hipMalloc(&d_vr, nrtm * sizeof(double));
        if (e == hipSuccess) e = hipMalloc(&d_grid, (size_t)ncomp * nnod_rtp * sizeof(double));
        if (e == hipSuccess) e = hipMalloc(&d_flux, (size_t)nflux * nnod_rtp * sizeof(double));
        if (e == hipSuccess) e = hipMalloc(&d_slot, ncomp * sizeof(int));
        if (e == hipSuccess)
            e = hipMemcpyAsync(d_slot, slot, ncomp * sizeof(int), hipMemcpyHostToDevice, s);
        if (e != hipSuccess) {
            *err = std::string("energy fluxes: ") + hipGetErrorString(e);
            break;
        }
        hipLaunchKernelGGL(k_gather_rlm, dim3(blocks_for(nrlm)), dim3(256), 0, s, c->p, ncomp,
                           d_slot, d_sp);
        if (leg_backward_dev(&c->leg, ncomp, 4, 1, d_sp, d_vr, s, err)) break;
        if (fft_backward_dev(&c->fft, ncomp, d_vr, d_grid, s, err)) break;
        hipLaunchKernelGGL(k_energy_flux, dim3(blocks_for(nnod_rtp)), dim3(256), 0, s, nnod_rtp,
                           c->nr, d_grid, c->d_radius, c->co.coef_lor, c->co.coef_buo, fid,
                           d_flux);
        // the forward transform reuses the backward's buffers (nflux <= 13 components)
        if (fft_forward_dev(&c->fft, nflux, d_flux, d_vr, s, err)) break;
        if (leg_forward_dev(&c->leg, nflux, 0, nflux, d_vr, d_sp, s, err)) break;
        hipLaunchKernelGGL(k_scatter_rlm, dim3(blocks_for(nflux * nj)), dim3(256), 0, s, c->p,
                           nflux, d_sp, d_grid);
        e = hipGetLastError();
        std::vector<double> own((size_t)nflux * nj);
        if (e == hipSuccess)
            e = hipMemcpyAsync(own.data(), d_grid, own.size() * sizeof(double),
                               hipMemcpyDeviceToHost, s);
        if (e == hipSuccess) e = hipStreamSynchronize(s);
        if (e != hipSuccess) {
            *err = std::string("energy fluxes: ") + hipGetErrorString(e);
            break;
        }
        const int J = c->J;
        for (int q = 0; q < nflux; ++q)
            for (int k = 0; k < c->nr; ++k)
                for (int j = 0; j < J; ++j)
                    host[(size_t)q * nj + (size_t)j + (size_t)J * k] =
                        own[(size_t)q * nj + (size_t)(c->rj_order.empty() ? j : c->rj_order[j]) +
                            (size_t)J * k];
        rc = 0;
    } while (false);
    release();
    return rc;
}

// dynamo_benchmark_1 initial fields (set_initial_velo_sph,
// set_ini_reference_temp_sph, set_initial_temp_sph,
// initial_b_dynamobench_1): full temperature, as a restart holds it.
namespace {

// The benchmark start (sph_initial_data_control, set_initial_sph_dynamo.f90:
// set_initial_velo_sph, set_ini_reference_temp_sph, set_initial_temp_sph,
// initial_b_dynamobench_1 / _2): zero velocity, the full temperature (the
// reference profile at every point plus the (m, m) perturbation of isig in the
// shell), as a restart holds it, and benchmark 1's field in the shell or
// benchmark 2's from the first point to the CMB.
int set_benchmark_ic(MhdContext *c, int isig, int bench, std::string *err)
{
    const int nr = c->nr, J = c->J, ki = c->p.ki, ko = c->p.ko;
    const size_t n = (size_t)nr * J;
    const double pi = 4.0 * std::atan(1.0);
    const double *r = c->radius.data();
    const double r_icb = r[ki], r_cmb = r[ko], shell = r_cmb - r_icb;
    std::vector<double> zero(n, 0.0), temp(n, 0.0), bp(n, 0.0), bd(n, 0.0), bt(n, 0.0),
        jp(n, 0.0), jd(n, 0.0), jt(n, 0.0);
    const int m = (isig % 1000) / 100;
    const int jj = m * (m + 1) + m;
    if (jj >= J) {
        *err = "initial temperature mode beyond the truncation";
        return -1;
    }
    for (int k = 0; k < nr; ++k) {
        temp[(size_t)J * k] = c->reft0[k];
        const double rr = r[k];
        const size_t is = 2 + (size_t)J * k, it = 6 + (size_t)J * k;
        if (k >= ki && k <= ko) {
            const double xr = 2.0 * r[k] - 1.0 * (r_cmb + r_icb) / shell;
            temp[jj + (size_t)J * k] = (1.0 - 3.0 * (xr * xr) + 3.0 * std::pow(xr, 4) - std::pow(xr, 6))
                                     * 0.1 * 3.0 / std::sqrt(2.0 * pi);
        }
        if (bench == 1 && k >= ki && k <= ko) {
            bp[is] = (5.0 / 8.0) * (-3.0 * std::pow(rr, 3) + (4.0 * r_cmb) * (rr * rr)
                                    - std::pow(r_icb, 4) / rr);
            bd[is] = (5.0 / 8.0) * (-9.0 * (rr * rr) + 8.0 * r_cmb * rr + std::pow(r_icb, 4) / (rr * rr));
            jt[is] = (5.0 * 3.0 / 2.0) * rr;
            bt[it] = (10.0 / 3.0) * rr * std::sin(pi * (rr - r_icb));
            jp[it] = bt[it];
            jd[it] = (10.0 / 3.0) * (std::sin(pi * (rr - r_icb)) + pi * rr * std::cos(pi * (rr - r_icb)));
        } else if (bench == 2 && k <= ko) {
            bp[is] = (5.0 / 2.0) * (rr * rr) * (4.0 * r_cmb - 3.0 * rr) / (r_cmb + 3.0);
            bd[is] = (5.0 / 2.0) * rr * (8.0 * r_cmb - 9.0 * rr) / (r_cmb + 3.0);
            jt[is] = 5.0 * 6.0 * rr / (3.0 + r_cmb);
            const double sn = std::sin(pi * rr / r_cmb), cs = std::cos(pi * rr / r_cmb);
            bt[it] = (10.0 / 3.0) * rr * sn;
            jp[it] = bt[it];
            jd[it] = (10.0 / 3.0) * (sn + (pi / r_cmb) * rr * cs);
        }
    }
    const std::pair<int, const std::vector<double> *> sets[] = {
        {kVelo, &zero}, {kVelo + 1, &zero}, {kVelo + 2, &zero}, {kTemp, &temp},
        {kMagne, &bp}, {kMagne + 1, &bd}, {kMagne + 2, &bt},
        {kCurrent, &jp}, {kCurrent + 1, &jd}, {kCurrent + 2, &jt}};
    for (const auto &sv : sets)
        if (mhd_field_io(c, sv.first, const_cast<double *>(sv.second->data()), false, err)) return -1;
    return 0;
}

}  // namespace

int mhd_set_dynamobench_ic(MhdContext *c, int isig, std::string *err)
{
    return set_benchmark_ic(c, isig, 1, err);
}

int mhd_set_dynamobench2_ic(MhdContext *c, int isig, std::string *err)
{
    if (!c->p.icore) {
        *err = "dynamo_benchmark_2 fields need the inner core (points below the ICB)";
        return -1;
    }
    return set_benchmark_ic(c, isig, 2, err);
}

// Seeded random initial fields: poloidal and toroidal velocity and magnetic
// field (l >= 1) and the temperature perturbation (l >= 0), each a
// counter-based Gaussian per (field, radius, mode) with amplitude (l+1)^-2
// times the wall envelope (seeded.hpp), so the walls see zero velocity,
// field and perturbation with zero radial derivatives.  The temperature is
// stored as the full temperature (reference profile in the l = 0 mode plus
// the perturbation), as for the benchmark start; the derived fields are
// formed by mhd_initialize.
int mhd_set_seeded_ic(MhdContext *c, unsigned long long seed, double amp_u, double amp_b,
                      double amp_t, std::string *err)
{
    if (c->p.icore) {
        *err = "seeded fields are defined for a shell without an inner core";
        return -1;
    }
    const int nr = c->nr, J = c->J, L = c->L;
    const size_t n = (size_t)nr * J;
    const double *r = c->radius.data();
    std::vector<double> zero(n, 0.0), up(n), ut(n), bp(n), bt(n), temp(n);
    seeded_spectrum(seed, kVelo, L, nr, r, amp_u, 2.0, true, 1, up.data());
    seeded_spectrum(seed, kVelo + 2, L, nr, r, amp_u, 2.0, true, 1, ut.data());
    seeded_spectrum(seed, kMagne, L, nr, r, amp_b, 2.0, true, 1, bp.data());
    seeded_spectrum(seed, kMagne + 2, L, nr, r, amp_b, 2.0, true, 1, bt.data());
    seeded_spectrum(seed, kTemp, L, nr, r, amp_t, 2.0, true, 0, temp.data());
    for (int k = 0; k < nr; ++k) temp[(size_t)J * k] = c->reft0[k] + temp[(size_t)J * k];
    const std::pair<int, const std::vector<double> *> sets[] = {
        {kVelo, &up}, {kVelo + 1, &zero}, {kVelo + 2, &ut}, {kTemp, &temp},
        {kMagne, &bp}, {kMagne + 1, &zero}, {kMagne + 2, &bt},
        {kCurrent, &zero}, {kCurrent + 1, &zero}, {kCurrent + 2, &zero}};
    for (const auto &sv : sets)
        if (mhd_field_io(c, sv.first, const_cast<double *>(sv.second->data()), false, err)) return -1;
    return 0;
}

namespace {

int nonlinear(MhdContext *c, bool temp_roundtrip, std::string *err)
{
    MhdDev &p = c->p;
    hipStream_t s = c->stream;
    const int nb = blocks_for(p.nnod);
    // radial levels per transpose work-group: as many as keep about 2048
    // non-empty tiles (about half the tile grid lies outside l >= |m|)
    const int tx = (2 * c->L + 1 + kTile - 1) / kTile, ty = (c->L + 1 + kTile - 1) / kTile;
    const int kz = std::max(1, std::min(8, (tx * ty / 2) * c->nr / 2048));
    const dim3 tiles(tx, ty, (c->nr + kz - 1) / kz);
    hipLaunchKernelGGL(k_pack, tiles, dim3(256), 0, s, p, kz);
    HIP_TRY(hipGetLastError());
    if (leg_backward_dev(&c->leg, 13, 4, 1, c->d_sp_b, c->d_vr_b, s, err)) return -1;
    const double coef[4] = {c->co.coef_velo, c->co.coef_lor, c->co.coef_induct, c->co.coef_temp};
    if (c->fused_fft) {
        if (fft_mhd_fused_dev(&c->fft, c->d_vr_b, c->d_vr_f, coef, s, err, kNFwd)) return -1;
    } else {
        const size_t nrtp = (size_t)c->nr * c->nth * c->nphi;
        if (fft_backward_dev(&c->fft, 13, c->d_vr_b, c->d_grid_b, s, err) ||
            products_mhd_dev(nrtp, c->d_grid_b, c->d_grid_f, coef, s, err, kNFwd) ||
            fft_forward_dev(&c->fft, kNFwd, c->d_grid_f, c->d_vr_f, s, err))
            return -1;
    }
    if (leg_forward_dev(&c->leg, kNFwd, 4, 0, c->d_vr_f, c->d_sp_f, s, err)) return -1;
    hipLaunchKernelGGL(k_unpack, tiles, dim3(256), 0, s, p, kz);
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_coriolis, dim3(nb), dim3(256), 0, s, p);
    HIP_TRY(hipGetLastError());
    if (p.icore) {
        hipLaunchKernelGGL(k_icore_torque, dim3(1), dim3(64), 0, s, p);
        HIP_TRY(hipGetLastError());
    }
    hipLaunchKernelGGL(k_rot_forces, dim3(nb), dim3(256), 0, s, p, temp_roundtrip ? 1 : 0);
    HIP_TRY(hipGetLastError());
    return 0;
}

// derived fields after new u, T, B (set_sph_field_to_start and the
// update_after_* calls of s_cal_sol_sph_MHD_crank)
int derived_fields(MhdContext *c, std::string *err)
{
    MhdDev &p = c->p;
    hipStream_t s = c->stream;
    const int nb = blocks_for(p.nnod), nj = (p.J + 127) / 128;
    hipLaunchKernelGGL(k_bc_rows, dim3(nj), dim3(128), 0, s, p);
    hipLaunchKernelGGL(k_grad_interior, dim3(nb), dim3(256), 0, s, p);
    hipLaunchKernelGGL(k_diffusion, dim3(nb), dim3(256), 0, s, p);
    hipLaunchKernelGGL(k_temp_bc, dim3(nj), dim3(128), 0, s, p);
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // namespace

int mhd_initialize(MhdContext *c, std::string *err)
{
    MhdDev &p = c->p;
    // sync_temp_by_per_temp_sph: the evolved temperature is the perturbation
    std::vector<double> t(p.nnod);
    if (mhd_field_io(c, kTemp, t.data(), true, err)) return -1;
    for (int k = 0; k < p.nr; ++k) t[(size_t)p.J * k] = t[(size_t)p.J * k] - c->reft0[k];
    if (mhd_field_io(c, kTemp, t.data(), false, err)) return -1;
    if (derived_fields(c, err)) return -1;
    if (nonlinear(c, false, err)) return -1;
    HIP_TRY(hipStreamSynchronize(c->stream));
    // a restart resumes its step count and time (the next step is then an
    // Adams-Bashforth step with the restart's previous explicit terms)
    c->i_step = c->restart_step >= 0 ? c->restart_step : 0;
    c->istep0 = c->i_step;
    c->time0 = c->restart_step >= 0 ? c->restart_time : 0.0;
    c->restart_step = -1;
    c->initialized = true;
    return 0;
}

// One time step (SPH_analyze_MHD without output): explicit terms, CN solves,
// derived fields, nonlinear terms; `first` selects the Euler start.
static int step_body(MhdContext *c, int first, std::string *err)
{
    MhdDev &p = c->p;
    hipStream_t s = c->stream;
    const int nb = blocks_for(p.nnod);
    const size_t J = p.J;
    hipLaunchKernelGGL(k_explicit, dim3(nb), dim3(256), 0, s, p, first, first);
    hipLaunchKernelGGL(k_cn_rhs, dim3(nb), dim3(256), 0, s, p);
    HIP_TRY(hipGetLastError());
    double *d = c->d_rj;
    // the four 3-band systems (u_tor, T, B_pol, B_tor) are independent: one launch
    const double *b3[4] = {c->d_lu[1], c->d_lu[2], c->d_lu[3], c->d_lu[4]};
    const int *p3[4] = {c->d_piv[1], c->d_piv[2], c->d_piv[3], c->d_piv[4]};
    double *x3[4] = {d + (kVelo + 2) * p.nnod, d + kTemp * p.nnod, d + kMagne * p.nnod,
                     d + (kMagne + 2) * p.nnod};
    if (lubksb_t_dev(5, p.nr, (int)J, c->d_lu[0], c->d_piv[0], d + kVelo * p.nnod, s, err) ||
        lubksb_t_batch_dev(3, p.nr, (int)J, 4, b3, p3, x3, s, err))
        return -1;
    if (derived_fields(c, err)) return -1;
    if (nonlinear(c, true, err)) return -1;
    return 0;
}

// Capture one Adams-Bashforth step into a HIP graph on the library's own
// stream (the caller's may be the legacy null stream, which cannot be
// captured); later steps replay it on the caller's stream.  The step holds
// kernel launches only: no memsets, copies or host synchronisation, and the
// Legendre plans already exist (made during the initialisation).
static int capture_step(MhdContext *c, std::string *err)
{
    if (!c->gstream) HIP_TRY(hipStreamCreateWithFlags(&c->gstream, hipStreamNonBlocking));
    hipStream_t user = c->stream;
    c->stream = c->gstream;
    HIP_TRY(hipStreamBeginCapture(c->gstream, hipStreamCaptureModeRelaxed));
    std::string e2;
    const int rc = step_body(c, 0, &e2);
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(c->gstream, &g);
    c->stream = user;
    if (rc || ec != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        *err = rc ? "capture of the step failed: " + e2
                  : std::string("hipStreamEndCapture failed: ") + hipGetErrorString(ec);
        return -1;
    }
    const hipError_t ei = hipGraphInstantiate(&c->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ei != hipSuccess) {
        c->gexec = nullptr;
        *err = std::string("hipGraphInstantiate failed: ") + hipGetErrorString(ei);
        return -1;
    }
    return 0;
}

static void drop_graph(MhdContext *c)
{
    if (c->gexec) (void)hipGraphExecDestroy(c->gexec);
    c->gexec = nullptr;
}

int mhd_steps(MhdContext *c, int nsteps, std::string *err)
{
    if (!c->initialized) {
        *err = "mhd_steps: call the initialisation first";
        return -1;
    }
    for (int it = 0; it < nsteps; ++it) {
        // the step count (and so the model time) advances only once the
        // step's work has been issued
        const int first = c->i_step + 1 == 1;
        if (c->use_graph && !first) {
            if (!c->gexec && capture_step(c, err)) return -1;
            HIP_TRY(hipGraphLaunch(c->gexec, c->stream));
        } else {
            if (step_body(c, first, err)) return -1;
        }
        c->i_step += 1;
    }
    return 0;
}

int mhd_energies(MhdContext *c, double *out7, std::string *err)
{
    MhdDev &p = c->p;
    hipStream_t s = c->stream;
    const int nb = blocks_for(p.nnod);
    double *tfull = c->d_scratch, *vol = c->d_scratch + p.nnod;
    hipLaunchKernelGGL(k_temp_full, dim3(nb), dim3(256), 0, s, p, tfull);
    HIP_TRY(hipGetLastError());
    const double *h_r = c->radius.data();
    const int kr_st = p.ki + 1, kr_ed = p.ko + 1;
    double h[3];
    const struct { int kind; const double *f; int o; } jobs[3] = {
        {1, c->d_rj + (size_t)kVelo * p.nnod, 0},
        {2, tfull, 3},
        {1, c->d_rj + (size_t)kMagne * p.nnod, 4}};
    for (const auto &jb : jobs) {
        if (mean_square_dev(jb.kind, p.nr, p.J, c->L, c->d_l_of_j, p.r, h_r, kr_st, kr_ed, jb.f,
                            vol, c->d_ms_out, s, err))
            return -1;
        HIP_TRY(hipMemcpyAsync(h, c->d_ms_out, sizeof(double) * 3, hipMemcpyDeviceToHost, s));
        HIP_TRY(hipStreamSynchronize(s));
        if (jb.kind == 2) {
            out7[3] = h[0];
        } else {
            out7[jb.o] = h[0];
            out7[jb.o + 1] = h[1];
            out7[jb.o + 2] = h[2];
        }
    }
    return 0;
}

// ------------------------------------------------------------- pressure --
// Right-hand side of the pressure Poisson equation at an output step
// (pressure_4_sph_mhd, lead_fields_4_sph_mhd.f90:70-95), one thread per node:
// -div(advection) + div(Lorentz force) + div(buoyancy) with the Coriolis
// term zero (the Coriolis vector is formed neither on the grid nor in rlm,
// cal_nonlinear_sph_MHD.f90:78-80, const_coriolis_sph_rlm.f90:92-95).
// Inside the shell the divergence cal_sph_nod_vect_div2
// (cal_sph_exp_rotation.f90:210-239) and the buoyancy's
// cal_div_buoyancy_sph_MHD (cal_div_buoyancies_sph_MHD.f90:124-145) over the
// perturbation temperature and its radial gradient (cal_sph_nod_gradient_2,
// cal_sph_exp_1st_diff.f90:40-75); on the walls the Lorentz force's
// divergence by the one-sided formulas cal_sph_nod_nobc_in/out_div2
// (cal_sph_exp_nod_none_bc.f90:133-160, 269-296), and the advection and
// buoyancy terms replaced by radial forces (s_const_radial_forces_on_bc,
// const_radial_forces_on_bc.f90:26-60: the advection's radial part is
// overwritten by the Lorentz force's; cal_r_buoyancy_on_sph,
// cal_r_buoyancies_on_sph.f90:97-113).  Zero outside the shell.
__global__ void k_press_rhs(MhdDev p, double coef_buo, double *rhs)
{
    const size_t J = p.J;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < p.nnod;
         e += (size_t)gridDim.x * blockDim.x) {
        const int j = (int)(e % J), k = (int)(e / J);
        if (k < p.ki || k > p.ko) {
            rhs[e] = 0.0;
            continue;
        }
        const double *t = p.d + (size_t)kTemp * p.nnod + j;
        const double *lp = p.d + (size_t)kLorentz * p.nnod + j;
        const double *ld = p.d + (size_t)(kLorentz + 1) * p.nnod + j;
        const double *ap = p.d + (size_t)kMAdvect * p.nnod + j;
        const double *ad = p.d + (size_t)(kMAdvect + 1) * p.nnod + j;
        const double g3 = p.g3[j];
        const double gh = fmax(g3, 0.5);
        double div_in, div_lor, div_buo;
        if (k == p.ki) {
            const double d1 = p.ffl_i[0][1] * lp[k * J] + p.ffl_i[1][1] * lp[(k + 1) * J] +
                              p.ffl_i[2][1] * lp[(k + 2) * J];
            div_lor = (d1 - ld[k * J]) * gh * p.ri[2];
        } else if (k == p.ko) {
            const double d1 = p.ffl_c[2][1] * lp[(k - 2) * J] + p.ffl_c[1][1] * lp[(k - 1) * J] +
                              p.ffl_c[0][1] * lp[k * J];
            div_lor = (d1 - ld[k * J]) * gh * p.rc[2];
        } else {
            const double *w = p.d1 + 3 * k;
            const double d1l = w[0] * lp[(k - 1) * J] + w[1] * lp[k * J] + w[2] * lp[(k + 1) * J];
            const double d1a = w[0] * ap[(k - 1) * J] + w[1] * ap[k * J] + w[2] * ap[(k + 1) * J];
            const double d1t = w[0] * t[(k - 1) * J] + w[1] * t[k * J] + w[2] * t[(k + 1) * J];
            div_lor = (d1l - ld[k * J]) * gh * p.ar2[k];
            div_in = (d1a - ad[k * J]) * gh * p.ar2[k];
            // poloidal part of grad T: dT/dr / (l(l+1)) r^2 (1 for l = 0)
            const double g13 = (j == 0) ? 1.0 : 1.0 / g3;
            const double gt = d1t * g13 * (p.r[k] * p.r[k]);
            div_buo = coef_buo * (3.0 * t[k * J] + gt * p.r[k] * g3 * p.ar1[k]);
        }
        if (k == p.ki || k == p.ko) {
            div_in = lp[k * J] * gh * p.ar2[k];
            div_buo = coef_buo * t[k * J] * p.r[k];
        }
        rhs[e] = -div_in + div_lor + div_buo;
    }
}

// adjust_by_ave_pressure_on_CMB (set_reference_sph_mhd.f90:53-73): the
// degree-0 pressure relative to its CMB value, on one thread.
__global__ void k_press_adjust(MhdDev p, double *press)
{
    if (blockIdx.x != 0 || threadIdx.x != 0) return;
    const double ref = press[(size_t)p.ko * p.J];
    for (int k = p.ki; k <= p.ko; ++k) press[(size_t)k * p.J] = press[(size_t)k * p.J] - ref;
}

// The pressure matrix of const_radial_mat_4_press_sph
// (const_r_mat_4_scalar_sph.f90:73-129): set_unit_mat_4_poisson and
// add_scalar_poisson_mat_sph (set_radial_mat_sph.f90:141-204) with
// coef_p = -coef_press, the walls' add_icb/cmb_scalar_poisson_mat
// (set_sph_scalar_mat_bc.f90:105-128, 179-202) by the fixed-gradient
// formulas, factored by ludcmp_3band on the device.
static int pressure_matrix(MhdContext *c, std::string *err)
{
    const int nr = c->nr, J = c->J, ki = c->p.ki, ko = c->p.ko;
    RadialFdm fd;
    radial_fdm(c->radius.data(), nr, ki, ko, &fd);
    const std::vector<double> &d1 = fd.d1, &d2 = fd.d2;
    std::vector<double> ar1(nr), ar2(nr), g3(J);
    for (int k = 0; k < nr; ++k) {
        ar1[k] = 1.0 / c->radius[k];
        ar2[k] = ar1[k] * ar1[k];
    }
    for (int j = 0; j < J; ++j) {
        const int l = (int)std::sqrt((double)j + 0.5);
        g3[j] = (double)(l * (l + 1));
    }
    std::vector<double> m((size_t)3 * nr * J, 0.0);
    auto M3 = [&](int j, int k, int b) -> double & { return m[((size_t)j * nr + k) * 3 + b]; };
    const double coef_p = -c->co.coef_press;
    for (int j = 0; j < J; ++j) {
        for (int k = 0; k < nr; ++k) M3(j, k, 1) = (k < ki || k > ko) ? 1.0 : 0.0;
        for (int k = ki + 1; k < ko; ++k) {
            M3(j, k - 1, 2) = M3(j, k - 1, 2) - coef_p * (d2[3 * k + 0] + 2.0 * ar1[k] * d1[3 * k + 0]);
            M3(j, k, 1) = M3(j, k, 1) -
                          coef_p * (d2[3 * k + 1] + 2.0 * ar1[k] * d1[3 * k + 1] - g3[j] * ar2[k]);
            M3(j, k + 1, 0) = M3(j, k + 1, 0) - coef_p * (d2[3 * k + 2] + 2.0 * ar1[k] * d1[3 * k + 2]);
        }
        M3(j, ki, 1) = M3(j, ki, 1) -
                       coef_p * (fd.fdr_i[1][2] + 2.0 * fd.ri[1] * fd.fdr_i[1][1] - g3[j] * fd.ri[2]);
        M3(j, ki + 1, 0) = M3(j, ki + 1, 0) - coef_p * (fd.fdr_i[2][2] + 2.0 * fd.ri[1] * fd.fdr_i[2][1]);
        M3(j, ko - 1, 2) = M3(j, ko - 1, 2) - coef_p * (fd.fdr_c[0][2] + 2.0 * fd.rc[1] * fd.fdr_c[0][1]);
        M3(j, ko, 1) = M3(j, ko, 1) -
                       coef_p * (fd.fdr_c[1][2] + 2.0 * fd.rc[1] * fd.fdr_c[1][1] - g3[j] * fd.rc[2]);
    }
    if (factor_band_for_solve(3, nr, J, m.data(), &c->d_lu_p, &c->d_piv_p, c->stream, err)) {
        *err = "pressure matrix: " + *err;
        return -1;
    }
    return 0;
}

int mhd_pressure(MhdContext *c, double *host, std::string *err)
{
    if (!(c->co.coef_press != 0.0)) {
        *err = "pressure: coef_press is not set (coef_4_press, E^-1 for the benchmarks)";
        return -1;
    }
    if (!c->initialized) {
        *err = "pressure: the state is not initialised (calypso_mhd_initialize)";
        return -1;
    }
    MhdDev &p = c->p;
    hipStream_t s = c->stream;
    if (!c->d_press) {
        HIP_TRY(hipMalloc(&c->d_press, sizeof(double) * p.nnod));
        if (pressure_matrix(c, err)) return -1;
    }
    hipLaunchKernelGGL(k_press_rhs, dim3(blocks_for(p.nnod)), dim3(256), 0, s, p, c->co.coef_buo,
                       c->d_press);
    HIP_TRY(hipGetLastError());
    if (lubksb_t_dev(3, p.nr, p.J, c->d_lu_p, c->d_piv_p, c->d_press, s, err)) return -1;
    hipLaunchKernelGGL(k_press_adjust, dim3(1), dim3(64), 0, s, p, c->d_press);
    HIP_TRY(hipGetLastError());
    const int J = c->J, nr = c->nr;
    std::vector<double> own((size_t)J * nr);
    HIP_TRY(hipMemcpyAsync(own.data(), c->d_press, sizeof(double) * p.nnod, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int k = 0; k < nr; ++k)
        for (int j = 0; j < J; ++j)
            host[(size_t)j + (size_t)J * k] =
                own[(size_t)(c->rj_order.empty() ? j : c->rj_order[j]) + (size_t)J * k];
    return 0;
}

}  // namespace calypso
