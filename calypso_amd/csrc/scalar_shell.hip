// scalar_shell.hip -- scalars (temperature, composition, ...) diffusing in a
// spherical shell at rest, each wall holding a prescribed value or a
// prescribed flux per spherical-harmonic mode: the model of the reference's
// tests/heterogineous_temp, whose boundary data come from a spectral file
// (bc_spectr.btx, read_boundary_spectr_file,
// m_sph_boundary_input_data.f90:147-300).  FP64, device resident.
//
// One step per scalar (SPH_analyze_MHD with only scalars evolved):
//   k_sc_explicit   scalar_diff_advect_euler on step 1 (starting the Adams
//                   history, set_ini_adams_scalar), scalar_diff_advect_adams
//                   after (cal_diff_adv_src_explicit.f90), then the wall rows
//                   of cal_sol_scalar_sph_crank (cal_sol_sph_fluid_crank.f90:
//                   309-391): set_fixed_scalar_sph at a fixed-value wall,
//                   adjust_in/out_fixed_flux_sph at a flux wall
//                   (set_scalar_boundary_sph.f90)
//   lubksb          3-band Crank-Nicolson solve of all scalars in one launch
//                   (lubksb_3band_mul)
//   k_sc_diffusion  update_after_heat_sph / update_after_composit_sph
//                   (cal_sol_sph_MHD_crank.f90:191-236): a fixed-value wall
//                   takes its value, then the diffusion term
//                   (cal_sph_nod_scalar_diffuse2, cal_sph_exp_diffusion.f90;
//                   wall rows cal_sph_fix_scalar_in/out_diffuse2,
//                   cal_sph_exp_fixed_scalar.f90, and
//                   cal_sph_in/out_fix_flux_diffuse2, cal_sph_exp_fixed_flux.f90)
// With the fluid at rest the advection term is zero except in the rows of a
// flux wall, where the divergence formula takes the prescribed flux as the
// wall flux (cal_div_sph_in/out_fix_flux_2, cal_sph_exp_fixed_flux.f90):
// (q_j - 0) max(l(l+1), 1/2) / r^2 -- a constant of the run, set at create.
// Layout: mode j = l(l+1)+m fastest, radial point k: node j + J k.  The
// stencils are HBM-bound; FMA contraction is off so their arithmetic follows
// the reference's expressions operation by operation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "fortran_fmt.hpp"
#include "hip_check.hpp"
#include "mhd_internal.hpp"
#include "radial_internal.hpp"
#include "scalar_internal.hpp"

#pragma clang fp contract(off)

namespace calypso {

namespace {

// what every scalar shares: grid, stencils, wall formulas
struct ScShared {
    int nr, J, ki, ko;
    size_t nnod;
    const double *ar1, *ar2, *d1, *d2, *g3;
    double ffl_i[3][3], fdr_i[3][3], ffl_c[3][3], fdr_c[3][3];
    double ri[3], rc[3];
    double dt;
};

// one scalar's state and coefficients
struct ScOne {
    double *x, *diff, *adv, *pre;
    const double *v_i, *v_c;          // wall values [J] (inner flux negated)
    int kind_i, kind_c;
    double coef_d, coef_exp, coef_bc;   // coef_bc = dt coef_imp coef_d
};

struct ScBatch {
    ScOne s[kScalarMax];
};

__global__ void k_sc_explicit(ScShared g, ScBatch b, int first)
{
    const ScOne &p = b.s[blockIdx.y];
    const double adam_0 = 3.0 / 2.0, adam_1 = -1.0 / 2.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < g.nnod;
         i += (size_t)gridDim.x * blockDim.x) {
        const double D = p.diff[i], F = p.adv[i];
        double x = p.x[i];
        if (first) {
            x = x + g.dt * (p.coef_exp * D - F);
        } else {
            x = x + g.dt * (p.coef_exp * D - adam_0 * F + adam_1 * p.pre[i]);
        }
        p.pre[i] = -F;
        const int k = (int)(i / g.J), j = (int)(i - (size_t)k * g.J);
        if (k == g.ko) {
            const double v = p.v_c[j];
            x = (p.kind_c == kBcFixed) ? v : x + p.coef_bc * (g.fdr_c[2][2] + 2.0 * g.rc[1]) * v;
        }
        if (k == g.ki) {
            const double v = p.v_i[j];
            x = (p.kind_i == kBcFixed) ? v : x + p.coef_bc * (g.fdr_i[0][2] + 2.0 * g.ri[1]) * v;
        }
        p.x[i] = x;
    }
}

__global__ void k_sc_diffusion(ScShared g, ScBatch b)
{
    const ScOne &p = b.s[blockIdx.y];
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < g.nnod;
         i += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(i / g.J), j = (int)(i - (size_t)k * g.J);
        const int J = g.J, ki = g.ki, ko = g.ko;
        const double vi = p.v_i[j], vc = p.v_c[j];
        // the field with the fixed-value walls in place
        auto xv = [&](int kk) -> double {
            if (kk == ki && p.kind_i == kBcFixed) return vi;
            if (kk == ko && p.kind_c == kBcFixed) return vc;
            return p.x[j + (size_t)J * kk];
        };
        const double g3 = g.g3[j];
        double out;
        if (k == ki) {
            if (p.kind_i == kBcFixed) {
                const double d1t = g.ffl_i[0][1] * vi + g.ffl_i[1][1] * xv(ki + 1) + g.ffl_i[2][1] * xv(ki + 2);
                const double d2t = g.ffl_i[0][2] * vi + g.ffl_i[1][2] * xv(ki + 1) + g.ffl_i[2][2] * xv(ki + 2);
                out = p.coef_d * (d2t + 2.0 * g.ri[1] * d1t - g3 * g.ri[2] * vi);
                p.x[i] = vi;
            } else {
                const double x0 = xv(ki);
                const double d2t = g.fdr_i[0][2] * vi + g.fdr_i[1][2] * x0 + g.fdr_i[2][2] * xv(ki + 1);
                out = p.coef_d * (d2t + 2.0 * g.ri[1] * vi - g3 * g.ri[2] * x0);
            }
        } else if (k == ko) {
            if (p.kind_c == kBcFixed) {
                const double d1t = g.ffl_c[2][1] * xv(ko - 2) + g.ffl_c[1][1] * xv(ko - 1) + g.ffl_c[0][1] * vc;
                const double d2t = g.ffl_c[2][2] * xv(ko - 2) + g.ffl_c[1][2] * xv(ko - 1) + g.ffl_c[0][2] * vc;
                out = p.coef_d * (d2t + 2.0 * g.rc[1] * d1t - g3 * g.rc[2] * vc);
                p.x[i] = vc;
            } else {
                const double x0 = xv(ko);
                const double d2t = g.fdr_c[0][2] * xv(ko - 1) + g.fdr_c[1][2] * x0 + g.fdr_c[2][2] * vc;
                out = p.coef_d * (d2t + 2.0 * g.rc[1] * vc - g3 * g.rc[2] * x0);
            }
        } else {
            const double xm = xv(k - 1), x0 = p.x[i], xp = xv(k + 1);
            const double *c1 = g.d1 + 3 * k, *c2 = g.d2 + 3 * k;
            const double d1s = c1[0] * xm + c1[1] * x0 + c1[2] * xp;
            const double d2s = c2[0] * xm + c2[1] * x0 + c2[2] * xp;
            out = p.coef_d * (d2s + 2.0 * g.ar1[k] * d1s - g3 * g.ar2[k] * x0);
        }
        p.diff[i] = out;
    }
}

// set_noize_scalar_sph: the radial profile prof[k] on the modes with m >= 0
// and l >= 1, zero elsewhere.
__global__ void k_sc_noise(int nr, int J, const int *l_of_j, const double *prof, double *x)
{
    const size_t nnod = (size_t)nr * J;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nnod;
         i += (size_t)gridDim.x * blockDim.x) {
        const int k = (int)(i / J), j = (int)(i - (size_t)k * J);
        const int l = l_of_j[j], m = j - l * (l + 1);
        x[i] = (l >= 1 && m >= 0) ? prof[k] : 0.0;
    }
}

int blocks_for(size_t n) { return (int)std::min<size_t>((n + 255) / 256, 16384); }

// device pointers of scalar q: x, diff, adv, pre
double *field_ptr(const ScalarShell *c, int q, int which)
{
    const size_t nnod = (size_t)c->nr * c->J;
    return c->d_fields + ((size_t)q * 4 + which) * nnod;
}

void shared_and_batch(const ScalarShell *c, ScShared *g, ScBatch *b)
{
    const int nr = c->nr, J = c->J;
    RadialFdm fd;
    radial_fdm(c->radius.data(), nr, &fd);
    *g = ScShared();
    g->nr = nr;
    g->J = J;
    g->ki = 0;
    g->ko = nr - 1;
    g->nnod = (size_t)nr * J;
    g->ar1 = c->d_const + nr;
    g->ar2 = c->d_const + 2 * nr;
    g->d1 = c->d_const + 3 * nr;
    g->d2 = c->d_const + 6 * nr;
    g->g3 = c->d_const + 9 * nr;
    std::memcpy(g->ffl_i, fd.ffl_i, sizeof g->ffl_i);
    std::memcpy(g->fdr_i, fd.fdr_i, sizeof g->fdr_i);
    std::memcpy(g->ffl_c, fd.ffl_c, sizeof g->ffl_c);
    std::memcpy(g->fdr_c, fd.fdr_c, sizeof g->fdr_c);
    std::memcpy(g->ri, fd.ri, sizeof g->ri);
    std::memcpy(g->rc, fd.rc, sizeof g->rc);
    g->dt = c->dt;
    *b = ScBatch();
    for (int q = 0; q < c->nsc; ++q) {
        const ScalarSpec &sp = c->spec[q];
        ScOne &o = b->s[q];
        o.x = field_ptr(c, q, 0);
        o.diff = field_ptr(c, q, 1);
        o.adv = field_ptr(c, q, 2);
        o.pre = field_ptr(c, q, 3);
        o.v_i = c->d_walls + (size_t)(2 * q) * J;
        o.v_c = c->d_walls + (size_t)(2 * q + 1) * J;
        o.kind_i = sp.bc_icb;
        o.kind_c = sp.bc_cmb;
        o.coef_d = sp.coef_d;
        o.coef_exp = 1.0 - sp.coef_imp;
        o.coef_bc = c->dt * sp.coef_imp * sp.coef_d;
    }
}

int diffusion(ScalarShell *c, std::string *err)
{
    ScShared g;
    ScBatch b;
    shared_and_batch(c, &g, &b);
    hipLaunchKernelGGL(k_sc_diffusion, dim3(blocks_for(g.nnod), c->nsc), dim3(256), 0, c->stream,
                       g, b);
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // namespace

void scalar_shell_free(ScalarShell *c)
{
    (void)hipFree(c->d_fields);
    (void)hipFree(c->d_walls);
    for (int q = 0; q < kScalarMax; ++q) {
        (void)hipFree(c->d_lu[q]);
        (void)hipFree(c->d_piv[q]);
        c->d_lu[q] = nullptr;
        c->d_piv[q] = nullptr;
    }
    (void)hipFree(c->d_const);
    (void)hipFree(c->d_l_of_j);
    (void)hipFree(c->d_ms);
    c->d_fields = c->d_walls = c->d_const = c->d_ms = nullptr;
    c->d_l_of_j = nullptr;
}

int scalar_shell_create(ScalarShell *c, int L, int nr, const double *radius, double dt,
                        const std::vector<ScalarSpec> &spec, const double *val_icb,
                        const double *val_cmb, hipStream_t s, std::string *err)
{
    const int nsc = (int)spec.size();
    if (L < 1 || nr < 5 || nsc < 1 || nsc > kScalarMax || !(dt > 0.0)) {
        *err = "scalar shell: need L >= 1, nr >= 5, 1..8 scalars and dt > 0";
        return -1;
    }
    if ((long long)nr * (L + 1) * (L + 1) >= (1LL << 31)) {
        *err = "scalar shell: nr (L+1)^2 must stay below 2^31 spectral nodes";
        return -1;
    }
    for (int k = 1; k < nr; ++k)
        if (!(radius[k] > radius[k - 1]) || !(radius[0] > 0.0)) {
            *err = "scalar shell: the radial grid must be positive and increasing";
            return -1;
        }
    for (const ScalarSpec &sp : spec)
        if ((sp.bc_icb != kBcFixed && sp.bc_icb != kBcFlux) ||
            (sp.bc_cmb != kBcFixed && sp.bc_cmb != kBcFlux)) {
            *err = "scalar shell: boundary kind must be 0 (fixed value) or 1 (fixed flux)";
            return -1;
        }
    c->L = L;
    c->J = (L + 1) * (L + 1);
    c->nr = nr;
    c->nsc = nsc;
    c->dt = dt;
    c->i_step = 0;
    c->stream = s;
    c->spec = spec;
    c->radius.assign(radius, radius + nr);
    const int J = c->J, ki = 0, ko = nr - 1;
    const size_t nnod = (size_t)nr * J;
    RadialFdm fd;
    radial_fdm(radius, nr, &fd);
    std::vector<double> ar1(nr), ar2(nr), g3(J);
    std::vector<int> l_of_j(J);
    for (int k = 0; k < nr; ++k) {
        ar1[k] = 1.0 / radius[k];
        ar2[k] = ar1[k] * ar1[k];
    }
    for (int j = 0; j < J; ++j) {
        int l = (int)std::sqrt((double)j);
        while ((l + 1) * (l + 1) <= j) ++l;
        while (l * l > j) --l;
        l_of_j[j] = l;
        g3[j] = (double)(l * (l + 1));
    }
    // constants: r, 1/r, 1/r^2, d1, d2, l(l+1)
    {
        std::vector<double> h;
        h.insert(h.end(), radius, radius + nr);
        h.insert(h.end(), ar1.begin(), ar1.end());
        h.insert(h.end(), ar2.begin(), ar2.end());
        h.insert(h.end(), fd.d1.begin(), fd.d1.end());
        h.insert(h.end(), fd.d2.begin(), fd.d2.end());
        h.insert(h.end(), g3.begin(), g3.end());
        HIP_TRY(hipMalloc(&c->d_const, sizeof(double) * h.size()));
        HIP_TRY(hipMemcpy(c->d_const, h.data(), sizeof(double) * h.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMalloc(&c->d_l_of_j, sizeof(int) * J));
        HIP_TRY(hipMemcpy(c->d_l_of_j, l_of_j.data(), sizeof(int) * J, hipMemcpyHostToDevice));
        HIP_TRY(hipMalloc(&c->d_ms, sizeof(double) * (J + 3)));
    }
    // wall values (the inner flux negated: inner_sph_bc_scalar_sph,
    // set_bc_sph_scalars.f90) and the flux-wall advection rows
    std::vector<double> walls(2 * (size_t)nsc * J), adv(nnod);
    HIP_TRY(hipMalloc(&c->d_fields, sizeof(double) * 4 * nnod * nsc));
    HIP_TRY(hipMemset(c->d_fields, 0, sizeof(double) * 4 * nnod * nsc));
    for (int q = 0; q < nsc; ++q) {
        const ScalarSpec &sp = spec[q];
        double *vi = walls.data() + (size_t)(2 * q) * J, *vc = vi + J;
        for (int j = 0; j < J; ++j) {
            vi[j] = (sp.bc_icb == kBcFlux) ? -val_icb[(size_t)q * J + j] : val_icb[(size_t)q * J + j];
            vc[j] = val_cmb[(size_t)q * J + j];
        }
        std::fill(adv.begin(), adv.end(), 0.0);
        for (int j = 0; j < J; ++j) {
            const double gm = std::max(g3[j], 0.5);
            if (sp.bc_icb == kBcFlux) adv[(size_t)ki * J + j] = (vi[j] - 0.0) * gm * fd.ri[2];
            if (sp.bc_cmb == kBcFlux) adv[(size_t)ko * J + j] = (vc[j] - 0.0) * gm * fd.rc[2];
        }
        HIP_TRY(hipMemcpy(field_ptr(c, q, 2), adv.data(), sizeof(double) * nnod,
                          hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMalloc(&c->d_walls, sizeof(double) * walls.size()));
    HIP_TRY(hipMemcpy(c->d_walls, walls.data(), sizeof(double) * walls.size(),
                      hipMemcpyHostToDevice));
    // Crank-Nicolson matrices (const_radial_mat_4_scalar_sph,
    // add_scalar_poisson_mat_sph, set_fix_fld_*_poisson_mat,
    // add_fix_flux_*_poisson_mat), band m[(j nr + k) 3 + b]
    std::vector<double> m(3 * nnod);
    auto M3 = [&](int j, int k, int b) -> double & { return m[((size_t)j * nr + k) * 3 + b]; };
    const std::vector<double> &d1 = fd.d1, &d2 = fd.d2;
    for (int q = 0; q < nsc; ++q) {
        const ScalarSpec &sp = spec[q];
        const double coef = sp.coef_imp * sp.coef_d * dt;
        std::fill(m.begin(), m.end(), 0.0);
        for (int j = 0; j < J; ++j)
            for (int k = 0; k < nr; ++k) M3(j, k, 1) = 1.0;
        for (int k = ki + 1; k < ko; ++k)
            for (int j = 0; j < J; ++j) {
                M3(j, k - 1, 2) = M3(j, k - 1, 2) - coef * (d2[3 * k + 0] + 2.0 * ar1[k] * d1[3 * k + 0]);
                M3(j, k, 1) = M3(j, k, 1)
                            - coef * (d2[3 * k + 1] + 2.0 * ar1[k] * d1[3 * k + 1] - g3[j] * ar2[k]);
                M3(j, k + 1, 0) = M3(j, k + 1, 0) - coef * (d2[3 * k + 2] + 2.0 * ar1[k] * d1[3 * k + 2]);
            }
        for (int j = 0; j < J; ++j) {
            if (sp.bc_icb == kBcFlux) {
                M3(j, ki, 1) = M3(j, ki, 1) - coef * (fd.fdr_i[1][2] - g3[j] * fd.ri[2]);
                M3(j, ki + 1, 0) = M3(j, ki + 1, 0) - coef * fd.fdr_i[2][2];
            } else {
                M3(j, ki, 1) = 1.0;
                M3(j, ki + 1, 0) = 0.0;
            }
            if (sp.bc_cmb == kBcFlux) {
                M3(j, ko - 1, 2) = M3(j, ko - 1, 2) - coef * fd.fdr_c[0][2];
                M3(j, ko, 1) = M3(j, ko, 1) - coef * (fd.fdr_c[1][2] - g3[j] * fd.rc[2]);
            } else {
                M3(j, ko - 1, 2) = 0.0;
                M3(j, ko, 1) = 1.0;
            }
        }
        if (factor_band_for_solve(3, nr, J, m.data(), &c->d_lu[q], &c->d_piv[q], s, err)) {
            *err = "scalar shell: " + *err;
            return -1;
        }
    }
    HIP_TRY(hipDeviceSynchronize());
    return 0;
}

int scalar_shell_set_noise(ScalarShell *c, std::string *err)
{
    const int nr = c->nr, ki = 0, ko = nr - 1;
    const std::vector<double> &r = c->radius;
    const double pi = 4.0 * std::atan(1.0);
    const double shell = r[ko] - r[ki];
    std::vector<double> prof(nr, 0.0);
    for (int k = ki + 2; k <= ko - 2; ++k) {
        const double xr = 2.0 * r[k] - (r[ki + 2] + r[ko - 2]) / shell;
        prof[k] = (1.0 - 3.0 * std::pow(xr, 2.0) + 3.0 * std::pow(xr, 4.0) - std::pow(xr, 6.0))
                * 1.0e-4 * 6.0 / std::sqrt(pi);
    }
    DevBuf d_prof;
    HIP_TRY(hipMalloc(&d_prof.p, sizeof(double) * nr));
    HIP_TRY(hipMemcpyAsync(d_prof.p, prof.data(), sizeof(double) * nr, hipMemcpyHostToDevice,
                           c->stream));
    const size_t nnod = (size_t)nr * c->J;
    for (int q = 0; q < c->nsc; ++q) {
        hipLaunchKernelGGL(k_sc_noise, dim3(blocks_for(nnod)), dim3(256), 0, c->stream, nr, c->J,
                           c->d_l_of_j, static_cast<const double *>(d_prof.p), field_ptr(c, q, 0));
        HIP_TRY(hipGetLastError());
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int scalar_shell_set_field(ScalarShell *c, int q, const double *host, std::string *err)
{
    if (q < 0 || q >= c->nsc) {
        *err = "scalar shell: scalar index out of range";
        return -1;
    }
    HIP_TRY(hipMemcpyAsync(field_ptr(c, q, 0), host, sizeof(double) * c->nr * (size_t)c->J,
                           hipMemcpyHostToDevice, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int scalar_shell_get_field(ScalarShell *c, int q, double *host, std::string *err)
{
    if (q < 0 || q >= 4 * c->nsc) {
        *err = "scalar shell: field index out of range (4 per scalar: value, diffusion, "
               "advection, Adams history)";
        return -1;
    }
    HIP_TRY(hipMemcpyAsync(host, field_ptr(c, q % c->nsc, q / c->nsc),
                           sizeof(double) * c->nr * (size_t)c->J, hipMemcpyDeviceToHost,
                           c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int scalar_shell_initialize(ScalarShell *c, std::string *err)
{
    c->i_step = 0;
    if (diffusion(c, err)) return -1;
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

int scalar_shell_step(ScalarShell *c, int nsteps, std::string *err)
{
    ScShared g;
    ScBatch b;
    shared_and_batch(c, &g, &b);
    const double *band[kScalarMax];
    const int *piv[kScalarMax];
    double *x[kScalarMax];
    for (int q = 0; q < c->nsc; ++q) {
        band[q] = c->d_lu[q];
        piv[q] = c->d_piv[q];
        x[q] = b.s[q].x;
    }
    for (int n = 0; n < nsteps; ++n) {
        const int first = c->i_step == 0;
        hipLaunchKernelGGL(k_sc_explicit, dim3(blocks_for(g.nnod), c->nsc), dim3(256), 0,
                           c->stream, g, b, first);
        HIP_TRY(hipGetLastError());
        if (lubksb_t_batch_dev(3, c->nr, c->J, c->nsc, band, piv, x, c->stream, err)) return -1;
        hipLaunchKernelGGL(k_sc_diffusion, dim3(blocks_for(g.nnod), c->nsc), dim3(256), 0,
                           c->stream, g, b);
        HIP_TRY(hipGetLastError());
        c->i_step += 1;
    }
    return 0;
}

int scalar_shell_mean_square(ScalarShell *c, double *out, std::string *err)
{
    const int nr = c->nr, J = c->J;
    const double *r = c->d_const;
    for (int q = 0; q < c->nsc; ++q) {
        if (mean_square_dev(2, nr, J, c->L, c->d_l_of_j, r, c->radius.data(), 1, nr,
                            field_ptr(c, q, 0), c->d_ms, c->d_ms + J, c->stream, err))
            return -1;
        HIP_TRY(hipMemcpyAsync(out + q, c->d_ms + J, sizeof(double), hipMemcpyDeviceToHost,
                               c->stream));
    }
    HIP_TRY(hipStreamSynchronize(c->stream));
    return 0;
}

// The volume mean square file of the evolved scalars, header and rows as
// write_sph_mean_sq_header / write_sph_vol_ms_file (m_rms_4_sph_spectr.f90:
// 245-280, output_sph_m_square_file.f90:134-156).
int scalar_shell_write_volume_ms(ScalarShell *c, const std::string &path, std::string *err)
{
    std::vector<double> e(c->nsc);
    if (scalar_shell_mean_square(c, e.data(), err)) return -1;
    bool append;
    {
        std::ifstream probe(path);
        append = probe.good();
    }
    std::ofstream f(path, append ? std::ios::app : std::ios::trunc);
    if (!f) {
        *err = "cannot open " + path;
        return -1;
    }
    if (!append) {
        f << "radial_layers, truncation\n" << fort_i(0, 16) << fort_i(c->L, 16) << "\n";
        f << "ICB_id, CMB_id\n" << fort_i(1, 16) << fort_i(c->nr, 16) << "\n";
        f << "number of components\n" << fort_i(c->nsc, 16) << fort_i(c->nsc, 16) << "\n";
        for (int q = 0; q < c->nsc; ++q) f << fort_i(1, 5);
        f << "\n" << "t_step    time    ";
        for (int q = 0; q < c->nsc; ++q)
            f << (q < (int)c->names.size() ? c->names[q] : "scalar_" + std::to_string(q + 1))
              << "    ";
        f << "\n";
    }
    f << fort_i(c->i_step, 16) << fort_e(c->time0 + c->i_step * c->dt, 23, 14);
    for (double v : e) f << fort_e(v, 23, 14);
    f << "\n";
    if (!f) {
        *err = "write error on " + path;
        return -1;
    }
    return 0;
}

}  // namespace calypso
