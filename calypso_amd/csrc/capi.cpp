// capi.cpp -- extern "C" boundary of libcalypso_amd.so (include/calypso_amd.h).
#include "../../include/calypso_amd.h"

#include <hip/hip_runtime.h>

#include <cctype>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "fft_internal.hpp"
#include "leg_internal.hpp"
#include "mhd_internal.hpp"
#include "radial_internal.hpp"
#include "scalar_internal.hpp"
#include "seeded.hpp"

using calypso::LegContext;

struct calypso_leg {
    LegContext c;
};

struct calypso_fft {
    calypso::FftContext f;
};

struct calypso_scalar_shell {
    calypso::ScalarShell c;
};
struct calypso_mhd {
    calypso::MhdContext c;
};

namespace {

thread_local std::string g_err;
calypso_leg *g_sel = nullptr;   // process-wide selector context

int fail(const std::string &m)
{
    g_err = m;
    return -1;
}

#define CHECK_HIP(expr)                                                   \
    do {                                                                  \
        hipError_t e_ = (expr);                                           \
        if (e_ != hipSuccess) return fail(std::string(#expr ": ") + hipGetErrorString(e_)); \
    } while (0)

int ensure(double **buf, size_t *cap, size_t n)
{
    if (*cap >= n) return 0;
    if (*buf) (void)hipFree(*buf);
    *buf = nullptr;
    *cap = 0;
    hipError_t e = hipMalloc(buf, n * sizeof(double));
    if (e != hipSuccess) return fail(std::string("hipMalloc staging: ") + hipGetErrorString(e));
    *cap = n;
    return 0;
}

int host_transform(calypso_leg *h, bool fwd, int ncomp, int nvec, int nscl, int n_WR,
                   int n_WS, const double *WR, double *WS)
{
    LegContext &c = h->c;
    const size_t nrlm = (size_t)ncomp * c.g.J * c.g.nr;
    const size_t nrtm = (size_t)ncomp * c.g.nth * c.g.nr * c.g.M;
    const size_t nin = fwd ? nrtm : nrlm, nout = fwd ? nrlm : nrtm;
    if ((size_t)n_WR < nin) return fail("n_WR smaller than ncomp * nnod of the input space");
    if ((size_t)n_WS < nout) return fail("n_WS smaller than ncomp * nnod of the output space");
    if (ensure(&c.d_rlm, &c.cap_rlm, nrlm) || ensure(&c.d_rtm, &c.cap_rtm, nrtm)) return -1;
    double *din = fwd ? c.d_rtm : c.d_rlm, *dout = fwd ? c.d_rlm : c.d_rtm;
    CHECK_HIP(hipMemcpyAsync(din, WR, nin * sizeof(double), hipMemcpyHostToDevice, c.stream));
    std::string err;
    int rc = fwd ? calypso::leg_forward_dev(&c, ncomp, nvec, nscl, din, dout, c.stream, &err)
                 : calypso::leg_backward_dev(&c, ncomp, nvec, nscl, din, dout, c.stream, &err);
    if (rc) return fail(err);
    CHECK_HIP(hipMemcpyAsync(WS, dout, nout * sizeof(double), hipMemcpyDeviceToHost, c.stream));
    CHECK_HIP(hipStreamSynchronize(c.stream));
    return 0;
}

// Process-wide FFT selector (init_MHD_FFT_select analogue): context, its
// stream and the device copies of one host call's buffers
struct FftSel {
    calypso_fft *f = nullptr;
    hipStream_t s = nullptr;
    double *d_rtm = nullptr, *d_rtp = nullptr;
    size_t cap_rtm = 0, cap_rtp = 0;
    size_t nnod_rtm = 0, nnod_rtp = 0;
    int ncomp = 0;
};
FftSel g_fsel;

// host band solves: device copies of one call's factors, pivots and modes
struct LuSel {
    double *d_lu = nullptr, *d_piv = nullptr, *d_x = nullptr;   // d_piv holds ints
    size_t cap_lu = 0, cap_piv = 0, cap_x = 0;
};
LuSel g_lusel;

// host grid products: device copies of one call's fields and forces
struct ProdSel {
    double *d_fld = nullptr, *d_frc = nullptr;
    size_t cap_fld = 0, cap_frc = 0;
};
ProdSel g_prodsel;

void fsel_free()
{
    if (g_fsel.f) {
        calypso::fft_context_free(&g_fsel.f->f);
        delete g_fsel.f;
    }
    if (g_fsel.d_rtm) (void)hipFree(g_fsel.d_rtm);
    if (g_fsel.d_rtp) (void)hipFree(g_fsel.d_rtp);
    if (g_fsel.s) (void)hipStreamDestroy(g_fsel.s);
    g_fsel = FftSel();
}

// lubksb_3band_mul / lubksb_5band_mul on host arrays: modes
// msmp_stack[0]+1 .. msmp_stack[msmp] (1-based) of x(mcomp, n), the union of
// the reference's per-thread ranges; the other modes are not touched
int host_lubksb(int nb, int msmp, const int *msmp_stack, int mcomp, int n,
                const double *band_lu, const int *i_pivot, double *x)
{
    if (msmp < 1 || !msmp_stack || mcomp < 1 || !band_lu || !i_pivot || !x)
        return fail("lubksb_band_mul: bad argument");
    for (int p = 0; p < msmp; ++p)
        if (msmp_stack[p] > msmp_stack[p + 1]) return fail("lubksb_band_mul: Msmp_stack not ascending");
    const int mst = msmp_stack[0], med = msmp_stack[msmp];
    if (mst < 0 || med > mcomp) return fail("lubksb_band_mul: Msmp_stack outside 0..mcomp");
    const int cnt = med - mst;
    if (cnt == 0) return 0;
    // a pivot outside 1..n would move a row from outside the matrix
    for (size_t q = (size_t)mst * n; q < (size_t)med * n; ++q)
        if (i_pivot[q] < 1 || i_pivot[q] > n) return fail("lubksb_band_mul: i_pivot outside 1..n");
    const size_t w = 2 * (size_t)nb - 1, nlu = w * n * cnt, nx = (size_t)n * cnt;
    LuSel &g = g_lusel;
    if (ensure(&g.d_lu, &g.cap_lu, nlu) || ensure(&g.d_piv, &g.cap_piv, (nx + 1) / 2) ||
        ensure(&g.d_x, &g.cap_x, nx))
        return -1;
    int *d_piv = reinterpret_cast<int *>(g.d_piv);
    CHECK_HIP(hipMemcpy(g.d_lu, band_lu + w * n * mst, nlu * sizeof(double), hipMemcpyHostToDevice));
    CHECK_HIP(hipMemcpy(d_piv, i_pivot + (size_t)n * mst, nx * sizeof(int), hipMemcpyHostToDevice));
    // x(mst+1:med, 1:n): rows of cnt modes at host pitch mcomp
    CHECK_HIP(hipMemcpy2D(g.d_x, cnt * sizeof(double), x + mst, mcomp * sizeof(double),
                          cnt * sizeof(double), n, hipMemcpyHostToDevice));
    std::string err;
    if (calypso::lubksb_dev(nb, n, cnt, g.d_lu, d_piv, g.d_x, nullptr, &err)) return fail(err);
    CHECK_HIP(hipMemcpy2D(x + mst, mcomp * sizeof(double), g.d_x, cnt * sizeof(double),
                          cnt * sizeof(double), n, hipMemcpyDeviceToHost));
    return 0;
}

}  // namespace

extern "C" {

const char *calypso_last_error(void) { return g_err.c_str(); }

int calypso_legendre_mode_id(const char *name)
{
    if (!name) return -1;
    std::string s;
    for (const char *p = name; *p; ++p) {
        if (std::isspace((unsigned char)*p)) continue;
        s.push_back((char)std::tolower((unsigned char)*p));
    }
    if (s == "hip_fp64") return CALYPSO_LEG_FP64;
    if (s == "hip_fp32") return CALYPSO_LEG_FP32;
    if (s == "hip_bf16x3") return CALYPSO_LEG_BF16X3;
    if (s == "hip_int8x6") return CALYPSO_LEG_INT8X6;
    return -1;
}

int calypso_gauss_points(int n, double *point, double *weight, double *colat)
{
    if (n < 1 || !point || !weight || !colat) return fail("calypso_gauss_points: bad arguments");
    calypso::gauss_points_host(n, point, weight, colat);
    return 0;
}

int calypso_leg_create(int l_truncation, int nth, int nr, const double *radius,
                       const double *colat, const double *weight, int ncomp, int nvector,
                       int nscalar, int mode, calypso_leg_t **out)
{
    if (!out || !radius) return fail("calypso_leg_create: null argument");
    *out = nullptr;
    calypso_leg *h = new (std::nothrow) calypso_leg();
    if (!h) return fail("calypso_leg_create: out of host memory");
    std::string err;
    if (calypso::leg_context_build(&h->c, l_truncation, nth, nr, radius, colat, weight, ncomp,
                                   nvector, nscalar, mode, &err)) {
        calypso::leg_context_free(&h->c);
        delete h;
        return fail(err);
    }
    *out = h;
    return 0;
}

int calypso_leg_destroy(calypso_leg_t *ctx)
{
    if (!ctx) return 0;
    calypso::leg_context_free(&ctx->c);
    delete ctx;
    return 0;
}

int calypso_leg_backward_device(calypso_leg_t *ctx, int ncomp, int nvector, int nscalar,
                                const double *d_sp_rlm, double *d_vr_rtm, void *stream)
{
    if (!ctx || !d_sp_rlm || !d_vr_rtm) return fail("calypso_leg_backward_device: null argument");
    std::string err;
    hipStream_t s = (hipStream_t)stream;
    if (calypso::leg_backward_dev(&ctx->c, ncomp, nvector, nscalar, d_sp_rlm, d_vr_rtm, s, &err))
        return fail(err);
    return 0;
}

int calypso_leg_forward_device(calypso_leg_t *ctx, int ncomp, int nvector, int nscalar,
                               const double *d_vr_rtm, double *d_sp_rlm, void *stream)
{
    if (!ctx || !d_sp_rlm || !d_vr_rtm) return fail("calypso_leg_forward_device: null argument");
    std::string err;
    hipStream_t s = (hipStream_t)stream;
    if (calypso::leg_forward_dev(&ctx->c, ncomp, nvector, nscalar, d_vr_rtm, d_sp_rlm, s, &err))
        return fail(err);
    return 0;
}

int calypso_leg_backward(calypso_leg_t *ctx, int ncomp, int nvector, int nscalar, int n_WR,
                         int n_WS, const double *WR, double *WS)
{
    if (!ctx || !WR || !WS) return fail("calypso_leg_backward: null argument");
    return host_transform(ctx, false, ncomp, nvector, nscalar, n_WR, n_WS, WR, WS);
}

int calypso_leg_forward(calypso_leg_t *ctx, int ncomp, int nvector, int nscalar, int n_WR,
                        int n_WS, const double *WR, double *WS)
{
    if (!ctx || !WR || !WS) return fail("calypso_leg_forward: null argument");
    return host_transform(ctx, true, ncomp, nvector, nscalar, n_WR, n_WS, WR, WS);
}

int calypso_leg_tables(calypso_leg_t *ctx, double *P, double *dP)
{
    if (!ctx || !P || !dP) return fail("calypso_leg_tables: null argument");
    std::string err;
    if (calypso::leg_tables_to_host(&ctx->c, P, dP, &err)) return fail(err);
    return 0;
}

int calypso_leg_info(calypso_leg_t *ctx, long long *info)
{
    if (!ctx || !info) return fail("calypso_leg_info: null argument");
    LegContext &c = ctx->c;
    std::string err;
    const calypso::LegPlan *pb = calypso::leg_plan(&c, c.ncomp, c.nvec, c.nscl, false, &err);
    const calypso::LegPlan *pf = calypso::leg_plan(&c, c.ncomp, c.nvec, c.nscl, true, &err);
    if (!pb || !pf) return fail(err);
    info[0] = c.g.L;
    info[1] = c.g.nth;
    info[2] = c.g.nr;
    info[3] = c.prec;
    info[4] = (long long)c.tab_bytes;
    info[5] = pb->KT;
    info[6] = pf->KT;
    info[7] = pb->n_items;
    return 0;
}

void *calypso_leg_stream(calypso_leg_t *ctx) { return ctx ? (void *)ctx->c.stream : nullptr; }

int calypso_leg_plan_info(calypso_leg_t *ctx, int ncomp, int nvector, int nscalar, int fwd,
                          long long *info)
{
    if (!ctx || !info) return fail("calypso_leg_plan_info: null argument");
    std::string err;
    const calypso::LegPlan *p = calypso::leg_plan(&ctx->c, ncomp, nvector, nscalar, fwd != 0, &err);
    if (!p) return fail(err);
    info[0] = p->KT;
    info[1] = p->mct1;
    info[2] = p->mct2;
    info[3] = p->rpw;
    info[4] = p->pair;
    info[5] = p->tc;
    info[6] = p->tuned_ms < 0 ? -1 : (long long)(p->tuned_ms * 1000.0f);
    info[7] = p->n_items;
    info[8] = p->ver;
    info[9] = p->nw;
    info[10] = p->pd;
    return 0;
}

int calypso_leg_plan_count(calypso_leg_t *ctx, int ncomp, int nvector, int nscalar, int fwd)
{
    if (!ctx) return fail("calypso_leg_plan_count: null argument");
    std::string err;
    const int n = calypso::leg_plan_count(&ctx->c, ncomp, nvector, nscalar, fwd != 0, &err);
    if (n < 0) return fail(err);
    return n;
}

int calypso_leg_force_plan(calypso_leg_t *ctx, int ncomp, int nvector, int nscalar, int fwd,
                           int idx)
{
    if (!ctx) return fail("calypso_leg_force_plan: null argument");
    std::string err;
    if (calypso::leg_force_plan(&ctx->c, ncomp, nvector, nscalar, fwd != 0, idx, &err))
        return fail(err);
    return 0;
}

int calypso_leg_set_autotune(calypso_leg_t *ctx, int on)
{
    if (!ctx) return fail("calypso_leg_set_autotune: null argument");
    ctx->c.autotune = on != 0;
    return 0;
}

int calypso_sel_init_legendre_trans(int mode, int l_truncation, int nth, int nr,
                                    const double *radius, const double *colat,
                                    const double *weight, int ncomp, int nvector, int nscalar)
{
    if (g_sel) calypso_leg_destroy(g_sel);
    g_sel = nullptr;
    return calypso_leg_create(l_truncation, nth, nr, radius, colat, weight, ncomp, nvector,
                              nscalar, mode, &g_sel);
}

int calypso_sel_backward_legendre_trans(int ncomp, int nvector, int nscalar, int n_WR, int n_WS,
                                        const double *WR, double *WS)
{
    if (!g_sel) return fail("sel_backward_legendre_trans: selector not initialised");
    return calypso_leg_backward(g_sel, ncomp, nvector, nscalar, n_WR, n_WS, WR, WS);
}

int calypso_sel_forward_legendre_trans(int ncomp, int nvector, int nscalar, int n_WR, int n_WS,
                                       const double *WR, double *WS)
{
    if (!g_sel) return fail("sel_forward_legendre_trans: selector not initialised");
    if (ncomp <= 0) return 0;  // legendre_transform_select.f90:391
    return calypso_leg_forward(g_sel, ncomp, nvector, nscalar, n_WR, n_WS, WR, WS);
}

int calypso_sel_finalize_legendre_trans(void)
{
    int rc = calypso_leg_destroy(g_sel);
    g_sel = nullptr;
    return rc;
}

int calypso_fft_create(int l_truncation, int nth, int nr, int nphi, calypso_fft_t **out)
{
    if (!out) return fail("calypso_fft_create: null argument");
    *out = nullptr;
    calypso_fft *h = new (std::nothrow) calypso_fft();
    if (!h) return fail("calypso_fft_create: out of host memory");
    std::string err;
    if (calypso::fft_context_build(&h->f, l_truncation, nth, nr, nphi, &err)) {
        calypso::fft_context_free(&h->f);
        delete h;
        return fail(err);
    }
    *out = h;
    return 0;
}

int calypso_fft_destroy(calypso_fft_t *fft)
{
    if (!fft) return 0;
    calypso::fft_context_free(&fft->f);
    delete fft;
    return 0;
}

int calypso_fft_backward_device(calypso_fft_t *fft, int ncomp, const double *d_vr_rtm,
                                double *d_fld_rtp, void *stream)
{
    if (!fft || !d_vr_rtm || !d_fld_rtp || ncomp < 1) return fail("calypso_fft_backward_device: bad argument");
    std::string err;
    if (calypso::fft_backward_dev(&fft->f, ncomp, d_vr_rtm, d_fld_rtp, (hipStream_t)stream, &err))
        return fail(err);
    return 0;
}

int calypso_fft_forward_device(calypso_fft_t *fft, int ncomp, const double *d_frc_rtp,
                               double *d_vr_rtm, void *stream)
{
    if (!fft || !d_vr_rtm || !d_frc_rtp || ncomp < 1) return fail("calypso_fft_forward_device: bad argument");
    std::string err;
    if (calypso::fft_forward_dev(&fft->f, ncomp, d_frc_rtp, d_vr_rtm, (hipStream_t)stream, &err))
        return fail(err);
    return 0;
}

int calypso_nonlinear_products_device(long long nnod, const double *d_fld, double *d_frc,
                                      const double *coef, void *stream)
{
    if (nnod < 1 || !d_fld || !d_frc || !coef) return fail("calypso_nonlinear_products_device: bad argument");
    std::string err;
    if (calypso::products_mhd_dev((size_t)nnod, d_fld, d_frc, coef, (hipStream_t)stream, &err))
        return fail(err);
    return 0;
}

int calypso_fft_mhd_fused_device(calypso_fft_t *fft, const double *d_bwd_rtm, double *d_fwd_rtm,
                                 const double *coef, void *stream)
{
    if (!fft || !d_bwd_rtm || !d_fwd_rtm || !coef) return fail("calypso_fft_mhd_fused_device: bad argument");
    std::string err;
    if (calypso::fft_mhd_fused_dev(&fft->f, d_bwd_rtm, d_fwd_rtm, coef, (hipStream_t)stream, &err))
        return fail(err);
    return 0;
}

int calypso_ludcmp_band_device(int nb, int n, int mcomp, const double *d_band_a, double *d_band_lu,
                               int *d_i_pivot, int *d_ierr, void *stream)
{
    if (!d_band_a || !d_band_lu || !d_i_pivot) return fail("calypso_ludcmp_band_device: null argument");
    std::string err;
    if (calypso::ludcmp_dev(nb, n, mcomp, d_band_a, d_band_lu, d_i_pivot, d_ierr, (hipStream_t)stream, &err))
        return fail(err);
    return 0;
}

int calypso_lubksb_band_mul_device(int nb, int n, int mcomp, const double *d_band_lu,
                                   const int *d_i_pivot, double *d_x, void *stream)
{
    if (!d_band_lu || !d_i_pivot || !d_x) return fail("calypso_lubksb_band_mul_device: null argument");
    std::string err;
    if (calypso::lubksb_dev(nb, n, mcomp, d_band_lu, d_i_pivot, d_x, (hipStream_t)stream, &err))
        return fail(err);
    return 0;
}

int calypso_explicit_update_device(int kind, int scheme, long long n, double dt, double coef_exp,
                                   double *d_x, const double *d_diffuse, const double *d_force,
                                   double *d_pre, void *stream)
{
    if (n < 1 || !d_x || !d_diffuse || !d_force || (!d_pre && scheme != 1))
        return fail("calypso_explicit_update_device: bad argument");
    std::string err;
    if (calypso::explicit_update_dev(kind, scheme, (size_t)n, dt, coef_exp, d_x, d_diffuse, d_force,
                                     d_pre, (hipStream_t)stream, &err))
        return fail(err);
    return 0;
}

int calypso_sph_mean_square(int kind, int nr, int J, const int *l_of_j, const double *radius,
                            int kr_st, int kr_ed, const double *d_f, double *out, void *stream)
{
    if (nr < 2 || J < 1 || !l_of_j || !radius || !d_f || !out)
        return fail("calypso_sph_mean_square: bad argument");
    int lmax = 0;
    for (int j = 0; j < J; ++j) lmax = l_of_j[j] > lmax ? l_of_j[j] : lmax;
    hipStream_t s = (hipStream_t)stream;
    int *d_l = nullptr;
    double *d_r = nullptr, *d_vol = nullptr, *d_out = nullptr;
    CHECK_HIP(hipMallocAsync(&d_l, sizeof(int) * J, s));
    CHECK_HIP(hipMallocAsync(&d_r, sizeof(double) * nr, s));
    CHECK_HIP(hipMallocAsync(&d_vol, sizeof(double) * 3 * (size_t)J, s));
    CHECK_HIP(hipMallocAsync(&d_out, sizeof(double) * 3, s));
    CHECK_HIP(hipMemcpyAsync(d_l, l_of_j, sizeof(int) * J, hipMemcpyHostToDevice, s));
    CHECK_HIP(hipMemcpyAsync(d_r, radius, sizeof(double) * nr, hipMemcpyHostToDevice, s));
    std::string err;
    int rc = calypso::mean_square_dev(kind, nr, J, lmax, d_l, d_r, radius, kr_st, kr_ed, d_f, d_vol,
                                      d_out, s, &err);
    double h[3] = {0, 0, 0};
    if (!rc) CHECK_HIP(hipMemcpyAsync(h, d_out, sizeof(double) * 3, hipMemcpyDeviceToHost, s));
    (void)hipFreeAsync(d_l, s);
    (void)hipFreeAsync(d_r, s);
    (void)hipFreeAsync(d_vol, s);
    (void)hipFreeAsync(d_out, s);
    CHECK_HIP(hipStreamSynchronize(s));
    if (rc) return fail(err);
    for (int c = 0; c < 3; ++c) out[c] = h[c];
    return 0;
}

int calypso_sph_mean_square_noalloc(int kind, int nr, int J, const int *l_of_j,
                                    const double *radius, int kr_st, int kr_ed, const double *d_f,
                                    double *out, void *stream)
{
    if (nr < 2 || J < 1 || !l_of_j || !radius || !d_f || !out)
        return fail("calypso_sph_mean_square_noalloc: bad argument");
    std::string err;
    if (calypso::mean_square_noalloc_dev(kind, nr, J, l_of_j, radius, kr_st, kr_ed, d_f, out,
                                         (hipStream_t)stream, &err))
        return fail("calypso_sph_mean_square_noalloc: " + err);
    return 0;
}

int calypso_mhd_create(const calypso_mhd_params *pp, const double *radius, void *stream,
                       calypso_mhd_t **out)
{
    if (!pp || !radius || !out) return fail("calypso_mhd_create: null argument");
    calypso::MhdCoefs co;
    co.dt = pp->dt;
    co.coef_imp_v = pp->coef_imp_v;
    co.coef_imp_t = pp->coef_imp_t;
    co.coef_imp_b = pp->coef_imp_b;
    co.coef_velo = pp->coef_velo;
    co.coef_d_velo = pp->coef_d_velo;
    co.coef_buo = pp->coef_buo;
    co.coef_cor = pp->coef_cor;
    co.coef_lor = pp->coef_lor;
    co.coef_temp = pp->coef_temp;
    co.coef_d_temp = pp->coef_d_temp;
    co.coef_magne = pp->coef_magne;
    co.coef_d_magne = pp->coef_d_magne;
    co.coef_induct = pp->coef_induct;
    co.depth_high_t = pp->depth_high_t;
    co.depth_low_t = pp->depth_low_t;
    co.high_temp = pp->high_temp;
    co.low_temp = pp->low_temp;
    co.bc_velo_icb = pp->bc_velo_icb;
    co.bc_velo_cmb = pp->bc_velo_cmb;
    co.bc_magne_icb = pp->bc_magne_icb;
    co.coef_press = pp->coef_press;
    co.kr_icb = pp->kr_icb > 0 ? pp->kr_icb - 1 : 0;
    co.kr_cmb = pp->kr_cmb > 0 ? pp->kr_cmb - 1 : -1;
    if (co.dt <= 0.0) return fail("calypso_mhd_create: dt must be positive");
    if (pp->legendre_mode < CALYPSO_LEG_FP64 || pp->legendre_mode > CALYPSO_LEG_INT8X6)
        return fail("calypso_mhd_create: unknown Legendre mode");
    if (pp->kr_icb < 0 || pp->kr_cmb < 0)
        return fail("calypso_mhd_create: kr_icb / kr_cmb must be 0 or a 1-based point");
    {
        std::string err;
        if (pp->nr < 5) return fail("calypso_mhd_create: nr must be at least 5");
        if (calypso::mhd_check_model(pp->nr, radius, co, &err)) return fail(err);
    }
    calypso_mhd *h = new (std::nothrow) calypso_mhd;
    if (!h) return fail("calypso_mhd_create: out of host memory");
    std::string err;
    if (calypso::mhd_create(&h->c, pp->l_truncation, pp->nth, pp->nphi, pp->nr, radius, co,
                            pp->legendre_mode, (hipStream_t)stream, &err)) {
        calypso::mhd_free(&h->c);
        delete h;
        return fail(err);
    }
    *out = h;
    return 0;
}

int calypso_mhd_destroy(calypso_mhd_t *mhd)
{
    if (!mhd) return 0;
    calypso::mhd_free(&mhd->c);
    delete mhd;
    return 0;
}

int calypso_mhd_set_field(calypso_mhd_t *mhd, int comp, const double *host)
{
    if (!mhd || !host) return fail("calypso_mhd_set_field: null argument");
    std::string err;
    if (calypso::mhd_field_io_caller(&mhd->c, comp, const_cast<double *>(host), false, &err))
        return fail(err);
    return 0;
}

int calypso_mhd_snapshot_rtp(calypso_mhd_t *mhd, int nvec, const int *vec_comp, int nscl,
                             const int *scl_comp, double *host_rtp)
{
    if (!mhd) return fail("calypso_mhd_snapshot_rtp: null argument");
    std::string err;
    if (calypso::mhd_snapshot_rtp(&mhd->c, nvec, vec_comp, nscl, scl_comp, host_rtp, &err))
        return fail(err);
    return 0;
}

int calypso_mhd_pressure(calypso_mhd_t *mhd, double *host)
{
    if (!mhd || !host) return fail("calypso_mhd_pressure: null argument");
    std::string err;
    if (calypso::mhd_pressure(&mhd->c, host, &err)) return fail(err);
    return 0;
}

int calypso_mhd_energy_fluxes(calypso_mhd_t *mhd, int nflux, const int *flux_ids, double *host)
{
    if (!mhd || !flux_ids || !host) return fail("calypso_mhd_energy_fluxes: null argument");
    std::string err;
    if (calypso::mhd_energy_fluxes(&mhd->c, nflux, flux_ids, host, &err)) return fail(err);
    return 0;
}

int calypso_mhd_get_field(calypso_mhd_t *mhd, int comp, double *host)
{
    if (!mhd || !host) return fail("calypso_mhd_get_field: null argument");
    std::string err;
    if (calypso::mhd_field_io_caller(&mhd->c, comp, host, true, &err)) return fail(err);
    return 0;
}

int calypso_mhd_set_dynamobench_ic(calypso_mhd_t *mhd, int isig)
{
    if (!mhd) return fail("calypso_mhd_set_dynamobench_ic: null handle");
    std::string err;
    if (calypso::mhd_set_dynamobench_ic(&mhd->c, isig, &err)) return fail(err);
    return 0;
}

int calypso_mhd_set_dynamobench2_ic(calypso_mhd_t *mhd, int isig)
{
    if (!mhd) return fail("calypso_mhd_set_dynamobench2_ic: null handle");
    std::string err;
    if (calypso::mhd_set_dynamobench2_ic(&mhd->c, isig, &err)) return fail(err);
    return 0;
}

int calypso_mhd_set_rj_order(calypso_mhd_t *mhd, int J, const int *j_global)
{
    if (!mhd || !j_global) return fail("calypso_mhd_set_rj_order: null argument");
    std::string err;
    if (calypso::mhd_set_rj_order(&mhd->c, J, j_global, &err)) return fail(err);
    return 0;
}

int calypso_mhd_set_seeded_ic(calypso_mhd_t *mhd, unsigned long long seed, double amp_u,
                              double amp_b, double amp_t)
{
    if (!mhd) return fail("calypso_mhd_set_seeded_ic: null handle");
    std::string err;
    if (calypso::mhd_set_seeded_ic(&mhd->c, seed, amp_u, amp_b, amp_t, &err)) return fail(err);
    return 0;
}

int calypso_sph_write_grid_files(const char *prefix, int l_truncation, int nth, int nphi,
                                 int nr, const double *radius)
{
    if (!prefix || !radius) return fail("calypso_sph_write_grid_files: null argument");
    std::string err;
    if (calypso::sph_write_grid_files(prefix, l_truncation, nth, nphi, nr, radius, &err))
        return fail("calypso_sph_write_grid_files: " + err);
    return 0;
}

int calypso_sph_rj_mode_order(int l_truncation, int *j_global)
{
    if (l_truncation < 0 || !j_global) return fail("calypso_sph_rj_mode_order: bad argument");
    calypso::sph_rj_mode_order(l_truncation, j_global);
    return 0;
}

int calypso_seeded_spectrum(unsigned long long seed, int field, int l_truncation, int nr,
                            const double *radius, double amplitude, double power, int envelope,
                            int l_min, double *out)
{
    if (l_truncation < 0 || nr < 1 || !radius || !out || (envelope && nr < 2))
        return fail("calypso_seeded_spectrum: bad argument");
    calypso::seeded_spectrum(seed, field, l_truncation, nr, radius, amplitude, power,
                             envelope != 0, l_min, out);
    return 0;
}

int calypso_mhd_initialize(calypso_mhd_t *mhd)
{
    if (!mhd) return fail("calypso_mhd_initialize: null handle");
    std::string err;
    if (calypso::mhd_initialize(&mhd->c, &err)) return fail(err);
    return 0;
}

int calypso_mhd_step(calypso_mhd_t *mhd, int nsteps)
{
    if (!mhd || nsteps < 0) return fail("calypso_mhd_step: bad argument");
    std::string err;
    if (calypso::mhd_steps(&mhd->c, nsteps, &err)) return fail(err);
    return 0;
}

int calypso_mhd_energies(calypso_mhd_t *mhd, double *out)
{
    if (!mhd || !out) return fail("calypso_mhd_energies: null argument");
    std::string err;
    if (calypso::mhd_energies(&mhd->c, out, &err)) return fail(err);
    return 0;
}

int calypso_mhd_info(calypso_mhd_t *mhd, long long *info)
{
    if (!mhd || !info) return fail("calypso_mhd_info: null argument");
    info[0] = mhd->c.i_step;
    info[1] = mhd->c.nr;
    info[2] = mhd->c.J;
    info[3] = calypso::kMhdNumComp;
    return 0;
}

int calypso_mhd_time(calypso_mhd_t *mhd, double *time)
{
    if (!mhd || !time) return fail("calypso_mhd_time: null argument");
    *time = calypso::mhd_time(&mhd->c);
    return 0;
}

int calypso_mhd_write_volume_ms(calypso_mhd_t *mhd, const char *path)
{
    if (!mhd || !path) return fail("calypso_mhd_write_volume_ms: null argument");
    std::string err;
    if (calypso::mhd_write_volume_ms(&mhd->c, path, &err)) return fail(err);
    return 0;
}

int calypso_mhd_write_restart(calypso_mhd_t *mhd, const char *prefix, int istep_fld,
                              char *name_out, int name_len)
{
    if (!mhd || !prefix) return fail("calypso_mhd_write_restart: null argument");
    const std::string name = calypso::mhd_restart_name(prefix, istep_fld);
    if (name_out && name_len > 0) {
        std::strncpy(name_out, name.c_str(), (size_t)name_len - 1);
        name_out[name_len - 1] = 0;
    }
    std::string err;
    if (calypso::mhd_write_restart(&mhd->c, name, &err)) return fail(err);
    return 0;
}

int calypso_mhd_read_restart(calypso_mhd_t *mhd, const char *path)
{
    if (!mhd || !path) return fail("calypso_mhd_read_restart: null argument");
    std::string err;
    if (calypso::mhd_read_restart(&mhd->c, path, &err)) return fail(err);
    return 0;
}

int calypso_mhd_set_option(calypso_mhd_t *mhd, const char *name, long long value)
{
    if (!mhd || !name) return fail("calypso_mhd_set_option: null argument");
    std::string err;
    if (calypso::mhd_set_option(&mhd->c, name, value, &err)) return fail(err);
    return 0;
}

int calypso_mhd_coriolis_table(int l_truncation, double *cor, int *idx)
{
    if (l_truncation < 1 || !cor || !idx) return fail("calypso_mhd_coriolis_table: bad argument");
    const int J = (l_truncation + 1) * (l_truncation + 1);
    calypso::coriolis_table(l_truncation, cor, idx, idx + J, idx + 2 * J);
    return 0;
}

int calypso_init_MHD_FFT_select(int l_truncation, int nth, int nr, int nphi, int ncomp,
                                int ncomp_fwd, int ncomp_bwd)
{
    fsel_free();
    if (ncomp < 1 || ncomp_fwd < 1 || ncomp_bwd < 1 || ncomp_fwd > ncomp || ncomp_bwd > ncomp)
        return fail("init_MHD_FFT_select: need 1 <= ncomp_fwd, ncomp_bwd <= ncomp");
    calypso_fft *h = nullptr;
    if (calypso_fft_create(l_truncation, nth, nr, nphi, &h)) return -1;
    g_fsel.f = h;
    g_fsel.ncomp = ncomp;
    g_fsel.nnod_rtm = (size_t)nth * nr * (2 * l_truncation + 1);
    g_fsel.nnod_rtp = (size_t)nth * nr * nphi;
    hipError_t e = hipStreamCreateWithFlags(&g_fsel.s, hipStreamNonBlocking);
    if (e != hipSuccess || ensure(&g_fsel.d_rtm, &g_fsel.cap_rtm, g_fsel.nnod_rtm * ncomp) ||
        ensure(&g_fsel.d_rtp, &g_fsel.cap_rtp, g_fsel.nnod_rtp * ncomp)) {
        if (e != hipSuccess) g_err = std::string("hipStreamCreateWithFlags: ") + hipGetErrorString(e);
        fsel_free();
        return -1;
    }
    return 0;
}

int calypso_back_MHD_FFT_sel_from_recv(int ncomp_bwd, int n_WR, const double *WR,
                                       double *fld_rtp)
{
    if (!g_fsel.f) return fail("back_MHD_FFT_sel_from_recv: FFT selector not initialised");
    if (ncomp_bwd < 1 || ncomp_bwd > g_fsel.ncomp || !WR || !fld_rtp)
        return fail("back_MHD_FFT_sel_from_recv: bad argument");
    const size_t nin = (size_t)ncomp_bwd * g_fsel.nnod_rtm, nout = (size_t)ncomp_bwd * g_fsel.nnod_rtp;
    if (n_WR < 0 || (size_t)n_WR < nin) return fail("back_MHD_FFT_sel_from_recv: n_WR smaller than ncomp_bwd * nnod_rtm");
    hipStream_t s = g_fsel.s;
    CHECK_HIP(hipMemcpyAsync(g_fsel.d_rtm, WR, nin * sizeof(double), hipMemcpyHostToDevice, s));
    std::string err;
    if (calypso::fft_backward_dev(&g_fsel.f->f, ncomp_bwd, g_fsel.d_rtm, g_fsel.d_rtp, s, &err))
        return fail(err);
    CHECK_HIP(hipMemcpyAsync(fld_rtp, g_fsel.d_rtp, nout * sizeof(double), hipMemcpyDeviceToHost, s));
    CHECK_HIP(hipStreamSynchronize(s));
    return 0;
}

int calypso_fwd_MHD_FFT_sel_from_recv(int ncomp_fwd, int n_WS, const double *frc_rtp,
                                      double *WS)
{
    if (!g_fsel.f) return fail("fwd_MHD_FFT_sel_from_recv: FFT selector not initialised");
    if (ncomp_fwd < 1 || ncomp_fwd > g_fsel.ncomp || !frc_rtp || !WS)
        return fail("fwd_MHD_FFT_sel_from_recv: bad argument");
    const size_t nin = (size_t)ncomp_fwd * g_fsel.nnod_rtp, nout = (size_t)ncomp_fwd * g_fsel.nnod_rtm;
    if (n_WS < 0 || (size_t)n_WS < nout) return fail("fwd_MHD_FFT_sel_from_recv: n_WS smaller than ncomp_fwd * nnod_rtm");
    hipStream_t s = g_fsel.s;
    CHECK_HIP(hipMemcpyAsync(g_fsel.d_rtp, frc_rtp, nin * sizeof(double), hipMemcpyHostToDevice, s));
    std::string err;
    if (calypso::fft_forward_dev(&g_fsel.f->f, ncomp_fwd, g_fsel.d_rtp, g_fsel.d_rtm, s, &err))
        return fail(err);
    CHECK_HIP(hipMemcpyAsync(WS, g_fsel.d_rtm, nout * sizeof(double), hipMemcpyDeviceToHost, s));
    CHECK_HIP(hipStreamSynchronize(s));
    return 0;
}

int calypso_nonlinear_products(long long nnod, const double *fld_rtp, double *frc_rtp,
                               const double *coef)
{
    if (nnod < 1 || !fld_rtp || !frc_rtp || !coef) return fail("calypso_nonlinear_products: bad argument");
    ProdSel &g = g_prodsel;
    const size_t nin = (size_t)13 * nnod, nout = (size_t)19 * nnod;
    if (ensure(&g.d_fld, &g.cap_fld, nin) || ensure(&g.d_frc, &g.cap_frc, nout)) return -1;
    CHECK_HIP(hipMemcpy(g.d_fld, fld_rtp, nin * sizeof(double), hipMemcpyHostToDevice));
    std::string err;
    if (calypso::products_mhd_dev((size_t)nnod, g.d_fld, g.d_frc, coef, nullptr, &err)) return fail(err);
    CHECK_HIP(hipMemcpy(frc_rtp, g.d_frc, nout * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int calypso_finalize_MHD_FFT_select(void)
{
    fsel_free();
    return 0;
}

int calypso_lubksb_3band_mul(int msmp, const int *msmp_stack, int mcomp, int n,
                             const double *band_lu, const int *i_pivot, double *x)
{
    return host_lubksb(3, msmp, msmp_stack, mcomp, n, band_lu, i_pivot, x);
}

int calypso_lubksb_5band_mul(int msmp, const int *msmp_stack, int mcomp, int n,
                             const double *band_lu, const int *i_pivot, double *x)
{
    return host_lubksb(5, msmp, msmp_stack, mcomp, n, band_lu, i_pivot, x);
}

int calypso_sph_write_field_file(const char *path, int rank, int istep, double time,
                                 double dt, long long nnod, int nfld, const int *ncomp,
                                 const char *const *names, const double *data)
{
    if (!path || (nfld > 0 && (!ncomp || !names || !data)) || nnod < 0 || nfld < 0)
        return fail("calypso_sph_write_field_file: null argument or negative size");
    std::vector<const double *> comp;
    long long ntot = 0;
    for (int i = 0; i < nfld; ++i) {
        if (ncomp[i] < 1 || ncomp[i] > 9 || !names[i])
            return fail("calypso_sph_write_field_file: 1..9 components and a name per field");
        ntot += ncomp[i];
    }
    for (long long c = 0; c < ntot; ++c) comp.push_back(data + c * nnod);
    std::string err;
    if (calypso::sph_write_field_file(path, rank, istep, time, dt, nnod, nfld, ncomp, names,
                                      comp.data(), &err))
        return fail(err);
    return 0;
}

// ---- scalars diffusing in a shell at rest -----------------------------------
int calypso_scalar_shell_create(int l_truncation, int nr, const double *radius, double dt,
                                int nscalar, const calypso_scalar_bc *scalars,
                                const double *val_icb, const double *val_cmb, void *stream,
                                calypso_scalar_shell_t **out)
{
    if (!radius || !scalars || !val_icb || !val_cmb || !out)
        return fail("calypso_scalar_shell_create: null argument");
    if (nscalar < 1 || nscalar > calypso::kScalarMax)
        return fail("calypso_scalar_shell_create: 1..8 scalars");
    std::vector<calypso::ScalarSpec> spec((size_t)nscalar);
    std::vector<std::string> names((size_t)nscalar);
    for (int q = 0; q < nscalar; ++q) {
        spec[q].coef_d = scalars[q].coef_d;
        spec[q].coef_imp = scalars[q].coef_imp;
        spec[q].bc_icb = scalars[q].bc_icb;
        spec[q].bc_cmb = scalars[q].bc_cmb;
        names[q] = scalars[q].name ? scalars[q].name : "scalar_" + std::to_string(q + 1);
    }
    calypso_scalar_shell *h = new (std::nothrow) calypso_scalar_shell;
    if (!h) return fail("calypso_scalar_shell_create: out of host memory");
    std::string err;
    if (calypso::scalar_shell_create(&h->c, l_truncation, nr, radius, dt, spec, val_icb, val_cmb,
                                     (hipStream_t)stream, &err)) {
        calypso::scalar_shell_free(&h->c);
        delete h;
        return fail(err);
    }
    h->c.names = names;
    *out = h;
    return 0;
}

int calypso_scalar_shell_destroy(calypso_scalar_shell_t *sh)
{
    if (!sh) return 0;
    calypso::scalar_shell_free(&sh->c);
    delete sh;
    return 0;
}

int calypso_scalar_shell_set_noise(calypso_scalar_shell_t *sh)
{
    if (!sh) return fail("calypso_scalar_shell_set_noise: null argument");
    std::string err;
    if (calypso::scalar_shell_set_noise(&sh->c, &err)) return fail(err);
    return 0;
}

int calypso_scalar_shell_set_field(calypso_scalar_shell_t *sh, int scalar, const double *host)
{
    if (!sh || !host) return fail("calypso_scalar_shell_set_field: null argument");
    std::string err;
    if (calypso::scalar_shell_set_field(&sh->c, scalar, host, &err)) return fail(err);
    return 0;
}

int calypso_scalar_shell_get_field(calypso_scalar_shell_t *sh, int field, double *host)
{
    if (!sh || !host) return fail("calypso_scalar_shell_get_field: null argument");
    std::string err;
    if (calypso::scalar_shell_get_field(&sh->c, field, host, &err)) return fail(err);
    return 0;
}

int calypso_scalar_shell_initialize(calypso_scalar_shell_t *sh)
{
    if (!sh) return fail("calypso_scalar_shell_initialize: null argument");
    std::string err;
    if (calypso::scalar_shell_initialize(&sh->c, &err)) return fail(err);
    return 0;
}

int calypso_scalar_shell_step(calypso_scalar_shell_t *sh, int nsteps)
{
    if (!sh) return fail("calypso_scalar_shell_step: null argument");
    if (nsteps < 0) return fail("calypso_scalar_shell_step: nsteps must be >= 0");
    std::string err;
    if (calypso::scalar_shell_step(&sh->c, nsteps, &err)) return fail(err);
    return 0;
}

int calypso_scalar_shell_mean_square(calypso_scalar_shell_t *sh, double *out)
{
    if (!sh || !out) return fail("calypso_scalar_shell_mean_square: null argument");
    std::string err;
    if (calypso::scalar_shell_mean_square(&sh->c, out, &err)) return fail(err);
    return 0;
}

int calypso_scalar_shell_write_volume_ms(calypso_scalar_shell_t *sh, const char *path)
{
    if (!sh || !path) return fail("calypso_scalar_shell_write_volume_ms: null argument");
    std::string err;
    if (calypso::scalar_shell_write_volume_ms(&sh->c, path, &err)) return fail(err);
    return 0;
}

}  // extern "C"
