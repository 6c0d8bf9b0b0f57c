// radial.hip -- FP64 radial-direction kernels of the spectral time step
// (gfx950): banded LU factorisation and back substitution of the
// Crank-Nicolson matrices, one thread per spherical-harmonic mode.
//
// The factorisations follow the reference's partial-pivoting rules
//   ludcmp_3band  (direct_solver/m_ludcmp_3band.f90:76-173)
//   ludcmp_band   (direct_solver/m_ludcmp_band.f90:83-207), nband = 5
// and the solves
//   lubksb_3band_mul / lubksb_5band_mul (direct_solver/lubksb_357band_mul.f90:73-195)
// with the same operation order and no FMA contraction, so results are
// bitwise those of the reference.  Arrays keep the reference shapes:
// band_a(nb, n, mcomp), band_lu(2nb-1, n, mcomp), i_pivot(n, mcomp),
// x(mcomp, n) (mode fastest, as d_rj stores a field).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <mutex>
#include <string>
#include <vector>

#include "hip_check.hpp"
#include "radial_internal.hpp"

namespace calypso {


#pragma clang fp contract(off)

// 1-based accessors matching the reference indexing
#define A3(k, i) ba[(size_t)((k) - 1) + 3 * (size_t)((i) - 1)]
#define LU5(k, i) lu[(size_t)((k) - 1) + 5 * (size_t)((i) - 1)]

__global__ void k_ludcmp_3band(int n, int mcomp, const double *band_a, double *band_lu,
                               int *i_pivot, int *ierr, double *wk_all, int *iwk_all)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= mcomp) return;
    const double TINY = 1.0e-30;
    const double *ba = band_a + (size_t)3 * n * m;
    double *lu = band_lu + (size_t)5 * n * m;
    int *ipiv = i_pivot + (size_t)n * m;
    double *vwk = wk_all + (size_t)(n + 1) * m;   // 1-based
    int *idx = iwk_all + (size_t)(n + 1) * m;
    for (int k = 0; k < 5 * n; ++k) lu[k] = 0.0;
    LU5(3, 1) = A3(2, 1);
    LU5(4, 1) = A3(3, 1);
    for (int j = 2; j <= n - 1; ++j) {
        LU5(2, j) = A3(1, j);
        LU5(3, j) = A3(2, j);
        LU5(4, j) = A3(3, j);
    }
    LU5(2, n) = A3(1, n);
    LU5(3, n) = A3(2, n);
    for (int i = 1; i <= n; ++i) {
        double aamax = fabs(LU5(3, i));
        if (i < n) aamax = fmax(aamax, fabs(LU5(2, i + 1)));
        if (i > 1) aamax = fmax(aamax, fabs(LU5(4, i - 1)));
        if (aamax == 0.0) {
            ierr[m] = 100;   // singular, as the reference reports it
            return;
        }
        vwk[i] = 1.0 / aamax;
        idx[i] = i;
    }
    for (int j = 1; j <= n; ++j) {
        if (j == 2) {
            LU5(3, j) = LU5(3, j) - LU5(4, j - 1) * LU5(2, j);
        } else if (j > 2) {
            LU5(2, j) = LU5(2, j) - LU5(4, j - 2) * LU5(1, j);
            LU5(3, j) = LU5(3, j) - LU5(5, j - 2) * LU5(1, j) - LU5(4, j - 1) * LU5(2, j);
        }
        ipiv[j - 1] = j;
        if (j < n) {
            double aamax = vwk[j] * fabs(LU5(3, j));
            if (j > 1) LU5(4, j) = LU5(4, j) - LU5(5, j - 1) * LU5(2, j);
            double dum = vwk[j + 1] * fabs(LU5(4, j));
            if (dum >= aamax && j <= idx[j]) {
                int k = idx[j + 1];
                idx[j + 1] = idx[j];
                idx[j] = k;
                ipiv[j - 1] = j + 1;
                const int kst = j - 1 > 1 ? j - 1 : 1;
                const int ked = j + 2 < n ? j + 2 : n;
                for (k = kst; k <= ked; ++k) {
                    double t = LU5(j + 4 - k, k);
                    LU5(j + 4 - k, k) = LU5(j + 3 - k, k);
                    LU5(j + 3 - k, k) = t;
                }
                vwk[j + 1] = vwk[j];
            }
        }
        if (LU5(3, n) == 0.0) LU5(3, n) = TINY;
        double dum = 1.0 / LU5(3, j);
        if (j < n) LU5(4, j) = LU5(4, j) * dum;
    }
    ierr[m] = 0;
}
#undef A3
#undef LU5

#define BA(k, i) ba[(size_t)((k) - 1) + 5 * (size_t)((i) - 1)]
#define BL(k, i) bl[(size_t)((k) - 1) + 9 * (size_t)((i) - 1)]

__global__ void k_ludcmp_5band(int n, int mcomp, const double *band_a, double *band_lu,
                               int *indx_all, double *wk_all, int *iwk_all)
{
    const int mc = blockIdx.x * blockDim.x + threadIdx.x;
    if (mc >= mcomp) return;
    const double TINY = 1.0e-30;
    const int nb1 = 2, nb2 = 4;
    const double *ba = band_a + (size_t)5 * n * mc;
    double *bl = band_lu + (size_t)9 * n * mc;
    int *indx = indx_all + (size_t)n * mc;
    double *vv = wk_all + (size_t)(n + 1) * mc;
    int *idx = iwk_all + (size_t)(n + 1) * mc;
    for (int i = 1; i <= n; ++i) {
        const int jst = i - nb1 > 1 ? i - nb1 : 1;
        const int jed = i + nb1 < n ? i + nb1 : n;
        double aamax = fabs(BA(i - jst + nb1 + 1, jst));
        for (int j = jst + 1; j <= jed; ++j) aamax = fmax(aamax, fabs(BA(i - j + nb1 + 1, j)));
        vv[i] = 1.0 / aamax;
        idx[i] = i;
    }
    for (int k = 0; k < 9 * n; ++k) bl[k] = 0.0;
    for (int i = 1; i <= n; ++i) {
        const int jst = i - nb1 > 1 ? i - nb1 : 1;
        const int jed = i + nb1 < n ? i + nb1 : n;
        for (int j = jst; j <= jed; ++j) BL(i - j + nb2 + 1, j) = BA(i - j + nb1 + 1, j);
    }
    for (int j = 1; j <= n; ++j) {
        const int ist = j - nb2 > 1 ? j - nb2 : 1;
        const int ied = j + nb1 < n ? j + nb1 : n;
        for (int i = ist; i <= j - 1; ++i) {
            const int l = i - j + nb2 + 1;
            for (int k = ist; k <= i - 1; ++k)
                BL(l, j) = BL(l, j) - BL(i - k + nb2 + 1, k) * BL(k - j + nb2 + 1, j);
        }
        double aamax = 0.0;
        int imax = j;
        for (int i = j; i <= ied; ++i) {
            const int l = i - j + nb2 + 1;
            const int kst = i - nb2 > 1 ? i - nb2 : 1;
            for (int k = kst; k <= j - 1; ++k)
                BL(l, j) = BL(l, j) - BL(i - k + nb2 + 1, k) * BL(k - j + nb2 + 1, j);
            const double dum = vv[i] * fabs(BL(l, j));
            if (dum >= aamax && i <= (idx[j] + nb1)) {
                imax = i;
                aamax = dum;
            }
        }
        if (j != imax) {
            int k = idx[imax];
            idx[imax] = idx[j];
            idx[j] = k;
            const int kst = j - nb2 > 1 ? j - nb2 : 1;
            const int ked = imax + nb2 < n ? imax + nb2 : n;
            for (k = kst; k <= ked; ++k) {
                double t = BL(imax - k + nb2 + 1, k);
                BL(imax - k + nb2 + 1, k) = BL(j - k + nb2 + 1, k);
                BL(j - k + nb2 + 1, k) = t;
            }
            vv[imax] = vv[j];
        }
        indx[j - 1] = imax;
        if (BL(nb2 + 1, j) == 0.0) BL(nb2 + 1, j) = TINY;
        if (j != n) {
            const double dum = 1.0 / BL(nb2 + 1, j);
            const int ied2 = j + nb2 < n ? j + nb2 : n;
            for (int i = j + 1; i <= ied2; ++i) BL(i - j + nb2 + 1, j) = BL(i - j + nb2 + 1, j) * dum;
        }
    }
}
#undef BA
#undef BL

// lubksb_3band_mul / lubksb_5band_mul for x(mcomp, n), one thread per mode;
// x is read and written with the mode index fastest (coalesced).
template <int NB>
__global__ void k_lubksb_mul(int n, int mcomp, const double *band_lu, const int *i_pivot,
                             double *x)
{
    constexpr int W = (NB == 3) ? 5 : 9;
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= mcomp) return;
    const double *bl = band_lu + (size_t)W * n * m;
#define BL(k, i) bl[(size_t)((k) - 1) + (size_t)W * ((i) - 1)]
#define X(i) x[(size_t)m + (size_t)mcomp * ((i) - 1)]
    for (int i = 1; i <= n; ++i) {
        const int ll = i_pivot[(size_t)(i - 1) + (size_t)n * m];
        const double s = X(ll);
        X(ll) = X(i);
        X(i) = s;
    }
    if (NB == 3) {
        X(2) = X(2) - BL(4, 1) * X(1);
        for (int i = 3; i <= n; ++i) X(i) = X(i) - BL(5, i - 2) * X(i - 2) - BL(4, i - 1) * X(i - 1);
        X(n) = X(n) / BL(3, n);
        X(n - 1) = (X(n - 1) - BL(2, n) * X(n)) / BL(3, n - 1);
        for (int i = n - 2; i >= 1; --i)
            X(i) = (X(i) - BL(2, i + 1) * X(i + 1) - BL(1, i + 2) * X(i + 2)) / BL(3, i);
    } else {
        X(2) = X(2) - BL(6, 1) * X(1);
        X(3) = X(3) - BL(7, 1) * X(1) - BL(6, 2) * X(2);
        X(4) = X(4) - BL(8, 1) * X(1) - BL(7, 2) * X(2) - BL(6, 3) * X(3);
        for (int i = 5; i <= n; ++i)
            X(i) = X(i) - BL(9, i - 4) * X(i - 4) - BL(8, i - 3) * X(i - 3)
                 - BL(7, i - 2) * X(i - 2) - BL(6, i - 1) * X(i - 1);
        X(n) = X(n) / BL(5, n);
        X(n - 1) = (X(n - 1) - BL(4, n) * X(n)) / BL(5, n - 1);
        X(n - 2) = (X(n - 2) - BL(4, n - 1) * X(n - 1) - BL(3, n) * X(n)) / BL(5, n - 2);
        X(n - 3) = (X(n - 3) - BL(4, n - 2) * X(n - 2) - BL(3, n - 1) * X(n - 1)
                    - BL(2, n) * X(n)) / BL(5, n - 3);
        for (int i = n - 4; i >= 1; --i)
            X(i) = (X(i) - BL(4, i + 1) * X(i + 1) - BL(3, i + 2) * X(i + 2)
                    - BL(2, i + 3) * X(i + 3) - BL(1, i + 4) * X(i + 4)) / BL(5, i);
    }
#undef BL
#undef X
}

// The same substitution on factors stored mode-fastest: band_t[((i-1) W +
// (k-1)) mcomp + m] = BL(k, i) of mode m and ipiv_t[(i-1) mcomp + m], so that
// every load of a wavefront is one contiguous run (the reference layout
// band_lu(W, n, mcomp) puts W n doubles between neighbouring modes).
// The recurrences run in registers: the row interchanges (the reference
// applies them all first) are applied in a window of the next H = (NB-1)/2
// rows during the forward sweep, which is the same because interchange i
// only touches rows i..i+H (checked on the host when the factors are made);
// the loads of the next block of KB rows are issued before the current
// block is computed.  Operation order per row is the reference's, so the
// results are bitwise those of k_lubksb_mul.
constexpr int kLuBlock = 8;
constexpr int kLuBatchMax = 8;

template <int NB>
__device__ __forceinline__ void lubksb_t_mode(int n, int mcomp, int m,
                                              const double *__restrict__ band_t,
                                              const int *__restrict__ ipiv_t,
                                              double *__restrict__ x)
{
    constexpr int W = 2 * NB - 1, KL = NB - 1, H = (NB - 1) / 2, KB = kLuBlock;
    const size_t mc = (size_t)mcomp;
    auto BL = [&](int k, int i) { return band_t[((size_t)(i - 1) * W + (k - 1)) * mc + m]; };
    auto XP = [&](int i) { return x + (size_t)m + mc * (i - 1); };

    // ---- interchanges + forward substitution --------------------------
    double win[H + 1];          // rows i..i+H after the interchanges before i
#pragma unroll
    for (int h = 0; h <= H; ++h) win[h] = (1 + h <= n) ? *XP(1 + h) : 0.0;
    double yh[KL];              // y(i-1), y(i-2), ...
#pragma unroll
    for (int d = 0; d < KL; ++d) yh[d] = 0.0;
    int pv[2][KB];
    double lb[2][KB][KL], xr[2][KB];
    auto load_f = [&](int buf, int i0) {
#pragma unroll
        for (int u = 0; u < KB; ++u) {
            const int i = i0 + u;
            pv[buf][u] = (i <= n) ? ipiv_t[(size_t)(i - 1) * mc + m] : i;
            xr[buf][u] = (i + 1 + H <= n) ? *XP(i + 1 + H) : 0.0;
#pragma unroll
            for (int d = 1; d <= KL; ++d) lb[buf][u][d - 1] = (i <= n && i - d >= 1) ? BL(NB + d, i - d) : 0.0;
        }
    };
    auto step_f = [&](int buf, int i0) {
#pragma unroll
        for (int u = 0; u < KB; ++u) {
            const int i = i0 + u;
            if (i > n) break;
            const int dl = pv[buf][u] - i;   // 0..H
            if (H >= 1 && dl == 1) { const double t = win[0]; win[0] = win[1]; win[1] = t; }
            if (H >= 2 && dl == 2) { const double t = win[0]; win[0] = win[H]; win[H] = t; }
            double y = win[0];
#pragma unroll
            for (int d = KL; d >= 1; --d)
                if (d <= i - 1) y = y - lb[buf][u][d - 1] * yh[d - 1];
            *XP(i) = y;
#pragma unroll
            for (int d = KL - 1; d >= 1; --d) yh[d] = yh[d - 1];
            yh[0] = y;
#pragma unroll
            for (int h = 0; h < H; ++h) win[h] = win[h + 1];
            win[H] = xr[buf][u];
        }
    };
    // two blocks per iteration, so that both buffers are indexed statically
    load_f(0, 1);
    for (int i0 = 1; i0 <= n; i0 += 2 * KB) {
        if (i0 + KB <= n) load_f(1, i0 + KB);
        step_f(0, i0);
        if (i0 + KB > n) break;
        if (i0 + 2 * KB <= n) load_f(0, i0 + 2 * KB);
        step_f(1, i0 + KB);
    }

    // ---- back substitution ----------------------------------------------
    double zh[KL];              // z(i+1), z(i+2), ...
#pragma unroll
    for (int d = 0; d < KL; ++d) zh[d] = 0.0;
    double ub[2][KB][KL], dg[2][KB], yb[2][KB];
    auto load_b = [&](int buf, int i0) {   // rows i0, i0-1, ...
#pragma unroll
        for (int u = 0; u < KB; ++u) {
            const int i = i0 - u;
            yb[buf][u] = (i >= 1) ? *XP(i) : 0.0;
            dg[buf][u] = (i >= 1) ? BL(NB, i) : 1.0;
#pragma unroll
            for (int d = 1; d <= KL; ++d) ub[buf][u][d - 1] = (i >= 1 && i + d <= n) ? BL(NB - d, i + d) : 0.0;
        }
    };
    auto step_b = [&](int buf, int i0) {
#pragma unroll
        for (int u = 0; u < KB; ++u) {
            const int i = i0 - u;
            if (i < 1) break;
            double z = yb[buf][u];
#pragma unroll
            for (int d = 1; d <= KL; ++d)
                if (i + d <= n) z = z - ub[buf][u][d - 1] * zh[d - 1];
            z = z / dg[buf][u];
            *XP(i) = z;
#pragma unroll
            for (int d = KL - 1; d >= 1; --d) zh[d] = zh[d - 1];
            zh[0] = z;
        }
    };
    load_b(0, n);
    for (int i0 = n; i0 >= 1; i0 -= 2 * KB) {
        if (i0 - KB >= 1) load_b(1, i0 - KB);
        step_b(0, i0);
        if (i0 - KB < 1) break;
        if (i0 - 2 * KB >= 1) load_b(0, i0 - 2 * KB);
        step_b(1, i0 - KB);
    }
}

template <int NB>
__global__ __launch_bounds__(128) void k_lubksb_mul_t(int n, int mcomp, const double *band_t,
                                                      const int *ipiv_t, double *x)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= mcomp) return;
    lubksb_t_mode<NB>(n, mcomp, m, band_t, ipiv_t, x);
}

// several independent systems of the same band width in one launch
// (blockIdx.y selects the system)
struct LuBatch {
    const double *band[kLuBatchMax];
    const int *ipiv[kLuBatchMax];
    double *x[kLuBatchMax];
};

template <int NB>
__global__ __launch_bounds__(128) void k_lubksb_mul_t_batch(int n, int mcomp, LuBatch b)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= mcomp) return;
    const int q = blockIdx.y;
    lubksb_t_mode<NB>(n, mcomp, m, b.band[q], b.ipiv[q], b.x[q]);
}

// reference layout band_lu(W, n, mcomp), i_pivot(n, mcomp) -> mode-fastest
__global__ void k_lu_transpose(int W, int n, int mcomp, const double *band_lu, const int *ipiv,
                               double *band_t, int *ipiv_t)
{
    const size_t tot = (size_t)W * n * mcomp;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < tot;
         e += (size_t)gridDim.x * blockDim.x) {
        const int m = (int)(e % mcomp);
        const size_t r = e / mcomp;   // (i-1) W + (k-1)
        band_t[e] = band_lu[r + (size_t)W * n * m];
        if (r < (size_t)n) ipiv_t[e] = ipiv[r + (size_t)n * m];
    }
}

// ----------------------------------------------------------------- host --

int lu_transpose_dev(int nb, int n, int mcomp, const double *d_band_lu, const int *d_ipiv,
                     double *d_band_t, int *d_ipiv_t, hipStream_t s, std::string *err)
{
    if (nb != 3 && nb != 5) { *err = "lu_transpose: nb must be 3 or 5"; return -1; }
    const int W = 2 * nb - 1;
    const size_t tot = (size_t)W * n * mcomp;
    const int blocks = (int)std::min<size_t>((tot + 255) / 256, 65536);
    hipLaunchKernelGGL(k_lu_transpose, dim3(blocks), dim3(256), 0, s, W, n, mcomp, d_band_lu,
                       d_ipiv, d_band_t, d_ipiv_t);
    HIP_TRY(hipGetLastError());
    return 0;
}

int lubksb_t_dev(int nb, int n, int mcomp, const double *d_band_t, const int *d_ipiv_t,
                 double *d_x, hipStream_t s, std::string *err)
{
    if ((nb != 3 && nb != 5) || n < 5 || mcomp < 1) {
        *err = "lubksb: nb must be 3 or 5, n >= 5";
        return -1;
    }
    const int blocks = (mcomp + 127) / 128;
    if (nb == 3)
        hipLaunchKernelGGL(k_lubksb_mul_t<3>, dim3(blocks), dim3(128), 0, s, n, mcomp, d_band_t,
                           d_ipiv_t, d_x);
    else
        hipLaunchKernelGGL(k_lubksb_mul_t<5>, dim3(blocks), dim3(128), 0, s, n, mcomp, d_band_t,
                           d_ipiv_t, d_x);
    HIP_TRY(hipGetLastError());
    return 0;
}

int lubksb_t_batch_dev(int nb, int n, int mcomp, int count, const double *const *d_band_t,
                       const int *const *d_ipiv_t, double *const *d_x, hipStream_t s,
                       std::string *err)
{
    if ((nb != 3 && nb != 5) || n < 5 || mcomp < 1 || count < 1 || count > kLuBatchMax) {
        *err = "lubksb batch: nb must be 3 or 5, n >= 5, 1..8 systems";
        return -1;
    }
    LuBatch b = {};
    for (int q = 0; q < count; ++q) {
        b.band[q] = d_band_t[q];
        b.ipiv[q] = d_ipiv_t[q];
        b.x[q] = d_x[q];
    }
    const dim3 grid((mcomp + 127) / 128, count);
    if (nb == 3)
        hipLaunchKernelGGL(k_lubksb_mul_t_batch<3>, grid, dim3(128), 0, s, n, mcomp, b);
    else
        hipLaunchKernelGGL(k_lubksb_mul_t_batch<5>, grid, dim3(128), 0, s, n, mcomp, b);
    HIP_TRY(hipGetLastError());
    return 0;
}

int ludcmp_dev(int nb, int n, int mcomp, const double *d_band_a, double *d_band_lu,
               int *d_ipiv, int *d_ierr, hipStream_t s, std::string *err)
{
    if ((nb != 3 && nb != 5) || n < 5 || mcomp < 1) {
        *err = "ludcmp: nb must be 3 or 5, n >= 5";
        return -1;
    }
    double *wk = nullptr;
    int *iwk = nullptr;
    HIP_TRY(hipMallocAsync(&wk, sizeof(double) * (size_t)(n + 1) * mcomp, s));
    HIP_TRY(hipMallocAsync(&iwk, sizeof(int) * (size_t)(n + 1) * mcomp, s));
    const int blocks = (mcomp + 127) / 128;
    if (nb == 3) {
        hipLaunchKernelGGL(k_ludcmp_3band, dim3(blocks), dim3(128), 0, s, n, mcomp, d_band_a,
                           d_band_lu, d_ipiv, d_ierr, wk, iwk);
    } else {
        hipLaunchKernelGGL(k_ludcmp_5band, dim3(blocks), dim3(128), 0, s, n, mcomp, d_band_a,
                           d_band_lu, d_ipiv, wk, iwk);
        if (d_ierr) HIP_TRY(hipMemsetAsync(d_ierr, 0, sizeof(int) * mcomp, s));
    }
    HIP_TRY(hipGetLastError());
    HIP_TRY(hipFreeAsync(wk, s));
    HIP_TRY(hipFreeAsync(iwk, s));
    return 0;
}

// LU factors of nb-band matrices h_band[(j n + k) nb + b] (Fortran
// band(b+1, k+1, j+1)), ready for lubksb_t_dev / lubksb_t_batch_dev: factored
// on the device with the reference's pivoting, checked (no singular mode, no
// row interchange beyond the solver's register window of (nb-1)/2 rows) and
// transposed to the mode-fastest layout.  *d_lu_t / *d_piv_t are allocated
// here and owned by the caller; nothing is left allocated on failure.
int factor_band_for_solve(int nb, int n, int mcomp, const double *h_band, double **d_lu_t,
                          int **d_piv_t, hipStream_t s, std::string *err)
{
    const size_t nnod = (size_t)n * mcomp;
    DevBuf a, lu, piv, ierr_d, lu_t, piv_t;
    HIP_TRY(hipMalloc(&a.p, sizeof(double) * nb * nnod));
    HIP_TRY(hipMalloc(&lu.p, sizeof(double) * (2 * nb - 1) * nnod));
    HIP_TRY(hipMalloc(&piv.p, sizeof(int) * nnod));
    HIP_TRY(hipMalloc(&ierr_d.p, sizeof(int) * mcomp));
    HIP_TRY(hipMemcpy(a.p, h_band, sizeof(double) * nb * nnod, hipMemcpyHostToDevice));
    if (ludcmp_dev(nb, n, mcomp, static_cast<double *>(a.p), static_cast<double *>(lu.p),
                   static_cast<int *>(piv.p), static_cast<int *>(ierr_d.p), s, err))
        return -1;
    HIP_TRY(hipStreamSynchronize(s));
    std::vector<int> ierr(mcomp), pv(nnod);
    HIP_TRY(hipMemcpy(ierr.data(), ierr_d.p, sizeof(int) * mcomp, hipMemcpyDeviceToHost));
    for (int j = 0; j < mcomp; ++j)
        if (ierr[j]) {
            *err = "singular radial matrix";
            return -1;
        }
    HIP_TRY(hipMemcpy(pv.data(), piv.p, sizeof(int) * nnod, hipMemcpyDeviceToHost));
    for (int j = 0; j < mcomp; ++j)
        for (int i = 1; i <= n; ++i) {
            const int d = pv[(size_t)(i - 1) + (size_t)n * j] - i;
            if (d < 0 || d > (nb - 1) / 2) {
                *err = "row interchange outside the band window";
                return -1;
            }
        }
    HIP_TRY(hipMalloc(&lu_t.p, sizeof(double) * (2 * nb - 1) * nnod));
    HIP_TRY(hipMalloc(&piv_t.p, sizeof(int) * nnod));
    if (lu_transpose_dev(nb, n, mcomp, static_cast<double *>(lu.p), static_cast<int *>(piv.p),
                         static_cast<double *>(lu_t.p), static_cast<int *>(piv_t.p), s, err))
        return -1;
    HIP_TRY(hipStreamSynchronize(s));
    *d_lu_t = static_cast<double *>(lu_t.release());
    *d_piv_t = static_cast<int *>(piv_t.release());
    return 0;
}

int lubksb_dev(int nb, int n, int mcomp, const double *d_band_lu, const int *d_ipiv, double *d_x,
               hipStream_t s, std::string *err)
{
    if ((nb != 3 && nb != 5) || n < 5 || mcomp < 1) {
        *err = "lubksb: nb must be 3 or 5, n >= 5";
        return -1;
    }
    const int blocks = (mcomp + 127) / 128;
    if (nb == 3)
        hipLaunchKernelGGL(k_lubksb_mul<3>, dim3(blocks), dim3(128), 0, s, n, mcomp, d_band_lu, d_ipiv, d_x);
    else
        hipLaunchKernelGGL(k_lubksb_mul<5>, dim3(blocks), dim3(128), 0, s, n, mcomp, d_band_lu, d_ipiv, d_x);
    HIP_TRY(hipGetLastError());
    return 0;
}

// ------------------------------------------------------------------------
// Explicit integration (pointwise over d_rj), reference operation order:
// kind 0 x += dt (c D + 3/2 F - 1/2 P), P <- F      (cal_vorticity_eq_adams,
//        cal_vorticity_terms_adams.f90:31-51; cal_diff_induction_MHD_adams,
//        cal_explicit_terms.f90:22-42)
// kind 1 x += dt (c D - 3/2 A - 1/2 P), P <- -A     (scalar_diff_advect_adams,
//        cal_diff_adv_src_explicit.f90:129-154)
// scheme 0 Adams-Bashforth, 1 Euler, 2 Euler + start of the history (step 1).
__global__ void k_explicit_update(int kind, int scheme, size_t n, double dt, double coef_exp,
                                  double *x, const double *diff, const double *force,
                                  double *pre)
{
    const double adam_0 = 3.0 / 2.0, adam_1 = -1.0 / 2.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const double D = diff[i], F = force[i];
        if (kind == 0) {
            if (scheme == 0) {
                x[i] = x[i] + dt * (coef_exp * D + adam_0 * F + adam_1 * pre[i]);
                pre[i] = F;
            } else {
                x[i] = x[i] + dt * (coef_exp * D + F);
                if (scheme == 2) pre[i] = F;
            }
        } else {
            if (scheme == 0) {
                x[i] = x[i] + dt * (coef_exp * D - adam_0 * F + adam_1 * pre[i]);
                pre[i] = -F;
            } else {
                x[i] = x[i] + dt * (coef_exp * D - F);
                if (scheme == 2) pre[i] = -F;
            }
        }
    }
}

// Volume mean square, stage 1: per mode, the radial trapezoid integral of
// the layer mean square over kr_st..kr_ed (cal_rms_each_vector_sph_spec /
// _scalar_sph_spec, cal_rms_by_sph_spectr.f90:63-154;
// radial_int_by_trapezoid, radial_int_for_sph_spec.f90:66-94).
__global__ void k_ms_modes(int kind, int nr, int J, const int *l_of_j, const double *radius,
                           int kr_st, int kr_ed, const double *f, double *vol)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= J) return;
    const int l = l_of_j[j];
    const int ncp = (kind == 2) ? 1 : 3;
    double s[3] = {0.0, 0.0, 0.0};
    double prev[3] = {0.0, 0.0, 0.0};
    for (int k = kr_st; k <= kr_ed; ++k) {
        const size_t in = (size_t)j + (size_t)J * (k - 1);
        const double r = radius[k - 1], ar = 1.0 / r;
        double cur[3];
        if (kind == 2) {
            const double g11 = 1.0 / (double)(2 * l + 1);
            const double d = f[in];
            cur[0] = d * d * g11 * r * r;
            cur[1] = cur[2] = 0.0;
        } else {
            const double S = f[in], dS = f[in + (size_t)J * nr], T = f[in + 2 * (size_t)J * nr];
            double r1, r2;
            if (l == 0) {
                r1 = (0.5 * S) * (0.5 * S) * ar * ar;
                r2 = 0.0;
            } else {
                const double g3 = (double)(l * (l + 1));
                const double g12 = (double)(l * (l + 1)) / (double)(2 * l + 1);
                r1 = g12 * (g3 * ar * ar * S * S + dS * dS);
                r2 = g12 * T * T;
            }
            double r3 = r1 + r2;
            if (kind == 1) {
                r1 = 0.5 * r1;
                r2 = 0.5 * r2;
                r3 = 0.5 * r3;
            }
            cur[0] = r1;
            cur[1] = r2;
            cur[2] = r3;
        }
        if (k > kr_st) {
            const double dr1 = radius[k - 1] - radius[k - 2];
            for (int c = 0; c < ncp; ++c) s[c] = s[c] + 0.5 * dr1 * (prev[c] + cur[c]);
        }
        for (int c = 0; c < 3; ++c) prev[c] = cur[c];
    }
    for (int c = 0; c < ncp; ++c) vol[(size_t)j + (size_t)J * c] = s[c];
}

// stage 2 (one work-group): sum the modes of each degree in mode order, then
// the degrees in order, times 3 / (r_ed^3 - r_st^3)
__global__ void k_ms_reduce(int kind, int J, int lmax, const int *l_of_j, const double *vol,
                            double rs, double re, double *out)
{
    extern __shared__ double sl[];   // [ncp][lmax+1]
    const int ncp = (kind == 2) ? 1 : 3;
    for (int e = threadIdx.x; e < ncp * (lmax + 1); e += blockDim.x) {
        const int c = e / (lmax + 1), l = e % (lmax + 1);
        double s = 0.0;
        for (int j = 0; j < J; ++j)
            if (l_of_j[j] == l) s = s + vol[(size_t)j + (size_t)J * c];
        sl[e] = s;
    }
    __syncthreads();
    if (threadIdx.x < ncp) {
        const int c = threadIdx.x;
        double tot = sl[c * (lmax + 1)];
        for (int l = 1; l <= lmax; ++l) tot = tot + sl[c * (lmax + 1) + l];
        const double avol = 3.0 / (re * re * re - rs * rs * rs);
        out[c] = avol * tot;
    }
}

int explicit_update_dev(int kind, int scheme, size_t n, double dt, double coef_exp, double *d_x,
                        const double *d_diff, const double *d_force, double *d_pre,
                        hipStream_t s, std::string *err)
{
    if (kind < 0 || kind > 1 || scheme < 0 || scheme > 2) {
        *err = "explicit update: kind 0|1, scheme 0|1|2";
        return -1;
    }
    const int blocks = (int)std::min<size_t>((n + 255) / 256, 65536);
    hipLaunchKernelGGL(k_explicit_update, dim3(blocks), dim3(256), 0, s, kind, scheme, n, dt,
                       coef_exp, d_x, d_diff, d_force, d_pre);
    HIP_TRY(hipGetLastError());
    return 0;
}

int mean_square_dev(int kind, int nr, int J, int lmax, const int *d_l_of_j,
                    const double *d_radius, const double *h_radius, int kr_st, int kr_ed,
                    const double *d_f, double *d_vol, double *d_out, hipStream_t s,
                    std::string *err)
{
    if (kind < 0 || kind > 2 || kr_st < 1 || kr_ed > nr || kr_ed <= kr_st) {
        *err = "mean square: kind 0|1|2 and 1 <= kr_st < kr_ed <= nr";
        return -1;
    }
    hipLaunchKernelGGL(k_ms_modes, dim3((J + 127) / 128), dim3(128), 0, s, kind, nr, J, d_l_of_j,
                       d_radius, kr_st, kr_ed, d_f, d_vol);
    HIP_TRY(hipGetLastError());
    const int ncp = (kind == 2) ? 1 : 3;
    hipLaunchKernelGGL(k_ms_reduce, dim3(1), dim3(256), sizeof(double) * ncp * (lmax + 1), s, kind,
                       J, lmax, d_l_of_j, d_vol, h_radius[kr_st - 1], h_radius[kr_ed - 1], d_out);
    HIP_TRY(hipGetLastError());
    return 0;
}

// ---- allocation-free volume mean square -----------------------------------
// The same two kernels on a file-scope device array instead of buffers
// allocated per call: the per-mode volume integrals (3 x kMsMaxJ doubles), the
// result (3, padded to 32), the radial grid (kMsMaxNr) and the degrees of the
// modes (kMsMaxJ int32), 7.4 MB on every GPU the library runs on, held until
// the process ends.  Every element the kernels read is written earlier in the
// same call (the grid and degrees by the copies, the integrals by k_ms_modes,
// the result by k_ms_reduce), so nothing needs clearing between calls; calls
// that overlap on one GPU race on this array.
constexpr int kMsMaxJ = 262144;   // (L+1)^2 at L = 511
constexpr int kMsMaxNr = 4096;
constexpr int kMsMaxL = 2047;     // k_ms_reduce's LDS: 3 x 2048 doubles
constexpr size_t kMsVolOff = 0;
constexpr size_t kMsOutOff = kMsVolOff + 3 * (size_t)kMsMaxJ;
constexpr size_t kMsRadOff = kMsOutOff + 32;
constexpr size_t kMsLOff = kMsRadOff + kMsMaxNr;
constexpr size_t kMsScratchDoubles = kMsLOff + kMsMaxJ / 2;

static __device__ __align__(256) double scratch[kMsScratchDoubles];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (each GPU holds its own copy of the array at its own address).
int scratch_on_device(double **out, std::string *err)
{
    int dev = -1;
    HIP_TRY(hipGetDevice(&dev));
    if (dev < 0 || dev >= kMaxDevices) {
        *err = "mean square (no-alloc): device index out of range";
        return -1;
    }
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        HIP_TRY(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) {
            *err = "mean square (no-alloc): no device address for the scratch array";
            return -1;
        }
        g_scratch_ptr[dev] = static_cast<double *>(p);
    }
    *out = g_scratch_ptr[dev];
    return 0;
}
}  // namespace

int mean_square_noalloc_dev(int kind, int nr, int J, const int *h_l_of_j, const double *h_radius,
                            int kr_st, int kr_ed, const double *d_f, double *h_out,
                            hipStream_t s, std::string *err)
{
    if (J < 1 || J > kMsMaxJ || nr < 2 || nr > kMsMaxNr) {
        *err = "mean square (no-alloc): needs 1 <= J <= " + std::to_string(kMsMaxJ) +
               " and 2 <= nr <= " + std::to_string(kMsMaxNr) + " (got J=" + std::to_string(J) +
               ", nr=" + std::to_string(nr) + ")";
        return -1;
    }
    int lmax = 0;
    for (int j = 0; j < J; ++j) {
        if (h_l_of_j[j] < 0 || h_l_of_j[j] > kMsMaxL) {
            *err = "mean square (no-alloc): degrees must lie in 0.." + std::to_string(kMsMaxL);
            return -1;
        }
        lmax = std::max(lmax, h_l_of_j[j]);
    }
    double *base = nullptr;
    if (scratch_on_device(&base, err)) return -1;
    double *d_vol = base + kMsVolOff, *d_out = base + kMsOutOff, *d_r = base + kMsRadOff;
    int *d_l = reinterpret_cast<int *>(base + kMsLOff);
    HIP_TRY(hipMemcpyAsync(d_l, h_l_of_j, sizeof(int) * J, hipMemcpyHostToDevice, s));
    HIP_TRY(hipMemcpyAsync(d_r, h_radius, sizeof(double) * nr, hipMemcpyHostToDevice, s));
    if (mean_square_dev(kind, nr, J, lmax, d_l, d_r, h_radius, kr_st, kr_ed, d_f, d_vol, d_out, s,
                        err))
        return -1;
    double h[3] = {0.0, 0.0, 0.0};
    const int ncp = (kind == 2) ? 1 : 3;
    HIP_TRY(hipMemcpyAsync(h, d_out, sizeof(double) * ncp, hipMemcpyDeviceToHost, s));
    HIP_TRY(hipStreamSynchronize(s));
    for (int c = 0; c < 3; ++c) h_out[c] = c < ncp ? h[c] : 0.0;
    return 0;
}

}  // namespace calypso
