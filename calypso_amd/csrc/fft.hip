// fft.hip -- longitude real FFTs, grid-space nonlinear products and the
// rtm <-> rtp exchange for one GPU (gfx950).
//
// Semantics follow the reference's FFTPACK5 path
//   sph_RFFTMF_to_send_t / sph_RFFTMB_from_recv_t (t_sph_FFTPACK5.f90:183-331)
//   forward  a0 = (1/N) sum x, a_m = (2/N) sum x cos(m phi), b_m = (2/N) sum x sin(m phi)
//   backward x  = a0 + sum_{m<=L} a_m cos(m phi) + b_m sin(m phi)
// with a_m in rtm slot L+m, b_m in slot L-m, a0 in slot L
// (set_sph_tranform_ordering.f90:43-159), so the rtm<->rtp transposes the
// reference does with MPI (calypso_sph_comm_rtm_2_rtp_N / _rtp_2_rtm_N,
// spherical_SRs_N.f90:333-369) become the gather/scatter of these kernels.
//
// One work-group owns one (radius, colatitude) pencil and all of its
// components: it stages the pencil's 2L+1 coefficients of every component in
// LDS (reads of ncomp contiguous doubles per order), runs the FFTs of pairs of
// real components as complex Stockham FFTs (radix 4/2/3, FP64, twiddles from
// an FP64 table in LDS), and writes the other side.  The fused MHD kernel
// keeps the grid values in LDS between the inverse FFT, the products
// (cal_nonlinear_sph_MHD.f90:35-75) and the forward FFT, so the physical grid
// never touches HBM.
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>
#include <vector>

#include "fft_internal.hpp"

namespace calypso {

#define HIP_TRY(expr)                                                         \
    do {                                                                      \
        hipError_t e_ = (expr);                                               \
        if (e_ != hipSuccess) {                                               \
            if (err) *err = std::string(#expr " failed: ") + hipGetErrorString(e_); \
            return -1;                                                        \
        }                                                                     \
    } while (0)

struct cplx {
    double x, y;
};

__device__ inline cplx cadd(cplx a, cplx b) { return {a.x + b.x, a.y + b.y}; }
__device__ inline cplx csub(cplx a, cplx b) { return {a.x - b.x, a.y - b.y}; }
__device__ inline cplx cmul(cplx a, cplx b) { return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
// multiply by s*i (s = +-1)
__device__ inline cplx cmul_si(cplx a, double s) { return {-s * a.y, s * a.x}; }

struct FftArgs {
    const double *in;
    double *out;
    int L, nth, nr, N, ncomp_in, ncomp_out;
    int nstage;
    int radix[FFT_MAX_STAGES];
    const double *tw;   // N complex twiddles exp(-2 pi i t / N)
    int pairs_per_batch;
    double coef[4];     // coef_velo, coef_lor, coef_induct, coef_temp
};

// One Stockham stage of radix R over `nseq` complex sequences of length N:
// data0 -> data1 (Govindaraju et al., SC'08, mixed-radix autosort form).
template <int R>
__device__ void stockham_stage(const cplx *d0, cplx *d1, const cplx *tw, int N, int Ns,
                               int nseq, double sgn)
{
    const int nb = N / R;
    for (int idx = threadIdx.x; idx < nseq * nb; idx += blockDim.x) {
        const int s = idx / nb, j = idx % nb;
        const int kk = j % Ns;
        const int tstep = N / (Ns * R);
        const cplx *src = d0 + (size_t)s * N;
        cplx v[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            v[r] = src[j + r * nb];
            if (r) {
                cplx w = tw[(r * kk * tstep) % N];
                w.y *= -sgn;  // table holds exp(-i...); sgn=+1 (inverse) conjugates
                v[r] = cmul(v[r], w);
            }
        }
        cplx o[R];
        if constexpr (R == 2) {
            o[0] = cadd(v[0], v[1]);
            o[1] = csub(v[0], v[1]);
        } else if constexpr (R == 4) {
            cplx a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
            cplx b0 = cadd(v[1], v[3]), b1 = cmul_si(csub(v[1], v[3]), sgn);
            o[0] = cadd(a0, b0);
            o[1] = cadd(a1, b1);
            o[2] = csub(a0, b0);
            o[3] = csub(a1, b1);
        } else {  // R == 3
            const double c = -0.5, sn = sgn * 0.86602540378443864676;
            cplx t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
            cplx m1 = {v[0].x + c * t1.x, v[0].y + c * t1.y};
            cplx m2 = {-sn * t2.y, sn * t2.x};
            o[0] = cadd(v[0], t1);
            o[1] = cadd(m1, m2);
            o[2] = csub(m1, m2);
        }
        cplx *dst = d1 + (size_t)s * N + (j / Ns) * Ns * R + kk;
#pragma unroll
        for (int r = 0; r < R; ++r) dst[r * Ns] = o[r];
    }
}

// full FFT over nseq sequences; returns the buffer holding the result
__device__ cplx *fft_lds(cplx *b0, cplx *b1, const cplx *tw, const FftArgs &a, int nseq,
                         double sgn)
{
    int Ns = 1;
    cplx *src = b0, *dst = b1;
    for (int st = 0; st < a.nstage; ++st) {
        const int R = a.radix[st];
        if (R == 4) stockham_stage<4>(src, dst, tw, a.N, Ns, nseq, sgn);
        else if (R == 2) stockham_stage<2>(src, dst, tw, a.N, Ns, nseq, sgn);
        else stockham_stage<3>(src, dst, tw, a.N, Ns, nseq, sgn);
        __syncthreads();
        Ns *= R;
        cplx *t = src;
        src = dst;
        dst = t;
    }
    return src;
}

enum PencilMode { kPencilBwd = 0, kPencilFwd = 1, kPencilMhd = 2 };

// LDS: spec[ncs][M] doubles | tw[N] cplx | buf0[PB][N] cplx | buf1[PB][N] cplx
template <int MODE>
__global__ __launch_bounds__(256) void k_pencil(FftArgs a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int M = 2 * a.L + 1, N = a.N, L = a.L;
    const int ncs = a.ncomp_in > a.ncomp_out ? a.ncomp_in : a.ncomp_out;
    double *spec = reinterpret_cast<double *>(smem_raw);
    const size_t spec_bytes = ((size_t)ncs * M * sizeof(double) + 15) & ~(size_t)15;
    cplx *tw = reinterpret_cast<cplx *>(smem_raw + spec_bytes);
    cplx *buf0 = tw + N;
    cplx *buf1 = buf0 + (size_t)a.pairs_per_batch * N;

    const int pencil = blockIdx.x;            // t + nth*k (rtm pencil order)
    const int t = pencil % a.nth, k = pencil / a.nth;
    const size_t sl = (size_t)a.nth * a.nr;   // rtm slot stride in pencils
    const size_t nnod = (size_t)a.nr * a.nth * N;
    const size_t gbase = (size_t)k + (size_t)a.nr * t;  // rtp node at phi = 0

    for (int i = threadIdx.x; i < N; i += blockDim.x)
        tw[i] = {a.tw[2 * i], a.tw[2 * i + 1]};

    // ---- input side --------------------------------------------------
    const int nin = a.ncomp_in;
    const int npair_in = (nin + 1) / 2;
    if (MODE == kPencilBwd || MODE == kPencilMhd) {
        for (int e = threadIdx.x; e < nin * M; e += blockDim.x) {
            const int nd = e % nin, mp = e / nin;
            spec[nd * M + mp] = a.in[nd + (size_t)nin * (pencil + sl * mp)];
        }
    }
    __syncthreads();

    // grid values live in the result buffer of the inverse FFT (pairs of
    // components as real/imaginary parts); MHD mode needs all 7 pairs at once
    cplx *grid = nullptr;
    if (MODE == kPencilBwd || MODE == kPencilMhd) {
        for (int p0 = 0; p0 < npair_in; p0 += a.pairs_per_batch) {
            const int np = min(a.pairs_per_batch, npair_in - p0);
            // Hermitian spectra of z = x_a + i x_b
            for (int e = threadIdx.x; e < np * N; e += blockDim.x) {
                const int p = e / N, kf = e % N;
                const int ca = 2 * (p0 + p), cb = ca + 1;
                const bool hb = cb < nin;
                cplx z = {0.0, 0.0};
                if (kf == 0) {
                    z.x = spec[ca * M + L];
                    z.y = hb ? spec[cb * M + L] : 0.0;
                } else if (kf <= L) {
                    const int m = kf;
                    double aa = spec[ca * M + L + m], ba = spec[ca * M + L - m];
                    double ab = hb ? spec[cb * M + L + m] : 0.0;
                    double bb = hb ? spec[cb * M + L - m] : 0.0;
                    z.x = 0.5 * (aa + bb);
                    z.y = 0.5 * (ab - ba);
                } else if (kf >= N - L) {
                    const int m = N - kf;
                    double aa = spec[ca * M + L + m], ba = spec[ca * M + L - m];
                    double ab = hb ? spec[cb * M + L + m] : 0.0;
                    double bb = hb ? spec[cb * M + L - m] : 0.0;
                    z.x = 0.5 * (aa - bb);
                    z.y = 0.5 * (ba + ab);
                }
                buf0[(size_t)p * N + kf] = z;
            }
            __syncthreads();
            grid = fft_lds(buf0, buf1, tw, a, np, +1.0);
            if (MODE == kPencilBwd) {
                for (int e = threadIdx.x; e < np * N; e += blockDim.x) {
                    const int p = e / N, i = e % N;
                    const int ca = 2 * (p0 + p), cb = ca + 1;
                    const cplx z = grid[(size_t)p * N + i];
                    a.out[gbase + (size_t)a.nr * a.nth * i + nnod * ca] = z.x;
                    if (cb < nin) a.out[gbase + (size_t)a.nr * a.nth * i + nnod * cb] = z.y;
                }
                __syncthreads();
            }
        }
    }
    if (MODE == kPencilBwd) return;

    // ---- products (MHD) or grid load (fwd) ---------------------------
    // forward input pairs: fwd mode = component pairs of the grid array;
    // MHD mode = the 12 product components in f_trns order 0-2, 9-17
    cplx *fin = (grid == buf0) ? buf1 : buf0;
    const int nout = a.ncomp_out;
    int npair_f;
    if (MODE == kPencilMhd) {
        npair_f = 6;
        for (int i = threadIdx.x; i < N; i += blockDim.x) {
            // b_trns: u 0-2, w 3-5, B 6-8, J 9-11, T 12 -> pairs (0,1)...(12,-)
            const cplx c01 = grid[0 * N + i], c23 = grid[1 * N + i], c45 = grid[2 * N + i];
            const cplx c67 = grid[3 * N + i], c89 = grid[4 * N + i], c1011 = grid[5 * N + i];
            const cplx c12 = grid[6 * N + i];
            const double u1 = c01.x, u2 = c01.y, u3 = c23.x;
            const double w1 = c23.y, w2 = c45.x, w3 = c45.y;
            const double b1 = c67.x, b2 = c67.y, b3 = c89.x;
            const double j1 = c89.y, j2 = c1011.x, j3 = c1011.y;
            const double T = c12.x;
            const double cv = a.coef[0], cl = a.coef[1], ci = a.coef[2], ct = a.coef[3];
            double f[12];
            {
            // same operation order and no contraction: bitwise equal to the
            // reference formulas
#pragma clang fp contract(off)
            f[0] = (w2 * u3 - w3 * u2) * cv;
            f[1] = (w3 * u1 - w1 * u3) * cv;
            f[2] = (w1 * u2 - w2 * u1) * cv;
            f[3] = (j2 * b3 - j3 * b2) * cl;
            f[4] = (j3 * b1 - j1 * b3) * cl;
            f[5] = (j1 * b2 - j2 * b1) * cl;
            f[6] = (u2 * b3 - u3 * b2) * ci;
            f[7] = (u3 * b1 - u1 * b3) * ci;
            f[8] = (u1 * b2 - u2 * b1) * ci;
            f[9] = u1 * T * ct;
            f[10] = u2 * T * ct;
            f[11] = u3 * T * ct;
            }
#pragma unroll
            for (int p = 0; p < 6; ++p) fin[(size_t)p * N + i] = {f[2 * p], f[2 * p + 1]};
        }
        __syncthreads();
    } else {
        npair_f = (nin + 1) / 2;
    }

    // ---- forward FFTs ---------------------------------------------------
    for (int e = threadIdx.x; e < nout * M; e += blockDim.x) spec[e] = 0.0;
    const double invN = 1.0 / (double)N;
    for (int p0 = 0; p0 < npair_f; p0 += a.pairs_per_batch) {
        const int np = min(a.pairs_per_batch, npair_f - p0);
        cplx *src = fin;
        if (MODE == kPencilFwd) {
            src = buf0;
            for (int e = threadIdx.x; e < np * N; e += blockDim.x) {
                const int p = e / N, i = e % N;
                const int ca = 2 * (p0 + p), cb = ca + 1;
                const size_t g = gbase + (size_t)a.nr * a.nth * i;
                src[(size_t)p * N + i] = {a.in[g + nnod * ca], cb < nin ? a.in[g + nnod * cb] : 0.0};
            }
            __syncthreads();
        }
        cplx *other = (src == buf0) ? buf1 : buf0;
        cplx *Z = fft_lds(src + (MODE == kPencilMhd ? (size_t)p0 * N : 0),
                          other + (MODE == kPencilMhd ? (size_t)p0 * N : 0), tw, a, np, -1.0);
        // unpack two real spectra; component ids in the output array
        for (int e = threadIdx.x; e < np * (L + 1); e += blockDim.x) {
            const int p = e / (L + 1), m = e % (L + 1);
            const cplx zk = Z[(size_t)p * N + m];
            const cplx zn = Z[(size_t)p * N + (m == 0 ? 0 : N - m)];
            // X_a = (Z_k + conj Z_{N-k})/2, X_b = (Z_k - conj Z_{N-k})/(2i)
            const double xar = 0.5 * (zk.x + zn.x), xai = 0.5 * (zk.y - zn.y);
            const double xbr = 0.5 * (zk.y + zn.y), xbi = -0.5 * (zk.x - zn.x);
            int ca, cb;
            if (MODE == kPencilMhd) {
                const int q = 2 * (p0 + p);              // product index 0..11
                ca = q < 3 ? q : q + 6;                  // f_trns slot
                cb = (q + 1) < 3 ? q + 1 : q + 7;
            } else {
                ca = 2 * (p0 + p);
                cb = ca + 1;
            }
            if (m == 0) {
                spec[ca * M + L] = xar * invN;
                if (cb < nout) spec[cb * M + L] = xbr * invN;
            } else {
                spec[ca * M + L + m] = 2.0 * xar * invN;
                spec[ca * M + L - m] = -2.0 * xai * invN;
                if (cb < nout) {
                    spec[cb * M + L + m] = 2.0 * xbr * invN;
                    spec[cb * M + L - m] = -2.0 * xbi * invN;
                }
            }
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < nout * M; e += blockDim.x) {
        const int nd = e % nout, mp = e / nout;
        a.out[nd + (size_t)nout * (pencil + sl * mp)] = spec[nd * M + mp];
    }
}

// pointwise products on the reference grid layout fld(nnod, 13) -> frc(nnod, 19)
__global__ void k_products_mhd(size_t nnod, const double *fld, double *frc, double cv,
                               double cl, double ci, double ct)
{
#pragma clang fp contract(off)
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < nnod;
         i += (size_t)gridDim.x * blockDim.x) {
        const double u1 = fld[i], u2 = fld[nnod + i], u3 = fld[2 * nnod + i];
        const double w1 = fld[3 * nnod + i], w2 = fld[4 * nnod + i], w3 = fld[5 * nnod + i];
        const double b1 = fld[6 * nnod + i], b2 = fld[7 * nnod + i], b3 = fld[8 * nnod + i];
        const double j1 = fld[9 * nnod + i], j2 = fld[10 * nnod + i], j3 = fld[11 * nnod + i];
        const double T = fld[12 * nnod + i];
        frc[i] = (w2 * u3 - w3 * u2) * cv;
        frc[nnod + i] = (w3 * u1 - w1 * u3) * cv;
        frc[2 * nnod + i] = (w1 * u2 - w2 * u1) * cv;
        for (int c = 3; c < 9; ++c) frc[c * nnod + i] = 0.0;
        frc[9 * nnod + i] = (j2 * b3 - j3 * b2) * cl;
        frc[10 * nnod + i] = (j3 * b1 - j1 * b3) * cl;
        frc[11 * nnod + i] = (j1 * b2 - j2 * b1) * cl;
        frc[12 * nnod + i] = (u2 * b3 - u3 * b2) * ci;
        frc[13 * nnod + i] = (u3 * b1 - u1 * b3) * ci;
        frc[14 * nnod + i] = (u1 * b2 - u2 * b1) * ci;
        frc[15 * nnod + i] = u1 * T * ct;
        frc[16 * nnod + i] = u2 * T * ct;
        frc[17 * nnod + i] = u3 * T * ct;
        frc[18 * nnod + i] = 0.0;
    }
}

// ----------------------------------------------------------------- host --

int fft_context_build(FftContext *f, int L, int nth, int nr, int nphi, std::string *err)
{
    if (L < 1 || nth < 1 || nr < 1) { *err = "invalid grid size"; return -1; }
    if (nphi < 2 * L + 2) { *err = "nphi must be at least 2*(L+1)"; return -1; }
    int n = nphi, ns = 0;
    while (n % 4 == 0) { f->radix[ns++] = 4; n /= 4; }
    while (n % 2 == 0) { f->radix[ns++] = 2; n /= 2; }
    while (n % 3 == 0) { f->radix[ns++] = 3; n /= 3; }
    if (n != 1 || ns > FFT_MAX_STAGES) {
        *err = "nphi must be of the form 2^a 3^b";
        return -1;
    }
    f->L = L;
    f->nth = nth;
    f->nr = nr;
    f->N = nphi;
    f->nstage = ns;
    std::vector<double> tw(2 * (size_t)nphi);
    const double pi = 3.14159265358979323846;
    for (int i = 0; i < nphi; ++i) {
        long double ang = -2.0L * (long double)pi * (long double)i / (long double)nphi;
        tw[2 * i] = (double)std::cos(ang);
        tw[2 * i + 1] = (double)std::sin(ang);
    }
    HIP_TRY(hipMalloc(&f->d_tw, sizeof(double) * tw.size()));
    HIP_TRY(hipMemcpy(f->d_tw, tw.data(), sizeof(double) * tw.size(), hipMemcpyHostToDevice));
    return 0;
}

void fft_context_free(FftContext *f)
{
    if (f && f->d_tw) (void)hipFree(f->d_tw);
    if (f) f->d_tw = nullptr;
}

static size_t pencil_lds(const FftContext *f, int ncs, int pb)
{
    const size_t M = 2 * (size_t)f->L + 1;
    size_t s = ((size_t)ncs * M * sizeof(double) + 15) & ~(size_t)15;
    return s + (size_t)f->N * 16 + 2 * (size_t)pb * f->N * 16;
}

static int launch_pencil(const FftContext *f, int mode, int nin, int nout, const double *in,
                         double *out, const double *coef, hipStream_t s, std::string *err)
{
    FftArgs a;
    a.in = in;
    a.out = out;
    a.L = f->L;
    a.nth = f->nth;
    a.nr = f->nr;
    a.N = f->N;
    a.ncomp_in = nin;
    a.ncomp_out = nout;
    a.nstage = f->nstage;
    for (int i = 0; i < FFT_MAX_STAGES; ++i) a.radix[i] = f->radix[i];
    a.tw = f->d_tw;
    for (int i = 0; i < 4; ++i) a.coef[i] = coef ? coef[i] : 0.0;
    const int ncs = nin > nout ? nin : nout;
    const size_t limit = 160 * 1024;
    int pb;
    if (mode == kPencilMhd) {
        pb = 7;
        if (pencil_lds(f, ncs, pb) > limit) { *err = "pencil too long for the fused MHD kernel"; return -1; }
    } else {
        const int npair = ((mode == kPencilBwd ? nin : nin) + 1) / 2;
        pb = npair;
        while (pb > 1 && pencil_lds(f, ncs, pb) > limit) --pb;
        if (pencil_lds(f, ncs, pb) > limit) { *err = "pencil too long for LDS"; return -1; }
    }
    a.pairs_per_batch = pb;
    const size_t lds = pencil_lds(f, ncs, pb);
    const int grid = f->nr * f->nth;
    const void *fn = mode == kPencilBwd ? reinterpret_cast<const void *>(&k_pencil<kPencilBwd>)
                   : mode == kPencilFwd ? reinterpret_cast<const void *>(&k_pencil<kPencilFwd>)
                                        : reinterpret_cast<const void *>(&k_pencil<kPencilMhd>);
    if (lds > 64 * 1024)
        HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (mode == kPencilBwd)
        hipLaunchKernelGGL(k_pencil<kPencilBwd>, dim3(grid), dim3(256), lds, s, a);
    else if (mode == kPencilFwd)
        hipLaunchKernelGGL(k_pencil<kPencilFwd>, dim3(grid), dim3(256), lds, s, a);
    else
        hipLaunchKernelGGL(k_pencil<kPencilMhd>, dim3(grid), dim3(256), lds, s, a);
    HIP_TRY(hipGetLastError());
    return 0;
}

int fft_backward_dev(const FftContext *f, int ncomp, const double *d_vr_rtm, double *d_fld_rtp,
                     hipStream_t s, std::string *err)
{
    return launch_pencil(f, kPencilBwd, ncomp, ncomp, d_vr_rtm, d_fld_rtp, nullptr, s, err);
}

int fft_forward_dev(const FftContext *f, int ncomp, const double *d_frc_rtp, double *d_vr_rtm,
                    hipStream_t s, std::string *err)
{
    return launch_pencil(f, kPencilFwd, ncomp, ncomp, d_frc_rtp, d_vr_rtm, nullptr, s, err);
}

bool fft_mhd_fused_fits(const FftContext *f)
{
    return pencil_lds(f, 19, 7) <= 160 * 1024;
}

int fft_mhd_fused_dev(const FftContext *f, const double *d_bwd_rtm, double *d_fwd_rtm,
                      const double *coef, hipStream_t s, std::string *err)
{
    return launch_pencil(f, kPencilMhd, 13, 19, d_bwd_rtm, d_fwd_rtm, coef, s, err);
}

int products_mhd_dev(size_t nnod, const double *d_fld, double *d_frc, const double *coef,
                     hipStream_t s, std::string *err)
{
    const int blocks = (int)std::min<size_t>((nnod + 255) / 256, 65536);
    hipLaunchKernelGGL(k_products_mhd, dim3(blocks), dim3(256), 0, s, nnod, d_fld, d_frc, coef[0],
                       coef[1], coef[2], coef[3]);
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // namespace calypso
