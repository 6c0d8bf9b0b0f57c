// legendre.hip -- backward / forward spherical-harmonic Legendre transforms
// for gfx950 (CDNA4) as per-order MFMA contractions.
//
// Arithmetic follows the reference org loops
//   backward: legendre_b_trans_vector_org / _scalar_org
//             (SPH_SHELL_src/legendre_bwd_trans_org.f90:45-197)
//   forward:  legendre_f_trans_vector_org / _scalar_org
//             (SPH_SHELL_src/legendre_fwd_trans_org.f90:45-160)
// in the equatorially folded form of the reference's symmetric variants
// (legendre_bwd_sym_matmul.f90, legendre_fwd_sym_matmul.f90):
// P_l^m(pi-theta) = (-1)^(l-m) P_l^m(theta) and dP/dtheta flips the other way,
// so each order m needs only the hemisphere rows and the degree list split
// by the parity of l-m.
//
// Data layouts are the reference's single-domain ones (component fastest):
//   sp_rlm[nd + ncomp*(j + J*k)]              j in rlm order (m-major)
//   vr_rtm[nd + ncomp*(t + nth*k + nth*nr*mp)] mp = m + L
//
// MFMA accumulator layouts differ by instruction: row = 4*(lane/16)+v for
// the f32 16x16x4 and bf16 16x16x32 forms, row = lane/16 + 4v for the f64
// 16x16x4 form (column = lane%16 for all); T::row encodes it.
//
// Kernels (one work-group = one |m|, a tile of KT radial points, a group of
// 16-row tiles; the +m and -m orders share every P tile):
//   backward: rows = colatitude, K = degree;  out(t) = S+A, out(pi-t) = S-A
//   forward : rows = degree,     K = colatitude, input folded N+-S on load
// Two forms: the first reads P, dP and mP/sin(theta) tiles; the third reads P
// and dP only, stages the B operand GEMM-ready and keeps its K loop free of
// branches (see its section).  The autotuner picks per direction.
// Precisions: FP64 MFMA 16x16x4, FP32 MFMA 16x16x4, split-BF16x3 MFMA
// 16x16x32 (three bf16 slices per operand, six slice products, FP32
// accumulation); the FP64-accurate int8 digit form lives in leg_ozaki.hip
// (kernel form 5) and shares this file's tables, plans and autotuner.
// P_l^m tables are computed on the GPU in FP64 and packed in
// the lane order the MFMA A operand is read in (one 64-lane unit per
// 16 x KSTEP tile), so every A fetch is a contiguous 512 B - 1 KiB load.
#include <array>
#include <map>
#include <mutex>

#include "hip_check.hpp"
#include "leg_internal.hpp"

#include <hip/hip_bf16.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace calypso {


typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));
typedef __bf16 bf8v __attribute__((ext_vector_type(8)));
typedef unsigned short us8v __attribute__((ext_vector_type(8)));

// ------------------------------------------------------------ traits --

struct TF64 {
    using elem = double;
    using frag = double;
    using acc = d4v;
    static constexpr int KSTEP = 4, NS = 1, NSA = 1, EPL = 1;
    __device__ static inline void store(elem *dst, double x, int stride) { dst[0] = x; }
    __device__ static inline frag load(const elem *p) { return *p; }
    __device__ static inline frag neg(frag a) { return -a; }
    __device__ static inline void mma(acc &c, const frag *a, const frag *b) {
        c = __builtin_amdgcn_mfma_f64_16x16x4f64(a[0], b[0], c, 0, 0, 0);
    }
    __device__ static inline double get(const acc &c, int v) { return c[v]; }
    // accumulator row of value v in lane `lane` (column = lane % 16):
    // the f64 16x16x4 MFMA interleaves rows across the four lane groups
    __device__ static inline int row(int lane, int v) { return (lane >> 4) + 4 * v; }
};

struct TF32 {
    using elem = float;
    using frag = float;
    using acc = f4v;
    static constexpr int KSTEP = 4, NS = 1, NSA = 1, EPL = 1;
    __device__ static inline void store(elem *dst, double x, int stride) { dst[0] = (float)x; }
    __device__ static inline frag load(const elem *p) { return *p; }
    __device__ static inline frag neg(frag a) { return -a; }
    __device__ static inline void mma(acc &c, const frag *a, const frag *b) {
        c = __builtin_amdgcn_mfma_f32_16x16x4f32(a[0], b[0], c, 0, 0, 0);
    }
    __device__ static inline double get(const acc &c, int v) { return (double)c[v]; }
    __device__ static inline int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

__device__ inline unsigned short bf16_bits(double x) {
    __bf16 h = (__bf16)(float)x;
    return __builtin_bit_cast(unsigned short, h);
}
__device__ inline double bf16_val(unsigned short b) {
    return (double)(float)__builtin_bit_cast(__bf16, b);
}

struct TBF3 {
    using elem = unsigned short;  // bf16 bit pattern
    using frag = bf8v;
    using acc = f4v;
    static constexpr int KSTEP = 32, NS = 3, NSA = 3, EPL = 8;
    // three-slice split x = hi + mid + lo; each residual is exact in FP64
    __device__ static inline void store(elem *dst, double x, int stride) {
        unsigned short h = bf16_bits(x);
        double r1 = x - bf16_val(h);
        unsigned short m = bf16_bits(r1);
        double r2 = r1 - bf16_val(m);
        dst[0] = h;
        dst[stride] = m;
        dst[2 * stride] = bf16_bits(r2);
    }
    __device__ static inline frag load(const elem *p) {
        us8v u = *reinterpret_cast<const us8v *>(p);
        return __builtin_bit_cast(bf8v, u);
    }
    __device__ static inline frag neg(frag a) {
        us8v u = __builtin_bit_cast(us8v, a);
        u ^= (unsigned short)0x8000;
        return __builtin_bit_cast(bf8v, u);
    }
    // six slice products, smallest first, FP32 accumulation
    __device__ static inline void mma(acc &c, const frag *a, const frag *b) {
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[2], b[0], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[1], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[2], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[1], b[0], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[1], c, 0, 0, 0);
        c = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[0], b[0], c, 0, 0, 0);
    }
    __device__ static inline double get(const acc &c, int v) { return (double)c[v]; }
    __device__ static inline int row(int lane, int v) { return 4 * (lane >> 4) + v; }
};

// unit = one 16 x KSTEP operand tile in lane order: NSA slices x 64 lanes x EPL
template <class T> constexpr int unit_elems() { return T::NSA * 64 * T::EPL; }

// ------------------------------------------------- P_l^m recursion (FP64) --

// Running Schmidt semi-normalised P_l^q(cos theta) for l = 0,1,2,...
// (schmidt_legendre_m, schmidt_fix_m.f90:89-141), evaluated one l at a time.
struct SchmidtSeq {
    int q;
    double c, s, sq, df_q, d1, d2, p1, p2;
    __device__ void init(int q_, double c_, double s_) {
        q = q_; c = c_; s = s_;
        if (q <= 0) return;
        double prod = 1.0;
        for (int k = 1; k <= q; ++k) prod = prod * (double)(2 * k - 1) / (double)(2 * k);
        df_q = prod;
        sq = 1.0;
        for (int k = 1; k <= q; ++k) sq = sq * s;
    }
    // value at degree l; must be called for l = 0,1,2,... in order
    __device__ double at(int l) {
        if (q < 0) return 0.0;
        if (q == 0) {
            double v;
            if (l == 0) v = 1.0;
            else if (l == 1) v = c;
            else v = p1 * (double)(2 * l - 1) / (double)l * c - p2 * (double)(l - 1) / (double)l;
            p2 = p1; p1 = v;
            return v;
        }
        if (l < q) return 0.0;
        double df;
        if (l == q) {
            df = sqrt(2.0 * df_q);
            d1 = df;
        } else if (l == q + 1) {
            df = sqrt(2.0 * df_q * (double)(2 * q + 1)) * c;
            d2 = d1; d1 = df;
        } else {
            df = (c * (double)(2 * l - 1) * d1 - sqrt((double)((l - 1) * (l - 1) - q * q)) * d2)
                 / sqrt((double)(l * l - q * q));
            d2 = d1; d1 = df;
        }
        return df * sq;
    }
};

// raw tables P, dP/dtheta and m P / sin(theta) at the hemisphere rows:
// raw[(pos(m) + l - m) * th_h + t], pos(m) = sum_{m'<m} (L - m' + 1).
// dP follows diff_schmidt_legendre_m (schmidt_fix_m.f90:145-175).
__global__ void k_schmidt_raw(int L, int th_h, const double *colat,
                              double *P, double *dP, double *PM)
{
    const int m = blockIdx.y;
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= th_h) return;
    const double th = colat[t];
    const double c = cos(th), s = sin(th), as = 1.0 / s;
    const long long pos = (long long)m * (L + 1) - (long long)m * (m - 1) / 2;
    SchmidtSeq sn, sm, sp;
    sn.init(m - 1, c, s);
    sm.init(m, c, s);
    sp.init(m + 1, c, s);
    for (int l = 0; l <= L; ++l) {
        double pn = sn.at(l), pm = sm.at(l), pp = sp.at(l);
        if (l < m) continue;
        double d;
        if (m == 0) {
            d = (l == 0) ? 0.0 : -sqrt((double)(l * (l + 1) / 2)) * pp;
        } else if (m == 1) {
            d = (l == 1) ? pn
                         : 0.5 * (sqrt((double)(2 * l * (l + 1))) * pn
                                  - sqrt((double)((l - 1) * (l + 2))) * pp);
        } else {
            d = (l == m) ? 0.5 * sqrt((double)(2 * m)) * pn
                         : 0.5 * (sqrt((double)((l + m) * (l - m + 1))) * pn
                                  - sqrt((double)((l - m) * (l + m + 1))) * pp);
        }
        const long long o = (pos + (l - m)) * th_h + t;
        P[o] = pm;
        dP[o] = d;
        PM[o] = pm * (double)m * as;
    }
}

// bwd units: for m: for rt < RT: for ks < KS(m): for w < 6
// (P even, P odd, dP even, dP odd, PM even, PM odd); element (s, lane, e):
// row t = rt*16 + lane%16, degree index i = ks*KSTEP + (KSTEP/4)*(lane/16) + e
template <class T>
__global__ void k_pack_bwd(int L, int th_h, int RT, const double *P,
                           const double *dP, const double *PM,
                           const long long *off, typename T::elem *tab)
{
    const int m = blockIdx.y;
    const long long u0 = off[m], nu = off[m + 1] - off[m];
    const int K = L - m + 1, K0 = (K + 1) / 2;
    const int KS = (K0 + T::KSTEP - 1) / T::KSTEP;
    const long long pos = (long long)m * (L + 1) - (long long)m * (m - 1) / 2;
    constexpr int UE = unit_elems<T>() / T::NS;  // per slice
    for (long long x = blockIdx.x * (long long)blockDim.x + threadIdx.x; x < nu * UE;
         x += (long long)gridDim.x * blockDim.x) {
        const long long u = x / UE;
        const int r = (int)(x % UE);
        const int lane = r / T::EPL, e = r % T::EPL;
        const int w = (int)(u % 6);
        const long long u2 = u / 6;
        const int ks = (int)(u2 % KS), rt = (int)(u2 / KS);
        const int par = w & 1;
        const int t = rt * 16 + (lane & 15);
        const int i = ks * T::KSTEP + (T::KSTEP / 4) * (lane >> 4) + e;
        const int l = m + par + 2 * i;
        double v = 0.0;
        if (t < th_h && l <= L) {
            const double *src = (w < 2) ? P : (w < 4) ? dP : PM;
            v = src[(pos + (l - m)) * th_h + t];
        }
        typename T::elem *dst = tab + (u0 + u) * unit_elems<T>() + lane * T::EPL + e;
        T::store(dst, v, 64 * T::EPL);
    }
}

// fwd units: for m: for par: for rt < RT_par(m): for ks < KSt: for w < 3
// (P, dP, PM); element (s, lane, e): degree index i = rt*16 + lane%16,
// row t = ks*KSTEP + (KSTEP/4)*(lane/16) + e
template <class T>
__global__ void k_pack_fwd(int L, int th_h, int KSt, const double *P,
                           const double *dP, const double *PM,
                           const long long *off, typename T::elem *tab)
{
    const int mp = blockIdx.y;  // 2*m + par
    const int m = mp >> 1, par = mp & 1;
    const long long u0 = off[mp], nu = off[mp + 1] - off[mp];
    const long long pos = (long long)m * (L + 1) - (long long)m * (m - 1) / 2;
    constexpr int UE = unit_elems<T>() / T::NS;
    for (long long x = blockIdx.x * (long long)blockDim.x + threadIdx.x; x < nu * UE;
         x += (long long)gridDim.x * blockDim.x) {
        const long long u = x / UE;
        const int r = (int)(x % UE);
        const int lane = r / T::EPL, e = r % T::EPL;
        const int w = (int)(u % 3);
        const long long u2 = u / 3;
        const int ks = (int)(u2 % KSt), rt = (int)(u2 / KSt);
        const int i = rt * 16 + (lane & 15);
        const int t = ks * T::KSTEP + (T::KSTEP / 4) * (lane >> 4) + e;
        const int l = m + par + 2 * i;
        double v = 0.0;
        if (t < th_h && l <= L) {
            const double *src = (w == 0) ? P : (w == 1) ? dP : PM;
            v = src[(pos + (l - m)) * th_h + t];
        }
        typename T::elem *dst = tab + (u0 + u) * unit_elems<T>() + lane * T::EPL + e;
        T::store(dst, v, 64 * T::EPL);
    }
}

// ------------------------------------------------------ kernel params --

struct LegArgs {
    const double *in;
    double *out;
    const void *tab;
    const long long *off;
    const int *lstack;
    const double *radius;
    const double *weight;
    const WorkItem *items;
    const double *rsin;     // 1/sin(theta) at the hemisphere rows
    const double *gfac;     // [L+1] g7 and [L+1] g6 of order m > 0
    int L, nth, nr, th_h, J;
    int ncomp, nvec, nscl;
    int KT, RT, KSt;
    int TC;     // colatitudes per LDS chunk (fwd), multiple of KSTEP
    int pair;   // fwd: the item also carries order L-m
};

constexpr int kThreads = 256;
constexpr int kChunkB = 32;   // degree pairs staged per LDS chunk (bwd)
constexpr int kChunkF = 64;   // largest colatitude chunk staged in LDS (fwd)

// Row padding of the LDS operand tiles.  A B fragment read has 16 columns x 4
// K values per wavefront; a row stride of 2 (mod 32) doubles or 4 (mod 64)
// floats puts the 64 lanes of one read on distinct banks.
template <class T> __host__ __device__ constexpr int lds_pad()
{
    return T::EPL == 8 ? 8 : (sizeof(typename T::elem) == 8 ? 2 : 4);
}
// third form: an odd FP64 column pitch puts the 16 columns of a B fragment
// read on 16 distinct bank pairs
template <class T> __host__ __device__ constexpr int lds_pad3()
{
    return sizeof(typename T::elem) == 8 ? 1 : lds_pad<T>();
}

// ------------------------------------------------------- backward ------

// Column map of the backward kernel for one work item.
//  GEMM1 (A = P):       c = (kk*2+sg)*(nv+ns) + q
//      q <  nv: v_r of vector q = sum P * l(l+1)/r^2 S_pol
//      q >= nv: scalar q-nv     = sum P * S
//  GEMM2 (A = dP | PM): c = ((kk*2+sg)*nv + q)*2 + a, a = 0 theta, 1 phi
//      v_theta(sm) = sum dP dS(sm)/r + (sm/sin) sum P T(-sm)/r
//      v_phi(sm)   = -sum dP T(sm)/r + (sm/sin) sum P dS(-sm)/r
// (legendre_bwd_trans_org.f90:71-151 with the +-m partner mn = M-mp+1).
template <class T, int RPW, int MCT1, int MCT2, int PD>
__global__ __launch_bounds__(kThreads) void k_leg_bwd(LegArgs a)
{
    using elem = typename T::elem;
    using frag = typename T::frag;
    using acc = typename T::acc;
    constexpr int KS_ = T::KSTEP, NS = T::NS, EPL = T::EPL;
    constexpr int KC = (kChunkB / KS_ > 0 ? kChunkB : KS_);
    const int KCP = KC + lds_pad<T>();

    const WorkItem it = a.items[blockIdx.x];
    if (it.m < 0) return;   // padding of the XCD interleave
    const int m = it.m, kt0 = it.ktile * a.KT;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int K = a.L - m + 1, K0 = (K + 1) / 2;
    const int KS = (K0 + KS_ - 1) / KS_;
    const int nkk = min(a.KT, a.nr - kt0);
    const int nv = a.nvec, ns = a.nscl, ncomp = a.ncomp;
    const int ncol = a.KT * 2 * ncomp;
    const int c1n = a.KT * 2 * (nv + ns), c2n = a.KT * 2 * 2 * nv;
    const int nct1 = (c1n + 15) / 16, nct2 = (c2n + 15) / 16;
    const int j0p = a.lstack[a.L + m], j0n = a.lstack[a.L - m];

    extern __shared__ __align__(16) unsigned char smem_raw[];
    elem *X = reinterpret_cast<elem *>(smem_raw);  // [NS][2 par][ncol][KCP]
    const int slice = 2 * ncol * KCP;
    __shared__ double rinv[8];                     // 1/r of the item's radii
    if (threadIdx.x < a.KT && threadIdx.x < 8)
        rinv[threadIdx.x] = (kt0 + (int)threadIdx.x < a.nr) ? 1.0 / a.radius[kt0 + threadIdx.x] : 0.0;

    // per-lane B column descriptors
    int b1col[MCT1];
    int b2x[MCT2], b2y[MCT2];
    bool b2xn[MCT2], b2yn[MCT2];
#pragma unroll
    for (int ct = 0; ct < MCT1; ++ct) {
        int c = ct * 16 + (lane & 15);
        b1col[ct] = 0;
        if (c < c1n) {
            int per = nv + ns, kk = c / (2 * per), rem = c % (2 * per);
            int sg = rem / per, q = rem % per;
            int comp = q < nv ? 3 * q : 3 * nv + (q - nv);
            b1col[ct] = (kk * 2 + sg) * ncomp + comp;
        }
    }
#pragma unroll
    for (int ct = 0; ct < MCT2; ++ct) {
        int c = ct * 16 + (lane & 15);
        b2x[ct] = b2y[ct] = 0;
        b2xn[ct] = b2yn[ct] = false;
        if (c < c2n) {
            int kk = c / (4 * nv), rem = c % (4 * nv);
            int sg = rem / (2 * nv), r2 = rem % (2 * nv);
            int q = r2 >> 1, ax = r2 & 1;
            int xs = (kk * 2 + sg) * ncomp, ys = (kk * 2 + (1 - sg)) * ncomp;
            if (ax == 0) {  // theta: X = dS(sg), Y = +-T(-sg)
                b2x[ct] = xs + 3 * q + 1; b2xn[ct] = false;
                b2y[ct] = ys + 3 * q + 2; b2yn[ct] = (sg == 1);
            } else {        // phi: X = -T(sg), Y = +-dS(-sg)
                b2x[ct] = xs + 3 * q + 2; b2xn[ct] = true;
                b2y[ct] = ys + 3 * q + 1; b2yn[ct] = (sg == 1);
            }
        }
    }

    acc a1[RPW][MCT1][2], a2[RPW][MCT2][2];
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
#pragma unroll
        for (int ct = 0; ct < MCT1; ++ct) { a1[r][ct][0] = acc{}; a1[r][ct][1] = acc{}; }
#pragma unroll
        for (int ct = 0; ct < MCT2; ++ct) { a2[r][ct][0] = acc{}; a2[r][ct][1] = acc{}; }
    }

    const elem *tab = reinterpret_cast<const elem *>(a.tab);
    const long long ubase = a.off[m];
    constexpr int UE = unit_elems<T>();

    // P fragments run PD K steps ahead of the MFMAs in a register ring
    // (slot = step % PD; chunks start on multiples of PD)
    frag Aq[NS == 1 ? PD : 1][RPW][6];
    auto load_a = [&](frag (&A)[RPW][6], int ks) {
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            const int rt = (it.rgroup * 4 + wave) * RPW + r;
            if (rt >= a.RT) continue;
            const elem *ap = tab + (ubase + ((long long)rt * KS + ks) * 6) * UE + lane * EPL;
#pragma unroll
            for (int w = 0; w < 6; ++w) A[r][w] = T::load(ap + w * UE);
        }
    };
    if constexpr (NS == 1) {
#pragma unroll
        for (int d = 0; d < PD; ++d)
            if (d < KS) load_a(Aq[d], d);
    }

    for (int c0 = 0; c0 < KS * KS_; c0 += KC) {
        __syncthreads();
        // stage degree pairs [c0, c0+KC) of both parities, both +-m, KT radii
        // (degree, component) pairs over the threads; the (radius, sign)
        // columns in an inner loop, so each thread divides once per pair
        for (int q = threadIdx.x; q < 2 * KC * ncomp; q += kThreads) {
            const int comp = q % ncomp, li = q / ncomp;
            const int l = m + 2 * c0 + li;
            const int par = li & 1, il = li >> 1;
            const bool lok = l <= a.L;
            const double g3 = (l == 0) ? 0.5 : (double)(l * (l + 1));
            const int kind = comp < 3 * nv ? (comp % 3 == 0 ? 0 : 1) : 2;
            for (int g = 0; g < 2 * a.KT; ++g) {
                const int sg = g & 1, kk = g >> 1;
                double v = 0.0;
                if (lok && kk < nkk && (sg == 0 || m > 0)) {
                    const int k = kt0 + kk;
                    const int j = (sg ? j0n : j0p) + (l - m);
                    v = a.in[comp + (size_t)ncomp * (j + (size_t)a.J * k)];
                    const double a1r = rinv[kk];
                    if (kind == 0) v = v * (a1r * a1r) * g3;
                    else if (kind == 1) v = v * a1r;
                }
                T::store(X + (par * ncol + g * ncomp + comp) * KCP + il, v, slice);
            }
        }
        __syncthreads();

        const int ks_end = min(KS, (c0 + KC) / KS_);
        if constexpr (NS == 1) {
            for (int kb = c0 / KS_; kb < ks_end; kb += PD) {
#pragma unroll
                for (int d = 0; d < PD; ++d) {
                    const int ks = kb + d;
                    if (ks >= ks_end) break;
                    const int kl = (ks * KS_ - c0) + (KS_ / 4) * (lane >> 4);
                    frag be[MCT1], bo[MCT1], xe[MCT2], xo[MCT2], ye[MCT2], yo[MCT2];
#pragma unroll
                    for (int ct = 0; ct < MCT1; ++ct) {
                        be[ct] = T::load(X + (0 * ncol + b1col[ct]) * KCP + kl);
                        bo[ct] = T::load(X + (1 * ncol + b1col[ct]) * KCP + kl);
                    }
#pragma unroll
                    for (int ct = 0; ct < MCT2; ++ct) {
                        xe[ct] = T::load(X + (0 * ncol + b2x[ct]) * KCP + kl);
                        xo[ct] = T::load(X + (1 * ncol + b2x[ct]) * KCP + kl);
                        ye[ct] = T::load(X + (0 * ncol + b2y[ct]) * KCP + kl);
                        yo[ct] = T::load(X + (1 * ncol + b2y[ct]) * KCP + kl);
                        if (b2xn[ct]) { xe[ct] = T::neg(xe[ct]); xo[ct] = T::neg(xo[ct]); }
                        if (b2yn[ct]) { ye[ct] = T::neg(ye[ct]); yo[ct] = T::neg(yo[ct]); }
                    }
#pragma unroll
                    for (int r = 0; r < RPW; ++r) {
                        const int rt = (it.rgroup * 4 + wave) * RPW + r;
                        if (rt >= a.RT) continue;
#pragma unroll
                        for (int ct = 0; ct < MCT1; ++ct) {
                            if (ct >= nct1) break;
                            T::mma(a1[r][ct][0], &Aq[d][r][0], &be[ct]);
                            T::mma(a1[r][ct][1], &Aq[d][r][1], &bo[ct]);
                        }
#pragma unroll
                        for (int ct = 0; ct < MCT2; ++ct) {
                            if (ct >= nct2) break;
                            // symmetric set: dP odd, PM even; antisymmetric: dP even, PM odd
                            T::mma(a2[r][ct][0], &Aq[d][r][3], &xo[ct]);
                            T::mma(a2[r][ct][0], &Aq[d][r][4], &ye[ct]);
                            T::mma(a2[r][ct][1], &Aq[d][r][2], &xe[ct]);
                            T::mma(a2[r][ct][1], &Aq[d][r][5], &yo[ct]);
                        }
                    }
                    if (ks + PD < KS) load_a(Aq[d], ks + PD);
                }
            }
        } else {
            for (int ks = c0 / KS_; ks < ks_end; ++ks) {
                const int kl = (ks * KS_ - c0) + (KS_ / 4) * (lane >> 4);
    #pragma unroll
                for (int r = 0; r < RPW; ++r) {
                    const int rt = (it.rgroup * 4 + wave) * RPW + r;
                    if (rt >= a.RT) continue;
                    frag A[6][NS];
                    const elem *ap = tab + (ubase + ((long long)rt * KS + ks) * 6) * UE + lane * EPL;
    #pragma unroll
                    for (int w = 0; w < 6; ++w)
    #pragma unroll
                        for (int s = 0; s < NS; ++s) A[w][s] = T::load(ap + w * UE + s * 64 * EPL);
    #pragma unroll
                    for (int ct = 0; ct < MCT1; ++ct) {
                        if (ct >= nct1) break;
                        frag be[NS], bo[NS];
    #pragma unroll
                        for (int s = 0; s < NS; ++s) {
                            be[s] = T::load(X + s * slice + (0 * ncol + b1col[ct]) * KCP + kl);
                            bo[s] = T::load(X + s * slice + (1 * ncol + b1col[ct]) * KCP + kl);
                        }
                        T::mma(a1[r][ct][0], A[0], be);
                        T::mma(a1[r][ct][1], A[1], bo);
                    }
    #pragma unroll
                    for (int ct = 0; ct < MCT2; ++ct) {
                        if (ct >= nct2) break;
                        frag xe[NS], xo[NS], ye[NS], yo[NS];
    #pragma unroll
                        for (int s = 0; s < NS; ++s) {
                            xe[s] = T::load(X + s * slice + (0 * ncol + b2x[ct]) * KCP + kl);
                            xo[s] = T::load(X + s * slice + (1 * ncol + b2x[ct]) * KCP + kl);
                            ye[s] = T::load(X + s * slice + (0 * ncol + b2y[ct]) * KCP + kl);
                            yo[s] = T::load(X + s * slice + (1 * ncol + b2y[ct]) * KCP + kl);
                            if (b2xn[ct]) { xe[s] = T::neg(xe[s]); xo[s] = T::neg(xo[s]); }
                            if (b2yn[ct]) { ye[s] = T::neg(ye[s]); yo[s] = T::neg(yo[s]); }
                        }
                        // symmetric set: dP odd, PM even; antisymmetric: dP even, PM odd
                        T::mma(a2[r][ct][0], A[3], xo);
                        T::mma(a2[r][ct][0], A[4], ye);
                        T::mma(a2[r][ct][1], A[2], xe);
                        T::mma(a2[r][ct][1], A[5], yo);
                    }
                }
            }
        }
    }

    // epilogue: hemisphere row t -> out(t) = S + A, mirror nth-1-t -> S - A
    const size_t sl = (size_t)a.nth * a.nr;
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
        const int rt = (it.rgroup * 4 + wave) * RPW + r;
        if (rt >= a.RT) continue;
#pragma unroll
        for (int ct = 0; ct < MCT1; ++ct) {
            if (ct >= nct1) break;
            int c = ct * 16 + (lane & 15);
            if (c >= c1n) continue;
            int per = nv + ns, kk = c / (2 * per), rem = c % (2 * per);
            int sg = rem / per, q = rem % per;
            if (kk >= nkk || (sg == 1 && m == 0)) continue;
            int comp = q < nv ? 3 * q : 3 * nv + (q - nv);
            int mp = a.L + (sg ? -m : m), k = kt0 + kk;
            double *o = a.out + comp + (size_t)ncomp * ((size_t)a.nth * k + sl * mp);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                int t = rt * 16 + T::row(lane, v);
                if (t >= a.th_h) continue;
                double S = T::get(a1[r][ct][0], v), An = T::get(a1[r][ct][1], v);
                int tm = a.nth - 1 - t;
                o[(size_t)ncomp * t] = S + An;
                if (tm != t) o[(size_t)ncomp * tm] = S - An;
            }
        }
#pragma unroll
        for (int ct = 0; ct < MCT2; ++ct) {
            if (ct >= nct2) break;
            int c = ct * 16 + (lane & 15);
            if (c >= c2n) continue;
            int kk = c / (4 * nv), rem = c % (4 * nv);
            int sg = rem / (2 * nv), r2 = rem % (2 * nv);
            int q = r2 >> 1, ax = r2 & 1;
            if (kk >= nkk || (sg == 1 && m == 0)) continue;
            int comp = 3 * q + 1 + ax;
            int mp = a.L + (sg ? -m : m), k = kt0 + kk;
            double *o = a.out + comp + (size_t)ncomp * ((size_t)a.nth * k + sl * mp);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                int t = rt * 16 + T::row(lane, v);
                if (t >= a.th_h) continue;
                double S = T::get(a2[r][ct][0], v), An = T::get(a2[r][ct][1], v);
                int tm = a.nth - 1 - t;
                o[(size_t)ncomp * t] = S + An;
                if (tm != t) o[(size_t)ncomp * tm] = S - An;
            }
        }
    }
}

// -------------------------------------------------------- forward ------

// Column map of the forward kernel (rows = degree index of one parity):
//  GEMM1 (A = P):  c = (kk*2+sg)*(nv+ns) + q
//      S_pol(sm) = r^2 g7 sum w P v_r(sm);  scalar = g6 sum w P f(sm)
//  GEMM2 (A = dP | PM): c = ((kk*2+sg)*nv + q)*2 + a
//      dS(sm) = r g7 [ sum w dP v_th(sm) - (sm/sin) sum w P v_ph(-sm) ]
//      T(sm)  = -r g7 [ sum w dP v_ph(sm) + (sm/sin) sum w P v_th(-sm) ]
// (legendre_fwd_trans_org.f90:70-115).  The sphere sum is folded onto the
// hemisphere: P-terms of an even (l-m) row use V(t)+V(pi-t), dP-terms use
// V(t)-V(pi-t); odd rows swap the two.
//
// A work-group takes the orders m and L-m together (the pairing of the
// reference's order distribution, set_sph_tranform_ordering.f90:69-159), so
// every group has ~(L+2)/16 degree rows whatever m is.
template <class T, int RPW, int MCT1, int MCT2, int PD>
__global__ __launch_bounds__(kThreads) void k_leg_fwd(LegArgs a)
{
    using elem = typename T::elem;
    using frag = typename T::frag;
    using acc = typename T::acc;
    constexpr int KS_ = T::KSTEP, NS = T::NS, EPL = T::EPL;
    const int TC = a.TC;
    const int TCP = TC + lds_pad<T>();

    const WorkItem it = a.items[blockIdx.x];
    if (it.m < 0) return;   // padding of the XCD interleave
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    // orders of this item: m_0 = it.m, m_1 = L - it.m (if different)
    const int nm = (a.pair && a.L - it.m != it.m) ? 2 : 1;
    int mm[2], K0v[2], K1v[2], RTe[2], RTo[2], j0p[2], j0n[2];
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int m = (i == 0) ? it.m : a.L - it.m;
        mm[i] = m;
        const int K = a.L - m + 1;
        K0v[i] = (K + 1) / 2;
        K1v[i] = K / 2;
        RTe[i] = (K0v[i] + 15) / 16;
        RTo[i] = (K1v[i] + 15) / 16;
        j0p[i] = a.lstack[a.L + m];
        j0n[i] = a.lstack[a.L - m];
    }
    const int rows0 = RTe[0] + RTo[0];
    const int nrows = rows0 + (nm == 2 ? RTe[1] + RTo[1] : 0);
    const int kt0 = it.ktile * a.KT;
    const int nkk = min(a.KT, a.nr - kt0);
    const int nv = a.nvec, ns = a.nscl, ncomp = a.ncomp;
    const int ncolm = a.KT * 2 * ncomp;      // columns of one order
    const int ncol = nm * ncolm;
    const int c1n = a.KT * 2 * (nv + ns), c2n = a.KT * 2 * 2 * nv;
    const int nct1 = (c1n + 15) / 16, nct2 = (c2n + 15) / 16;

    extern __shared__ __align__(16) unsigned char smem_raw[];
    elem *V = reinterpret_cast<elem *>(smem_raw);  // [NS][2 fold][ncol][TCP]
    const int slice = 2 * ncol * TCP;

    int b1col[MCT1];
    int b2x[MCT2], b2y[MCT2];
    bool b2xn[MCT2], b2yn[MCT2];
#pragma unroll
    for (int ct = 0; ct < MCT1; ++ct) {
        int c = ct * 16 + (lane & 15);
        b1col[ct] = 0;
        if (c < c1n) {
            int per = nv + ns, kk = c / (2 * per), rem = c % (2 * per);
            int sg = rem / per, q = rem % per;
            int comp = q < nv ? 3 * q : 3 * nv + (q - nv);
            b1col[ct] = (kk * 2 + sg) * ncomp + comp;
        }
    }
#pragma unroll
    for (int ct = 0; ct < MCT2; ++ct) {
        int c = ct * 16 + (lane & 15);
        b2x[ct] = b2y[ct] = 0;
        b2xn[ct] = b2yn[ct] = false;
        if (c < c2n) {
            int kk = c / (4 * nv), rem = c % (4 * nv);
            int sg = rem / (2 * nv), r2 = rem % (2 * nv);
            int q = r2 >> 1, ax = r2 & 1;
            int xs = (kk * 2 + sg) * ncomp, ys = (kk * 2 + (1 - sg)) * ncomp;
            if (ax == 0) {  // dS: X = v_th(sg), Y = -+ v_ph(-sg)
                b2x[ct] = xs + 3 * q + 1; b2xn[ct] = false;
                b2y[ct] = ys + 3 * q + 2; b2yn[ct] = (sg == 0);
            } else {        // T: X = -v_ph(sg), Y = -+ v_th(-sg)
                b2x[ct] = xs + 3 * q + 2; b2xn[ct] = true;
                b2y[ct] = ys + 3 * q + 1; b2yn[ct] = (sg == 0);
            }
        }
    }

    // row tile r of this wave -> (order slot i, parity, tile index)
    int ri[RPW], rpar[RPW], rtile[RPW];
    bool rok[RPW];
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
        const int rr = (it.rgroup * 4 + wave) * RPW + r;
        rok[r] = rr < nrows;
        int i = (rr < rows0) ? 0 : 1;
        int x = (i == 0) ? rr : rr - rows0;
        int p = (x < RTe[i]) ? 0 : 1;
        ri[r] = i;
        rpar[r] = p;
        rtile[r] = p ? x - RTe[i] : x;
        if (!rok[r]) { ri[r] = 0; rpar[r] = 0; rtile[r] = 0; }
    }

    acc a1[RPW][MCT1], a2[RPW][MCT2];
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
#pragma unroll
        for (int ct = 0; ct < MCT1; ++ct) a1[r][ct] = acc{};
#pragma unroll
        for (int ct = 0; ct < MCT2; ++ct) a2[r][ct] = acc{};
    }

    const elem *tab = reinterpret_cast<const elem *>(a.tab);
    constexpr int UE = unit_elems<T>();
    const size_t sl = (size_t)a.nth * a.nr;

    auto aptr = [&](int r, int ks) -> const elem * {
        const int m = mm[ri[r]];
        return tab + (a.off[2 * m + rpar[r]] + ((long long)rtile[r] * a.KSt + ks) * 3) * UE
               + lane * EPL;
    };

    auto load_a = [&](frag (&A)[RPW][3][NS], int ks) {
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            if (!rok[r]) continue;
            const elem *ap = aptr(r, ks);
#pragma unroll
            for (int w = 0; w < 3; ++w)
#pragma unroll
                for (int s2 = 0; s2 < NS; ++s2) A[r][w][s2] = T::load(ap + w * UE + s2 * 64 * EPL);
        }
    };
    auto step = [&](const frag (&A)[RPW][3][NS], int ks, int t0) {
        const int kl = (ks * KS_ - t0) + (KS_ / 4) * (lane >> 4);
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            if (!rok[r]) continue;
            const int par = rpar[r];
            const int cofs = ri[r] * ncolm;
            const int fP = par, fD = 1 - par;  // fold index: 0 sym, 1 anti
#pragma unroll
            for (int ct = 0; ct < MCT1; ++ct) {
                if (ct >= nct1) break;
                frag b[NS];
#pragma unroll
                for (int s2 = 0; s2 < NS; ++s2)
                    b[s2] = T::load(V + s2 * slice + (fP * ncol + cofs + b1col[ct]) * TCP + kl);
                T::mma(a1[r][ct], A[r][0], b);
            }
#pragma unroll
            for (int ct = 0; ct < MCT2; ++ct) {
                if (ct >= nct2) break;
                frag x[NS], y[NS];
#pragma unroll
                for (int s2 = 0; s2 < NS; ++s2) {
                    x[s2] = T::load(V + s2 * slice + (fD * ncol + cofs + b2x[ct]) * TCP + kl);
                    y[s2] = T::load(V + s2 * slice + (fP * ncol + cofs + b2y[ct]) * TCP + kl);
                    if (b2xn[ct]) x[s2] = T::neg(x[s2]);
                    if (b2yn[ct]) y[s2] = T::neg(y[s2]);
                }
                T::mma(a2[r][ct], A[r][1], x);
                T::mma(a2[r][ct], A[r][2], y);
            }
        }
    };
    // single-slice P fragments run PD K steps ahead in a register ring
    // (slot = step % PD; chunks start on multiples of PD)
    frag Aq[NS == 1 ? PD : 1][RPW][3][NS];
    if constexpr (NS == 1) {
#pragma unroll
        for (int d = 0; d < PD; ++d)
            if (d < a.KSt) load_a(Aq[d], d);
    }

    for (int t0 = 0; t0 < a.KSt * KS_; t0 += TC) {
        __syncthreads();
        // (colatitude, component) pairs over the threads; the (order,
        // radius, sign) columns in inner loops
        for (int q = threadIdx.x; q < TC * ncomp; q += kThreads) {
            const int comp = q % ncomp, tl = q / ncomp;
            const int t = t0 + tl;
            const bool tok = t < a.th_h;
            const int tm = a.nth - 1 - t;
            const double w = tok ? a.weight[t] : 0.0;
            for (int i = 0; i < nm; ++i) {
                const int m = mm[i];
                for (int kk = 0; kk < a.KT; ++kk)
                    for (int sg = 0; sg < 2; ++sg) {
                        double vs = 0.0, va = 0.0;
                        if (tok && kk < nkk && (sg == 0 || m > 0)) {
                            const int mp = a.L + (sg ? -m : m), k = kt0 + kk;
                            const double *src = a.in + comp + (size_t)ncomp * ((size_t)a.nth * k + sl * mp);
                            const double v1 = src[(size_t)ncomp * t];
                            if (tm != t) {
                                const double v2 = src[(size_t)ncomp * tm];
                                vs = w * (v1 + v2);
                                va = w * (v1 - v2);
                            } else {
                                vs = w * v1;
                            }
                        }
                        const int col = i * ncolm + (kk * 2 + sg) * ncomp + comp;
                        T::store(V + (0 * ncol + col) * TCP + tl, vs, slice);
                        T::store(V + (1 * ncol + col) * TCP + tl, va, slice);
                    }
            }
        }
        __syncthreads();

        const int ks_end = min(a.KSt, (t0 + TC) / KS_);
        if constexpr (NS == 1) {
            for (int kb = t0 / KS_; kb < ks_end; kb += PD) {
#pragma unroll
                for (int d = 0; d < PD; ++d) {
                    const int ks = kb + d;
                    if (ks >= ks_end) break;
                    step(Aq[d], ks, t0);
                    if (ks + PD < a.KSt) load_a(Aq[d], ks + PD);
                }
            }
        } else {
            for (int ks = t0 / KS_; ks < ks_end; ++ks) {
                frag A[RPW][3][NS];
                load_a(A, ks);
                step(A, ks, t0);
            }
        }
    }

    // epilogue: scale by r^2 g7, r g7 or g6 (sph_normalizations,
    // spherical_harmonics.f90:121-139) and store in rlm order
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
        if (!rok[r]) continue;
        const int i_m = ri[r], par = rpar[r], rt = rtile[r];
        const int m = mm[i_m];
        const int Kp = par ? K1v[i_m] : K0v[i_m];
#pragma unroll
        for (int ct = 0; ct < MCT1 + MCT2; ++ct) {
            const bool g1 = ct < MCT1;
            const int cc = g1 ? ct : ct - MCT1;
            if (g1 ? (cc >= nct1) : (cc >= nct2)) continue;
            int c = cc * 16 + (lane & 15);
            int kk, sg, comp, kind;  // kind 0: r^2 g7, 1: r g7, 2: g6
            if (g1) {
                if (c >= c1n) continue;
                int per = nv + ns;
                kk = c / (2 * per);
                int rem = c % (2 * per);
                sg = rem / per;
                int q = rem % per;
                comp = q < nv ? 3 * q : 3 * nv + (q - nv);
                kind = q < nv ? 0 : 2;
            } else {
                if (c >= c2n) continue;
                kk = c / (4 * nv);
                int rem = c % (4 * nv);
                sg = rem / (2 * nv);
                int r2 = rem % (2 * nv);
                comp = 3 * (r2 >> 1) + 1 + (r2 & 1);
                kind = 1;
            }
            if (kk >= nkk || (sg == 1 && m == 0)) continue;
            const int k = kt0 + kk;
            const double rad = a.radius[k];
            const int j0 = sg ? j0n[i_m] : j0p[i_m];
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = rt * 16 + T::row(lane, v);
                if (i >= Kp) continue;
                const int l = m + par + 2 * i;
                double fac;
                if (kind == 2) {
                    fac = (l == 0) ? 0.5 : (m == 0 ? (double)(2 * l + 1) / 2.0
                                                   : (double)(2 * l + 1) / 4.0);
                } else {
                    double g7 = (l == 0) ? 1.0
                              : (m == 0 ? (double)(2 * l + 1) / (double)(2 * l * (l + 1))
                                        : (double)(2 * l + 1) / (double)(4 * l * (l + 1)));
                    fac = (kind == 0) ? rad * rad * g7 : rad * g7;
                }
                double val = (g1 ? T::get(a1[r][cc], v) : T::get(a2[r][cc], v)) * fac;
                a.out[comp + (size_t)ncomp * ((j0 + (l - m)) + (size_t)a.J * k)] = val;
            }
        }
    }
}

// ====================================================== third form ====
//
// Both directions, restructured so that the K loop is nothing but operand
// loads and MFMAs:
//  * two P tables (P, dP) instead of three: the m/sin(theta) factor of the
//    PM terms goes onto the staged input (forward, where theta is the K
//    index) or onto the accumulated rows in the epilogue (backward, where
//    theta is the row index), so no PM tile is ever read;
//  * the LDS tile holds the B operand column by column exactly as the MFMA
//    reads it (GEMM1 columns, then the dP columns X, then the P columns Y of
//    the theta/phi GEMM), with signs and factors applied while staging, plus
//    one zero column that the lanes beyond the last column read;
//  * a wave holds row tiles of one parity only, so each B fragment feeds all
//    of its row tiles; a row tile past the end repeats a valid one and is
//    not stored, so there are no per-tile branches in the K loop;
//  * table pointers are formed once; the K steps of every order are padded
//    with zero tiles to a multiple of the prefetch depth;
//  * the global loads of the next LDS chunk are issued before the MFMAs of
//    the current one and written to LDS after them.

constexpr int kSlotF = 6;    // (row, mirror row) pairs staged per thread per chunk (fwd)
// sched_barrier mask: VALU, SALU and LDS may be moved across, MFMA and VMEM not
constexpr int kSchedNoVmem = 0x2 | 0x4 | 0x80 | 0x100 | 0x200;
constexpr int kSlotB = 6;    // spectral values staged per thread per chunk (bwd)

// Marks per-lane values as redefined here, so that quantities derived from
// them are recomputed after this point instead of being hoisted out of the
// enclosing loops and held in registers across the MFMA loop.
template <int N> __device__ inline void opaque(int (&x)[N])
{
#pragma unroll
    for (int i = 0; i < N; ++i) asm volatile("" : "+v"(x[i]));
}

// K-step padding of the third-form tables (backward, forward): the
// prefetch depth of a plan must divide it
template <class T> __host__ __device__ constexpr int kpad3b() { return T::NS == 1 ? 2 : 1; }
template <class T> __host__ __device__ constexpr int kpad3f() { return T::NS == 1 ? 4 : 1; }

// fwd3 units: for m: for par: for rt < RT_par(m): for ks < KSt3: for w < 2
// (P, dP); element (s, lane, e): degree index i = rt*16 + lane%16,
// row t = ks*KSTEP + (KSTEP/4)*(lane/16) + e
template <class T>
__global__ void k_pack_fwd3(int L, int th_h, int KSt, const double *P, const double *dP,
                            const long long *off, typename T::elem *tab)
{
    const int mp = blockIdx.y;  // 2*m + par
    const int m = mp >> 1, par = mp & 1;
    const long long u0 = off[mp], nu = off[mp + 1] - off[mp];
    const long long pos = (long long)m * (L + 1) - (long long)m * (m - 1) / 2;
    constexpr int UE = unit_elems<T>() / T::NS;
    for (long long x = blockIdx.x * (long long)blockDim.x + threadIdx.x; x < nu * UE;
         x += (long long)gridDim.x * blockDim.x) {
        const long long u = x / UE;
        const int r = (int)(x % UE);
        const int lane = r / T::EPL, e = r % T::EPL;
        const int w = (int)(u & 1);
        const long long u2 = u >> 1;
        const int ks = (int)(u2 % KSt), rt = (int)(u2 / KSt);
        const int i = rt * 16 + (lane & 15);
        const int t = ks * T::KSTEP + (T::KSTEP / 4) * (lane >> 4) + e;
        const int l = m + par + 2 * i;
        double v = 0.0;
        if (t < th_h && l <= L) v = (w ? dP : P)[(pos + (l - m)) * th_h + t];
        typename T::elem *dst = tab + (u0 + u) * unit_elems<T>() + lane * T::EPL + e;
        T::store(dst, v, 64 * T::EPL);
    }
}

// bwd3 units: for m: for rt < RT: for ks < KS3(m): for w < 4 (P even,
// P odd, dP even, dP odd); element (s, lane, e): row t = rt*16 + lane%16,
// degree index i = ks*KSTEP + (KSTEP/4)*(lane/16) + e
template <class T>
__global__ void k_pack_bwd3(int L, int th_h, int RT, const double *P, const double *dP,
                            const long long *off, typename T::elem *tab)
{
    const int m = blockIdx.y;
    const long long u0 = off[m], nu = off[m + 1] - off[m];
    const int K0 = (L - m + 2) / 2;
    const int KS = ((K0 + T::KSTEP - 1) / T::KSTEP + kpad3b<T>() - 1) / kpad3b<T>() * kpad3b<T>();
    const long long pos = (long long)m * (L + 1) - (long long)m * (m - 1) / 2;
    constexpr int UE = unit_elems<T>() / T::NS;
    for (long long x = blockIdx.x * (long long)blockDim.x + threadIdx.x; x < nu * UE;
         x += (long long)gridDim.x * blockDim.x) {
        const long long u = x / UE;
        const int r = (int)(x % UE);
        const int lane = r / T::EPL, e = r % T::EPL;
        const int w = (int)(u & 3);
        const long long u2 = u >> 2;
        const int ks = (int)(u2 % KS), rt = (int)(u2 / KS);
        const int par = w & 1;
        const int t = rt * 16 + (lane & 15);
        const int i = ks * T::KSTEP + (T::KSTEP / 4) * (lane >> 4) + e;
        const int l = m + par + 2 * i;
        double v = 0.0;
        if (t < th_h && l <= L) v = (w < 2 ? P : dP)[(pos + (l - m)) * th_h + t];
        typename T::elem *dst = tab + (u0 + u) * unit_elems<T>() + lane * T::EPL + e;
        T::store(dst, v, 64 * T::EPL);
    }
}

__global__ void k_rsin(int th_h, const double *colat, double *rsin)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < th_h) rsin[t] = 1.0 / sin(colat[t]);
}

// Staged column of one input component (both directions).  GEMM1 column of
// v_r / scalar q: (kk*2+sg)*(nv+ns) + q.  GEMM2 column (kk, sg, q, ax):
// CX or CY + ((kk*2+sg)*nv + q)*2 + ax.  A theta/phi component x (1, 2) of
// sign sg is the dP operand X of its own column (kk, sg, q, x-1) and the P
// operand Y of the partner column (kk, 1-sg, q, 2-x).
struct StageCols {
    int ca, cb;   // cb < 0: no second column
};
__device__ inline StageCols stage_cols(int comp, int kk, int sg, int nv, int ns, int CX, int CY)
{
    StageCols s;
    const int gb = kk * 2 + sg;
    s.cb = -1;
    if (comp >= 3 * nv) {
        s.ca = gb * (nv + ns) + nv + (comp - 3 * nv);
    } else {
        const int q3 = comp / 3, x = comp - 3 * q3;
        if (x == 0) {
            s.ca = gb * (nv + ns) + q3;
        } else {
            s.ca = CX + (gb * nv + q3) * 2 + (x - 1);
            s.cb = CY + ((kk * 2 + (1 - sg)) * nv + q3) * 2 + (2 - x);
        }
    }
    return s;
}

// Forward, third form.  Work item = (m, first radial tile, row group, number
// of radial tiles); wave w takes rows of parity w & 1, row tiles
// (rgroup*2 + w/2)*RPW + r, and the work-group runs its radial tiles one
// after the other (everything that depends only on m is formed once).
//   c1 = P  * B1          (S_pol, scalars)
//   c2 = dP * X + P * Y   (dS, T), Y staged with m/sin(theta)
template <class T, int RPW, int NC1, int NC2, int PD, int NW = 4>
__global__ __launch_bounds__(NW * 64) void k_leg_fwd3(LegArgs a)
{
    // NW waves: NW/2 degree tiles of each parity per row group share one
    // staged input
    constexpr int NT = NW * 64;
    using elem = typename T::elem;
    using frag = typename T::frag;
    using acc = typename T::acc;
    constexpr int KS_ = T::KSTEP, NS = T::NS, EPL = T::EPL;
    constexpr int NB = 16 * (NC1 + 2 * NC2) + 1;
    constexpr int CX = 16 * NC1, CY = 16 * (NC1 + NC2), CZ = NB - 1;
    constexpr int UE = unit_elems<T>();
    constexpr int kNone = 0xFFFF;
    const int TC = a.TC, TCP = TC + lds_pad3<T>();

    const WorkItem it = a.items[blockIdx.x];
    if (it.m < 0) return;   // padding of the XCD interleave
    const int m = it.m;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int K = a.L - m + 1;
    const int par = wave & 1;
    const int Kp = par ? K / 2 : (K + 1) / 2;
    const int RTp = (Kp + 15) / 16;
    const int nv = a.nvec, ns = a.nscl, ncomp = a.ncomp;
    const int c1n = a.KT * 2 * (nv + ns), c2n = a.KT * 4 * nv;
    const int nks = a.KSt;
    const int j0p = a.lstack[a.L + m], j0n = a.lstack[a.L - m];

    extern __shared__ __align__(16) unsigned char smem_raw[];
    elem *V = reinterpret_cast<elem *>(smem_raw);   // [NS][2 fold][NB][TCP]
    const int slice = 2 * NB * TCP;
    // per-row factors (w, w m/sin(theta)) after the tile, zero past th_h; the
    // equator row of an odd grid is its own mirror and gets half the weight
    double2 *WF = reinterpret_cast<double2 *>(
        smem_raw + ((NS * slice * sizeof(elem) + 15) & ~(size_t)15));
    for (int x = threadIdx.x; x < 2 * TCP; x += NT)
        T::store(V + ((x / TCP) * NB + CZ) * TCP + x % TCP, 0.0, slice);
    for (int t = threadIdx.x; t < nks * KS_; t += NT) {
        double w = 0.0, y = 0.0;
        if (t < a.th_h) {
            w = a.weight[t];
            if (a.nth - 1 - t == t) w *= 0.5;
            y = w * ((double)m * a.rsin[t]);
        }
        WF[t] = make_double2(w, y);
    }

    int rtile[RPW];
    bool rok[RPW];
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
        const int rt = (it.rgroup * (NW / 2) + (wave >> 1)) * RPW + r;
        rok[r] = rt < RTp;
        rtile[r] = rok[r] ? rt : 0;
    }
    // the last row group of an order can leave a wave without a degree tile:
    // it still stages its share of the input but issues no MFMAs
    bool anyrow = false;
#pragma unroll
    for (int r = 0; r < RPW; ++r) anyrow |= rok[r];
    // an order with no odd rows (m = L) lets its odd waves read even tiles
    const int pt = RTp > 0 ? par : 0;
    const elem *tab = reinterpret_cast<const elem *>(a.tab);
    const elem *ap[RPW];
#pragma unroll
    for (int r = 0; r < RPW; ++r)
        ap[r] = tab + (a.off[2 * m + pt] + (long long)rtile[r] * nks * 2) * UE + lane * EPL;

    // B fragment offsets of this lane (fold 0: V(t)+V(pi-t), 1: V(t)-V(pi-t));
    // P terms of parity par read fold par, dP terms fold 1-par
    const int kq = (KS_ / 4) * (lane >> 4);
    int o1[NC1], ox[NC2], oy[NC2];
#pragma unroll
    for (int ct = 0; ct < NC1; ++ct) {
        const int c = ct * 16 + (lane & 15);
        o1[ct] = (par * NB + (c < c1n ? c : CZ)) * TCP + kq;
    }
#pragma unroll
    for (int ct = 0; ct < NC2; ++ct) {
        const int c = ct * 16 + (lane & 15);
        ox[ct] = ((1 - par) * NB + (c < c2n ? CX + c : CZ)) * TCP + kq;
        oy[ct] = (par * NB + (c < c2n ? CY + c : CZ)) * TCP + kq;
    }
    // output column of this lane per column tile: comp | kk << 8 | sg << 11 |
    // kind << 12 (kind 0: r^2 g7, 1: r g7, 2: g6), or -1
    int ocol[NC1 + NC2];
#pragma unroll
    for (int ct = 0; ct < NC1 + NC2; ++ct) {
        const bool g1c = ct < NC1;
        const int c = (g1c ? ct : ct - NC1) * 16 + (lane & 15);
        int kk, sg, comp, kind;
        if (g1c) {
            const int per = nv + ns;
            kk = c / (2 * per);
            const int rem = c % (2 * per);
            sg = rem / per;
            const int q = rem % per;
            comp = q < nv ? 3 * q : 3 * nv + (q - nv);
            kind = q < nv ? 0 : 2;
        } else {
            kk = c / (4 * nv);
            const int rem = c % (4 * nv);
            sg = rem / (2 * nv);
            const int r2 = rem % (2 * nv);
            comp = 3 * (r2 >> 1) + 1 + (r2 & 1);
            kind = 1;
        }
        const bool ok = (g1c ? c < c1n : c < c2n) && !(sg == 1 && m == 0);
        ocol[ct] = ok ? comp | (kk << 8) | (sg << 11) | (kind << 12) : -1;
    }

    // staging: (component, row, radius/sign) triples over the threads,
    // component fastest so a wave reads contiguous runs of vr_rtm.  Each
    // slot's source offset (radial tile 0, chunk 0), LDS destinations and
    // signs are formed once; per chunk a slot costs two loads, one LDS read of
    // the row factors, the fold and its stores.
    const int nst = TC * a.KT * 2 * ncomp;
    const int nthnr = a.nth * a.nr;
    int s_src[kSlotF], s_dst[kSlotF], s_meta[kSlotF];
#pragma unroll
    for (int i = 0; i < kSlotF; ++i) {
        const int q = threadIdx.x + i * NT;
        s_src[i] = -2;   // no slot
        s_dst[i] = 0;
        s_meta[i] = 0;
        if (q < nst) {
            const int comp = q % ncomp, rest = q / ncomp;
            const int tl = rest % TC, g = rest / TC;
            const int kk = g >> 1, sg = g & 1;
            const StageCols sc = stage_cols(comp, kk, sg, nv, ns, CX, CY);
            s_src[i] = (sg == 0 || m > 0)
                     ? comp + ncomp * (a.nth * kk + nthnr * (a.L + (sg ? -m : m)) + tl) : -1;
            s_dst[i] = (sc.ca * TCP + tl) | ((sc.cb >= 0 ? sc.cb * TCP + tl : kNone) << 16);
            // X = +v_theta | -v_phi; Y = -+ (m/sin) v_(phi|theta)(-sg)
            s_meta[i] = tl | (kk << 8) | ((sc.cb >= 0 && comp % 3 == 2) ? 0x10000 : 0)
                      | (sg == 1 ? 0x20000 : 0);
        }
    }

    auto load_a = [&](frag (&A)[RPW][2][NS], int ks) {
#pragma unroll
        for (int r = 0; r < RPW; ++r)
#pragma unroll
            for (int w = 0; w < 2; ++w)
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    A[r][w][s] = T::load(ap[r] + (ks * 2 + w) * UE + s * 64 * EPL);
    };

    for (int kt = it.ktile; kt < it.ktile + it.nkt; ++kt) {
        const int kt0 = kt * a.KT;
        const int nkk = min(a.KT, a.nr - kt0);
        const int koff = ncomp * a.nth * kt0;
        double g1[kSlotF], g2[kSlotF];
        auto fetch = [&](int t0) {
            opaque(s_src);
            opaque(s_meta);
#pragma unroll
            for (int i = 0; i < kSlotF; ++i) {
                g1[i] = 0.0;
                g2[i] = 0.0;
                if (s_src[i] >= 0 && ((s_meta[i] >> 8) & 7) < nkk) {
                    const int t = t0 + (s_meta[i] & 0xFF);
                    if (t < a.th_h) {
                        const int o = s_src[i] + koff + ncomp * t0;
                        g1[i] = a.in[o];
                        g2[i] = a.in[o + ncomp * (a.nth - 1 - 2 * t)];   // row nth-1-t
                    }
                }
            }
        };
        auto put = [&](int t0) {
            opaque(s_src);
            opaque(s_dst);
            opaque(s_meta);
#pragma unroll
            for (int i = 0; i < kSlotF; ++i) {
                if (s_src[i] == -2) continue;
                const double2 wf = WF[t0 + (s_meta[i] & 0xFF)];
                const double vs = g1[i] + g2[i], va = g1[i] - g2[i];
                const double wx = (s_meta[i] & 0x10000) ? -wf.x : wf.x;
                const int da = s_dst[i] & 0xFFFF, db = (int)((unsigned)s_dst[i] >> 16);
                T::store(V + da, wx * vs, slice);
                T::store(V + NB * TCP + da, wx * va, slice);
                if (db != kNone) {
                    const double fy = (s_meta[i] & 0x20000) ? -wf.y : wf.y;
                    T::store(V + db, fy * vs, slice);
                    T::store(V + NB * TCP + db, fy * va, slice);
                }
            }
        };

        acc c1[RPW][NC1], c2[RPW][NC2];
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
#pragma unroll
            for (int ct = 0; ct < NC1; ++ct) c1[r][ct] = acc{};
#pragma unroll
            for (int ct = 0; ct < NC2; ++ct) c2[r][ct] = acc{};
        }
        auto mstep = [&](const frag (&A)[RPW][2][NS], int kl) {
            frag b1[NC1][NS], bx[NC2][NS], by[NC2][NS];
#pragma unroll
            for (int ct = 0; ct < NC1; ++ct)
#pragma unroll
                for (int s = 0; s < NS; ++s) b1[ct][s] = T::load(V + s * slice + o1[ct] + kl);
#pragma unroll
            for (int ct = 0; ct < NC2; ++ct)
#pragma unroll
                for (int s = 0; s < NS; ++s) {
                    bx[ct][s] = T::load(V + s * slice + ox[ct] + kl);
                    by[ct][s] = T::load(V + s * slice + oy[ct] + kl);
                }
#pragma unroll
            for (int r = 0; r < RPW; ++r) {
#pragma unroll
                for (int ct = 0; ct < NC1; ++ct) T::mma(c1[r][ct], A[r][0], b1[ct]);
#pragma unroll
                for (int ct = 0; ct < NC2; ++ct) {
                    T::mma(c2[r][ct], A[r][1], bx[ct]);
                    T::mma(c2[r][ct], A[r][0], by[ct]);
                }
            }
        };

        frag Aq[PD][RPW][2][NS];
        fetch(0);
#pragma unroll
        for (int d = 0; d < PD; ++d) load_a(Aq[d], min(d, nks - 1));
        for (int t0 = 0; t0 < nks * KS_; t0 += TC) {
            __syncthreads();
            put(t0);
            __syncthreads();
            if (t0 + TC < nks * KS_) fetch(t0 + TC);
            const int kse = anyrow ? min(nks, (t0 + TC) / KS_) : 0;
            for (int kb = t0 / KS_; kb < kse; kb += PD) {
#pragma unroll
                for (int d = 0; d < PD; ++d) {
                    const int ks = kb + d;
                    mstep(Aq[d], ks * KS_ - t0);
                    load_a(Aq[d], min(ks + PD, nks - 1));
                    // keep the table prefetch right after the MFMAs that free
                    // its registers (only VMEM may not cross)
                    __builtin_amdgcn_sched_barrier(kSchedNoVmem);
                }
            }
        }

        // epilogue: scale by r^2 g7, r g7 or g6 (sph_normalizations,
        // spherical_harmonics.f90:121-139; gfac holds the m > 0 values, m = 0
        // doubles them) and store in rlm order
        opaque(ocol);
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            if (!rok[r]) continue;
            const int rt = rtile[r];
#pragma unroll
            for (int ct = 0; ct < NC1 + NC2; ++ct) {
                const int oc = ocol[ct];
                if (oc < 0) continue;
                const int comp = oc & 0xFF, kk = (oc >> 8) & 7, sg = (oc >> 11) & 1;
                const int kind = (oc >> 12) & 3;
                if (kk >= nkk) continue;
                const int k = kt0 + kk;
                const double rad = a.radius[k];
                const int j0 = sg ? j0n : j0p;
                const double m2 = (m == 0) ? 2.0 : 1.0;
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int i = rt * 16 + T::row(lane, v);
                    if (i >= Kp) continue;
                    const int l = m + par + 2 * i;
                    double fac;
                    if (kind == 2) {
                        fac = (l == 0) ? 0.5 : m2 * a.gfac[a.L + 1 + l];
                    } else {
                        const double g7 = (l == 0) ? 1.0 : m2 * a.gfac[l];
                        fac = (kind == 0) ? rad * rad * g7 : rad * g7;
                    }
                    const double val = (ct < NC1 ? T::get(c1[r][ct], v) : T::get(c2[r][ct - NC1], v)) * fac;
                    a.out[comp + (size_t)ncomp * ((j0 + (l - m)) + (size_t)a.J * k)] = val;
                }
            }
        }
    }
}

// Backward, third form.  Work item = (m, first radial tile, row group, number
// of radial tiles); rows are the hemisphere colatitudes, wave w takes row
// tiles (rgroup*4 + w)*RPW + r; K runs over the degrees of both parities
// (folded output S +- A).
//   a1 = P * B1                         (v_r, scalars)
//   ax = dP * X,  ay = P * Y;  v_theta, v_phi = ax + (m/sin(theta)) ay
template <class T, int RPW, int NC1, int NC2, int PD, int NW = 4>
__global__ __launch_bounds__(NW * 64) void k_leg_bwd3(LegArgs a)
{
    // NW waves: NW colatitude tiles per row group share one staged input
    constexpr int NT = NW * 64;
    using elem = typename T::elem;
    using frag = typename T::frag;
    using acc = typename T::acc;
    constexpr int KS_ = T::KSTEP, NS = T::NS, EPL = T::EPL;
    constexpr int NB = 16 * (NC1 + 2 * NC2) + 1;
    constexpr int CX = 16 * NC1, CY = 16 * (NC1 + NC2), CZ = NB - 1;
    constexpr int UE = unit_elems<T>();
    constexpr int kNone = 0xFFFF;
    const int KC = a.TC, KCP = KC + lds_pad3<T>();

    const WorkItem it = a.items[blockIdx.x];
    if (it.m < 0) return;   // padding of the XCD interleave
    const int m = it.m;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int K0 = (a.L - m + 2) / 2;
    const int KS = ((K0 + KS_ - 1) / KS_ + kpad3b<T>() - 1) / kpad3b<T>() * kpad3b<T>();
    const int nv = a.nvec, ns = a.nscl, ncomp = a.ncomp;
    const int c1n = a.KT * 2 * (nv + ns), c2n = a.KT * 4 * nv;
    const int j0p = a.lstack[a.L + m], j0n = a.lstack[a.L - m];

    extern __shared__ __align__(16) unsigned char smem_raw[];
    elem *X = reinterpret_cast<elem *>(smem_raw);   // [NS][2 par][NB][KCP]
    const int slice = 2 * NB * KCP;
    __shared__ double rinv[8];                      // 1/r of the current radial tile
    for (int x = threadIdx.x; x < 2 * KCP; x += NT)
        T::store(X + ((x / KCP) * NB + CZ) * KCP + x % KCP, 0.0, slice);

    int rtile[RPW];
    bool rok[RPW];
#pragma unroll
    for (int r = 0; r < RPW; ++r) {
        const int rt = (it.rgroup * NW + wave) * RPW + r;
        rok[r] = rt < a.RT;
        rtile[r] = rok[r] ? rt : 0;
    }
    // a wave past the last colatitude tile stages its share but issues no MFMAs
    bool anyrow = false;
#pragma unroll
    for (int r = 0; r < RPW; ++r) anyrow |= rok[r];
    const elem *tab = reinterpret_cast<const elem *>(a.tab);
    const elem *ap[RPW];
#pragma unroll
    for (int r = 0; r < RPW; ++r)
        ap[r] = tab + (a.off[m] + (long long)rtile[r] * KS * 4) * UE + lane * EPL;

    const int kq = (KS_ / 4) * (lane >> 4);
    int o1[NC1], ox[NC2], oy[NC2];   // parity-even offsets; odd = + NB*KCP
#pragma unroll
    for (int ct = 0; ct < NC1; ++ct) {
        const int c = ct * 16 + (lane & 15);
        o1[ct] = (c < c1n ? c : CZ) * KCP + kq;
    }
#pragma unroll
    for (int ct = 0; ct < NC2; ++ct) {
        const int c = ct * 16 + (lane & 15);
        ox[ct] = (c < c2n ? CX + c : CZ) * KCP + kq;
        oy[ct] = (c < c2n ? CY + c : CZ) * KCP + kq;
    }
    const int po = NB * KCP;
    // output column of this lane per column tile: comp | kk << 8 | sg << 11, or -1
    int ocol[NC1 + NC2];
#pragma unroll
    for (int ct = 0; ct < NC1 + NC2; ++ct) {
        const bool g1c = ct < NC1;
        const int c = (g1c ? ct : ct - NC1) * 16 + (lane & 15);
        int kk, sg, comp;
        if (g1c) {
            const int per = nv + ns, rem = c % (2 * per), q = rem % per;
            kk = c / (2 * per);
            sg = rem / per;
            comp = q < nv ? 3 * q : 3 * nv + (q - nv);
        } else {
            const int rem = c % (4 * nv), r2 = rem % (2 * nv);
            kk = c / (4 * nv);
            sg = rem / (2 * nv);
            comp = 3 * (r2 >> 1) + 1 + (r2 & 1);
        }
        const bool ok = (g1c ? c < c1n : c < c2n) && !(sg == 1 && m == 0);
        ocol[ct] = ok ? comp | (kk << 8) | (sg << 11) : -1;
    }

    // staging: (component, degree, radius/sign) triples, component fastest
    // (contiguous runs of sp_rlm); degrees [m + 2 c0, m + 2 (c0 + KC)).  Slot
    // descriptors (source offset at radial tile 0 and chunk 0, LDS
    // destinations, kind and signs) are formed once.
    const int nst = 2 * KC * ncomp * a.KT * 2;
    int s_src[kSlotB], s_dst[kSlotB], s_meta[kSlotB];
#pragma unroll
    for (int i = 0; i < kSlotB; ++i) {
        const int q = threadIdx.x + i * NT;
        s_src[i] = -2;   // no slot
        s_dst[i] = 0;
        s_meta[i] = 0;
        if (q < nst) {
            const int comp = q % ncomp, rest = q / ncomp;
            const int li = rest % (2 * KC), gg = rest / (2 * KC);
            const int kk = gg >> 1, sg = gg & 1;
            const int p = li & 1, il = li >> 1;
            const StageCols sc = stage_cols(comp, kk, sg, nv, ns, CX, CY);
            s_src[i] = (sg == 0 || m > 0) ? comp + ncomp * ((sg ? j0n : j0p) + li + a.J * kk) : -1;
            s_dst[i] = (p * NB * KCP + sc.ca * KCP + il)
                     | ((sc.cb >= 0 ? p * NB * KCP + sc.cb * KCP + il : kNone) << 16);
            const int kind = comp < 3 * nv ? (comp % 3 == 0 ? 0 : 1) : 2;
            // X = +dS | -T; Y = +- T | dS of the partner sign
            s_meta[i] = li | (kk << 8) | (kind << 12) | ((sc.cb >= 0 && comp % 3 == 2) ? 1 << 14 : 0)
                      | (sg == 0 ? 1 << 15 : 0);
        }
    }

    auto load_a = [&](frag (&A)[RPW][4][NS], int ks) {
#pragma unroll
        for (int r = 0; r < RPW; ++r)
#pragma unroll
            for (int w = 0; w < 4; ++w)
#pragma unroll
                for (int s = 0; s < NS; ++s)
                    A[r][w][s] = T::load(ap[r] + (ks * 4 + w) * UE + s * 64 * EPL);
    };

    const size_t sl = (size_t)a.nth * a.nr;
    for (int kt = it.ktile; kt < it.ktile + it.nkt; ++kt) {
        const int kt0 = kt * a.KT;
        const int nkk = min(a.KT, a.nr - kt0);
        const int koff = ncomp * a.J * kt0;
        // the previous tile's last reads of rinv were before its final barrier
        if (threadIdx.x < 8)
            rinv[threadIdx.x] = ((int)threadIdx.x < nkk) ? 1.0 / a.radius[kt0 + threadIdx.x] : 0.0;
        double g[kSlotB];
        auto fetch = [&](int c0) {
            opaque(s_src);
            opaque(s_meta);
#pragma unroll
            for (int i = 0; i < kSlotB; ++i) {
                g[i] = 0.0;
                if (s_src[i] >= 0 && ((s_meta[i] >> 8) & 7) < nkk &&
                    (s_meta[i] & 0xFF) + 2 * c0 <= a.L - m)
                    g[i] = a.in[s_src[i] + koff + 2 * ncomp * c0];
            }
        };
        auto put = [&](int c0) {
            opaque(s_src);
            opaque(s_dst);
            opaque(s_meta);
#pragma unroll
            for (int i = 0; i < kSlotB; ++i) {
                if (s_src[i] == -2) continue;
                const int meta = s_meta[i];
                const int kind = (meta >> 12) & 3;
                double v = g[i];
                if (kind != 2) {
                    // S_pol l(l+1)/r^2 (g_sph_rlm(j,3) = 1/2 at l = 0), dS/r, T/r
                    const double ri = rinv[(meta >> 8) & 7];
                    if (kind == 0) {
                        const int l = m + 2 * c0 + (meta & 0xFF);
                        v = v * (ri * ri) * ((l == 0) ? 0.5 : (double)(l * (l + 1)));
                    } else {
                        v = v * ri;
                    }
                }
                const int da = s_dst[i] & 0xFFFF, db = (int)((unsigned)s_dst[i] >> 16);
                T::store(X + da, (meta & (1 << 14)) ? -v : v, slice);
                if (db != kNone) T::store(X + db, (meta & (1 << 15)) ? -v : v, slice);
            }
        };

        acc a1[RPW][NC1][2], ax[RPW][NC2][2], ay[RPW][NC2][2];
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
#pragma unroll
            for (int ct = 0; ct < NC1; ++ct) a1[r][ct][0] = a1[r][ct][1] = acc{};
#pragma unroll
            for (int ct = 0; ct < NC2; ++ct) {
                ax[r][ct][0] = ax[r][ct][1] = acc{};
                ay[r][ct][0] = ay[r][ct][1] = acc{};
            }
        }
        auto mstep = [&](const frag (&A)[RPW][4][NS], int kl) {
            frag b1[NC1][2][NS], bx[NC2][2][NS], by[NC2][2][NS];
#pragma unroll
            for (int p = 0; p < 2; ++p) {
#pragma unroll
                for (int ct = 0; ct < NC1; ++ct)
#pragma unroll
                    for (int s = 0; s < NS; ++s)
                        b1[ct][p][s] = T::load(X + s * slice + p * po + o1[ct] + kl);
#pragma unroll
                for (int ct = 0; ct < NC2; ++ct)
#pragma unroll
                    for (int s = 0; s < NS; ++s) {
                        bx[ct][p][s] = T::load(X + s * slice + p * po + ox[ct] + kl);
                        by[ct][p][s] = T::load(X + s * slice + p * po + oy[ct] + kl);
                    }
            }
#pragma unroll
            for (int r = 0; r < RPW; ++r) {
#pragma unroll
                for (int ct = 0; ct < NC1; ++ct) {
                    T::mma(a1[r][ct][0], A[r][0], b1[ct][0]);
                    T::mma(a1[r][ct][1], A[r][1], b1[ct][1]);
                }
#pragma unroll
                for (int ct = 0; ct < NC2; ++ct) {
                    // symmetric part: dP odd, P even; antisymmetric: dP even, P odd
                    T::mma(ax[r][ct][0], A[r][3], bx[ct][1]);
                    T::mma(ax[r][ct][1], A[r][2], bx[ct][0]);
                    T::mma(ay[r][ct][0], A[r][0], by[ct][0]);
                    T::mma(ay[r][ct][1], A[r][1], by[ct][1]);
                }
            }
        };

        frag Aq[PD][RPW][4][NS];
        fetch(0);
#pragma unroll
        for (int d = 0; d < PD; ++d) load_a(Aq[d], min(d, KS - 1));
        for (int c0 = 0; c0 < KS * KS_; c0 += KC) {
            __syncthreads();
            put(c0);
            __syncthreads();
            if (c0 + KC < KS * KS_) fetch(c0 + KC);
            const int kse = anyrow ? min(KS, (c0 + KC) / KS_) : 0;
            for (int kb = c0 / KS_; kb < kse; kb += PD) {
#pragma unroll
                for (int d = 0; d < PD; ++d) {
                    const int ks = kb + d;
                    mstep(Aq[d], ks * KS_ - c0);
                    load_a(Aq[d], min(ks + PD, KS - 1));
                    __builtin_amdgcn_sched_barrier(kSchedNoVmem);
                }
            }
        }

        // epilogue: hemisphere row t -> out(t) = S + A, mirror nth-1-t -> S - A
        opaque(ocol);
#pragma unroll
        for (int r = 0; r < RPW; ++r) {
            if (!rok[r]) continue;
            const int rt = rtile[r];
#pragma unroll
            for (int ct = 0; ct < NC1 + NC2; ++ct) {
                const int oc = ocol[ct];
                if (oc < 0) continue;
                const int comp = oc & 0xFF, kk = (oc >> 8) & 7, sg = (oc >> 11) & 1;
                if (kk >= nkk) continue;
                const int mp = a.L + (sg ? -m : m), k = kt0 + kk;
                double *o = a.out + comp + (size_t)ncomp * ((size_t)a.nth * k + sl * mp);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int t = rt * 16 + T::row(lane, v);
                    if (t >= a.th_h) continue;
                    double S, An;
                    if (ct < NC1) {
                        S = T::get(a1[r][ct][0], v);
                        An = T::get(a1[r][ct][1], v);
                    } else {
                        const double f = (double)m * a.rsin[t];
                        S = T::get(ax[r][ct - NC1][0], v) + f * T::get(ay[r][ct - NC1][0], v);
                        An = T::get(ax[r][ct - NC1][1], v) + f * T::get(ay[r][ct - NC1][1], v);
                    }
                    const int tm = a.nth - 1 - t;
                    o[(size_t)ncomp * t] = S + An;
                    if (tm != t) o[(size_t)ncomp * tm] = S - An;
                }
            }
        }
    }
}

// ------------------------------------------------------------- host ----

// Gauss-Legendre points on [-1,1] by Newton iteration on [0,1] with the
// same start values, stopping rule (3e-12) and mapping as the reference
// construct_gauss_coefs / set_gauss_colatitude (m_gauss_points.f90:117-195),
// so colatitudes and weights agree with the reference grid.
void gauss_points_host(int n, double *point, double *weight, double *colat)
{
    const double pi = 4.0 * std::atan(1.0);
    if (n == 1) {
        point[0] = 0.5;
        weight[0] = 1.0;
    } else {
        const int half_n = (n + 1 - ((n + 1) % 2)) / 2;
        for (int i = 1; i <= half_n; ++i) {
            double z = std::cos(pi * ((double)i - 0.25) / ((double)n + 0.5));
            double z1, pp, p1, p2, p3;
            for (;;) {
                p1 = 1.0;
                p2 = 0.0;
                for (int jj = 1; jj <= n; ++jj) {
                    p3 = p2;
                    p2 = p1;
                    p1 = ((2.0 * (double)jj - 1.0) * z * p2 - ((double)jj - 1.0) * p3) / (double)jj;
                }
                pp = (double)n * (z * p1 - p2) / (z * z - 1.0);
                z1 = z;
                z = z1 - p1 / pp;
                if (std::fabs(z - z1) < 3.0e-12) break;
            }
            point[i - 1] = 0.5 - 0.5 * z;
            point[n - i] = 0.5 + 0.5 * z;
            weight[i - 1] = 2.0 * 0.5 / ((1.0 - z * z) * pp * pp);
            weight[n - i] = weight[i - 1];
        }
    }
    for (int i = 0; i < n; ++i) {
        point[i] = point[i] * 2.0 - 1.0;
        weight[i] = weight[i] * 2.0;
        colat[i] = std::acos(point[i]);
    }
}

template <class T>
static size_t elem_size() { return sizeof(typename T::elem); }

template <class T>
static int build_tables(LegContext *c, const double *dP_raw_P, const double *dP_raw_dP,
                        const double *dP_raw_PM, std::string *err)
{
    const LegGeometry &g = c->g;
    const int L = g.L, th_h = g.th_h;
    c->kstep = T::KSTEP;
    c->rt_bwd = (th_h + 15) / 16;
    c->ks_fwd = (th_h + T::KSTEP - 1) / T::KSTEP;

    std::vector<long long> ob(L + 2), of(2 * (L + 1) + 1);
    ob[0] = 0;
    for (int m = 0; m <= L; ++m) {
        int K = L - m + 1, K0 = (K + 1) / 2;
        int KS = (K0 + T::KSTEP - 1) / T::KSTEP;
        ob[m + 1] = ob[m] + (long long)c->rt_bwd * KS * 6;
    }
    of[0] = 0;
    for (int m = 0; m <= L; ++m) {
        int K = L - m + 1, K0 = (K + 1) / 2, K1 = K / 2;
        of[2 * m + 1] = of[2 * m] + (long long)((K0 + 15) / 16) * c->ks_fwd * 3;
        of[2 * m + 2] = of[2 * m + 1] + (long long)((K1 + 15) / 16) * c->ks_fwd * 3;
    }
    const size_t ue = (size_t)unit_elems<T>();
    const size_t bb = (size_t)ob[L + 1] * ue * sizeof(typename T::elem);
    const size_t bf = (size_t)of[2 * (L + 1)] * ue * sizeof(typename T::elem);
    HIP_TRY(hipMalloc(&c->d_off_bwd, sizeof(long long) * ob.size()));
    HIP_TRY(hipMalloc(&c->d_off_fwd, sizeof(long long) * of.size()));
    HIP_TRY(hipMemcpy(c->d_off_bwd, ob.data(), sizeof(long long) * ob.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->d_off_fwd, of.data(), sizeof(long long) * of.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&c->d_tab_bwd, bb));
    HIP_TRY(hipMalloc(&c->d_tab_fwd, bf));
    c->tab_bytes = bb + bf;

    dim3 gb(256, L + 1), gf(256, 2 * (L + 1));
    hipLaunchKernelGGL(k_pack_bwd<T>, gb, dim3(256), 0, c->stream, L, th_h, c->rt_bwd,
                       dP_raw_P, dP_raw_dP, dP_raw_PM, c->d_off_bwd,
                       reinterpret_cast<typename T::elem *>(c->d_tab_bwd));
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_pack_fwd<T>, gf, dim3(256), 0, c->stream, L, th_h, c->ks_fwd,
                       dP_raw_P, dP_raw_dP, dP_raw_PM, c->d_off_fwd,
                       reinterpret_cast<typename T::elem *>(c->d_tab_fwd));
    HIP_TRY(hipGetLastError());

    // third form: P and dP only, K steps padded with zero tiles
    constexpr int pb = kpad3b<T>(), pf = kpad3f<T>();
    c->ks_fwd3 = ((th_h + T::KSTEP - 1) / T::KSTEP + pf - 1) / pf * pf;
    std::vector<long long> ob3(L + 2), of3(2 * (L + 1) + 1);
    ob3[0] = 0;
    for (int m = 0; m <= L; ++m) {
        const int K0 = (L - m + 2) / 2;
        const int KS = ((K0 + T::KSTEP - 1) / T::KSTEP + pb - 1) / pb * pb;
        ob3[m + 1] = ob3[m] + (long long)c->rt_bwd * KS * 4;
    }
    of3[0] = 0;
    for (int m = 0; m <= L; ++m) {
        const int K = L - m + 1, K0 = (K + 1) / 2, K1 = K / 2;
        of3[2 * m + 1] = of3[2 * m] + (long long)((K0 + 15) / 16) * c->ks_fwd3 * 2;
        of3[2 * m + 2] = of3[2 * m + 1] + (long long)((K1 + 15) / 16) * c->ks_fwd3 * 2;
    }
    const size_t bb3 = (size_t)ob3[L + 1] * ue * sizeof(typename T::elem);
    const size_t bf3 = (size_t)of3[2 * (L + 1)] * ue * sizeof(typename T::elem);
    HIP_TRY(hipMalloc(&c->d_off_bwd3, sizeof(long long) * ob3.size()));
    HIP_TRY(hipMalloc(&c->d_off_fwd3, sizeof(long long) * of3.size()));
    HIP_TRY(hipMemcpy(c->d_off_bwd3, ob3.data(), sizeof(long long) * ob3.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->d_off_fwd3, of3.data(), sizeof(long long) * of3.size(), hipMemcpyHostToDevice));
    HIP_TRY(hipMalloc(&c->d_tab_bwd3, bb3));
    HIP_TRY(hipMalloc(&c->d_tab_fwd3, bf3));
    c->tab_bytes += bb3 + bf3;
    hipLaunchKernelGGL(k_pack_bwd3<T>, gb, dim3(256), 0, c->stream, L, th_h, c->rt_bwd,
                       dP_raw_P, dP_raw_dP, c->d_off_bwd3,
                       reinterpret_cast<typename T::elem *>(c->d_tab_bwd3));
    HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(k_pack_fwd3<T>, gf, dim3(256), 0, c->stream, L, th_h, c->ks_fwd3,
                       dP_raw_P, dP_raw_dP, c->d_off_fwd3,
                       reinterpret_cast<typename T::elem *>(c->d_tab_fwd3));
    HIP_TRY(hipGetLastError());
    return 0;
}

// Candidate tile plans for one field set and direction.  Kernel instances
// (column tiles of the P GEMM, of the dP|PM GEMM): (1,1), (1,2), (2,4); the
// radial tile KT is the largest that fits the instance; rows per wave 2, 3 or
// 5; the forward kernel optionally carries orders m and L-m together.
// rows3: padded colatitude rows of the third-form forward tables
static size_t plan_lds(const LegPlan &p, bool fwd, int prec, int rows3 = 0)
{
    const int ns = (prec == kBF16x3) ? 3 : 1;
    const size_t es = (prec == kF64) ? 8 : (prec == kF32 ? 4 : 2);
    const int pad = (prec == kBF16x3) ? 8 : (prec == kF64 ? 2 : 4);   // lds_pad<T>()
    if (p.ver == 3) {
        const int nb = 16 * (p.mct1 + 2 * p.mct2) + 1;
        const int pad3 = (prec == kF64) ? 1 : pad;   // lds_pad3<T>()
        return (((size_t)ns * 2 * nb * (p.tc + pad3) * es + 15) & ~(size_t)15)
               + (fwd ? (size_t)16 * rows3 : 0);   // + row factors (fwd)
    }
    const int ncol = p.KT * 2 * p.ncomp * ((fwd && p.pair) ? 2 : 1);
    const int kc = fwd ? p.tc : 32;
    return (size_t)ns * 2 * ncol * (kc + pad) * es;
}

// Third-form candidates: column-tile instances (GEMM1 tiles, GEMM2 tiles)
// with the largest radial tile KT each holds (an instance is offered only if
// no smaller one holds the same KT); rows per wave; prefetch depth; LDS chunk
// (colatitudes fwd, degrees per parity bwd) limited by the staging slots and
// an LDS budget that leaves room for two work-groups per CU.
static const int kInst3[4][2] = {{1, 1}, {1, 2}, {2, 3}, {2, 4}};

static void plan_candidates3(int ncomp, int nv, int ns, bool fwd, int prec,
                             std::vector<LegPlan> *out)
{
    const int per1 = 2 * (nv + ns), per2 = 4 * nv;
    const int kstep = (prec == kBF16x3) ? 32 : 4;
    int kt_prev = 0;
    for (int i = 0; i < 4; ++i) {
        int kt = 0;
        for (int k = 8; k >= 1; --k)
            if ((per1 * k + 15) / 16 <= kInst3[i][0] && (per2 * k + 15) / 16 <= kInst3[i][1]) {
                kt = k;
                break;
            }
        if (!kt || kt <= kt_prev) continue;
        kt_prev = kt;
        // three row tiles per wave in the backward kernel only with one
        // theta/phi column tile (registers)
        const std::vector<int> rpws = fwd ? std::vector<int>{1, 2}
                                    : kInst3[i][1] == 1 ? std::vector<int>{1, 3}
                                                        : std::vector<int>{1};
        std::vector<int> pds;
        if (prec == kBF16x3) pds = {1};
        else if (fwd) pds = {2, 4};
        else pds = {2};
        const std::vector<int> tcs = (prec == kBF16x3) ? std::vector<int>{64, 32}
                                     : fwd ? std::vector<int>{64, 32, 16}
                                           : std::vector<int>{64, 32, 16, 8};
        // waves per work-group: 4, or (one row tile per wave, not BF16x3) 8
        // and, backward, 12: more row tiles (four degree tiles of a parity,
        // or eight / twelve colatitude tiles) share one staged input, and the
        // staging slots of the extra waves allow longer chunks (fewer
        // barriers per MFMA).  A 4-wave plan keeps its LDS to 80 KB (two
        // work-groups per CU); the larger ones run one work-group per CU.
        const std::vector<int> nws_all = fwd ? std::vector<int>{4, 8} : std::vector<int>{4, 8, 12};
        for (int rpw : rpws)
            for (int pd : pds)
                for (int nw : nws_all) {
                    if (nw != 4 && (rpw != 1 || prec == kBF16x3)) continue;
                    // twelve waves (three per SIMD, 168 registers each) only
                    // with one column tile per GEMM: more accumulators spill
                    if (nw == 12 && (kInst3[i][0] != 1 || kInst3[i][1] != 1)) continue;
                    int taken = 0;
                    for (int tc : tcs) {
                        if (taken == 2) break;
                        if (tc % (pd * kstep)) continue;
                        const long long nst = fwd ? (long long)tc * kt * 2 * ncomp
                                                  : (long long)2 * tc * ncomp * kt * 2;
                        if (nst > (long long)(fwd ? kSlotF : kSlotB) * 64 * nw) continue;
                        LegPlan p;
                        p.ncomp = ncomp;
                        p.nvec = nv;
                        p.nscl = ns;
                        p.fwd = fwd;
                        p.KT = kt;
                        p.mct1 = kInst3[i][0];
                        p.mct2 = kInst3[i][1];
                        p.rpw = rpw;
                        p.pd = pd;
                        p.tc = tc;
                        p.ver = 3;
                        p.nw = nw;
                        if (plan_lds(p, fwd, prec) > (nw == 4 ? 80 : 150) * 1024) continue;
                        out->push_back(p);
                        ++taken;
                    }
                }
    }
}

static std::vector<LegPlan> plan_candidates(int ncomp, int nv, int ns, bool fwd, int prec)
{
    const int per1 = 2 * (nv + ns), per2 = 4 * nv;
    static const int inst[3][2] = {{1, 2}, {2, 4}, {1, 1}};
    std::vector<LegPlan> out;
    if (prec == kInt8x6) {
        oz_plan_candidates(ncomp, nv, ns, fwd, &out);
        return out;
    }
    plan_candidates3(ncomp, nv, ns, fwd, prec, &out);
    for (int i = 0; i < 3; ++i) {
        int kt = 0;
        for (int k = 8; k >= 1; --k) {
            if ((per1 * k + 15) / 16 <= inst[i][0] && (per2 * k + 15) / 16 <= inst[i][1]) {
                kt = k;
                break;
            }
        }
        if (!kt) continue;
        std::vector<int> rpws;
        if (!fwd) rpws = {1, inst[i][0] == 2 ? 2 : 3};
        else if (prec == kBF16x3 || inst[i][0] == 2) rpws = {1, 3};
        else rpws = {1, 3, 5};
        for (int rpw : rpws)
            for (int pair = 0; pair < (fwd ? 2 : 1); ++pair) {
                LegPlan p;
                p.ncomp = ncomp;
                p.nvec = nv;
                p.nscl = ns;
                p.fwd = fwd;
                p.KT = kt;
                p.mct1 = inst[i][0];
                p.mct2 = inst[i][1];
                p.rpw = rpw;
                p.pair = pair;
                const int kmin = (prec == kBF16x3) ? 32 : 4;
                p.tc = 64;
                while (fwd && p.tc > kmin && plan_lds(p, fwd, prec) > 160 * 1024) p.tc /= 2;
                if (plan_lds(p, fwd, prec) > 160 * 1024) continue;
                out.push_back(p);
            }
    }
    if (prec != kBF16x3) {
        // deeper P prefetch for the small-tile plans (chunk steps must be a
        // multiple of the depth)
        const size_t n1 = out.size();
        for (size_t i = 0; i < n1; ++i) {
            LegPlan p = out[i];
            if (p.ver != 1 || p.mct1 != 1 || (p.rpw != 1 && p.rpw != 3)) continue;
            const int steps = fwd ? p.tc / 4 : 32 / 4;
            if (steps % 4) continue;
            p.pd = 4;
            out.push_back(p);
        }
    }
    return out;
}

// Order the work items for the 8 XCDs: work-group b runs on XCD b % 8, so
// the list is interleaved such that every item of one order pair (m, L-m)
// lands on the same XCD (its P tiles are then shared through that XCD's L2)
// while the pairs, which carry equal work, are dealt round-robin over the
// XCDs.  Shorter XCD lists are padded with empty items (m = -1).
static void xcd_interleave(int L, std::vector<WorkItem> *items)
{
    constexpr int NX = 8;
    std::vector<WorkItem> per[NX];
    for (const WorkItem &w : *items) {
        const int pr = std::min(w.m, L - w.m);
        per[pr % NX].push_back(w);
    }
    size_t len = 0;
    for (int x = 0; x < NX; ++x) len = std::max(len, per[x].size());
    std::vector<WorkItem> out(len * NX, WorkItem{-1, 0, 0, 0});
    for (int x = 0; x < NX; ++x)
        for (size_t i = 0; i < per[x].size(); ++i) out[i * NX + x] = per[x][i];
    items->swap(out);
}

// device memory a plan owns: its work items
static void plan_release(LegPlan *p)
{
    (void)hipFree(p->d_items);
    p->d_items = nullptr;
}

static int build_items(LegContext *c, LegPlan *p, bool fwd, std::string *err)
{
    if (p->ver == 5) return oz_build_items(c, p, fwd, err);
    const LegGeometry &g = c->g;
    const int nkt = (g.nr + p->KT - 1) / p->KT;
    std::vector<WorkItem> items;
    const int rows_per_group = 4 * p->rpw;
    auto rows = [&](int mm) {
        int K = g.L - mm + 1, K0 = (K + 1) / 2, K1 = K / 2;
        return (K0 + 15) / 16 + (K1 + 15) / 16;
    };
    if (p->ver == 3) {
        if (p->KT > 8) { *err = "radial tile above 8"; return -1; }
        // the third form addresses the rtm / rlm arrays with 32-bit offsets
        if ((double)p->ncomp * g.nth * g.nr * g.M >= 2147483647.0 ||
            (double)p->ncomp * g.J * g.nr >= 2147483647.0) {
            p->n_items = 0;
            p->d_items = nullptr;
            return 1;
        }
        auto ngroup = [&](int m) {
            const int K = g.L - m + 1, RTe = ((K + 1) / 2 + 15) / 16, RTo = (K / 2 + 15) / 16;
            const int per = (p->nw / 2) * p->rpw;   // degree tiles of a parity per group
            return fwd ? (std::max(RTe, RTo) + per - 1) / per
                       : (c->rt_bwd + p->nw * p->rpw - 1) / (p->nw * p->rpw);
        };
        // radial tiles per work-group: as many as keep >= 2048 work-groups
        long long n1 = 0;
        for (int m = 0; m <= g.L; ++m) n1 += (long long)nkt * ngroup(m);
        const int ntile = (int)std::max(1LL, std::min(8LL, n1 / 2048));
        for (int m = 0; m <= g.L; ++m) {
            // a work-group runs ntile consecutive radial tiles; the row groups
            // of one radial tile are adjacent (they stage the same input and
            // share it through L2)
            for (int kt = 0; kt < nkt; kt += ntile)
                for (int rg = 0; rg < ngroup(m); ++rg)
                    items.push_back({m, kt, rg, std::min(ntile, nkt - kt)});
        }
        xcd_interleave(g.L, &items);
        p->n_items = (int)items.size();
        HIP_TRY(hipMalloc(&p->d_items, sizeof(WorkItem) * items.size()));
        HIP_TRY(hipMemcpy(p->d_items, items.data(), sizeof(WorkItem) * items.size(),
                          hipMemcpyHostToDevice));
        return 0;
    }
    for (int m = 0; m <= g.L; ++m) {
        int nrow;
        if (fwd) {
            if (p->pair) {
                // orders m and L-m share a work item
                if (g.L - m < m) break;
                nrow = rows(m) + (g.L - m != m ? rows(g.L - m) : 0);
            } else {
                nrow = rows(m);
            }
        } else {
            nrow = c->rt_bwd;
        }
        int ngroup = (nrow + rows_per_group - 1) / rows_per_group;
        for (int rg = 0; rg < ngroup; ++rg)
            for (int kt = 0; kt < nkt; ++kt) items.push_back({m, kt, rg, 0});
    }
    xcd_interleave(g.L, &items);
    p->n_items = (int)items.size();
    HIP_TRY(hipMalloc(&p->d_items, sizeof(WorkItem) * items.size()));
    HIP_TRY(hipMemcpy(p->d_items, items.data(), sizeof(WorkItem) * items.size(),
                      hipMemcpyHostToDevice));
    return 0;
}

static int autotune_plan(LegContext *c, std::vector<LegPlan> &cand, bool fwd, std::string *err);

// autotuner picks of this process: (L, nth, nr, ncomp, nvec, nscl, fwd,
// precision) -> index into the candidate list (which depends on these only)
// and its timing
using PlanKey = std::array<int, 8>;
static std::map<PlanKey, std::pair<int, float>> &plan_cache()
{
    static std::map<PlanKey, std::pair<int, float>> m;
    return m;
}
static std::mutex &plan_cache_mutex()
{
    static std::mutex mu;
    return mu;
}

const LegPlan *leg_plan(LegContext *c, int ncomp, int nvec, int nscl, bool fwd,
                        std::string *err)
{
    if (ncomp != 3 * nvec + nscl || nvec < 0 || nscl < 0 || ncomp < 1) {
        *err = "ncomp must equal 3*nvector + nscalar";
        return nullptr;
    }
    for (const LegPlan &p : c->plans)
        if (p.ncomp == ncomp && p.nvec == nvec && p.nscl == nscl && p.fwd == fwd) return &p;
    std::vector<LegPlan> cand = plan_candidates(ncomp, nvec, nscl, fwd, c->prec);
    if (cand.empty()) {
        *err = "no Legendre tile plan fits this component count in LDS";
        return nullptr;
    }
    {
        std::vector<LegPlan> ok;
        for (LegPlan &p : cand) {
            const int rc = build_items(c, &p, fwd, err);
            if (rc < 0) return nullptr;
            if (rc == 0) ok.push_back(p);
        }
        cand.swap(ok);
    }
    if (cand.empty()) {
        *err = "no Legendre tile plan fits this grid";
        return nullptr;
    }
    int best = 0;
    if (c->autotune && cand.size() > 1) {
        // one timing per process and shape: later contexts of the same grid,
        // field set, direction and precision take the same plan, so they
        // compute bitwise the same (timing noise cannot pick another plan)
        const PlanKey key{c->g.L, c->g.nth, c->g.nr, ncomp, nvec, nscl, fwd ? 1 : 0, c->prec};
        std::lock_guard<std::mutex> lock(plan_cache_mutex());
        auto &cache = plan_cache();
        auto hit = cache.find(key);
        if (hit != cache.end() && hit->second.first < (int)cand.size()) {
            best = hit->second.first;
            cand[best].tuned_ms = hit->second.second;
        } else {
            best = autotune_plan(c, cand, fwd, err);
            if (best < 0) return nullptr;
            cache[key] = {best, cand[best].tuned_ms};
        }
    }
    for (size_t i = 0; i < cand.size(); ++i)
        if ((int)i != best) plan_release(&cand[i]);
    c->plans.push_back(cand[best]);
    return &c->plans.back();
}

// Candidate plans of a field set, in the autotuner's order, after the items
// are built (plans that cannot cover the grid are dropped).
static int valid_candidates(LegContext *c, int ncomp, int nvec, int nscl, bool fwd,
                            std::vector<LegPlan> *out, std::string *err)
{
    std::vector<LegPlan> cand = plan_candidates(ncomp, nvec, nscl, fwd, c->prec);
    for (LegPlan &p : cand) {
        const int rc = build_items(c, &p, fwd, err);
        if (rc < 0) return -1;
        if (rc == 0) out->push_back(p);
    }
    return 0;
}

int leg_plan_count(LegContext *c, int ncomp, int nvec, int nscl, bool fwd, std::string *err)
{
    if (ncomp != 3 * nvec + nscl || nvec < 0 || nscl < 0 || ncomp < 1) {
        *err = "ncomp must equal 3*nvector + nscalar";
        return -1;
    }
    std::vector<LegPlan> v;
    if (valid_candidates(c, ncomp, nvec, nscl, fwd, &v, err)) return -1;
    for (LegPlan &p : v) plan_release(&p);
    return (int)v.size();
}

int leg_force_plan(LegContext *c, int ncomp, int nvec, int nscl, bool fwd, int idx,
                   std::string *err)
{
    if (ncomp != 3 * nvec + nscl || nvec < 0 || nscl < 0 || ncomp < 1) {
        *err = "ncomp must equal 3*nvector + nscalar";
        return -1;
    }
    std::vector<LegPlan> v;
    if (valid_candidates(c, ncomp, nvec, nscl, fwd, &v, err)) return -1;
    if (idx < 0 || idx >= (int)v.size()) {
        for (LegPlan &p : v) plan_release(&p);
        *err = "plan index out of range";
        return -1;
    }
    for (int i = 0; i < (int)v.size(); ++i)
        if (i != idx) plan_release(&v[i]);
    for (LegPlan &p : c->plans)
        if (p.ncomp == ncomp && p.nvec == nvec && p.nscl == nscl && p.fwd == fwd) {
            plan_release(&p);
            p = v[idx];
            return 0;
        }
    c->plans.push_back(v[idx]);
    return 0;
}

int leg_context_build(LegContext *c, int L, int nth, int nr, const double *radius,
                      const double *colat, const double *weight, int ncomp, int nvec,
                      int nscl, int prec, std::string *err)
{
    if (L < 1 || nth < 2 || nr < 1) { *err = "invalid grid size"; return -1; }
    if (ncomp != 3 * nvec + nscl || nvec < 0 || nscl < 0 || ncomp < 1) {
        *err = "ncomp must equal 3*nvector + nscalar";
        return -1;
    }
    if (prec != kF64 && prec != kF32 && prec != kBF16x3 && prec != kInt8x6) {
        *err = "unknown precision";
        return -1;
    }
    c->g.L = L;
    c->g.nth = nth;
    c->g.nr = nr;
    c->g.th_h = (nth + 1) / 2;
    c->g.J = (L + 1) * (L + 1);
    c->g.M = 2 * L + 1;
    c->ncomp = ncomp;
    c->nvec = nvec;
    c->nscl = nscl;
    c->prec = prec;

    std::vector<double> pt(nth), wt(nth), ct(nth);
    gauss_points_host(nth, pt.data(), wt.data(), ct.data());
    if (colat) std::memcpy(ct.data(), colat, sizeof(double) * nth);
    if (weight) std::memcpy(wt.data(), weight, sizeof(double) * nth);
    // the folded kernels need a mirror-symmetric grid
    for (int t = 0; t < nth / 2; ++t) {
        if (std::fabs(wt[t] - wt[nth - 1 - t]) > 1e-14 * std::fabs(wt[t]) ||
            std::fabs(ct[t] + ct[nth - 1 - t] - 3.14159265358979323846) > 1e-12) {
            *err = "colatitude grid is not equatorially symmetric";
            return -1;
        }
    }
    std::vector<int> lstack(2 * L + 2);
    lstack[0] = 0;
    for (int mp = 0; mp < 2 * L + 1; ++mp) lstack[mp + 1] = lstack[mp] + (L - std::abs(mp - L) + 1);

    HIP_TRY(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIP_TRY(hipEventCreate(&c->ev0));
    HIP_TRY(hipEventCreate(&c->ev1));
    HIP_TRY(hipMalloc(&c->d_radius, sizeof(double) * nr));
    HIP_TRY(hipMalloc(&c->d_colat, sizeof(double) * nth));
    HIP_TRY(hipMalloc(&c->d_weight, sizeof(double) * nth));
    HIP_TRY(hipMalloc(&c->d_lstack, sizeof(int) * lstack.size()));
    HIP_TRY(hipMemcpy(c->d_radius, radius, sizeof(double) * nr, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->d_colat, ct.data(), sizeof(double) * nth, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->d_weight, wt.data(), sizeof(double) * nth, hipMemcpyHostToDevice));
    HIP_TRY(hipMemcpy(c->d_lstack, lstack.data(), sizeof(int) * lstack.size(), hipMemcpyHostToDevice));

    // FP64 recursion on the device, then pack into MFMA operand order
    const size_t nraw = (size_t)(L + 1) * (L + 2) / 2 * c->g.th_h;
    double *rP = nullptr, *rdP = nullptr, *rPM = nullptr;
    HIP_TRY(hipMalloc(&rP, sizeof(double) * nraw));
    HIP_TRY(hipMalloc(&rdP, sizeof(double) * nraw));
    HIP_TRY(hipMalloc(&rPM, sizeof(double) * nraw));
    dim3 grid((c->g.th_h + 127) / 128, L + 1);
    hipLaunchKernelGGL(k_schmidt_raw, grid, dim3(128), 0, c->stream, L, c->g.th_h, c->d_colat,
                       rP, rdP, rPM);
    HIP_TRY(hipGetLastError());
    {
        // g7 = (2l+1)/(4l(l+1)) and g6 = (2l+1)/4 of order m > 0
        // (spherical_harmonics.f90:121-139); order 0 doubles both exactly
        std::vector<double> gf(2 * (L + 1));
        for (int l = 0; l <= L; ++l) {
            gf[l] = (l == 0) ? 1.0 : (double)(2 * l + 1) / (double)(4 * l * (l + 1));
            gf[L + 1 + l] = (double)(2 * l + 1) / 4.0;
        }
        HIP_TRY(hipMalloc(&c->d_gfac, sizeof(double) * gf.size()));
        HIP_TRY(hipMemcpy(c->d_gfac, gf.data(), sizeof(double) * gf.size(), hipMemcpyHostToDevice));
    }
    HIP_TRY(hipMalloc(&c->d_rsin, sizeof(double) * c->g.th_h));
    hipLaunchKernelGGL(k_rsin, dim3((c->g.th_h + 127) / 128), dim3(128), 0, c->stream, c->g.th_h,
                       c->d_colat, c->d_rsin);
    HIP_TRY(hipGetLastError());
    int rc;
    if (prec == kF64) rc = build_tables<TF64>(c, rP, rdP, rPM, err);
    else if (prec == kF32) rc = build_tables<TF32>(c, rP, rdP, rPM, err);
    else if (prec == kInt8x6) rc = oz_build_tables(c, rP, rdP, rPM, err);
    else rc = build_tables<TBF3>(c, rP, rdP, rPM, err);
    HIP_TRY(hipStreamSynchronize(c->stream));
    (void)hipFree(rP);
    (void)hipFree(rdP);
    (void)hipFree(rPM);
    if (rc) return rc;

    if (!leg_plan(c, ncomp, nvec, nscl, false, err)) return -1;
    if (!leg_plan(c, ncomp, nvec, nscl, true, err)) return -1;
    return 0;
}

void leg_context_free(LegContext *c)
{
    if (!c) return;
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    (void)hipFree(c->d_radius);
    (void)hipFree(c->d_colat);
    (void)hipFree(c->d_weight);
    (void)hipFree(c->d_lstack);
    (void)hipFree(c->d_tab_bwd);
    (void)hipFree(c->d_tab_fwd);
    (void)hipFree(c->d_off_bwd);
    (void)hipFree(c->d_off_fwd);
    (void)hipFree(c->d_tab_bwd3);
    (void)hipFree(c->d_tab_fwd3);
    (void)hipFree(c->d_off_bwd3);
    (void)hipFree(c->d_off_fwd3);
    (void)hipFree(c->d_rsin);
    (void)hipFree(c->d_gfac);
    oz_free(c);
    for (LegPlan &p : c->plans) plan_release(&p);
    c->plans.clear();
    (void)hipFree(c->d_rlm);
    (void)hipFree(c->d_rtm);
    if (c->ev0) (void)hipEventDestroy(c->ev0);
    if (c->ev1) (void)hipEventDestroy(c->ev1);
    if (c->stream) (void)hipStreamDestroy(c->stream);
}

static LegArgs make_args(LegContext *c, const LegPlan &p, bool fwd, const double *in, double *out)
{
    LegArgs a;
    a.in = in;
    a.out = out;
    a.tab = fwd ? c->d_tab_fwd : c->d_tab_bwd;
    a.off = fwd ? c->d_off_fwd : c->d_off_bwd;
    a.lstack = c->d_lstack;
    a.radius = c->d_radius;
    a.weight = c->d_weight;
    a.items = p.d_items;
    a.L = c->g.L;
    a.nth = c->g.nth;
    a.nr = c->g.nr;
    a.th_h = c->g.th_h;
    a.J = c->g.J;
    a.ncomp = p.ncomp;
    a.nvec = p.nvec;
    a.nscl = p.nscl;
    a.KT = p.KT;
    a.RT = c->rt_bwd;
    a.KSt = c->ks_fwd;
    a.TC = p.tc;
    a.pair = p.pair;
    a.rsin = c->d_rsin;
    a.gfac = c->d_gfac;
    if (p.ver == 3) {
        a.tab = fwd ? c->d_tab_fwd3 : c->d_tab_bwd3;
        a.off = fwd ? c->d_off_fwd3 : c->d_off_bwd3;
        a.KSt = c->ks_fwd3;
    }
    return a;
}

template <class T>
static size_t lds_bytes(const LegContext *c, const LegPlan &p, bool fwd)
{
    return plan_lds(p, fwd, c->prec);
}

template <class T, int RPW, int MCT1, int MCT2, int PD>
static int launch_t(LegContext *c, const LegPlan &p, bool fwd, const double *in, double *out,
                    hipStream_t s, std::string *err)
{
    LegArgs a = make_args(c, p, fwd, in, out);
    size_t lds = lds_bytes<T>(c, p, fwd);
    if (lds > 160 * 1024) { *err = "LDS tile too large for this component count"; return -1; }
    if (lds > 64 * 1024) {
        const void *fn = fwd ? reinterpret_cast<const void *>(&k_leg_fwd<T, RPW, MCT1, MCT2, PD>)
                             : reinterpret_cast<const void *>(&k_leg_bwd<T, RPW, MCT1, MCT2, PD>);
        HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    }
    if (fwd)
        hipLaunchKernelGGL((k_leg_fwd<T, RPW, MCT1, MCT2, PD>), dim3(p.n_items), dim3(kThreads), lds, s, a);
    else
        hipLaunchKernelGGL((k_leg_bwd<T, RPW, MCT1, MCT2, PD>), dim3(p.n_items), dim3(kThreads), lds, s, a);
    HIP_TRY(hipGetLastError());
    return 0;
}

template <class T, int RPW, int NC1, int NC2, int PD, int NW = 4>
static int launch3_f(LegContext *c, const LegPlan &p, const double *in, double *out,
                     hipStream_t s, std::string *err)
{
    LegArgs a = make_args(c, p, true, in, out);
    const size_t lds = plan_lds(p, true, c->prec, c->ks_fwd3 * T::KSTEP);
    if (lds > 160 * 1024) { *err = "LDS tile too large for this grid"; return -1; }
    if (lds > 64 * 1024)
        HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_leg_fwd3<T, RPW, NC1, NC2, PD, NW>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((k_leg_fwd3<T, RPW, NC1, NC2, PD, NW>), dim3(p.n_items), dim3(NW * 64), lds, s, a);
    HIP_TRY(hipGetLastError());
    return 0;
}

template <class T, int RPW, int NC1, int NC2, int PD, int NW = 4>
static int launch3_b(LegContext *c, const LegPlan &p, const double *in, double *out,
                     hipStream_t s, std::string *err)
{
    LegArgs a = make_args(c, p, false, in, out);
    const size_t lds = plan_lds(p, false, c->prec);
    if (lds > 64 * 1024)
        HIP_TRY(hipFuncSetAttribute(reinterpret_cast<const void *>(&k_leg_bwd3<T, RPW, NC1, NC2, PD, NW>),
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL((k_leg_bwd3<T, RPW, NC1, NC2, PD, NW>), dim3(p.n_items), dim3(NW * 64), lds, s, a);
    HIP_TRY(hipGetLastError());
    return 0;
}

template <class T, int NC1, int NC2>
static int launch3_n(LegContext *c, const LegPlan &p, bool fwd, const double *in, double *out,
                     hipStream_t s, std::string *err)
{
    if constexpr (T::NS == 1) {
        if (fwd && p.nw == 8) {
            if (p.rpw == 1 && p.pd == 2) return launch3_f<T, 1, NC1, NC2, 2, 8>(c, p, in, out, s, err);
            if (p.rpw == 1 && p.pd == 4) return launch3_f<T, 1, NC1, NC2, 4, 8>(c, p, in, out, s, err);
        } else if (fwd) {
            if (p.rpw == 1 && p.pd == 2) return launch3_f<T, 1, NC1, NC2, 2>(c, p, in, out, s, err);
            if (p.rpw == 1 && p.pd == 4) return launch3_f<T, 1, NC1, NC2, 4>(c, p, in, out, s, err);
            if (p.rpw == 2 && p.pd == 2) return launch3_f<T, 2, NC1, NC2, 2>(c, p, in, out, s, err);
            if (p.rpw == 2 && p.pd == 4) return launch3_f<T, 2, NC1, NC2, 4>(c, p, in, out, s, err);
        } else if (p.nw == 8) {
            if (p.rpw == 1 && p.pd == 2) return launch3_b<T, 1, NC1, NC2, 2, 8>(c, p, in, out, s, err);
        } else if (p.nw == 12) {
            if constexpr (NC1 == 1 && NC2 == 1)
                if (p.rpw == 1 && p.pd == 2) return launch3_b<T, 1, NC1, NC2, 2, 12>(c, p, in, out, s, err);
        } else {
            if (p.rpw == 1 && p.pd == 2) return launch3_b<T, 1, NC1, NC2, 2>(c, p, in, out, s, err);
            if constexpr (NC2 == 1)
                if (p.rpw == 3 && p.pd == 2) return launch3_b<T, 3, NC1, NC2, 2>(c, p, in, out, s, err);
        }
    } else {
        if (p.pd == 1) {
            if (fwd && p.rpw == 1) return launch3_f<T, 1, NC1, NC2, 1>(c, p, in, out, s, err);
            if (fwd && p.rpw == 2) return launch3_f<T, 2, NC1, NC2, 1>(c, p, in, out, s, err);
            if (!fwd && p.rpw == 1) return launch3_b<T, 1, NC1, NC2, 1>(c, p, in, out, s, err);
            if constexpr (NC2 == 1)
                if (!fwd && p.rpw == 3) return launch3_b<T, 3, NC1, NC2, 1>(c, p, in, out, s, err);
        }
    }
    *err = "no third-form kernel instance for this plan";
    return -1;
}

template <class T>
static int launch3(LegContext *c, const LegPlan &p, bool fwd, const double *in, double *out,
                   hipStream_t s, std::string *err)
{
    if (p.KT > 8 || p.KT < 1) { *err = "radial tile outside 1..8"; return -1; }
    if (p.mct1 == 1 && p.mct2 == 1) return launch3_n<T, 1, 1>(c, p, fwd, in, out, s, err);
    if (p.mct1 == 1 && p.mct2 == 2) return launch3_n<T, 1, 2>(c, p, fwd, in, out, s, err);
    if (p.mct1 == 2 && p.mct2 == 3) return launch3_n<T, 2, 3>(c, p, fwd, in, out, s, err);
    if (p.mct1 == 2 && p.mct2 == 4) return launch3_n<T, 2, 4>(c, p, fwd, in, out, s, err);
    *err = "no third-form kernel instance for this column tiling";
    return -1;
}

template <class T>
static int launch_p(LegContext *c, const LegPlan &p, bool fwd, const double *in, double *out,
                    hipStream_t s, std::string *err)
{
    if (p.ver == 3) return launch3<T>(c, p, fwd, in, out, s, err);
    if constexpr (T::NS == 1) {
        if (p.pd == 4) {
            if (p.mct1 == 1 && p.mct2 == 1) {
                if (p.rpw == 1) return launch_t<T, 1, 1, 1, 4>(c, p, fwd, in, out, s, err);
                if (p.rpw == 3) return launch_t<T, 3, 1, 1, 4>(c, p, fwd, in, out, s, err);
            }
            if (p.mct1 == 1 && p.mct2 == 2) {
                if (p.rpw == 1) return launch_t<T, 1, 1, 2, 4>(c, p, fwd, in, out, s, err);
                if (p.rpw == 3) return launch_t<T, 3, 1, 2, 4>(c, p, fwd, in, out, s, err);
            }
            *err = "no kernel instance with this prefetch depth";
            return -1;
        }
    }
    if (p.mct1 == 1 && p.mct2 == 1) {
        if (p.rpw == 5) return launch_t<T, 5, 1, 1, 2>(c, p, fwd, in, out, s, err);
        if (p.rpw == 1) return launch_t<T, 1, 1, 1, 2>(c, p, fwd, in, out, s, err);
        return launch_t<T, 3, 1, 1, 2>(c, p, fwd, in, out, s, err);
    }
    if (p.mct1 == 1 && p.mct2 == 2) {
        if (p.rpw == 5) return launch_t<T, 5, 1, 2, 2>(c, p, fwd, in, out, s, err);
        if (p.rpw == 1) return launch_t<T, 1, 1, 2, 2>(c, p, fwd, in, out, s, err);
        return launch_t<T, 3, 1, 2, 2>(c, p, fwd, in, out, s, err);
    }
    if (p.rpw == 1) return launch_t<T, 1, 2, 4, 2>(c, p, fwd, in, out, s, err);
    if (p.rpw == 2) return launch_t<T, 2, 2, 4, 2>(c, p, fwd, in, out, s, err);
    return launch_t<T, 3, 2, 4, 2>(c, p, fwd, in, out, s, err);
}

static int launch_plan(LegContext *c, const LegPlan &p, bool fwd, const double *in, double *out,
                       hipStream_t s, std::string *err)
{
    if (c->prec == kInt8x6) return oz_launch(c, p, fwd, in, out, s, err);
    if (c->prec == kF64) return launch_p<TF64>(c, p, fwd, in, out, s, err);
    if (c->prec == kF32) return launch_p<TF32>(c, p, fwd, in, out, s, err);
    return launch_p<TBF3>(c, p, fwd, in, out, s, err);
}

static int launch(LegContext *c, int ncomp, int nvec, int nscl, bool fwd, const double *in,
                  double *out, hipStream_t s, std::string *err)
{
    const LegPlan *p = leg_plan(c, ncomp, nvec, nscl, fwd, err);
    if (!p) return -1;
    return launch_plan(c, *p, fwd, in, out, s, err);
}

// Time every candidate plan (3 launches each after one warm-up, on the
// context stream) and return the fastest: the GPU counterpart of the
// reference's init-time selection of the Legendre loop variant
// (select_legendre_transform, sph_transforms_4_MHD.f90:309-426).
static int autotune_plan(LegContext *c, std::vector<LegPlan> &cand, bool fwd, std::string *err)
{
    const LegPlan &p0 = cand[0];
    const size_t nrlm = (size_t)p0.ncomp * c->g.J * c->g.nr;
    const size_t nrtm = (size_t)p0.ncomp * c->g.nth * c->g.nr * c->g.M;
    DevBuf in_buf, out_buf;   // released on every exit path
    HIP_TRY(hipMalloc(&in_buf.p, sizeof(double) * (fwd ? nrtm : nrlm)));
    HIP_TRY(hipMalloc(&out_buf.p, sizeof(double) * (fwd ? nrlm : nrtm)));
    double *din = static_cast<double *>(in_buf.p), *dout = static_cast<double *>(out_buf.p);
    HIP_TRY(hipMemsetAsync(din, 0, sizeof(double) * (fwd ? nrtm : nrlm), c->stream));
    int best = -1;
    float best_ms = 0.f;
    for (size_t i = 0; i < cand.size(); ++i) {
        std::string e2;
        if (launch_plan(c, cand[i], fwd, din, dout, c->stream, &e2)) continue;
        HIP_TRY(hipEventRecord(c->ev0, c->stream));
        bool failed = false;
        for (int r = 0; r < 3 && !failed; ++r)
            failed = launch_plan(c, cand[i], fwd, din, dout, c->stream, &e2) != 0;
        HIP_TRY(hipEventRecord(c->ev1, c->stream));
        HIP_TRY(hipEventSynchronize(c->ev1));
        if (failed) continue;   // a partial timing never wins
        float ms = 0.f;
        HIP_TRY(hipEventElapsedTime(&ms, c->ev0, c->ev1));
        cand[i].tuned_ms = ms / 3.f;
        if (best < 0 || cand[i].tuned_ms < best_ms) {
            best = (int)i;
            best_ms = cand[i].tuned_ms;
        }
    }
    if (best < 0) *err = "no Legendre tile plan could be launched";
    return best;
}

int leg_backward_dev(LegContext *c, int ncomp, int nvec, int nscl, const double *d_sp,
                     double *d_vr, hipStream_t s, std::string *err)
{
    return launch(c, ncomp, nvec, nscl, false, d_sp, d_vr, s, err);
}

int leg_forward_dev(LegContext *c, int ncomp, int nvec, int nscl, const double *d_vr,
                    double *d_sp, hipStream_t s, std::string *err)
{
    return launch(c, ncomp, nvec, nscl, true, d_vr, d_sp, s, err);
}

// FP64 P and dP at all nth colatitudes in the reference's P_rtm(l_rtm, j)
// order (rlm modes, m-major), for checking the device recursion.
int leg_tables_to_host(LegContext *c, double *P, double *dP, std::string *err)
{
    const int L = c->g.L, th_h = c->g.th_h, nth = c->g.nth;
    const size_t nraw = (size_t)(L + 1) * (L + 2) / 2 * th_h;
    double *rP = nullptr, *rdP = nullptr, *rPM = nullptr;
    HIP_TRY(hipMalloc(&rP, sizeof(double) * nraw));
    HIP_TRY(hipMalloc(&rdP, sizeof(double) * nraw));
    HIP_TRY(hipMalloc(&rPM, sizeof(double) * nraw));
    dim3 grid((th_h + 127) / 128, L + 1);
    hipLaunchKernelGGL(k_schmidt_raw, grid, dim3(128), 0, c->stream, L, th_h, c->d_colat, rP, rdP,
                       rPM);
    HIP_TRY(hipGetLastError());
    std::vector<double> hP(nraw), hdP(nraw);
    HIP_TRY(hipMemcpyAsync(hP.data(), rP, sizeof(double) * nraw, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipMemcpyAsync(hdP.data(), rdP, sizeof(double) * nraw, hipMemcpyDeviceToHost, c->stream));
    HIP_TRY(hipStreamSynchronize(c->stream));
    (void)hipFree(rP);
    (void)hipFree(rdP);
    (void)hipFree(rPM);
    // expand hemisphere rows to the full grid by parity; rlm order of modes
    int j = 0;
    for (int mp = 0; mp < 2 * L + 1; ++mp) {
        int m = std::abs(mp - L);
        long long pos = (long long)m * (L + 1) - (long long)m * (m - 1) / 2;
        for (int l = m; l <= L; ++l, ++j) {
            double sgn = ((l - m) & 1) ? -1.0 : 1.0;
            for (int t = 0; t < th_h; ++t) {
                double p = hP[(pos + l - m) * th_h + t], d = hdP[(pos + l - m) * th_h + t];
                P[t + (size_t)nth * j] = p;
                dP[t + (size_t)nth * j] = d;
                int tm = nth - 1 - t;
                if (tm != t) {
                    P[tm + (size_t)nth * j] = sgn * p;
                    dP[tm + (size_t)nth * j] = -sgn * d;
                }
            }
        }
    }
    return 0;
}

}  // namespace calypso
