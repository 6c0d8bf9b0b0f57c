// leg_ozaki.hip -- FP64-accurate Legendre transforms on the int8 MFMA
// (gfx950), split-once form.
//
// Error-free splitting of both operands into int8 digit slices (Ozaki,
// Ogita, Oishi & Rump, Numer. Algorithms 59 (2012) 95-118; integer-MMA
// variant: Ootomo, Ozaki & Yokota, IJHPCA 38 (2024) 297-313).  Every table
// row and every staged column has one power-of-two scale 2^e with |x| < 2^e
// over its whole contraction length K, and each element is held as seven
// balanced base-128 digits
//     x = 2^(e-6) (d0 + d1/128 + ... + d6/128^6),   |d_i| <= 64,
// exact to 2^(e-55).  A row-column product is
//     2^(ea+eb-12) sum_{i+j<=6} 128^-(i+j) (d_i . d_j),
// each digit dot product an exact int32 MFMA sum over all of K
// (|.| <= 7 * K * 64 * 64 < 2^31 for K < 74898), so the seven digit groups
// are converted to FP64 once per output tile.  The pairs left out (i+j > 6)
// and the digits beyond the seventh change a term by about 2^(ea+eb-56).
//
// The contraction is arranged so that each output column is ONE product
// with no FP64 combination afterwards (the factor m/sin(theta) lives in the
// tables, whose rows carry their own scale):
//   backward, output symmetry s (0: even about the equator, 1: odd), rows
//   the hemisphere colatitudes t, K the degrees:
//     v_r, scalars:  P_s                       . [S l(l+1)/r^2 | s]_s
//     v_theta:       [dP_(1-s) | m P_s/sin]    . [dS/r (1-s) ; sg T'/r (s)]
//     v_phi:         [dP_(1-s) | m P_s/sin]    . [-T/r (1-s) ; sg dS'/r (s)]
//   (' = the partner order -m, sg the sign of the output order; P_s the
//   degrees l with l - m = s mod 2), out(t) = S + A, out(mirror t) = S - A
//   (legendre_b_trans_vector_org, legendre_bwd_trans_org.f90:45-209);
//   forward, degrees of parity p as rows, K the hemisphere colatitudes
//   (weights in the tables, halved on the equator), the grid folded about
//   the equator (f(t) +- f(mirror t)):
//     S_pol, scalar: (P_p w)           . fold(v_r | f)
//     dS:            [dP_p w | m P_p w/sin] . [fold-(v_theta) ; -sg fold+(v_phi')]
//     T:             [dP_p w | m P_p w/sin] . [-fold-(v_phi) ; -sg fold+(v_theta')]
//   scaled by r^2 g7, r g7 or g6 (legendre_f_trans_vector_org,
//   legendre_fwd_trans_org.f90:45-169).
// Backward tables are balanced before the split: each table column (degree)
// is divided by a power of two near its largest magnitude over the
// colatitudes and the staged spectral values of that degree are multiplied
// by it (both exact).  Without it, near the poles at small m the table row
// peaks at high degree while the spectrum peaks at low degree, the product
// of the two scales sits far above every term, and the digit pairs left out
// show at ~1e-11; with it the error stays near 1e-13.
// The tables are split into digits once, at set-up.  A work-group takes one
// order m, one column kind and a block of 16*NCT columns; it reads its
// spectral (backward) or grid (forward) values, finds the column scales,
// splits them into digits in LDS once, and runs the MFMAs for every row tile
// of the order with the table digits read from global memory (L2-resident).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <string>
#include <vector>

#include "hip_check.hpp"
#include "leg_internal.hpp"

namespace calypso {

typedef int oz_i4v __attribute__((ext_vector_type(4)));
typedef double oz_d4v __attribute__((ext_vector_type(4)));

namespace {

constexpr int kOzD = 7;             // digits per element
constexpr int kOzG = 6;             // digit-pair orders kept (i + j <= 6)
constexpr int kOzExpMin = -960;     // exponent floor: keeps 2^(6-e) finite
constexpr int kOzPlane = 1024;      // bytes of one digit plane of a 16 x 64 tile
constexpr int kOzUnit = kOzD * kOzPlane;
constexpr int kOzThreads = 256;

// degrees l in [m, L] with l - m = s (mod 2)
__host__ __device__ inline int oz_nk(int L, int m, int s)
{
    return (L - m >= s) ? (L - m - s) / 2 + 1 : 0;
}

// e with |x| < 2^e for normal x != 0; zero and subnormals give -1022
__device__ inline int oz_exp(double x) { return ((__double2hiint(x) >> 20) & 0x7FF) - 1022; }

__device__ inline double oz_pow2(int e) { return __hiloint2double((e + 1023) << 20, 0); }

// digit d of four values at scale e, packed into one word (byte q = value q)
__device__ inline void oz_split4(const double (&x)[4], int e, int (&w)[kOzD])
{
    const double sc = oz_pow2(6 - e);
    double y[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) y[q] = x[q] * sc;
#pragma unroll
    for (int d = 0; d < kOzD; ++d) {
        int word = 0;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double r = rint(y[q]);
            word |= ((int)r & 0xFF) << (8 * q);
            y[q] = (y[q] - r) * 128.0;
        }
        w[d] = word;
    }
}

// digit groups I_g += sum_{i+j=g} a_i . b_j of one K step
__device__ inline void oz_mma(oz_i4v (&I)[kOzG + 1], const oz_i4v (&a)[kOzD],
                              const oz_i4v (&b)[kOzD])
{
#pragma unroll
    for (int i = 0; i <= kOzG; ++i)
#pragma unroll
        for (int j = 0; i + j <= kOzG; ++j)
            I[i + j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[i], b[j], I[i + j], 0, 0, 0);
}

// 2^(ea+eb-12) sum_g 128^-g I_g for the lane's four accumulator rows
__device__ inline oz_d4v oz_flush(const oz_i4v (&I)[kOzG + 1], oz_i4v ea, int eb)
{
    oz_d4v c;
#pragma unroll
    for (int v = 0; v < 4; ++v) {
        double t = (double)I[kOzG][v];
#pragma unroll
        for (int g = kOzG - 1; g >= 0; --g) t = fma(t, 0x1p-7, (double)I[g][v]);
        const int E = min(max(ea[v] + eb + (1023 - 12), 0), 2046);   // 0: scale 0
        c[v] = t * __hiloint2double(E << 20, 0);
    }
    return c;
}

struct OzArgs {
    const double *in;
    double *out;
    const signed char *tab;     // digit units
    const int *tex;             // row exponents
    const long long *tunit;     // first unit of each table
    const int *texo;            // first exponent of each table
    const int *tcol;            // backward: column balancing exponents
    const int *tcolo;           // backward: first of them of each table
    const WorkItem *items;
    const int *lstack;
    const double *radius;
    const double *gfac;
    int L, nth, nr, th_h, J, ncomp, nvec, nscl;
    int RT;                     // backward row tiles (hemisphere)
    int cap;                    // LDS digit planes per (s, column tile)
};

// ---------------------------------------------------------------- tables --

struct OzTabDesc {
    long long unit0;
    int ex0, col0, rows, K, KS, m, sp, kind;
};

// backward table (m, s, kind) row t, position k
__device__ inline double oz_tab_bwd(const OzTabDesc &d, int L, int th_h, const double *P,
                                    const double *dP, const double *PM, int t, int k)
{
    if (t >= th_h || k >= d.K) return 0.0;
    const int m = d.m, s = d.sp;
    const long long pos = (long long)m * (L + 1) - (long long)m * (m - 1) / 2;
    if (d.kind == 0) return P[(pos + s + 2 * k) * th_h + t];
    const int k1 = oz_nk(L, m, 1 - s);
    if (k < k1) return dP[(pos + (1 - s) + 2 * k) * th_h + t];
    return PM[(pos + s + 2 * (k - k1)) * th_h + t];
}

// forward table (m, p, kind) row i (degree m + p + 2i), position k
__device__ inline double oz_tab_fwd(const OzTabDesc &d, int L, int th_h, int nth,
                                    const double *P, const double *dP, const double *PM,
                                    const double *w, int i, int k)
{
    if (i >= oz_nk(L, d.m, d.sp) || k >= d.K) return 0.0;
    const int m = d.m;
    const long long pos = (long long)m * (L + 1) - (long long)m * (m - 1) / 2;
    const long long row = (pos + d.sp + 2 * i) * th_h;
    const int t = k < th_h ? k : k - th_h;
    const double wq = (nth - 1 - t == t) ? 0.5 * w[t] : w[t];
    if (d.kind == 0) return P[row + t] * wq;
    return (k < th_h ? dP[row + t] : PM[row + t]) * wq;
}

// backward balancing exponents: one thread per table column k, the power of
// two at or below the column's largest magnitude (0 for an empty column)
__global__ void k_oz_colscale(const OzTabDesc *desc, int L, int th_h, const double *P,
                              const double *dP, const double *PM, int *tcol)
{
    const OzTabDesc d = desc[blockIdx.y];
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= d.K) return;
    double mx = 0.0;
    for (int t = 0; t < th_h; ++t) mx = fmax(mx, fabs(oz_tab_bwd(d, L, th_h, P, dP, PM, t, k)));
    tcol[d.col0 + k] = mx > 0.0 ? min(max(oz_exp(mx) - 1, -900), 900) : 0;
}

// one thread per table row: the row scale, then the digits, 16 K values at a
// time into the lane's 16 bytes of each digit plane
__global__ void k_oz_tables(bool fwd, const OzTabDesc *desc, int L, int th_h, int nth,
                            const double *P, const double *dP, const double *PM,
                            const double *w, const int *tcol, signed char *tab, int *tex)
{
    const OzTabDesc d = desc[blockIdx.y];
    const int row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= d.rows) return;
    auto val = [&](int k) {
        if (k >= d.K) return 0.0;   // K-step padding
        return fwd ? oz_tab_fwd(d, L, th_h, nth, P, dP, PM, w, row, k)
                   : oz_tab_bwd(d, L, th_h, P, dP, PM, row, k) * oz_pow2(-tcol[d.col0 + k]);
    };
    int e = kOzExpMin;
    for (int k = 0; k < d.K; ++k) e = max(e, oz_exp(val(k)));
    tex[d.ex0 + row] = e;
    const int rt = row >> 4, r16 = row & 15;
    for (int ks = 0; ks < d.KS; ++ks)
        for (int kg = 0; kg < 4; ++kg) {
            int words[4][kOzD];
#pragma unroll
            for (int q4 = 0; q4 < 4; ++q4) {
                double x[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) x[q] = val(ks * 64 + kg * 16 + q4 * 4 + q);
                oz_split4(x, e, words[q4]);
            }
            const int lane = r16 + 16 * kg;
            signed char *u = tab + (d.unit0 + (long long)rt * d.KS + ks) * kOzUnit + lane * 16;
#pragma unroll
            for (int dg = 0; dg < kOzD; ++dg)
                *reinterpret_cast<oz_i4v *>(u + dg * kOzPlane) =
                    oz_i4v{words[0][dg], words[1][dg], words[2][dg], words[3][dg]};
        }
}

// ------------------------------------------------------------- columns --
// A column is (radial point k, sign of the order, item); items of kind 0 are
// the vectors' radial parts then the scalars, of kind 1 the vectors'
// (theta, phi) parts: column c = item + nitem (sign + nsg k).
struct OzCol {
    bool ok;
    int k, sg, item, comp;   // comp: output component
};

__device__ inline OzCol oz_col(int c, int m, int kind, int nv, int ns, int nr)
{
    const int nsg = m == 0 ? 1 : 2;
    const int nitem = kind == 0 ? nv + ns : 2 * nv;
    OzCol o;
    o.ok = nitem > 0 && c < nitem * nsg * nr;
    const int cc = o.ok ? c : 0;
    o.item = nitem > 0 ? cc % nitem : 0;
    const int rest = nitem > 0 ? cc / nitem : 0;
    o.sg = rest % nsg;
    o.k = rest / nsg;
    if (kind == 0) o.comp = o.item < nv ? 3 * o.item : 3 * nv + (o.item - nv);
    else o.comp = 3 * (o.item >> 1) + 1 + (o.item & 1);
    return o;
}

// staged backward values of column o at positions k0..k0+3 (output symmetry s),
// times the table's balancing powers bal[]
__device__ inline void oz_bvals(const OzArgs &a, const OzCol &o, int m, int kind, int s,
                                int k0, const int *bal, double (&x)[4])
{
    const int L = a.L, nv = a.nvec, ncomp = a.ncomp;
    const int sgn = o.sg ? -1 : 1;
    const int j0p = a.lstack[L + sgn * m], j0n = a.lstack[L - sgn * m];
    const size_t base = (size_t)ncomp * a.J * o.k;
    const double ri = 1.0 / a.radius[o.k];
    const int ks_ = oz_nk(L, m, s), ko = oz_nk(L, m, 1 - s);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int kq = k0 + q;
        double v = 0.0;
        if (o.ok) {
            if (kind == 0) {
                if (kq < ks_) {
                    const int dl = s + 2 * kq;   // l - m
                    const double g = a.in[base + o.comp + (size_t)ncomp * (j0p + dl)];
                    if (o.item < nv) {
                        const int l = m + dl;
                        v = g * (ri * ri) * ((l == 0) ? 0.5 : (double)(l * (l + 1)));
                    } else {
                        v = g;
                    }
                }
            } else {
                const int nd = o.item >> 1, w = o.item & 1;
                if (kq < ko) {
                    const int dl = (1 - s) + 2 * kq;
                    const size_t ij = base + (size_t)ncomp * (j0p + dl);
                    v = w == 0 ? a.in[ij + 3 * nd + 1] * ri : -(a.in[ij + 3 * nd + 2] * ri);
                } else if (kq - ko < ks_ && m > 0) {   // m = 0: the table half is zero
                    const int dl = s + 2 * (kq - ko);
                    const size_t ij = base + (size_t)ncomp * (j0n + dl);
                    v = (double)sgn * (a.in[ij + 3 * nd + (w == 0 ? 2 : 1)] * ri);
                }
            }
            if (v != 0.0) v *= oz_pow2(bal[kq]);
        }
        x[q] = v;
    }
}

// staged forward values of column o at positions k0..k0+3 (degree parity p)
__device__ inline void oz_fvals(const OzArgs &a, const OzCol &o, int m, int kind, int p,
                                int k0, double (&x)[4])
{
    const int L = a.L, ncomp = a.ncomp, th_h = a.th_h, nth = a.nth;
    const int sgn = o.sg ? -1 : 1;
    const size_t sl = (size_t)nth * a.nr;
    const size_t bq = (size_t)ncomp * ((size_t)nth * o.k + sl * (L + sgn * m));
    const size_t bn = (size_t)ncomp * ((size_t)nth * o.k + sl * (L - sgn * m));
    const double pi_ = (p & 1) ? -1.0 : 1.0;   // parity of P of these degrees
    const int nd = o.item >> 1, w = o.item & 1;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const int kq = k0 + q;
        double v = 0.0;
        if (o.ok) {
            if (kind == 0) {
                if (kq < th_h) {
                    const int t = kq, tm = nth - 1 - t;
                    v = a.in[bq + o.comp + (size_t)ncomp * t]
                      + pi_ * a.in[bq + o.comp + (size_t)ncomp * tm];
                }
            } else if (kq < th_h) {
                const int t = kq, tm = nth - 1 - t;
                const int c = 3 * nd + 1 + w;   // v_theta (w = 0) or v_phi (w = 1)
                const double f = a.in[bq + c + (size_t)ncomp * t] - pi_ * a.in[bq + c + (size_t)ncomp * tm];
                v = w == 0 ? f : -f;
            } else if (kq < 2 * th_h && m > 0) {   // m = 0: the table half is zero
                const int t = kq - th_h, tm = nth - 1 - t;
                const int c = 3 * nd + 1 + (1 - w);   // partner v_phi (w = 0) or v_theta
                v = -(double)sgn * (a.in[bn + c + (size_t)ncomp * t]
                                    + pi_ * a.in[bn + c + (size_t)ncomp * tm]);
            }
        }
        x[q] = v;
    }
}

// ------------------------------------------------------------- kernels --
// LDS: digit planes [s][column tile][K step][digit] of 1 KB (byte (col, k) at
// lane (col + 16 (k/16)) * 16 + k%16, the MFMA B-operand order), then the
// column exponents [s][16 NCT].

// Stage the block's columns: scales (pass 1), then digits (pass 2).  Thread
// t owns word t of every plane: column (t/4)%16, K positions 16 (t/64) +
// 4 (t%4) .. +3.
template <int NCT, int NS, class Vals>
__device__ inline void oz_stage(const OzArgs &a, unsigned char *lds, const int (&KS)[2],
                                Vals vals)
{
    const int t = threadIdx.x;
    int *cex = reinterpret_cast<int *>(lds + (size_t)NS * NCT * a.cap * kOzUnit);
    for (int x = t; x < NS * NCT * 16; x += kOzThreads) cex[x] = kOzExpMin;
    __syncthreads();
    const int kk0 = 16 * (t >> 6) + 4 * (t & 3), c16 = (t >> 2) & 15;
    for (int s = 0; s < NS; ++s)
        for (int ct = 0; ct < NCT; ++ct) {
            int e = kOzExpMin;
            for (int ks = 0; ks < KS[s]; ++ks) {
                double x[4];
                vals(s, ct, ks * 64 + kk0, x);
#pragma unroll
                for (int q = 0; q < 4; ++q) e = max(e, oz_exp(x[q]));
            }
            e = max(e, __shfl_xor(e, 1));
            e = max(e, __shfl_xor(e, 2));
            if ((t & 3) == 0) atomicMax(cex + s * NCT * 16 + ct * 16 + c16, e);
        }
    __syncthreads();
    for (int s = 0; s < NS; ++s)
        for (int ct = 0; ct < NCT; ++ct) {
            const int e = cex[s * NCT * 16 + ct * 16 + c16];
            for (int ks = 0; ks < KS[s]; ++ks) {
                double x[4];
                vals(s, ct, ks * 64 + kk0, x);
                int wd[kOzD];
                oz_split4(x, e, wd);
                int *pl = reinterpret_cast<int *>(
                    lds + ((size_t)((s * NCT + ct) * a.cap + ks) * kOzD) * kOzPlane);
#pragma unroll
                for (int d = 0; d < kOzD; ++d) pl[d * (kOzPlane / 4) + t] = wd[d];
            }
        }
    __syncthreads();
}

// The wave's MFMA work: row tiles rt = wave, wave + 4, ... < nrt; for each,
// the K steps of symmetry 0 and then (NS = 2) of symmetry 1.  The table digits
// of the next step are loaded while the MFMAs of this one run.  done(rt, s, I)
// receives the digit sums of each (row tile, symmetry), in order; a symmetry
// without K steps gives zero sums.
template <int NCT, int NS, class Done>
__device__ inline void oz_gemm(const OzArgs &a, const unsigned char *lds, int nrt,
                               const int (&KS)[2], const long long (&tu)[2], Done done)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int KSe[2];
    for (int s = 0; s < 2; ++s) KSe[s] = (s < NS) ? (KS[s] > 0 ? KS[s] : 1) : 0;
    const int per = KSe[0] + KSe[1];
    const int nmine = nrt > wave ? (nrt - wave + 3) / 4 : 0;
    const int nstep = nmine * per;
    auto decode = [&](int i, int &rt, int &s, int &ks) {
        const int r = i / per, q = i - r * per;
        rt = wave + 4 * r;
        s = q >= KSe[0] ? 1 : 0;
        ks = s ? q - KSe[0] : q;
    };
    auto load_a = [&](oz_i4v (&A)[kOzD], int i) {
        int rt, s, ks;
        decode(i, rt, s, ks);
        if (ks >= KS[s]) return;
        const signed char *u = a.tab + (tu[s] + (long long)rt * KS[s] + ks) * kOzUnit + lane * 16;
#pragma unroll
        for (int d = 0; d < kOzD; ++d) A[d] = *reinterpret_cast<const oz_i4v *>(u + d * kOzPlane);
    };
    oz_i4v I[NCT][kOzG + 1];
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct)
#pragma unroll
        for (int g = 0; g <= kOzG; ++g) I[ct][g] = oz_i4v{0, 0, 0, 0};
    oz_i4v A[kOzD], An[kOzD];
#pragma unroll
    for (int d = 0; d < kOzD; ++d) A[d] = An[d] = oz_i4v{0, 0, 0, 0};
    if (nstep > 0) load_a(A, 0);
    for (int i = 0; i < nstep; ++i) {
        if (i + 1 < nstep) load_a(An, i + 1);
        int rt, s, ks;
        decode(i, rt, s, ks);
        if (ks < KS[s]) {
#pragma unroll
            for (int ct = 0; ct < NCT; ++ct) {
                oz_i4v B[kOzD];
                const unsigned char *b =
                    lds + ((size_t)((s * NCT + ct) * a.cap + ks) * kOzD) * kOzPlane + lane * 16;
#pragma unroll
                for (int d = 0; d < kOzD; ++d) B[d] = *reinterpret_cast<const oz_i4v *>(b + d * kOzPlane);
                oz_mma(I[ct], A, B);
            }
        }
        if (ks == KSe[s] - 1) {
            done(rt, s, I);
#pragma unroll
            for (int ct = 0; ct < NCT; ++ct)
#pragma unroll
                for (int g = 0; g <= kOzG; ++g) I[ct][g] = oz_i4v{0, 0, 0, 0};
        }
#pragma unroll
        for (int d = 0; d < kOzD; ++d) A[d] = An[d];
    }
}

// Backward: item (m, kind, column block); both output symmetries staged, then
// row tile by row tile S and A and the two hemispheres.
template <int NCT>
__global__ __launch_bounds__(kOzThreads) void k_oz_bwd(OzArgs a)
{
    const WorkItem it = a.items[blockIdx.x];
    if (it.m < 0) return;   // padding of the XCD interleave
    const int m = it.m, kind = it.ktile, cb = it.rgroup;
    extern __shared__ __align__(16) unsigned char lds[];
    const int L = a.L, nv = a.nvec, ns = a.nscl;
    int KS[2];
    long long tu[2];
    int te[2];
    const int *bal[2];
    for (int s = 0; s < 2; ++s) {
        const int K = kind == 0 ? oz_nk(L, m, s) : L - m + 1;
        KS[s] = (K + 63) / 64;
        const int tb = (m * 2 + s) * 2 + kind;
        tu[s] = a.tunit[tb];
        te[s] = a.texo[tb];
        bal[s] = a.tcol + a.tcolo[tb];
    }
    const int c16s = (threadIdx.x >> 2) & 15;
    OzCol sc[NCT];
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) sc[ct] = oz_col(cb * 16 * NCT + ct * 16 + c16s, m, kind, nv, ns, a.nr);
    oz_stage<NCT, 2>(a, lds, KS, [&](int s, int ct, int k0, double (&x)[4]) {
        oz_bvals(a, sc[ct], m, kind, s, k0, bal[s], x);
    });

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int *cex = reinterpret_cast<const int *>(lds + (size_t)2 * NCT * a.cap * kOzUnit);
    OzCol oc[NCT];
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) oc[ct] = oz_col(cb * 16 * NCT + ct * 16 + (lane & 15), m, kind, nv, ns, a.nr);
    const size_t sl = (size_t)a.nth * a.nr;
    oz_d4v res0[NCT];
    oz_gemm<NCT, 2>(a, lds, a.RT, KS, tu, [&](int rt, int s, const oz_i4v (&I)[NCT][kOzG + 1]) {
        const oz_i4v ea = *reinterpret_cast<const oz_i4v *>(a.tex + te[s] + rt * 16 + 4 * (lane >> 4));
        if (s == 0) {
#pragma unroll
            for (int ct = 0; ct < NCT; ++ct)
                res0[ct] = oz_flush(I[ct], ea, cex[ct * 16 + (lane & 15)]);
            return;
        }
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct) {
            const oz_d4v r1 = oz_flush(I[ct], ea, cex[NCT * 16 + ct * 16 + (lane & 15)]);
            const OzCol &o = oc[ct];
            if (!o.ok) continue;
            const int mp = L + (o.sg ? -m : m);
            double *out = a.out + o.comp + (size_t)a.ncomp * ((size_t)a.nth * o.k + sl * mp);
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int t = rt * 16 + 4 * (lane >> 4) + v;
                if (t >= a.th_h) continue;
                const int tm = a.nth - 1 - t;
                out[(size_t)a.ncomp * t] = res0[ct][v] + r1[v];
                if (tm != t) out[(size_t)a.ncomp * tm] = res0[ct][v] - r1[v];
            }
        }
    });
}

// Forward: item (m, 2 p + kind, column block); rows the degrees of parity p.
template <int NCT>
__global__ __launch_bounds__(kOzThreads) void k_oz_fwd(OzArgs a)
{
    const WorkItem it = a.items[blockIdx.x];
    if (it.m < 0) return;   // padding of the XCD interleave
    const int m = it.m, p = it.ktile >> 1, kind = it.ktile & 1, cb = it.rgroup;
    extern __shared__ __align__(16) unsigned char lds[];
    const int L = a.L, nv = a.nvec, ns = a.nscl;
    const int K = kind == 0 ? a.th_h : 2 * a.th_h;
    const int KS[2] = {(K + 63) / 64, 0};
    const int tb = (m * 2 + p) * 2 + kind;
    const long long tu = a.tunit[tb];
    const int te = a.texo[tb];
    const int nrow = oz_nk(L, m, p), RTp = (nrow + 15) / 16;
    const int c16s = (threadIdx.x >> 2) & 15;
    OzCol sc[NCT];
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) sc[ct] = oz_col(cb * 16 * NCT + ct * 16 + c16s, m, kind, nv, ns, a.nr);
    oz_stage<NCT, 1>(a, lds, KS, [&](int, int ct, int k0, double (&x)[4]) {
        oz_fvals(a, sc[ct], m, kind, p, k0, x);
    });

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int *cex = reinterpret_cast<const int *>(lds + (size_t)NCT * a.cap * kOzUnit);
    OzCol oc[NCT];
#pragma unroll
    for (int ct = 0; ct < NCT; ++ct) oc[ct] = oz_col(cb * 16 * NCT + ct * 16 + (lane & 15), m, kind, nv, ns, a.nr);
    const double m2 = (m == 0) ? 2.0 : 1.0;
    const long long tus[2] = {tu, 0};
    oz_gemm<NCT, 1>(a, lds, RTp, KS, tus, [&](int rt, int, const oz_i4v (&I)[NCT][kOzG + 1]) {
        const oz_i4v ea = *reinterpret_cast<const oz_i4v *>(a.tex + te + rt * 16 + 4 * (lane >> 4));
#pragma unroll
        for (int ct = 0; ct < NCT; ++ct) {
            const OzCol &o = oc[ct];
            const oz_d4v r = oz_flush(I[ct], ea, cex[ct * 16 + (lane & 15)]);
            if (!o.ok) continue;
            const int j0 = a.lstack[L + (o.sg ? -m : m)];
            const double rad = a.radius[o.k];
            double *out = a.out + o.comp + (size_t)a.ncomp * a.J * o.k;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int i = rt * 16 + 4 * (lane >> 4) + v;
                if (i >= nrow) continue;
                const int l = m + p + 2 * i;
                double fac;
                if (kind == 0 && o.item >= nv) {
                    fac = (l == 0) ? 0.5 : m2 * a.gfac[L + 1 + l];
                } else {
                    const double g7 = (l == 0) ? 1.0 : m2 * a.gfac[l];
                    fac = kind == 0 ? rad * rad * g7 : rad * g7;
                }
                out[(size_t)a.ncomp * (j0 + (l - m))] = r[v] * fac;
            }
        }
    });
}

}  // namespace

// ------------------------------------------------------------- host ----

struct OzTables {
    signed char *tab = nullptr;
    int *tex = nullptr;
    long long *tunit[2] = {nullptr, nullptr};   // [bwd, fwd] per table
    int *texo[2] = {nullptr, nullptr};
    int *tcol = nullptr, *tcolo = nullptr;      // backward balancing
    int cap[2] = {0, 0};   // largest K-step count of a column (bwd: per s)
};

void oz_free(LegContext *c)
{
    if (!c->oz) return;
    (void)hipFree(c->oz->tab);
    (void)hipFree(c->oz->tex);
    (void)hipFree(c->oz->tcol);
    (void)hipFree(c->oz->tcolo);
    for (int d = 0; d < 2; ++d) {
        (void)hipFree(c->oz->tunit[d]);
        (void)hipFree(c->oz->texo[d]);
    }
    delete c->oz;
    c->oz = nullptr;
}

int oz_build_tables(LegContext *c, const double *rP, const double *rdP, const double *rPM,
                    std::string *err)
{
    const int L = c->g.L, th_h = c->g.th_h, nth = c->g.nth;
    if (2 * th_h > 74000 || L + 1 > 74000) {
        *err = "int8 form: contraction length beyond the int32 digit sums";
        return -1;
    }
    c->oz = new OzTables;
    OzTables &z = *c->oz;
    const int RT = (th_h + 15) / 16;
    c->rt_bwd = RT;
    std::vector<OzTabDesc> desc[2];
    long long units = 0;
    int exps = 0, cols = 0;
    for (int dir = 0; dir < 2; ++dir)
        for (int m = 0; m <= L; ++m)
            for (int sp = 0; sp < 2; ++sp)
                for (int kind = 0; kind < 2; ++kind) {
                    OzTabDesc d;
                    d.m = m;
                    d.sp = sp;
                    d.kind = kind;
                    if (dir == 0) {
                        d.rows = RT * 16;
                        d.K = kind == 0 ? oz_nk(L, m, sp) : L - m + 1;
                    } else {
                        d.rows = (oz_nk(L, m, sp) + 15) / 16 * 16;
                        d.K = kind == 0 ? th_h : 2 * th_h;
                    }
                    d.KS = (d.K + 63) / 64;
                    d.unit0 = units;
                    d.ex0 = exps;
                    d.col0 = dir == 0 ? cols : 0;
                    units += (long long)(d.rows / 16) * d.KS;
                    exps += d.rows;
                    if (dir == 0) cols += d.K;
                    desc[dir].push_back(d);
                }
    z.cap[0] = (L + 1 + 63) / 64;
    z.cap[1] = (2 * th_h + 63) / 64;
    HIP_TRY(hipMalloc(&z.tab, (size_t)std::max(units, 1LL) * kOzUnit));
    HIP_TRY(hipMalloc(&z.tex, sizeof(int) * std::max(exps, 1)));
    HIP_TRY(hipMalloc(&z.tcol, sizeof(int) * std::max(cols, 1)));
    {
        std::vector<int> co;
        for (const OzTabDesc &d : desc[0]) co.push_back(d.col0);
        HIP_TRY(hipMalloc(&z.tcolo, sizeof(int) * co.size()));
        HIP_TRY(hipMemcpy(z.tcolo, co.data(), sizeof(int) * co.size(), hipMemcpyHostToDevice));
    }
    c->tab_bytes = (size_t)units * kOzUnit + sizeof(int) * exps;
    for (int dir = 0; dir < 2; ++dir) {
        std::vector<long long> tu;
        std::vector<int> te;
        int maxrows = 0;
        for (const OzTabDesc &d : desc[dir]) {
            tu.push_back(d.unit0);
            te.push_back(d.ex0);
            maxrows = std::max(maxrows, d.rows);
        }
        HIP_TRY(hipMalloc(&z.tunit[dir], sizeof(long long) * tu.size()));
        HIP_TRY(hipMalloc(&z.texo[dir], sizeof(int) * te.size()));
        HIP_TRY(hipMemcpy(z.tunit[dir], tu.data(), sizeof(long long) * tu.size(), hipMemcpyHostToDevice));
        HIP_TRY(hipMemcpy(z.texo[dir], te.data(), sizeof(int) * te.size(), hipMemcpyHostToDevice));
        DevBuf dd;
        HIP_TRY(hipMalloc(&dd.p, sizeof(OzTabDesc) * desc[dir].size()));
        HIP_TRY(hipMemcpy(dd.p, desc[dir].data(), sizeof(OzTabDesc) * desc[dir].size(),
                          hipMemcpyHostToDevice));
        if (dir == 0) {
            int maxk = 0;
            for (const OzTabDesc &d : desc[0]) maxk = std::max(maxk, d.K);
            if (maxk > 0) {
                const dim3 grid((maxk + 63) / 64, (unsigned)desc[0].size());
                hipLaunchKernelGGL(k_oz_colscale, grid, dim3(64), 0, c->stream,
                                   static_cast<const OzTabDesc *>(dd.p), L, th_h, rP, rdP, rPM,
                                   z.tcol);
                HIP_TRY(hipGetLastError());
            }
        }
        if (maxrows > 0) {
            const dim3 grid((maxrows + 63) / 64, (unsigned)desc[dir].size());
            hipLaunchKernelGGL(k_oz_tables, grid, dim3(64), 0, c->stream, dir == 1,
                               static_cast<const OzTabDesc *>(dd.p), L, th_h, nth, rP, rdP, rPM,
                               c->d_weight, z.tcol, z.tab, z.tex);
            HIP_TRY(hipGetLastError());
        }
        HIP_TRY(hipStreamSynchronize(c->stream));
    }
    return 0;
}

static size_t oz_lds(const LegContext *c, int nct, bool fwd)
{
    const int ns = fwd ? 1 : 2;
    return (size_t)ns * nct * c->oz->cap[fwd ? 1 : 0] * kOzUnit + sizeof(int) * ns * nct * 16;
}

void oz_plan_candidates(int ncomp, int nv, int ns, bool fwd, std::vector<LegPlan> *out)
{
    for (int nct : {1, 2}) {
        LegPlan p;
        p.ncomp = ncomp;
        p.nvec = nv;
        p.nscl = ns;
        p.fwd = fwd;
        p.ver = 5;
        p.mct1 = nct;
        p.mct2 = 0;
        p.KT = 1;
        p.rpw = 0;
        p.tc = 64;
        out->push_back(p);
    }
}

int oz_build_items(LegContext *c, LegPlan *p, bool fwd, std::string *err)
{
    if (!c->oz) {
        *err = "int8 form: tables missing";
        return -1;
    }
    if (oz_lds(c, p->mct1, fwd) > 160 * 1024) {
        p->n_items = 0;
        p->d_items = nullptr;
        return 1;   // does not fit: not a candidate
    }
    const LegGeometry &g = c->g;
    const int L = g.L, nv = p->nvec, ns = p->nscl, cols = 16 * p->mct1;
    // items of one order that read the same radial points are adjacent (so
    // they run together and share the input through L2): sorted by (m, first
    // radial point of the block, kind, parity)
    struct Keyed {
        int m, k0, kind, par;
        WorkItem w;
    };
    std::vector<Keyed> keyed;
    for (int m = 0; m <= L; ++m) {
        const int nsg = m == 0 ? 1 : 2;
        for (int kind = 0; kind < 2; ++kind) {
            const int nitem = kind == 0 ? nv + ns : 2 * nv;
            if (nitem == 0) continue;
            const int ncol = nitem * nsg * g.nr;
            const int nb = (ncol + cols - 1) / cols;
            for (int b = 0; b < nb; ++b) {
                const int k0 = b * cols / (nitem * nsg);
                if (fwd) {
                    for (int par = 0; par < 2; ++par)
                        if (oz_nk(L, m, par) > 0) keyed.push_back({m, k0, kind, par, {m, 2 * par + kind, b, 0}});
                } else {
                    keyed.push_back({m, k0, kind, 0, {m, kind, b, 0}});
                }
            }
        }
    }
    std::stable_sort(keyed.begin(), keyed.end(), [](const Keyed &x, const Keyed &y) {
        if (x.m != y.m) return x.m < y.m;
        if (x.k0 != y.k0) return x.k0 < y.k0;
        if (x.kind != y.kind) return x.kind < y.kind;
        return x.par < y.par;
    });
    std::vector<WorkItem> items;
    for (const Keyed &x : keyed) items.push_back(x.w);
    // orders m and L-m on one XCD (their tables share its L2)
    constexpr int NX = 8;
    std::vector<WorkItem> per[NX];
    for (const WorkItem &w : items) per[std::min(w.m, L - w.m) % NX].push_back(w);
    size_t len = 0;
    for (int x = 0; x < NX; ++x) len = std::max(len, per[x].size());
    std::vector<WorkItem> out(len * NX, WorkItem{-1, 0, 0, 0});
    for (int x = 0; x < NX; ++x)
        for (size_t i = 0; i < per[x].size(); ++i) out[i * NX + x] = per[x][i];
    p->n_items = (int)out.size();
    HIP_TRY(hipMalloc(&p->d_items, sizeof(WorkItem) * std::max<size_t>(out.size(), 1)));
    HIP_TRY(hipMemcpy(p->d_items, out.data(), sizeof(WorkItem) * out.size(), hipMemcpyHostToDevice));
    return 0;
}

int oz_launch(LegContext *c, const LegPlan &p, bool fwd, const double *in, double *out,
              hipStream_t s, std::string *err)
{
    if (!c->oz || p.ver != 5) {
        *err = "int8 form: no tables or plan";
        return -1;
    }
    OzArgs a;
    a.in = in;
    a.out = out;
    a.tab = c->oz->tab;
    a.tex = c->oz->tex;
    a.tunit = c->oz->tunit[fwd ? 1 : 0];
    a.texo = c->oz->texo[fwd ? 1 : 0];
    a.tcol = c->oz->tcol;
    a.tcolo = c->oz->tcolo;
    a.items = p.d_items;
    a.lstack = c->d_lstack;
    a.radius = c->d_radius;
    a.gfac = c->d_gfac;
    a.L = c->g.L;
    a.nth = c->g.nth;
    a.nr = c->g.nr;
    a.th_h = c->g.th_h;
    a.J = c->g.J;
    a.ncomp = p.ncomp;
    a.nvec = p.nvec;
    a.nscl = p.nscl;
    a.RT = c->rt_bwd;
    a.cap = c->oz->cap[fwd ? 1 : 0];
    const size_t lds = oz_lds(c, p.mct1, fwd);
    if (p.n_items == 0) return 0;
    const void *fn;
    if (fwd) fn = p.mct1 == 2 ? reinterpret_cast<const void *>(&k_oz_fwd<2>)
                              : reinterpret_cast<const void *>(&k_oz_fwd<1>);
    else fn = p.mct1 == 2 ? reinterpret_cast<const void *>(&k_oz_bwd<2>)
                          : reinterpret_cast<const void *>(&k_oz_bwd<1>);
    HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (fwd) {
        if (p.mct1 == 2)
            hipLaunchKernelGGL(k_oz_fwd<2>, dim3(p.n_items), dim3(kOzThreads), lds, s, a);
        else
            hipLaunchKernelGGL(k_oz_fwd<1>, dim3(p.n_items), dim3(kOzThreads), lds, s, a);
    } else {
        if (p.mct1 == 2)
            hipLaunchKernelGGL(k_oz_bwd<2>, dim3(p.n_items), dim3(kOzThreads), lds, s, a);
        else
            hipLaunchKernelGGL(k_oz_bwd<1>, dim3(p.n_items), dim3(kOzThreads), lds, s, a);
    }
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // namespace calypso
