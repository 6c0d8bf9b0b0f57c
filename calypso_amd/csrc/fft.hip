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
// components: it stages the pencil's 2L+1 coefficients of a group of
// components in LDS (reads of contiguous doubles per order), runs the FFTs of
// pairs of real components as complex Stockham FFTs in place (radix 4/2/3,
// FP64, twiddles from an FP64 table in LDS, butterfly inputs held in
// registers across each stage's barrier), and writes the other side.  The
// fused MHD kernel keeps the grid values in LDS between the inverse FFT, the
// products (cal_nonlinear_sph_MHD.f90:35-75) and the forward FFT, so the
// physical grid never touches HBM; the spectra are staged in the same LDS
// buffer, so a pencil needs (pairs + 1) * nphi complex words.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>
#include <vector>

#include "hip_check.hpp"
#include "fft_internal.hpp"

namespace calypso {


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
    const double *tw;   // per-stage twiddle rows (N complex slots)
    int pairs_per_batch;
    double coef[4];     // coef_velo, coef_lor, coef_induct, coef_temp
};

constexpr int kFftMaxE = 12;   // complex values per thread per stage (any radix)
// butterflies per thread per stage
__host__ __device__ constexpr int fft_mb(int R) { return kFftMaxE / R; }
constexpr int kFftMaxZ = 4;    // frequencies per thread per pair (inverse side)
constexpr int kFftMaxU = 2;    // orders per thread per pair (forward side)
constexpr int kFftPre = 16;    // next pencil's staged spectra per thread (prefetch)

// n / d for 0 <= n < 2^23, 1 <= d, with rd = 1.0f / d: the float quotient is
// within one of the true one; one integer step corrects it (a few VALU
// instructions instead of an integer division sequence)
__device__ inline int qdiv(int n, int d, float rd)
{
    int q = (int)((float)n * rd);
    const int r = n - q * d;
    q += (r >= d) - (r < 0);
    return q;
}

// e = hi d + lo stepped by a constant: the quotient once, then carries
struct DivStep {
    int hi, lo, dhi, dlo, d;
    __device__ DivStep(int e0, int step, int d_) : d(d_)
    {
        hi = qdiv(e0, d, 1.0f / (float)d);
        lo = e0 - hi * d;
        dhi = step / d;
        dlo = step - dhi * d;
    }
    __device__ void next()
    {
        hi += dhi;
        lo += dlo;
        if (lo >= d) { lo -= d; ++hi; }
    }
};

// size-4 DFT X_k = sum_n x_n exp(sgn 2 pi i n k / 4)
__device__ inline void dft4(cplx x0, cplx x1, cplx x2, cplx x3, double sgn, cplx &X0, cplx &X1,
                            cplx &X2, cplx &X3)
{
    const cplx a0 = cadd(x0, x2), a1 = csub(x0, x2);
    const cplx b0 = cadd(x1, x3), b1 = cmul_si(csub(x1, x3), sgn);
    X0 = cadd(a0, b0);
    X1 = cadd(a1, b1);
    X2 = csub(a0, b0);
    X3 = csub(a1, b1);
}

// size-R DFT of one butterfly's (already twiddled) inputs, R = 2, 3, 4,
// X_k = sum_n x_n exp(sgn 2 pi i n k / R)
template <int R>
__device__ inline void dft(cplx (&v)[R], double sgn, cplx (&o)[R])
{
    if constexpr (R == 2) {
        o[0] = cadd(v[0], v[1]);
        o[1] = csub(v[0], v[1]);
    } else if constexpr (R == 4) {
        dft4(v[0], v[1], v[2], v[3], sgn, o[0], o[1], o[2], o[3]);
    } else {   // R == 3
        const double c = -0.5, sn = sgn * 0.86602540378443864676;
        cplx t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
        cplx m1 = {v[0].x + c * t1.x, v[0].y + c * t1.y};
        cplx m2 = {-sn * t2.y, sn * t2.x};
        o[0] = cadd(v[0], t1);
        o[1] = cadd(m1, m2);
        o[2] = csub(m1, m2);
    }
}

// One in-place Stockham stage of radix R over `nseq` complex sequences of
// length N held in buf (sequence s at buf + s N), mixed-radix autosort form
// (Govindaraju et al., SC'08): every thread reads the R inputs of its
// butterflies into registers, the work-group synchronises, and the outputs go
// to their autosorted positions in the same buffer.
template <int R, int NT, int SGN>
__device__ void stage_inplace(cplx *buf, const cplx *tw, int N, int Ns, int nseq)
{
    constexpr double sgn = SGN;
    constexpr int MB = fft_mb(R);   // butterflies per thread
    const int nb = N / R, total = nseq * nb;
    // butterfly b of this thread is idx = tid + b NT = s nb + j: (s, j) from
    // one quotient, then stepped by NT = ds nb + dj
    const int ds = NT / nb, dj = NT - ds * nb;
    const int s0 = qdiv(threadIdx.x, nb, 1.0f / (float)nb), j0 = threadIdx.x - s0 * nb;
    // Ns is a power of two whenever the threes come last (2^a 3 lengths)
    const bool ns2 = (Ns & (Ns - 1)) == 0;
    const int lns = __builtin_ctz(Ns);
    const float rns = 1.0f / (float)Ns;
    cplx v[MB][R];
    {
        int s = s0, j = j0;
#pragma unroll
        for (int b = 0; b < MB; ++b) {
            if (threadIdx.x + b * NT < total) {
                const cplx *src = buf + (size_t)s * N;
#pragma unroll
                for (int r = 0; r < R; ++r) v[b][r] = src[j + r * nb];
            }
            s += ds;
            j += dj;
            if (j >= nb) { j -= nb; ++s; }
        }
    }
    __syncthreads();
    {
        int s = s0, j = j0;
#pragma unroll
        for (int b = 0; b < MB; ++b) {
            if (threadIdx.x + b * NT < total) {
                const int jq = ns2 ? (j >> lns) : qdiv(j, Ns, rns);
                const int kk = j - jq * Ns;
                // this stage's twiddle rows: entry (r-1) Ns + kk = W_N^(r kk tstep);
                // the table holds exp(-i...), the inverse (SGN = +1) conjugates
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    const cplx w = tw[(r - 1) * Ns + kk];
                    v[b][r] = cmul(v[b][r], {w.x, -sgn * w.y});
                }
                cplx o[R];
                dft<R>(v[b], sgn, o);
                cplx *dst = buf + (size_t)s * N + jq * Ns * R + kk;
#pragma unroll
                for (int r = 0; r < R; ++r) dst[r * Ns] = o[r];
            }
            s += ds;
            j += dj;
            if (j >= nb) { j -= nb; ++s; }
        }
    }
    __syncthreads();
}

// full in-place FFT of nseq sequences; tw holds the stages' twiddle rows one
// after the other ((R-1) Ns entries per stage, N-1 in all)
template <int NT, int SGN>
__device__ void fft_inplace(cplx *buf, const cplx *tw, const FftArgs &a, int nseq)
{
    int Ns = 1, off = 0;
    for (int st = 0; st < a.nstage; ++st) {
        const int R = a.radix[st];
        if (R == 4) stage_inplace<4, NT, SGN>(buf, tw + off, a.N, Ns, nseq);
        else if (R == 2) stage_inplace<2, NT, SGN>(buf, tw + off, a.N, Ns, nseq);
        else stage_inplace<3, NT, SGN>(buf, tw + off, a.N, Ns, nseq);
        off += (R - 1) * Ns;
        Ns *= R;
    }
}

enum PencilMode { kPencilBwd = 0, kPencilFwd = 1, kPencilMhd = 2 };

// grid products at longitude i, in place: buf holds the 13 b_trns fields as
// pairs (0,1) ... (12,-); the 12 products go to pairs 0..5
// (cal_nonlinear_sph_MHD.f90:35-75, same operation order, no contraction:
// bitwise equal to the reference formulas)
__device__ inline void mhd_products(cplx *buf, int N, int i, const double *coef)
{
    // b_trns: u 0-2, w 3-5, B 6-8, J 9-11, T 12
    const cplx c01 = buf[0 * N + i], c23 = buf[1 * N + i], c45 = buf[2 * N + i];
    const cplx c67 = buf[3 * N + i], c89 = buf[4 * N + i], c1011 = buf[5 * N + i];
    const cplx c12 = buf[6 * N + i];
    const double u1 = c01.x, u2 = c01.y, u3 = c23.x;
    const double w1 = c23.y, w2 = c45.x, w3 = c45.y;
    const double b1 = c67.x, b2 = c67.y, b3 = c89.x;
    const double j1 = c89.y, j2 = c1011.x, j3 = c1011.y;
    const double T = c12.x;
    const double cv = coef[0], cl = coef[1], ci = coef[2], ct = coef[3];
    double f[12];
    {
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
    for (int p = 0; p < 6; ++p) buf[(size_t)p * N + i] = {f[2 * p], f[2 * p + 1]};
}

// Hermitian spectrum of z = x_a + i x_b at frequency kf from the real
// spectra of components a, b (rows ra, rb of M doubles, rtm order: a_m at
// L+m, b_m at L-m; rb null: x_b = 0)
__device__ inline cplx herm_in(const double *ra, const double *rb, int kf, int N, int L)
{
    cplx z = {0.0, 0.0};
    if (kf == 0) {
        z.x = ra[L];
        z.y = rb ? rb[L] : 0.0;
    } else if (kf <= L) {
        const double aa = ra[L + kf], ba = ra[L - kf];
        const double ab = rb ? rb[L + kf] : 0.0, bb = rb ? rb[L - kf] : 0.0;
        z.x = 0.5 * (aa + bb);
        z.y = 0.5 * (ab - ba);
    } else if (kf >= N - L) {
        const int m = N - kf;
        const double aa = ra[L + m], ba = ra[L - m];
        const double ab = rb ? rb[L + m] : 0.0, bb = rb ? rb[L - m] : 0.0;
        z.x = 0.5 * (aa - bb);
        z.y = 0.5 * (ba + ab);
    }
    return z;
}

// One work-group per (radius, colatitude) pencil.  LDS: tw[N] | buf[PB][N]
// complex; the spectra of the pencil are staged in buf (as doubles) before
// the inverse FFTs and after the forward ones, so a pencil needs (PB + 1) N
// complex words.  Components go in pairs (real / imaginary part of one
// complex FFT), PB pairs per group; the MHD mode holds all 7 input pairs.
// (256 threads: two work-groups per CU, i.e. at most 256 registers a lane)
template <int MODE, int NT>
__global__ __launch_bounds__(NT, NT == 256 ? 2 : 1) void k_pencil(FftArgs a)
{
    extern __shared__ __align__(16) unsigned char smem_raw[];
    const int M = 2 * a.L + 1, N = a.N, L = a.L;
    cplx *tw = reinterpret_cast<cplx *>(smem_raw);
    cplx *buf = tw + N;
    double *bd = reinterpret_cast<double *>(buf);
    const int PB = a.pairs_per_batch;

    const size_t sl = (size_t)a.nth * a.nr;   // rtm slot stride in pencils
    const size_t nnod = (size_t)a.nr * a.nth * N;
    const int npencil = a.nth * a.nr;
    const int nin = a.ncomp_in;
    const int npair_in = (nin + 1) / 2;
    // rtm words of one order across every pencil: the host keeps whole rtm
    // arrays below 2^32 words, so offsets within them are 32-bit
    const unsigned nsl = (unsigned)nin * (unsigned)sl;

    for (int i = threadIdx.x; i < N; i += NT) tw[i] = {a.tw[2 * i], a.tw[2 * i + 1]};

    // persistent work-groups: the first component group of the next pencil
    // is loaded into registers while the current pencil is transformed
    const int nc0 = min(2 * PB, nin);
    double pre[kFftPre];
    auto prefetch = [&](int pc) {
        // the slot arithmetic is redone per call rather than held in
        // registers across the pencil loop
        int tid = threadIdx.x;
        asm volatile("" : "+v"(tid));
        // e = mp nc0 + nd sits at nd + nin (pc + sl mp): a 32-bit offset from
        // the pencil's first word, stepped by the quotient's carries (no
        // multiplies per load)
        DivStep qp(tid, NT, nc0);
        unsigned off = (unsigned)qp.lo + nsl * (unsigned)qp.hi;
        const unsigned doff = (unsigned)qp.dlo + nsl * (unsigned)qp.dhi;
        const double *src = a.in + (size_t)nin * (pc < npencil ? pc : 0);
#pragma unroll
        for (int q = 0; q < kFftPre; ++q) {
            const int e = tid + q * NT;
            pre[q] = 0.0;
            if (pc < npencil && e < nc0 * M) pre[q] = src[off];
            off += doff;
            qp.lo += qp.dlo;
            if (qp.lo >= nc0) { qp.lo -= nc0; off += nsl - (unsigned)nc0; }
        }
    };
    // pencils of round i (offsets i G + o): chunks of 4 consecutive pencils
    // go to one XCD (work-groups b = x mod 8 run on XCD x) and consecutive
    // chunks to consecutive XCDs, so neighbouring colatitudes, which share
    // cache lines of every order's record, are read through one L2 while a
    // partial last round still spreads over all XCDs
    const int G = gridDim.x;
    const bool xm = (G & 31) == 0;
    const int slot = blockIdx.x >> 3;
    const int xbase = (((slot >> 2) << 3) + (blockIdx.x & 7)) * 4 + (slot & 3);
    auto pencil_of = [&](int i) { return xm ? i * G + xbase : (int)blockIdx.x + i * G; };
    if (MODE != kPencilFwd) prefetch(pencil_of(0));

    for (int it = 0, pencil = pencil_of(0); pencil < npencil; pencil = pencil_of(++it)) {
    const int t = pencil % a.nth, k = pencil / a.nth;   // pencil = t + nth*k
    const size_t gbase = (size_t)k + (size_t)a.nr * t;  // rtp node at phi = 0

    if constexpr (MODE == kPencilMhd) {
        // spectra in their own LDS rows after the 7 FFT sequences: 13 rows
        // of M doubles in, the 12 product rows out, so each unpack is one
        // pass with no barrier inside
        double *spec = reinterpret_cast<double *>(buf + (size_t)7 * N);
        __syncthreads();   // spec and buf free (previous pencil written out)
        {
            int tid = threadIdx.x;
            asm volatile("" : "+v"(tid));
            // e = mp 13 + nd goes to row nd, column mp: index stepped by carries
            DivStep qs(tid, NT, 13);
            int dst = qs.lo * M + qs.hi;
            const int ddst = qs.dlo * M + qs.dhi;
#pragma unroll
            for (int q = 0; q < kFftPre; ++q) {
                const int e = tid + q * NT;
                if (e < 13 * M) spec[dst] = pre[q];
                dst += ddst;
                qs.lo += qs.dlo;
                if (qs.lo >= 13) { qs.lo -= 13; dst += 1 - 13 * M; }
            }
        }
        prefetch(pencil_of(it + 1));
        __syncthreads();
        {
            DivStep q(threadIdx.x, NT, N);   // e = p N + kf
            for (int e = threadIdx.x; e < 7 * N; e += NT, q.next()) {
                const int p = q.hi;
                buf[e] = herm_in(spec + (size_t)2 * p * M,
                                 p < 6 ? spec + (size_t)(2 * p + 1) * M : nullptr, q.lo, N, L);
            }
        }
        __syncthreads();
        fft_inplace<NT, +1>(buf, tw, a, 7);
        for (int i = threadIdx.x; i < N; i += NT) mhd_products(buf, N, i, a.coef);
        __syncthreads();
        fft_inplace<NT, -1>(buf, tw, a, 6);
        // two real spectra per pair: X_a = (Z_k + conj Z_{N-k})/2,
        // X_b = (Z_k - conj Z_{N-k})/(2i), scaled as FFTPACK's forward
        const double invN = 1.0 / (double)N;
        DivStep qf(threadIdx.x, NT, L + 1);   // e = p (L+1) + m
        for (int e = threadIdx.x; e < 6 * (L + 1); e += NT, qf.next()) {
            const int p = qf.hi, m = qf.lo;
            const cplx zk = buf[(size_t)p * N + m];
            const cplx zn = buf[(size_t)p * N + (m == 0 ? 0 : N - m)];
            const double xar = 0.5 * (zk.x + zn.x), xai = 0.5 * (zk.y - zn.y);
            const double xbr = 0.5 * (zk.y + zn.y), xbi = -0.5 * (zk.x - zn.x);
            double *rows = spec + (size_t)2 * p * M;
            if (m == 0) {
                rows[L] = xar * invN;
                rows[M + L] = xbr * invN;
            } else {
                rows[L + m] = 2.0 * xar * invN;
                rows[L - m] = -2.0 * xai * invN;
                rows[M + L + m] = 2.0 * xbr * invN;
                rows[M + L - m] = -2.0 * xbi * invN;
            }
        }
        __syncthreads();
        // 19 outputs per order (the reference's slots; 3-8 and 18, which the
        // products do not fill, are zero) or the 12 products only
        // e = mp nout + nd: output offset nd + nout sl mp from the pencil's
        // first word and spectrum index nd M + mp, both stepped by carries
        const int nout = a.ncomp_out;
        const unsigned nslo = (unsigned)nout * (unsigned)sl;
        double *dstp = a.out + (size_t)nout * pencil;
        DivStep qo(threadIdx.x, NT, nout);
        unsigned oo = (unsigned)qo.lo + nslo * (unsigned)qo.hi;
        const unsigned doo = (unsigned)qo.dlo + nslo * (unsigned)qo.dhi;
        int si = qo.lo * M + qo.hi;
        const int dsi = qo.dlo * M + qo.dhi;
        for (int e = threadIdx.x; e < nout * M; e += NT) {
            const int nd = qo.lo;
            double v;
            if (nout == 12) v = spec[si];
            else v = nd < 3 ? spec[si] : (nd >= 9 && nd < 18 ? spec[si - 6 * M] : 0.0);
            dstp[oo] = v;
            oo += doo;
            si += dsi;
            qo.lo += qo.dlo;
            if (qo.lo >= nout) { qo.lo -= nout; oo += nslo - (unsigned)nout; si += 1 - nout * M; }
        }
        continue;
    }

    // ---- inverse side -----------------------------------------------
    if (MODE == kPencilBwd || MODE == kPencilMhd) {
        for (int p0 = 0; p0 < npair_in; p0 += PB) {
            const int np = min(PB, npair_in - p0);
            const int c0 = 2 * p0, nc = min(2 * np, nin - c0);
            __syncthreads();   // buffer free (previous group or pencil written out)
            // the group's spectra, nc contiguous doubles per order
            if (p0 == 0) {
                int tid = threadIdx.x;
                asm volatile("" : "+v"(tid));
#pragma unroll
                for (int q = 0; q < kFftPre; ++q) {
                    const int e = tid + q * NT;
                    if (e < nc * M) bd[(e % nc) * M + e / nc] = pre[q];
                }
                prefetch(pencil_of(it + 1));
            } else {
                for (int e = threadIdx.x; e < nc * M; e += NT) {
                    const int nd = e % nc, mp = e / nc;
                    bd[nd * M + mp] = a.in[c0 + nd + (size_t)nin * (pencil + sl * mp)];
                }
            }
            __syncthreads();
            // Hermitian spectra of z = x_a + i x_b, pair by pair from the last:
            // pair p's sequence covers doubles [2pN, 2pN+2N), which holds its
            // own staged rows and those of later pairs only, so each pair
            // passes through registers once, with one barrier
            for (int p = np - 1; p >= 0; --p) {
                const int ca = 2 * p, cb = ca + 1;
                const bool hb = cb < nc;
                cplx z[kFftMaxZ];
#pragma unroll
                for (int q = 0; q < kFftMaxZ; ++q) {
                    const int kf = threadIdx.x + q * NT;
                    z[q] = {0.0, 0.0};
                    if (kf < N) {
                        if (kf == 0) {
                            z[q].x = bd[ca * M + L];
                            z[q].y = hb ? bd[cb * M + L] : 0.0;
                        } else if (kf <= L) {
                            const int m = kf;
                            const double aa = bd[ca * M + L + m], ba = bd[ca * M + L - m];
                            const double ab = hb ? bd[cb * M + L + m] : 0.0;
                            const double bb = hb ? bd[cb * M + L - m] : 0.0;
                            z[q].x = 0.5 * (aa + bb);
                            z[q].y = 0.5 * (ab - ba);
                        } else if (kf >= N - L) {
                            const int m = N - kf;
                            const double aa = bd[ca * M + L + m], ba = bd[ca * M + L - m];
                            const double ab = hb ? bd[cb * M + L + m] : 0.0;
                            const double bb = hb ? bd[cb * M + L - m] : 0.0;
                            z[q].x = 0.5 * (aa - bb);
                            z[q].y = 0.5 * (ba + ab);
                        }
                    }
                }
                __syncthreads();
#pragma unroll
                for (int q = 0; q < kFftMaxZ; ++q) {
                    const int kf = threadIdx.x + q * NT;
                    if (kf < N) buf[(size_t)p * N + kf] = z[q];
                }
            }
            __syncthreads();
            fft_inplace<NT, +1>(buf, tw, a, np);
            if (MODE == kPencilBwd) {
                for (int e = threadIdx.x; e < np * N; e += NT) {
                    const int p = e / N, i = e % N;
                    const int ca = c0 + 2 * p, cb = ca + 1;
                    const cplx zz = buf[e];
                    a.out[gbase + (size_t)a.nr * a.nth * i + nnod * ca] = zz.x;
                    if (cb < nin) a.out[gbase + (size_t)a.nr * a.nth * i + nnod * cb] = zz.y;
                }
            }
        }
    }
    if (MODE == kPencilBwd) continue;

    // ---- products (MHD, in place per longitude) ---------------------
    const int nout = a.ncomp_out;
    int npair_f;
    if (MODE == kPencilMhd) {
        npair_f = 6;
        for (int i = threadIdx.x; i < N; i += NT) mhd_products(buf, N, i, a.coef);
        __syncthreads();
    } else {
        npair_f = (nin + 1) / 2;
    }

    // ---- forward FFTs -----------------------------------------------------
    const double invN = 1.0 / (double)N;
    for (int p0 = 0; p0 < npair_f; p0 += PB) {
        const int np = min(PB, npair_f - p0);
        if (MODE == kPencilFwd) {
            __syncthreads();   // buffer free
            for (int e = threadIdx.x; e < np * N; e += NT) {
                const int p = e / N, i = e % N;
                const int ca = 2 * (p0 + p), cb = ca + 1;
                const size_t g = gbase + (size_t)a.nr * a.nth * i;
                buf[e] = {a.in[g + nnod * ca], cb < nin ? a.in[g + nnod * cb] : 0.0};
            }
            __syncthreads();
        }
        fft_inplace<NT, -1>(buf, tw, a, np);
        // unpack two real spectra per pair, pair by pair: the spectra of pair
        // p (rows a, b of M doubles) go to the start of its own sequence,
        // doubles [2pN, 2pN + 2M), which no later pair reads
        for (int p = 0; p < np; ++p) {
            double u[kFftMaxU][4];
#pragma unroll
            for (int q = 0; q < kFftMaxU; ++q) {
                const int m = threadIdx.x + q * NT;
                if (m <= L) {
                    const cplx zk = buf[(size_t)p * N + m];
                    const cplx zn = buf[(size_t)p * N + (m == 0 ? 0 : N - m)];
                    // X_a = (Z_k + conj Z_{N-k})/2, X_b = (Z_k - conj Z_{N-k})/(2i)
                    const double xar = 0.5 * (zk.x + zn.x), xai = 0.5 * (zk.y - zn.y);
                    const double xbr = 0.5 * (zk.y + zn.y), xbi = -0.5 * (zk.x - zn.x);
                    if (m == 0) {
                        u[q][0] = xar * invN; u[q][1] = 0.0;
                        u[q][2] = xbr * invN; u[q][3] = 0.0;
                    } else {
                        u[q][0] = 2.0 * xar * invN; u[q][1] = -2.0 * xai * invN;
                        u[q][2] = 2.0 * xbr * invN; u[q][3] = -2.0 * xbi * invN;
                    }
                }
            }
            __syncthreads();
            double *rows = bd + (size_t)2 * p * N;
#pragma unroll
            for (int q = 0; q < kFftMaxU; ++q) {
                const int m = threadIdx.x + q * NT;
                if (m <= L) {
                    rows[L + m] = u[q][0];
                    rows[M + L + m] = u[q][2];
                    if (m > 0) {
                        rows[L - m] = u[q][1];
                        rows[M + L - m] = u[q][3];
                    }
                }
            }
        }
        __syncthreads();
        // write out, nout (MHD) or this group's components contiguous per order;
        // MHD slots the products do not fill (3-8, 18) are zero
        const int oc0 = (MODE == kPencilMhd) ? 0 : 2 * p0;
        const int onc = (MODE == kPencilMhd) ? nout : min(2 * np, nout - oc0);
        for (int e = threadIdx.x; e < onc * M; e += NT) {
            const int nd = e % onc, mp = e / onc;
            int q = nd;   // local component -> pair 2q, row q%2
            if (MODE == kPencilMhd) q = nd < 3 ? nd : (nd >= 9 && nd < 18 ? nd - 6 : -1);
            const double v = (q >= 0) ? bd[(size_t)2 * (q >> 1) * N + (q & 1) * M + mp] : 0.0;
            a.out[oc0 + nd + (size_t)nout * (pencil + sl * mp)] = v;
        }
    }
    }   // pencils
}

// pointwise products on the reference grid layout fld(nnod, 13) -> frc(nnod, 19)
// nout = 19: the reference's f_trns slots (0-2, 9-17; 3-8 and 18 zero);
// nout = 12: the twelve products only (advection, Lorentz, induction, heat flux)
__global__ void k_products_mhd(size_t nnod, const double *fld, double *frc, double cv,
                               double cl, double ci, double ct, int nout)
{
#pragma clang fp contract(off)
    const int o3 = nout == 12 ? 3 : 9;   // first slot after the advection
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
        if (nout == 19)
            for (int c = 3; c < 9; ++c) frc[c * nnod + i] = 0.0;
        frc[(o3 + 0) * nnod + i] = (j2 * b3 - j3 * b2) * cl;
        frc[(o3 + 1) * nnod + i] = (j3 * b1 - j1 * b3) * cl;
        frc[(o3 + 2) * nnod + i] = (j1 * b2 - j2 * b1) * cl;
        frc[(o3 + 3) * nnod + i] = (u2 * b3 - u3 * b2) * ci;
        frc[(o3 + 4) * nnod + i] = (u3 * b1 - u1 * b3) * ci;
        frc[(o3 + 5) * nnod + i] = (u1 * b2 - u2 * b1) * ci;
        frc[(o3 + 6) * nnod + i] = u1 * T * ct;
        frc[(o3 + 7) * nnod + i] = u2 * T * ct;
        frc[(o3 + 8) * nnod + i] = u3 * T * ct;
        if (nout == 19) frc[18 * nnod + i] = 0.0;
    }
}

// ----------------------------------------------------------------- host --

int fft_context_build(FftContext *f, int L, int nth, int nr, int nphi, std::string *err)
{
    if (L < 1 || nth < 1 || nr < 1) { *err = "invalid grid size"; return -1; }
    if (nphi < 2 * L + 2) { *err = "nphi must be at least 2*(L+1)"; return -1; }
    int n = nphi, ns = 0;
    while (n % 4 == 0 && ns < FFT_MAX_STAGES) { f->radix[ns++] = 4; n /= 4; }
    if (n % 2 == 0 && ns < FFT_MAX_STAGES) { f->radix[ns++] = 2; n /= 2; }
    while (n % 3 == 0 && ns < FFT_MAX_STAGES) { f->radix[ns++] = 3; n /= 3; }
    if (n != 1 || ns > FFT_MAX_STAGES) {
        *err = "nphi must be of the form 2^a 3^b";
        return -1;
    }
    f->L = L;
    f->nth = nth;
    f->nr = nr;
    f->N = nphi;
    f->nstage = ns;
    {
        int dev = 0;
        hipDeviceProp_t prop;
        HIP_TRY(hipGetDevice(&dev));
        HIP_TRY(hipGetDeviceProperties(&prop, dev));
        f->n_cu = prop.multiProcessorCount;
    }
    // per-stage twiddle rows: stage st (radix R, Ns = product of the earlier
    // radices) holds W_N^(r kk N/(Ns R)) = exp(-2 pi i r kk / (Ns R)) at
    // (r-1) Ns + kk, so consecutive lanes (consecutive kk) read consecutive
    // words; N-1 entries in all (the last slot stays zero)
    std::vector<double> tw(2 * (size_t)nphi, 0.0);
    const double pi = 3.14159265358979323846;
    {
        int Ns = 1, off = 0;
        for (int st = 0; st < ns; ++st) {
            const int R = f->radix[st];
            for (int r = 1; r < R; ++r)
                for (int kk = 0; kk < Ns; ++kk) {
                    // the angle of table entry i = r kk N/(Ns R) of exp(-2 pi i / N)
                    const long long i = (long long)r * kk * (nphi / (Ns * R));
                    long double ang = -2.0L * (long double)pi * (long double)i / (long double)nphi;
                    const size_t e = (size_t)off + (size_t)(r - 1) * Ns + kk;
                    tw[2 * e] = (double)std::cos(ang);
                    tw[2 * e + 1] = (double)std::sin(ang);
                }
            off += (R - 1) * Ns;
            Ns *= R;
        }
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

// LDS of a pencil work-group with pb component pairs per group: twiddles,
// pb sequences of N complex values; the fused MHD pencil adds its 13 rows of
// M doubles of spectra
static size_t pencil_lds(const FftContext *f, int pb, int mode)
{
    const size_t b = (size_t)(pb + 1) * f->N * 16;
    return mode == kPencilMhd ? b + (size_t)13 * (2 * f->L + 1) * 8 : b;
}

// pairs per group: everything the register staging, the prefetch of the
// first group and the LDS allow (the MHD mode needs its 7 input pairs in one
// group)
static int pencil_pairs(const FftContext *f, int mode, int nin, int nout, int nt)
{
    const int M = 2 * f->L + 1, N = f->N;
    const int need = (mode == kPencilMhd) ? 7 : (std::max(nin, nout) + 1) / 2;
    if (N > kFftMaxZ * nt || f->L + 1 > kFftMaxU * nt) return 0;
    auto fits = [&](int p) {
        const int np = std::min(p, need);
        for (int st = 0; st < f->nstage; ++st)
            if (np * (N / f->radix[st]) > fft_mb(f->radix[st]) * nt) return false;
        if (mode != kPencilFwd && std::min(2 * p, nin) * M > kFftPre * nt) return false;
        return pencil_lds(f, p, mode) <= 160 * 1024;
    };
    if (mode == kPencilMhd) return fits(need) ? need : 0;
    int pb = need;
    while (pb > 1 && !fits(pb)) --pb;
    return fits(pb) ? pb : 0;
}

template <int MODE>
static int launch_nt(const FftContext *f, const FftArgs &a, size_t lds, int nt, hipStream_t s,
                     std::string *err)
{
    // persistent: one round of work-groups, as many per CU as LDS allows
    const int per_cu = std::max(1, std::min(8, (int)((160 * 1024) / std::max<size_t>(lds, 1))));
    const int grid = std::min(f->nr * f->nth, std::max(1, f->n_cu) * per_cu);
    const void *fn = nt == 512 ? reinterpret_cast<const void *>(&k_pencil<MODE, 512>)
                               : reinterpret_cast<const void *>(&k_pencil<MODE, 256>);
    if (lds > 64 * 1024)
        HIP_TRY(hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    if (nt == 512)
        hipLaunchKernelGGL((k_pencil<MODE, 512>), dim3(grid), dim3(512), lds, s, a);
    else
        hipLaunchKernelGGL((k_pencil<MODE, 256>), dim3(grid), dim3(256), lds, s, a);
    HIP_TRY(hipGetLastError());
    return 0;
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
    // 256 threads while two or more work-groups fit on a CU, else 512
    int nt = 256, pb = pencil_pairs(f, mode, nin, nout, 256);
    if (!pb || pencil_lds(f, pb, mode) > 80 * 1024) {
        const int pb2 = pencil_pairs(f, mode, nin, nout, 512);
        if (pb2) { nt = 512; pb = pb2; }
    }
    if (!pb) {
        *err = mode == kPencilMhd ? "pencil too long for the fused MHD kernel"
                                  : "pencil too long for LDS";
        return -1;
    }
    // the kernels index an rtm array with 32-bit offsets
    const size_t rtm_words = (size_t)f->nth * f->nr * (2 * f->L + 1);
    if ((mode != kPencilFwd && (size_t)nin * rtm_words > 0xFFFFFFFFull) ||
        (mode == kPencilMhd && (size_t)nout * rtm_words > 0xFFFFFFFFull)) {
        *err = "rtm array of 2^32 words or more";
        return -1;
    }
    a.pairs_per_batch = pb;
    const size_t lds = pencil_lds(f, pb, mode);
    if (mode == kPencilBwd) return launch_nt<kPencilBwd>(f, a, lds, nt, s, err);
    if (mode == kPencilFwd) return launch_nt<kPencilFwd>(f, a, lds, nt, s, err);
    return launch_nt<kPencilMhd>(f, a, lds, nt, s, err);
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
    return pencil_pairs(f, kPencilMhd, 13, 19, 256) > 0 || pencil_pairs(f, kPencilMhd, 13, 19, 512) > 0;
}

int fft_mhd_fused_dev(const FftContext *f, const double *d_bwd_rtm, double *d_fwd_rtm,
                      const double *coef, hipStream_t s, std::string *err, int nout)
{
    if (nout != 19 && nout != 12) { *err = "fused pencil: 19 (reference slots) or 12 outputs"; return -1; }
    return launch_pencil(f, kPencilMhd, 13, nout, d_bwd_rtm, d_fwd_rtm, coef, s, err);
}

int products_mhd_dev(size_t nnod, const double *d_fld, double *d_frc, const double *coef,
                     hipStream_t s, std::string *err, int nout)
{
    if (nout != 19 && nout != 12) { *err = "products: 19 (reference slots) or 12 outputs"; return -1; }
    const int blocks = (int)std::min<size_t>((nnod + 255) / 256, 65536);
    hipLaunchKernelGGL(k_products_mhd, dim3(blocks), dim3(256), 0, s, nnod, d_fld, d_frc, coef[0],
                       coef[1], coef[2], coef[3], nout);
    HIP_TRY(hipGetLastError());
    return 0;
}

}  // namespace calypso
