// fft_internal.hpp -- longitude FFT context (gfx950).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>

namespace calypso {

constexpr int FFT_MAX_STAGES = 16;

struct FftContext {
    int L = 0, nth = 0, nr = 0, N = 0;
    int nstage = 0;
    int radix[FFT_MAX_STAGES] = {0};
    double *d_tw = nullptr;   // per-stage twiddle rows, N complex slots
    int n_cu = 256;           // compute units (persistent pencil grid)
};

int  fft_context_build(FftContext *f, int L, int nth, int nr, int nphi, std::string *err);
void fft_context_free(FftContext *f);
int  fft_backward_dev(const FftContext *f, int ncomp, const double *d_vr_rtm,
                      double *d_fld_rtp, hipStream_t s, std::string *err);
int  fft_forward_dev(const FftContext *f, int ncomp, const double *d_frc_rtp,
                     double *d_vr_rtm, hipStream_t s, std::string *err);
bool fft_mhd_fused_fits(const FftContext *f);   // 13 -> 19 pencil fits in LDS
// nout: 19 = the reference's forward slots (zeros where the products leave
// them), 12 = the products only (advection, Lorentz, induction, heat flux)
int  fft_mhd_fused_dev(const FftContext *f, const double *d_bwd_rtm, double *d_fwd_rtm,
                       const double *coef, hipStream_t s, std::string *err, int nout = 19);
int  products_mhd_dev(size_t nnod, const double *d_fld, double *d_frc, const double *coef,
                      hipStream_t s, std::string *err, int nout = 19);

}  // namespace calypso
