// radial_internal.hpp -- radial-direction kernels (gfx950).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>

namespace calypso {

int ludcmp_dev(int nb, int n, int mcomp, const double *d_band_a, double *d_band_lu, int *d_ipiv,
               int *d_ierr, hipStream_t s, std::string *err);
// mode-fastest copies of the factors for lubksb_t_dev (bitwise the same solve)
int lu_transpose_dev(int nb, int n, int mcomp, const double *d_band_lu, const int *d_ipiv,
                     double *d_band_t, int *d_ipiv_t, hipStream_t s, std::string *err);
int lubksb_t_dev(int nb, int n, int mcomp, const double *d_band_t, const int *d_ipiv_t,
                 double *d_x, hipStream_t s, std::string *err);
// up to 8 independent systems of one band width in one launch
int lubksb_t_batch_dev(int nb, int n, int mcomp, int count, const double *const *d_band_t,
                       const int *const *d_ipiv_t, double *const *d_x, hipStream_t s,
                       std::string *err);
// factor on the device, check, and transpose for lubksb_t_dev (allocates the
// two outputs, which the caller owns)
int factor_band_for_solve(int nb, int n, int mcomp, const double *h_band, double **d_lu_t,
                          int **d_piv_t, hipStream_t s, std::string *err);
int lubksb_dev(int nb, int n, int mcomp, const double *d_band_lu, const int *d_ipiv, double *d_x,
               hipStream_t s, std::string *err);

int explicit_update_dev(int kind, int scheme, size_t n, double dt, double coef_exp, double *d_x,
                        const double *d_diff, const double *d_force, double *d_pre,
                        hipStream_t s, std::string *err);
int mean_square_dev(int kind, int nr, int J, int lmax, const int *d_l_of_j,
                    const double *d_radius, const double *h_radius, int kr_st, int kr_ed,
                    const double *d_f, double *d_vol, double *d_out, hipStream_t s,
                    std::string *err);
// the same on a file-scope device scratch array: no device allocation
int mean_square_noalloc_dev(int kind, int nr, int J, const int *h_l_of_j, const double *h_radius,
                            int kr_st, int kr_ed, const double *d_f, double *h_out,
                            hipStream_t s, std::string *err);

}  // namespace calypso
