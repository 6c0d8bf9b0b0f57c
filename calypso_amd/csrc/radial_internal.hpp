// radial_internal.hpp -- radial-direction kernels (gfx950).  Not part of the C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <string>

namespace calypso {

int ludcmp_dev(int nb, int n, int mcomp, const double *d_band_a, double *d_band_lu, int *d_ipiv,
               int *d_ierr, hipStream_t s, std::string *err);
int lubksb_dev(int nb, int n, int mcomp, const double *d_band_lu, const int *d_ipiv, double *d_x,
               hipStream_t s, std::string *err);

}  // namespace calypso
