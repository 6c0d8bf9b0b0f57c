/*
 * calypso_oracle.h -- CPU FP64 restatement of Calypso's spherical-harmonic
 * time-step arithmetic, used ONLY as a test checker (tests/, the smoke check
 * and the CPU-baseline leg of bench.py).  Nothing in the product links it.
 *
 * Index conventions (0-based restatement of the reference's 1-based ones):
 *   L      truncation degree (l_truncation)
 *   J      (L+1)^2 spectral modes, rlm order: m = -L..L ascending, then
 *          l = |m|..L ascending  (set_sph_tranform_ordering.f90:69-206 with
 *          one domain)
 *   M      2L+1 zonal orders; rtm slot mp = m + L, its -m partner M-1-mp
 *   nth    Gauss colatitude points; nr radial points
 *   sp_rlm[nd + ncomp*(j + J*k)]            (set_local_sphere_param.f90:110-126)
 *   vr_rtm[nd + ncomp*(t + nth*k + nth*nr*mp)]  (set_local_sphere_param.f90:86-105)
 *   P[t + nth*j]   (P_rtm(l_rtm, j_rlm), m_schmidt_poly_on_rtm.f90)
 */
#ifndef CALYPSO_ORACLE_H
#define CALYPSO_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

int  orc_num_modes(int L);
void orc_rlm_index(int L, int *jlm, int *lstack);
void orc_gauss_points(int n, double *point, double *weight, double *colat);
void orc_schmidt_legendres_m(int ltr, int m, double theta, double *p,
                             double *dp);
void orc_sph_normalizations(int l, int m, double *g17);
void orc_g_sph_rlm(int L, double *g);
void orc_legendre_tables(int L, int nth, const double *colat, double *P,
                         double *dP);
void orc_leg_bwd(int L, int nth, int nr, int ncomp, int nvec, int nscl,
                 const double *radius, const double *colat, const double *P,
                 const double *dP, const double *sp_in, double *vr_out);
void orc_leg_fwd(int L, int nth, int nr, int ncomp, int nvec, int nscl,
                 const double *radius, const double *colat,
                 const double *weight, const double *P, const double *dP,
                 const double *vr_in, double *sp_out);

void orc_fft_bwd(int L, int nth, int nr, int nphi, int ncomp,
                 const double *vr_rtm, double *fld_rtp);
void orc_fft_fwd(int L, int nth, int nr, int nphi, int ncomp,
                 const double *fld_rtp, double *vr_rtm);
void orc_products_mhd(size_t nnod, const double *fld, double *frc,
                      double coef_velo, double coef_lor, double coef_induct,
                      double coef_temp);
void orc_ludcmp_3band(int n, int mcomp, const double *band_a, double *band_lu,
                      int *i_pivot, int *ierr);
void orc_ludcmp_5band(int n, int mcomp, const double *band_a, double *band_lu,
                      int *indx);
void orc_lubksb_mul(int nb, int n, int mcomp, const double *band_lu,
                    const int *i_pivot, double *x);
void orc_explicit_update(int kind, int scheme, size_t n, double dt, double coef_exp,
                         double *x, const double *diff, const double *force, double *pre);
void orc_sph_mean_square(int kind, int nr, int J, const int *l_of_j, int lmax,
                         const double *radius, int kr_st, int kr_ed, const double *f,
                         double *out);
uint64_t orc_splitmix64(uint64_t state);
double orc_box_muller(double u1, double u2);
double orc_seeded_gauss(uint64_t seed, uint64_t field, uint64_t k, uint64_t j);
void orc_seeded_spectrum(uint64_t seed, int field, int L, int nr, const double *radius,
                         double amplitude, double power, int envelope, int l_min, double *out);

#ifdef __cplusplus
}
#endif
#endif
