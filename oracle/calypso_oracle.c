/*
 * calypso_oracle.c -- TEST INFRASTRUCTURE, NOT PRODUCT CODE.
 *
 * A plain-C FP64 restatement of the reference arithmetic on the spherical-
 * harmonic transform path.  Each function cites the reference routine it
 * follows.  Only tests/, the smoke check and bench.py's CPU-baseline leg load
 * it, and only as the checker; the HIP library never calls into it.
 *
 * Parity pins (tests/test_oracle_*.py):
 *   orc_gauss_points        == compiled reference construct_gauss_coefs,
 *                              numpy leggauss, sum(w)=2, exact quadrature
 *   orc_schmidt_legendres_m == compiled reference schmidt_legendres_m,
 *                              scipy lpmv closed form, P_l^l closed form
 *   orc_sph_normalizations  closed forms of spherical_harmonics.f90:103-139
 *   orc_leg_bwd / fwd       == compiled reference legendre_*_trans_*_org,
 *                              fwd(bwd(x)) = x on band-limited data,
 *                              brute-force Y_l^m synthesis on tiny L
 */
#include "calypso_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

int orc_num_modes(int L) { return (L + 1) * (L + 1); }

/* Single-domain rlm ordering: set_zonal_wavenum_4_legendre with
 * ndomain_m = 1 lists m = -L..L ascending, and set_merged_index_4_sph_trans
 * (set_sph_tranform_ordering.f90:163-206) appends l = |m|..L for each m with
 * global index l(l+1)+m.  lstack[mp+1]-lstack[mp] = L-|m|+1
 * (set_params_sph_trans.f90:48-55). jlm[3j+0..2] = (l(l+1)+m, l, m). */
void orc_rlm_index(int L, int *jlm, int *lstack)
{
    int j = 0;
    lstack[0] = 0;
    for (int mp = 0; mp < 2 * L + 1; ++mp) {
        int m = mp - L;
        for (int l = abs(m); l <= L; ++l) {
            jlm[3 * j + 0] = l * (l + 1) + m;
            jlm[3 * j + 1] = l;
            jlm[3 * j + 2] = m;
            ++j;
        }
        lstack[mp + 1] = j;
    }
}

/* Gauss-Legendre points by Newton iteration on [0,1], mapped to [-1,1]:
 * construct_gauss_coefs and set_gauss_colatitude, m_gauss_points.f90:117-195.
 * The reference loops n = 2..n_point and overwrites every slot on the last
 * pass, so only n = n_point matters. */
void orc_gauss_points(int n, double *point, double *weight, double *colat)
{
    const double pi = 4.0 * atan(1.0);
    if (n == 1) {
        point[0] = 0.5;
        weight[0] = 1.0;
    } else {
        const double x1 = 0.0, x2 = 1.0;
        const double xm = 0.5 * (x2 + x1), xl = 0.5 * (x2 - x1);
        const int half_n = (n + 1 - ((n + 1) % 2)) / 2;
        for (int i = 1; i <= half_n; ++i) {
            double z = cos(pi * ((double)i - 0.25) / ((double)n + 0.5));
            double z1, pp, p1, p2, p3;
            for (;;) {
                p1 = 1.0;
                p2 = 0.0;
                for (int jj = 1; jj <= n; ++jj) {
                    p3 = p2;
                    p2 = p1;
                    p1 = ((2.0 * (double)jj - 1.0) * z * p2
                          - ((double)jj - 1.0) * p3) / (double)jj;
                }
                pp = (double)n * (z * p1 - p2) / (z * z - 1.0);
                z1 = z;
                z = z1 - p1 / pp;
                if (fabs(z - z1) < 3.0e-12) break;
            }
            point[i - 1] = xm - xl * z;
            point[n - i] = xm + xl * z;
            weight[i - 1] = 2.0 * xl / ((1.0 - z * z) * pp * pp);
            weight[n - i] = weight[i - 1];
        }
    }
    for (int i = 0; i < n; ++i) {
        point[i] = point[i] * 2.0 - 1.0;
        weight[i] = weight[i] * 2.0;
        colat[i] = acos(point[i]);
    }
}

/* Schmidt semi-normalised P_l^m(cos theta), no Condon-Shortley phase:
 * schmidt_legendre_m, schmidt_fix_m.f90:89-141.  p has ltr+1 entries; entries
 * below l = m are left untouched, as in the reference. */
static void schmidt_legendre_m(int ltr, int m, double theta, double *p,
                               double *df)
{
    if (m < 0) {
        for (int l = 0; l <= ltr; ++l) p[l] = 0.0;
    } else if (m == 0) {
        p[0] = 1.0;
        if (ltr >= 1) p[1] = cos(theta);
        for (int l = 2; l <= ltr; ++l)
            p[l] = p[l - 1] * (double)(2 * l - 1) / (double)l * cos(theta)
                 - p[l - 2] * (double)(l - 1) / (double)l;
    } else {
        df[m] = 1.0;
        for (int k = 1; k <= m; ++k)
            df[m] = df[m] * (double)(2 * k - 1) / (double)(2 * k);
        df[m + 1] = sqrt(2.0 * df[m] * (double)(2 * m + 1)) * cos(theta);
        df[m] = sqrt(2.0 * df[m]);
        if (m < ltr - 1) {
            for (int l = m + 2; l <= ltr; ++l)
                df[l] = (cos(theta) * (double)(2 * l - 1) * df[l - 1]
                         - sqrt((double)((l - 1) * (l - 1) - m * m)) * df[l - 2])
                      / sqrt((double)(l * l - m * m));
        }
        for (int l = m; l <= ltr; ++l) {
            p[l] = df[l];
            for (int m1 = 1; m1 <= m; ++m1) p[l] = p[l] * sin(theta);
        }
    }
}

/* dP_l^m/dtheta from P_l^{m-1} and P_l^{m+1}: diff_schmidt_legendre_m,
 * schmidt_fix_m.f90:145-175. */
static void diff_schmidt_legendre_m(int ltr, int m, const double *pmn1,
                                    const double *pmp1, double *dp)
{
    if (m == 0) {
        dp[0] = 0.0;
        for (int l = 1; l <= ltr; ++l)
            dp[l] = -sqrt((double)(l * (l + 1) / 2)) * pmp1[l];
    } else if (m == 1) {
        dp[1] = pmn1[1];
        for (int l = 2; l <= ltr; ++l)
            dp[l] = 0.5 * (sqrt((double)(2 * l * (l + 1))) * pmn1[l]
                           - sqrt((double)((l - 1) * (l + 2))) * pmp1[l]);
    } else {
        dp[m] = 0.5 * sqrt((double)(2 * m)) * pmn1[m];
        for (int l = m + 1; l <= ltr; ++l)
            dp[l] = 0.5 * (sqrt((double)((l + m) * (l - m + 1))) * pmn1[l]
                           - sqrt((double)((l - m) * (l + m + 1))) * pmp1[l]);
    }
}

/* schmidt_legendres_m, schmidt_fix_m.f90:62-85. */
void orc_schmidt_legendres_m(int ltr, int m, double theta, double *p,
                             double *dp)
{
    double *pmn1 = (double *)calloc((size_t)ltr + 1, sizeof(double));
    double *pmp1 = (double *)calloc((size_t)ltr + 1, sizeof(double));
    double *df = (double *)calloc((size_t)ltr + 3, sizeof(double));
    schmidt_legendre_m(ltr, m - 1, theta, pmn1, df);
    schmidt_legendre_m(ltr, m, theta, p, df);
    schmidt_legendre_m(ltr, m + 1, theta, pmp1, df);
    diff_schmidt_legendre_m(ltr, m, pmn1, pmp1, dp);
    free(pmn1);
    free(pmp1);
    free(df);
}

/* Normalisation constants g(1:17): sph_normalizations,
 * spherical_harmonics.f90:92-141 (stored 0-based: g17[i-1] = g(i)). */
void orc_sph_normalizations(int l, int m, double *g)
{
    const double pi = 4.0 * atan(1.0);
    g[0] = (double)l;
    g[1] = (double)m;
    g[2] = (double)(l * (l + 1));
    g[3] = (double)((l - 1) * l);
    g[4] = (double)((l + 1) * (l + 2));
    g[7] = 4.0 * pi * (double)(l * (l + 1)) / (double)(2 * l + 1);
    g[9] = 4.0 * pi / (double)(2 * l + 1);
    g[10] = 1.0 / (double)(2 * l + 1);
    g[11] = (double)(l * (l + 1)) / (double)(2 * l + 1);
    g[13] = 1.0;
    g[14] = 1.0;
    g[15] = (double)(2 * l + 1) / 4.0;
    if (l == 0 && m == 0) {
        g[5] = 1.0 / 2.0;
        g[6] = 1.0;
        g[8] = 1.0 / (4.0 * pi);
        g[12] = 1.0;
        g[16] = 0.0;
    } else {
        g[8] = (double)(2 * l + 1) / (4.0 * pi * (double)(l * (l + 1)));
        g[12] = 1.0 / (double)(l * (l + 1));
        g[16] = (double)(2 * l + 1) / (double)(4 * l * (l + 1));
        if (m == 0) {
            g[5] = (double)(2 * l + 1) / 2.0;
            g[6] = (double)(2 * l + 1) / (double)(2 * l * (l + 1));
        } else {
            g[5] = (double)(2 * l + 1) / 4.0;
            g[6] = (double)(2 * l + 1) / (double)(4 * l * (l + 1));
        }
    }
}

/* g_sph_rlm(j,1:17) with the l=m=0 override g(3) = 1/2:
 * copy_sph_normalization_2_rlm, schmidt_poly_on_rtm_grid.f90:92-116.
 * Layout g[17*j + i]. */
void orc_g_sph_rlm(int L, double *g)
{
    int J = orc_num_modes(L);
    int *jlm = (int *)malloc(sizeof(int) * 3 * (size_t)J);
    int *lstack = (int *)malloc(sizeof(int) * (size_t)(2 * L + 2));
    orc_rlm_index(L, jlm, lstack);
    for (int j = 0; j < J; ++j) {
        int l = jlm[3 * j + 1], m = jlm[3 * j + 2];
        orc_sph_normalizations(l, m, g + 17 * (size_t)j);
        if (l == 0 && m == 0) g[17 * (size_t)j + 2] = 0.5;
    }
    free(jlm);
    free(lstack);
}

/* P_rtm(l_rtm, j) and dPdt_rtm: set_lagender_4_rlm,
 * schmidt_poly_on_rtm_grid.f90:145-188 (the recursion uses |m|). */
void orc_legendre_tables(int L, int nth, const double *colat, double *P,
                         double *dP)
{
    /* colatitudes are independent: one per thread */
#pragma omp parallel
    {
        double *p = (double *)malloc(sizeof(double) * (size_t)(L + 1));
        double *dp = (double *)malloc(sizeof(double) * (size_t)(L + 1));
#pragma omp for schedule(dynamic)
        for (int t = 0; t < nth; ++t) {
            int j = 0;
            for (int mp = 0; mp < 2 * L + 1; ++mp) {
                int mm = abs(mp - L);
                orc_schmidt_legendres_m(L, mm, colat[t], p, dp);
                for (int l = mm; l <= L; ++l, ++j) {
                    P[t + (size_t)nth * j] = p[l];
                    dP[t + (size_t)nth * j] = dp[l];
                }
            }
        }
        free(p);
        free(dp);
    }
}

/* Backward transform, spectrum -> (r, theta, m) grid, following
 * legendre_b_trans_vector_org and legendre_b_trans_scalar_org
 * (legendre_bwd_trans_org.f90:45-197).  vr_out is fully overwritten
 * (the reference clears the work array first, legendre_transform_org.f90:63).
 * Vector components are (pol, dpol/dr, tor) -> (r, theta, phi). */
void orc_leg_bwd(int L, int nth, int nr, int ncomp, int nvec, int nscl,
                 const double *radius, const double *colat, const double *P,
                 const double *dP, const double *sp_in, double *vr_out)
{
    const int J = orc_num_modes(L), M = 2 * L + 1;
    int *jlm = (int *)malloc(sizeof(int) * 3 * (size_t)J);
    int *lstack = (int *)malloc(sizeof(int) * (size_t)(M + 1));
    double *g = (double *)malloc(sizeof(double) * 17 * (size_t)J);
    double *asin_t = (double *)malloc(sizeof(double) * (size_t)nth);
    double *sp = (double *)malloc(sizeof(double) * (size_t)ncomp * J * nr);
    orc_rlm_index(L, jlm, lstack);
    orc_g_sph_rlm(L, g);
    for (int t = 0; t < nth; ++t) asin_t[t] = 1.0 / sin(colat[t]);
    memcpy(sp, sp_in, sizeof(double) * (size_t)ncomp * J * nr);
    memset(vr_out, 0, sizeof(double) * (size_t)ncomp * nth * nr * M);

    for (int k = 0; k < nr; ++k) {
        double a1r = 1.0 / radius[k];
        double a2r = a1r * a1r;
        for (int j = 0; j < J; ++j)
            for (int nd = 0; nd < nvec; ++nd) {
                size_t i = 3 * (size_t)nd + (size_t)ncomp * (j + (size_t)J * k);
                sp[i + 0] = sp[i + 0] * a2r;
                sp[i + 1] = sp[i + 1] * a1r;
                sp[i + 2] = sp[i + 2] * a1r;
            }
    }
    /* radial levels are independent (every write is indexed by k): one per
     * thread, the order over mp within a level as above */
#pragma omp parallel for schedule(dynamic)
    for (int k = 0; k < nr; ++k)
    for (int mp = 0; mp < M; ++mp) {
        int mn = M - 1 - mp;
            for (int t = 0; t < nth; ++t) {
                for (int nd = 0; nd < nvec; ++nd) {
                    double v1 = 0.0, v2 = 0.0, v3 = 0.0, w2 = 0.0, w3 = 0.0;
                    for (int j = lstack[mp]; j < lstack[mp + 1]; ++j) {
                        size_t i = 3 * (size_t)nd + (size_t)ncomp * (j + (size_t)J * k);
                        double pg3 = P[t + (size_t)nth * j] * g[17 * (size_t)j + 2];
                        double dpt = dP[t + (size_t)nth * j];
                        double pgv = -P[t + (size_t)nth * j]
                                   * (double)jlm[3 * j + 2] * asin_t[t];
                        v1 = v1 + sp[i + 0] * pg3;
                        v2 = v2 + sp[i + 1] * dpt;
                        v3 = v3 - sp[i + 2] * dpt;
                        w2 = w2 + sp[i + 2] * pgv;
                        w3 = w3 + sp[i + 1] * pgv;
                    }
                    size_t ip = 3 * (size_t)nd + (size_t)ncomp
                              * (t + (size_t)nth * k + (size_t)nth * nr * mp);
                    size_t in = 3 * (size_t)nd + (size_t)ncomp
                              * (t + (size_t)nth * k + (size_t)nth * nr * mn);
                    vr_out[ip + 0] += v1;
                    vr_out[ip + 1] += v2;
                    vr_out[ip + 2] += v3;
                    vr_out[in + 1] += w2;
                    vr_out[in + 2] += w3;
                }
                for (int nd = 0; nd < nscl; ++nd) {
                    double v1 = 0.0;
                    for (int j = lstack[mp]; j < lstack[mp + 1]; ++j) {
                        size_t i = (size_t)nd + 3 * (size_t)nvec
                                 + (size_t)ncomp * (j + (size_t)J * k);
                        v1 = v1 + sp[i] * P[t + (size_t)nth * j];
                    }
                    size_t ip = (size_t)nd + 3 * (size_t)nvec + (size_t)ncomp
                              * (t + (size_t)nth * k + (size_t)nth * nr * mp);
                    vr_out[ip] = v1;
                }
            }
    }
    free(jlm);
    free(lstack);
    free(g);
    free(asin_t);
    free(sp);
}

/* Forward transform, (r, theta, m) grid -> spectrum, following
 * legendre_f_trans_vector_org and legendre_f_trans_scalar_org
 * (legendre_fwd_trans_org.f90:45-160).  sp_out is fully overwritten
 * (cleared first, legendre_transform_org.f90:86). */
void orc_leg_fwd(int L, int nth, int nr, int ncomp, int nvec, int nscl,
                 const double *radius, const double *colat,
                 const double *weight, const double *P, const double *dP,
                 const double *vr_in, double *sp_out)
{
    const int J = orc_num_modes(L), M = 2 * L + 1;
    int *jlm = (int *)malloc(sizeof(int) * 3 * (size_t)J);
    int *lstack = (int *)malloc(sizeof(int) * (size_t)(M + 1));
    int *mdx_p = (int *)malloc(sizeof(int) * (size_t)J);
    int *mdx_n = (int *)malloc(sizeof(int) * (size_t)J);
    double *g = (double *)malloc(sizeof(double) * 17 * (size_t)J);
    double *asin_t = (double *)malloc(sizeof(double) * (size_t)nth);
    orc_rlm_index(L, jlm, lstack);
    orc_g_sph_rlm(L, g);
    for (int t = 0; t < nth; ++t) asin_t[t] = 1.0 / sin(colat[t]);
    /* mdx_p_rlm_rtm / mdx_n_rlm_rtm: set_params_sph_trans.f90:57-68 */
    for (int j = 0; j < J; ++j) {
        mdx_p[j] = jlm[3 * j + 2] + L;
        mdx_n[j] = -jlm[3 * j + 2] + L;
    }
    memset(sp_out, 0, sizeof(double) * (size_t)ncomp * J * nr);

#pragma omp parallel for schedule(dynamic)
    for (int k = 0; k < nr; ++k) {
        double *pvw = (double *)malloc(sizeof(double) * (size_t)nth);
        double *dpvw = (double *)malloc(sizeof(double) * (size_t)nth);
        double *pgvw = (double *)malloc(sizeof(double) * (size_t)nth);
        double r2 = radius[k] * radius[k];
        for (int j = 0; j < J; ++j) {
            const double g7 = g[17 * (size_t)j + 6], g6 = g[17 * (size_t)j + 5];
            for (int t = 0; t < nth; ++t) {
                double p = P[t + (size_t)nth * j];
                pvw[t] = p * g7 * weight[t];
                dpvw[t] = dP[t + (size_t)nth * j] * g7 * weight[t];
                pgvw[t] = p * (double)jlm[3 * j + 2] * asin_t[t] * g7 * weight[t];
            }
            for (int nd = 0; nd < nvec; ++nd) {
                double s1 = 0.0, s2 = 0.0, s3 = 0.0;
                for (int t = 0; t < nth; ++t) {
                    size_t ip = 3 * (size_t)nd + (size_t)ncomp
                              * (t + (size_t)nth * k + (size_t)nth * nr * mdx_p[j]);
                    size_t in = 3 * (size_t)nd + (size_t)ncomp
                              * (t + (size_t)nth * k + (size_t)nth * nr * mdx_n[j]);
                    s1 = s1 + vr_in[ip + 0] * pvw[t];
                    s2 = s2 + (vr_in[ip + 1] * dpvw[t] - vr_in[in + 2] * pgvw[t]);
                    s3 = s3 - (vr_in[in + 1] * pgvw[t] + vr_in[ip + 2] * dpvw[t]);
                }
                size_t i = 3 * (size_t)nd + (size_t)ncomp * (j + (size_t)J * k);
                sp_out[i + 0] = sp_out[i + 0] + s1 * r2;
                sp_out[i + 1] = sp_out[i + 1] + s2 * radius[k];
                sp_out[i + 2] = sp_out[i + 2] + s3 * radius[k];
            }
            for (int t = 0; t < nth; ++t)
                pvw[t] = P[t + (size_t)nth * j] * g6 * weight[t];
            for (int nd = 0; nd < nscl; ++nd) {
                double s1 = 0.0;
                for (int t = 0; t < nth; ++t) {
                    size_t ip = (size_t)nd + 3 * (size_t)nvec + (size_t)ncomp
                              * (t + (size_t)nth * k + (size_t)nth * nr * mdx_p[j]);
                    s1 = s1 + vr_in[ip] * pvw[t];
                }
                size_t i = (size_t)nd + 3 * (size_t)nvec
                         + (size_t)ncomp * (j + (size_t)J * k);
                sp_out[i] = sp_out[i] + s1;
            }
        }
        free(pvw);
        free(dpvw);
        free(pgvw);
    }
    free(jlm);
    free(lstack);
    free(mdx_p);
    free(mdx_n);
    free(g);
    free(asin_t);
}

/* Longitude transforms with FFTPACK5 RFFTMF/RFFTMB semantics as used by
 * sph_RFFTMF_to_send_t / sph_RFFTMB_from_recv_t (t_sph_FFTPACK5.f90:183-331):
 *   forward  a0 = (1/N) sum x,  a_m = (2/N) sum x cos(m phi),
 *            b_m = (2/N) sum x sin(m phi),  phi_i = 2 pi i / N
 *   backward x = a0 + sum_{m=1..L} (a_m cos(m phi) + b_m sin(m phi))
 * and the exchange ordering of set_wavenumber_4_ispack_fft /
 * set_zonal_wavenum_4_legendre (set_sph_tranform_ordering.f90:43-159):
 * a_m goes to rtm slot L+m, b_m to slot L-m, a0 to slot L; m > L dropped.
 * Grid layout fld_rtp[(k + nr*t + nr*nth*i) + nnod_rtp*nd]
 * (set_local_sphere_param.f90:66-81), spectra in the rtm layout.
 * Evaluated as direct FP64 sums (the DFT definition FFTPACK computes). */
void orc_fft_bwd(int L, int nth, int nr, int nphi, int ncomp,
                 const double *vr_rtm, double *fld_rtp)
{
    const double pi = 4.0 * atan(1.0);
    const size_t nnod = (size_t)nr * nth * nphi;
    const size_t sl = (size_t)nth * nr;
    /* cos(m phi_i), sin(m phi_i) tabulated once: same values, same sums */
    double *ct = (double *)malloc(sizeof(double) * (size_t)(L + 1) * nphi);
    double *st = (double *)malloc(sizeof(double) * (size_t)(L + 1) * nphi);
    for (int i = 0; i < nphi; ++i) {
        double phi = 2.0 * pi * (double)i / (double)nphi;
        for (int m = 0; m <= L; ++m) {
            ct[(size_t)m * nphi + i] = cos((double)m * phi);
            st[(size_t)m * nphi + i] = sin((double)m * phi);
        }
    }
    /* one pencil (nd, t, k) per iteration; each sum keeps its order */
#pragma omp parallel
    {
        double *a = (double *)malloc(sizeof(double) * (size_t)(2 * L + 1));
#pragma omp for collapse(3) schedule(static)
        for (int nd = 0; nd < ncomp; ++nd)
            for (int t = 0; t < nth; ++t)
                for (int k = 0; k < nr; ++k) {
                    const size_t base = (size_t)t + (size_t)nth * k;
                    for (int mp = 0; mp < 2 * L + 1; ++mp)
                        a[mp] = vr_rtm[nd + (size_t)ncomp * (base + sl * mp)];
                    for (int i = 0; i < nphi; ++i) {
                        double x = a[L];
                        for (int m = 1; m <= L; ++m)
                            x += a[L + m] * ct[(size_t)m * nphi + i] + a[L - m] * st[(size_t)m * nphi + i];
                        fld_rtp[(size_t)k + (size_t)nr * t + (size_t)nr * nth * i + nnod * nd] = x;
                    }
                }
        free(a);
    }
    free(ct);
    free(st);
}

void orc_fft_fwd(int L, int nth, int nr, int nphi, int ncomp,
                 const double *fld_rtp, double *vr_rtm)
{
    const double pi = 4.0 * atan(1.0);
    const size_t nnod = (size_t)nr * nth * nphi;
    const size_t sl = (size_t)nth * nr;
    double *ct = (double *)malloc(sizeof(double) * (size_t)(L + 1) * nphi);
    double *st = (double *)malloc(sizeof(double) * (size_t)(L + 1) * nphi);
    for (int i = 0; i < nphi; ++i) {
        double phi = 2.0 * pi * (double)i / (double)nphi;
        for (int m = 0; m <= L; ++m) {
            ct[(size_t)m * nphi + i] = cos((double)m * phi);
            st[(size_t)m * nphi + i] = sin((double)m * phi);
        }
    }
#pragma omp parallel
    {
        double *xs = (double *)malloc(sizeof(double) * (size_t)nphi);
#pragma omp for collapse(3) schedule(static)
        for (int nd = 0; nd < ncomp; ++nd)
            for (int t = 0; t < nth; ++t)
                for (int k = 0; k < nr; ++k) {
                    const size_t base = (size_t)t + (size_t)nth * k;
                    const double *x = fld_rtp + (size_t)k + (size_t)nr * t + nnod * nd;
                    for (int i = 0; i < nphi; ++i) xs[i] = x[(size_t)nr * nth * i];
                    for (int m = 0; m <= L; ++m) {
                        double sc = 0.0, ss = 0.0;
                        for (int i = 0; i < nphi; ++i) {
                            sc += xs[i] * ct[(size_t)m * nphi + i];
                            ss += xs[i] * st[(size_t)m * nphi + i];
                        }
                        if (m == 0) {
                            vr_rtm[nd + (size_t)ncomp * (base + sl * L)] = sc / (double)nphi;
                        } else {
                            vr_rtm[nd + (size_t)ncomp * (base + sl * (L + m))] = 2.0 * sc / (double)nphi;
                            vr_rtm[nd + (size_t)ncomp * (base + sl * (L - m))] = 2.0 * ss / (double)nphi;
                        }
                    }
                }
        free(xs);
    }
    free(ct);
    free(st);
}

/* Grid-space nonlinear terms of s_cal_nonlinear_sph_MHD
 * (cal_nonlinear_sph_MHD.f90:35-75) with cal_cross_prod_w_coef_smp and
 * cal_vec_scalar_prod_w_coef_smp (cal_products_smp.f90:204-235,302-330):
 *   fld comps (b_trns): u 0-2, w 3-5, B 6-8, J 9-11, T 12
 *   frc comps (f_trns): w x u 0-2, Coriolis 3-5, rot Coriolis 6-8,
 *                       J x B 9-11, u x B 12-14, u T 15-17, div Coriolis 18
 * The Coriolis slots are filled in rlm space, so they leave the grid as 0. */
void orc_products_mhd(size_t nnod, const double *fld, double *frc,
                      double coef_velo, double coef_lor, double coef_induct,
                      double coef_temp)
{
#define F(c) (fld + (size_t)(c) * nnod)
#define G(c) (frc + (size_t)(c) * nnod)
    for (size_t i = 0; i < nnod; ++i) {
        const double u1 = F(0)[i], u2 = F(1)[i], u3 = F(2)[i];
        const double w1 = F(3)[i], w2 = F(4)[i], w3 = F(5)[i];
        const double b1 = F(6)[i], b2 = F(7)[i], b3 = F(8)[i];
        const double j1 = F(9)[i], j2 = F(10)[i], j3 = F(11)[i];
        const double T = F(12)[i];
        G(0)[i] = (w2 * u3 - w3 * u2) * coef_velo;
        G(1)[i] = (w3 * u1 - w1 * u3) * coef_velo;
        G(2)[i] = (w1 * u2 - w2 * u1) * coef_velo;
        for (int c = 3; c < 9; ++c) G(c)[i] = 0.0;
        G(9)[i] = (j2 * b3 - j3 * b2) * coef_lor;
        G(10)[i] = (j3 * b1 - j1 * b3) * coef_lor;
        G(11)[i] = (j1 * b2 - j2 * b1) * coef_lor;
        G(12)[i] = (u2 * b3 - u3 * b2) * coef_induct;
        G(13)[i] = (u3 * b1 - u1 * b3) * coef_induct;
        G(14)[i] = (u1 * b2 - u2 * b1) * coef_induct;
        G(15)[i] = u1 * T * coef_temp;
        G(16)[i] = u2 * T * coef_temp;
        G(17)[i] = u3 * T * coef_temp;
        G(18)[i] = 0.0;
    }
#undef F
#undef G
}

/* ---------------------------------------------------------------------
 * Banded LU with the reference's partial-pivoting rule, one matrix per
 * mode (mcomp modes, arrays in the reference's column-major shapes).
 * BA(k,i) = band_a(k,i), 1-based like the reference.
 * ------------------------------------------------------------------- */
#define A3(k, i) band_a[(size_t)((k) - 1) + 3 * (size_t)((i) - 1)]
#define LU5(k, i) lu[(size_t)((k) - 1) + 5 * (size_t)((i) - 1)]

/* ludcmp_3band, m_ludcmp_3band.f90:76-173 */
static int ludcmp_3band_one(int n, const double *band_a, int *ipiv, double *lu)
{
    const double TINY = 1.0e-30;
    double *vwk = (double *)calloc((size_t)n + 1, sizeof(double));
    int *idx = (int *)calloc((size_t)n + 1, sizeof(int));
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
            free(vwk);
            free(idx);
            return 100;
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
                int kst = j - 1 > 1 ? j - 1 : 1;
                int ked = j + 2 < n ? j + 2 : n;
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
    free(vwk);
    free(idx);
    return 0;
}
#undef A3
#undef LU5

void orc_ludcmp_3band(int n, int mcomp, const double *band_a, double *band_lu,
                      int *i_pivot, int *ierr)
{
    for (int m = 0; m < mcomp; ++m)
        ierr[m] = ludcmp_3band_one(n, band_a + (size_t)3 * n * m, i_pivot + (size_t)n * m,
                                   band_lu + (size_t)5 * n * m);
}

/* ludcmp_band with nband = 5 (nb1 = 2, nb2 = 4), m_ludcmp_band.f90:83-207 */
void orc_ludcmp_5band(int n, int mcomp, const double *band_a_all, double *band_lu_all,
                      int *indx_all)
{
    const double TINY = 1.0e-30;
    const int nband = 5, nb1 = 2, nb2 = 4;
#define BA(k, i) ba[(size_t)((k) - 1) + (size_t)nband * ((i) - 1)]
#define BL(k, i) bl[(size_t)((k) - 1) + 9 * (size_t)((i) - 1)]
    double *vv = (double *)calloc((size_t)n + 1, sizeof(double));
    int *idx = (int *)calloc((size_t)n + 1, sizeof(int));
    for (int mc = 0; mc < mcomp; ++mc) {
        const double *ba = band_a_all + (size_t)nband * n * mc;
        double *bl = band_lu_all + (size_t)9 * n * mc;
        int *indx = indx_all + (size_t)n * mc;
        for (int i = 1; i <= n; ++i) {
            int jst = i - nb1 > 1 ? i - nb1 : 1;
            int jed = i + nb1 < n ? i + nb1 : n;
            double aamax = fabs(BA(i - jst + nb1 + 1, jst));
            for (int j = jst + 1; j <= jed; ++j) aamax = fmax(aamax, fabs(BA(i - j + nb1 + 1, j)));
            vv[i] = 1.0 / aamax;
            idx[i] = i;
        }
        for (int k = 0; k < 9 * n; ++k) bl[k] = 0.0;
        for (int i = 1; i <= n; ++i) {
            int jst = i - nb1 > 1 ? i - nb1 : 1;
            int jed = i + nb1 < n ? i + nb1 : n;
            for (int j = jst; j <= jed; ++j) BL(i - j + nb2 + 1, j) = BA(i - j + nb1 + 1, j);
        }
        for (int j = 1; j <= n; ++j) {
            int ist = j - nb2 > 1 ? j - nb2 : 1;
            int ied = j + nb1 < n ? j + nb1 : n;
            for (int i = ist; i <= j - 1; ++i) {
                int l = i - j + nb2 + 1;
                for (int k = ist; k <= i - 1; ++k)
                    BL(l, j) = BL(l, j) - BL(i - k + nb2 + 1, k) * BL(k - j + nb2 + 1, j);
            }
            double aamax = 0.0;
            int imax = j;
            for (int i = j; i <= ied; ++i) {
                int l = i - j + nb2 + 1;
                int kst = i - nb2 > 1 ? i - nb2 : 1;
                for (int k = kst; k <= j - 1; ++k)
                    BL(l, j) = BL(l, j) - BL(i - k + nb2 + 1, k) * BL(k - j + nb2 + 1, j);
                double dum = vv[i] * fabs(BL(l, j));
                if (dum >= aamax && i <= (idx[j] + nb1)) {
                    imax = i;
                    aamax = dum;
                }
            }
            if (j != imax) {
                int k = idx[imax];
                idx[imax] = idx[j];
                idx[j] = k;
                int kst = j - nb2 > 1 ? j - nb2 : 1;
                int ked = imax + nb2 < n ? imax + nb2 : n;
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
                double dum = 1.0 / BL(nb2 + 1, j);
                int ied2 = j + nb2 < n ? j + nb2 : n;
                for (int i = j + 1; i <= ied2; ++i) BL(i - j + nb2 + 1, j) = BL(i - j + nb2 + 1, j) * dum;
            }
        }
    }
    free(vv);
    free(idx);
#undef BA
#undef BL
}

/* lubksb_3band_mul / lubksb_5band_mul, lubksb_357band_mul.f90:73-195;
 * x(mcomp, n) mode-fastest, solved in place. */
void orc_lubksb_mul(int nb, int n, int mcomp, const double *band_lu, const int *i_pivot,
                    double *x)
{
    const int w = (nb == 3) ? 5 : 9;
#define BL(k, i) bl[(size_t)((k) - 1) + (size_t)w * ((i) - 1)]
#define X(i) x[(size_t)m + (size_t)mcomp * ((i) - 1)]
    for (int m = 0; m < mcomp; ++m) {
        const double *bl = band_lu + (size_t)w * n * m;
        for (int i = 1; i <= n; ++i) {
            int ll = i_pivot[(size_t)(i - 1) + (size_t)n * m];
            double s = X(ll);
            X(ll) = X(i);
            X(i) = s;
        }
        if (nb == 3) {
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
    }
#undef BL
#undef X
}

/* ---------------------------------------------------------------------
 * Explicit time integration of the spectral fields (pointwise over d_rj):
 *   kind 0 (vorticity, induction; cal_vorticity_eq_adams,
 *     cal_vorticity_terms_adams.f90:31-51; cal_diff_induction_MHD_adams,
 *     cal_explicit_terms.f90:22-42):
 *       adams: x += dt (c D + 3/2 F - 1/2 P),  P <- F
 *       euler: x += dt (c D + F)               (and P <- F on step 1)
 *   kind 1 (scalar; scalar_diff_advect_adams/_euler,
 *     cal_diff_adv_src_explicit.f90:129-200):
 *       adams: x += dt (c D - 3/2 A - 1/2 P),  P <- -A
 *       euler: x += dt (c D - A)               (and P <- -A on step 1)
 * scheme: 0 adams, 1 euler, 2 euler + start of the Adams history.
 * ------------------------------------------------------------------- */
void orc_explicit_update(int kind, int scheme, size_t n, double dt, double coef_exp,
                         double *x, const double *diff, const double *force, double *pre)
{
    const double adam_0 = 3.0 / 2.0, adam_1 = -1.0 / 2.0;
    for (size_t i = 0; i < n; ++i) {
        if (kind == 0) {
            if (scheme == 0) {
                x[i] = x[i] + dt * (coef_exp * diff[i] + adam_0 * force[i] + adam_1 * pre[i]);
                pre[i] = force[i];
            } else {
                x[i] = x[i] + dt * (coef_exp * diff[i] + force[i]);
                if (scheme == 2) pre[i] = force[i];
            }
        } else {
            if (scheme == 0) {
                x[i] = x[i] + dt * (coef_exp * diff[i] - adam_0 * force[i] + adam_1 * pre[i]);
                pre[i] = -force[i];
            } else {
                x[i] = x[i] + dt * (coef_exp * diff[i] - force[i]);
                if (scheme == 2) pre[i] = -force[i];
            }
        }
    }
}

/* Volume mean square of one field in the shell kr_st..kr_ed (1-based):
 * cal_rms_each_vector_sph_spec / cal_rms_each_scalar_sph_spec
 * (cal_rms_by_sph_spectr.f90:63-154), set_sph_energies_by_rms (x 1/2 for
 * velocity and magnetic field), radial_int_by_trapezoid
 * (radial_int_for_sph_spec.f90:66-94), sum over modes of each degree, then
 * over degrees (sum_sph_rms_data.f90:290-320, cal_ave_4_rms_vector_sph.f90:
 * 127-148), times 3/(r_ed^3 - r_st^3) (cal_one_over_volume, :47-62).
 * f is d_rj of the field: f[j + J*k + J*nr*c], c = pol, dpol/dr, tor
 * (vector) or the scalar.  kind 0 vector, 1 vector energy, 2 scalar.
 * out[0..2] = pol, tor, total (vector) or out[0] (scalar). */
void orc_sph_mean_square(int kind, int nr, int J, const int *l_of_j, int lmax,
                         const double *radius, int kr_st, int kr_ed, const double *f,
                         double *out)
{
    const int ncp = (kind == 2) ? 1 : 3;
    double *vol = (double *)calloc((size_t)J * ncp, sizeof(double));
    double *lay = (double *)calloc((size_t)(nr + 1) * ncp, sizeof(double));
    for (int j = 0; j < J; ++j) {
        const int l = l_of_j[j];
        for (int k = 1; k <= nr; ++k) {
            const size_t in = (size_t)j + (size_t)J * (k - 1);
            const double r = radius[k - 1], ar = 1.0 / r;
            if (kind == 2) {
                const double g11 = 1.0 / (double)(2 * l + 1);
                const double d = f[in];
                lay[k] = d * d * g11 * r * r;
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
                lay[k] = r1;
                lay[k + (nr + 1)] = r2;
                lay[k + 2 * (nr + 1)] = r3;
            }
        }
        for (int c = 0; c < ncp; ++c) {
            double s = 0.0;
            for (int kr = kr_st; kr <= kr_ed - 1; ++kr) {
                const double dr1 = radius[kr] - radius[kr - 1];
                s = s + 0.5 * dr1 * (lay[kr + c * (nr + 1)] + lay[kr + 1 + c * (nr + 1)]);
            }
            vol[(size_t)j + (size_t)J * c] = s;
        }
    }
    const double rs = radius[kr_st - 1], re = radius[kr_ed - 1];
    const double avol = 3.0 / (re * re * re - rs * rs * rs);
    for (int c = 0; c < ncp; ++c) {
        double tot = 0.0;
        for (int l = 0; l <= lmax; ++l) {
            double sl = 0.0;
            for (int j = 0; j < J; ++j)
                if (l_of_j[j] == l) sl = sl + vol[(size_t)j + (size_t)J * c];
            tot = (l == 0) ? sl : tot + sl;
        }
        out[c] = avol * tot;
    }
    free(vol);
    free(lay);
}

/* Seeded synthetic spectra, restated from the published algorithms the
 * product's generator names: the SplitMix64 output function (Steele, Lea &
 * Flood, OOPSLA 2014) chained over (seed, field, k, j), two 53-bit uniforms
 * and the Box-Muller transform (Box & Muller 1958); amplitude (l+1)^-power
 * and the wall envelope ((r - r_0)(r_last - r))^2 / (d/2)^4.
 * Pinned by tests/test_seeded.py: SplitMix64's published output sequence
 * for seed 1234567, the Box-Muller closed form on chosen uniforms, Gaussian
 * moments over 10^6 draws, and the envelope's zeros at the walls. */
uint64_t orc_splitmix64(uint64_t state)
{
    uint64_t z = state + UINT64_C(0x9E3779B97F4A7C15);
    z ^= z >> 30;
    z *= UINT64_C(0xBF58476D1CE4E5B9);
    z ^= z >> 27;
    z *= UINT64_C(0x94D049BB133111EB);
    z ^= z >> 31;
    return z;
}

double orc_box_muller(double u1, double u2)
{
    return sqrt(-2.0 * log(u1)) * cos(8.0 * atan(1.0) * u2);
}

double orc_seeded_gauss(uint64_t seed, uint64_t field, uint64_t k, uint64_t j)
{
    uint64_t key = orc_splitmix64(seed);
    key = orc_splitmix64(key ^ field);
    key = orc_splitmix64(key ^ k);
    key = orc_splitmix64(key ^ j);
    const uint64_t a = orc_splitmix64(key) >> 11;
    const uint64_t b = orc_splitmix64(key ^ UINT64_C(0xD1B54A32D192ED03)) >> 11;
    const double u1 = ((double)a + 0.5) / 9007199254740992.0;   /* 2^53 */
    const double u2 = (double)b / 9007199254740992.0;
    return orc_box_muller(u1, u2);
}

void orc_seeded_spectrum(uint64_t seed, int field, int L, int nr, const double *radius,
                         double amplitude, double power, int envelope, int l_min, double *out)
{
    const int J = (L + 1) * (L + 1);
    const double half = 0.5 * (radius[nr - 1] - radius[0]);
    for (int k = 0; k < nr; ++k) {
        double e = 1.0;
        if (envelope) {
            const double q = (radius[k] - radius[0]) * (radius[nr - 1] - radius[k]);
            e = q * q / (half * half * half * half);
        }
        for (int j = 0; j < J; ++j) {
            int l = (int)sqrt((double)j);
            while (l * l > j) --l;
            while ((l + 1) * (l + 1) <= j) ++l;
            double v = 0.0;
            if (l >= l_min) {
                const double amp = amplitude * pow((double)(l + 1), -power) * e;
                v = amp * orc_seeded_gauss(seed, (uint64_t)field, (uint64_t)k, (uint64_t)j);
            }
            out[(size_t)j + (size_t)J * k] = v;
        }
    }
}
