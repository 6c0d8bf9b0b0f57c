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
    double *p = (double *)malloc(sizeof(double) * (size_t)(L + 1));
    double *dp = (double *)malloc(sizeof(double) * (size_t)(L + 1));
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
    for (int mp = 0; mp < M; ++mp) {
        int mn = M - 1 - mp;
        for (int k = 0; k < nr; ++k)
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
    double *pvw = (double *)malloc(sizeof(double) * (size_t)nth);
    double *dpvw = (double *)malloc(sizeof(double) * (size_t)nth);
    double *pgvw = (double *)malloc(sizeof(double) * (size_t)nth);
    orc_rlm_index(L, jlm, lstack);
    orc_g_sph_rlm(L, g);
    for (int t = 0; t < nth; ++t) asin_t[t] = 1.0 / sin(colat[t]);
    /* mdx_p_rlm_rtm / mdx_n_rlm_rtm: set_params_sph_trans.f90:57-68 */
    for (int j = 0; j < J; ++j) {
        mdx_p[j] = jlm[3 * j + 2] + L;
        mdx_n[j] = -jlm[3 * j + 2] + L;
    }
    memset(sp_out, 0, sizeof(double) * (size_t)ncomp * J * nr);

    for (int k = 0; k < nr; ++k) {
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
    }
    free(jlm);
    free(lstack);
    free(mdx_p);
    free(mdx_n);
    free(g);
    free(asin_t);
    free(pvw);
    free(dpvw);
    free(pgvw);
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
    for (int nd = 0; nd < ncomp; ++nd)
        for (int t = 0; t < nth; ++t)
            for (int k = 0; k < nr; ++k) {
                const size_t base = (size_t)t + (size_t)nth * k;
                for (int i = 0; i < nphi; ++i) {
                    double phi = 2.0 * pi * (double)i / (double)nphi;
                    double x = vr_rtm[nd + (size_t)ncomp * (base + sl * L)];
                    for (int m = 1; m <= L; ++m) {
                        double a = vr_rtm[nd + (size_t)ncomp * (base + sl * (L + m))];
                        double b = vr_rtm[nd + (size_t)ncomp * (base + sl * (L - m))];
                        x += a * cos((double)m * phi) + b * sin((double)m * phi);
                    }
                    fld_rtp[(size_t)k + (size_t)nr * t + (size_t)nr * nth * i + nnod * nd] = x;
                }
            }
}

void orc_fft_fwd(int L, int nth, int nr, int nphi, int ncomp,
                 const double *fld_rtp, double *vr_rtm)
{
    const double pi = 4.0 * atan(1.0);
    const size_t nnod = (size_t)nr * nth * nphi;
    const size_t sl = (size_t)nth * nr;
    for (int nd = 0; nd < ncomp; ++nd)
        for (int t = 0; t < nth; ++t)
            for (int k = 0; k < nr; ++k) {
                const size_t base = (size_t)t + (size_t)nth * k;
                const double *x = fld_rtp + (size_t)k + (size_t)nr * t + nnod * nd;
                for (int m = 0; m <= L; ++m) {
                    double sc = 0.0, ss = 0.0;
                    for (int i = 0; i < nphi; ++i) {
                        double phi = 2.0 * pi * (double)i / (double)nphi;
                        double xi = x[(size_t)nr * nth * i];
                        sc += xi * cos((double)m * phi);
                        ss += xi * sin((double)m * phi);
                    }
                    if (m == 0) {
                        vr_rtm[nd + (size_t)ncomp * (base + sl * L)] = sc / (double)nphi;
                    } else {
                        vr_rtm[nd + (size_t)ncomp * (base + sl * (L + m))] = 2.0 * sc / (double)nphi;
                        vr_rtm[nd + (size_t)ncomp * (base + sl * (L - m))] = 2.0 * ss / (double)nphi;
                    }
                }
            }
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
