/*
 * calypso_amd.h -- C ABI of the MI355X (gfx950) spherical-harmonic transform
 * library libcalypso_amd.so.  Plain pointers and sizes only; every entry
 * point returns 0 on success and a negative status on failure, with the
 * message available from calypso_last_error() (the reference aborts through
 * calypso_MPI_abort, calypso_mpi.f90:114-126; the Fortran shim does the same
 * on a nonzero status).
 *
 * Layouts are the reference's single-domain ones (1 MPI rank, identity
 * send/receive tables, set_local_sphere_param.f90:86-126):
 *   sp_rlm[nd + ncomp*(j + J*k)]                 J = (L+1)^2, rlm order
 *   vr_rtm[nd + ncomp*(t + nth*k + nth*nr*mp)]   mp = m + L, m = -L..L
 * with vectors first as (pol, dpol/dr, tor) <-> (r, theta, phi) triplets and
 * scalars after them (legendre_transform_select.f90:309-439).
 */
#ifndef CALYPSO_AMD_H
#define CALYPSO_AMD_H

#ifdef __cplusplus
extern "C" {
#endif

/* Precision / mode ids of the GPU Legendre transform. */
#define CALYPSO_LEG_FP64    1  /* FP64 MFMA 16x16x4                          */
#define CALYPSO_LEG_FP32    2  /* FP32 MFMA 16x16x4                          */
#define CALYPSO_LEG_BF16X3  3  /* split-BF16x3 MFMA 16x16x32, FP32 accumulate */
#define CALYPSO_LEG_INT8X6  4  /* each operand split into seven balanced base-128
                                  int8 digits at one power-of-two scale per table
                                  row and per data column (Ozaki splitting), the
                                  digit pairs i+j <= 6 on MFMA i32 16x16x64,
                                  exact int32 sums over the whole contraction,
                                  one FP64 conversion per tile: ~1e-13 of the
                                  FP64 results                                */

typedef struct calypso_leg calypso_leg_t;

/* Text of the last error on the calling thread ("" if none). */
const char *calypso_last_error(void);

/* Mode string -> id, the GPU counterpart of set_legendre_trans_mode_ctl
 * (legendre_transform_select.f90:181-224).  Accepts "hip_fp64", "hip_fp32",
 * "hip_bf16x3", "hip_int8x6" (case-insensitive); returns -1 for anything else. */
int calypso_legendre_mode_id(const char *name);

/* Gauss-Legendre points/weights/colatitudes of construct_gauss_coefs and
 * set_gauss_colatitude (m_gauss_points.f90:117-195); host-side, no GPU. */
int calypso_gauss_points(int n, double *point, double *weight, double *colat);

/* Create a transform context: the GPU counterpart of initialize_legendre_trans
 * + sel_init_legendre_trans (init_sph_trans.f90:64-92,
 * legendre_transform_select.f90:228-265).  P_l^m and dP_l^m/dtheta are
 * computed on the device in FP64 from the colatitudes.
 *   radius[nr]          radius_1d_rlm_r
 *   colat, weight[nth]  g_colat_rtm, weight_rtm (NULL: computed here)
 *   ncomp/nvector/nscalar  field set the plan is prepared for (others are
 *                       planned lazily on first use)                        */
int calypso_leg_create(int l_truncation, int nth, int nr, const double *radius,
                       const double *colat, const double *weight, int ncomp,
                       int nvector, int nscalar, int mode,
                       calypso_leg_t **out);

/* sel_finalize_legendre_trans (legendre_transform_select.f90:269-304). */
int calypso_leg_destroy(calypso_leg_t *ctx);

/* Device-resident transforms on HIP stream `stream` (a hipStream_t; NULL is
 * the null stream; calypso_leg_stream() gives the context's own stream).
 * Asynchronous.  d_sp_rlm / d_vr_rtm are device pointers.
 * backward = sel_backward_legendre_trans (legendre_transform_select.f90:309-371)
 * forward  = sel_forward_legendre_trans  (legendre_transform_select.f90:375-439)
 * The output array is fully overwritten. */
int calypso_leg_backward_device(calypso_leg_t *ctx, int ncomp, int nvector,
                                int nscalar, const double *d_sp_rlm,
                                double *d_vr_rtm, void *stream);
int calypso_leg_forward_device(calypso_leg_t *ctx, int ncomp, int nvector,
                               int nscalar, const double *d_vr_rtm,
                               double *d_sp_rlm, void *stream);

/* Host-buffer transforms with the reference signature
 * (ncomp, nvector, nscalar, n_WR, n_WS, WR, WS): WR is read, WS written;
 * synchronous (copies in, transform, copies out). */
int calypso_leg_backward(calypso_leg_t *ctx, int ncomp, int nvector,
                         int nscalar, int n_WR, int n_WS, const double *WR,
                         double *WS);
int calypso_leg_forward(calypso_leg_t *ctx, int ncomp, int nvector,
                        int nscalar, int n_WR, int n_WS, const double *WR,
                        double *WS);

/* Device P_l^m / dP_l^m/dtheta at all nth points in the reference's
 * P_rtm(l_rtm, j_rlm) order (m_schmidt_poly_on_rtm.f90), arrays nth*J. */
int calypso_leg_tables(calypso_leg_t *ctx, double *P, double *dP);

/* info[0..7]: L, nth, nr, mode, table bytes, radial tile (bwd), radial tile
 * (fwd), work-groups per backward launch (for the plan of the creation set). */
int calypso_leg_info(calypso_leg_t *ctx, long long *info);

/* The context's HIP stream (hipStream_t). */
void *calypso_leg_stream(calypso_leg_t *ctx);

/* Tile plan used for a field set and direction (fwd = 0 backward, 1
 * forward), planning it if needed: info[0..10] = radial tile, P-GEMM column
 * tiles, dP-GEMM column tiles, row tiles per wave, m/L-m pairing, colatitude
 * chunk, autotuned time per launch in microseconds (-1 if not timed),
 * work-groups per launch, kernel form (1 or 2), waves per work-group, P
 * prefetch depth in K steps.  Plans are chosen by timing the candidates on first
 * use (the GPU counterpart of select_legendre_transform,
 * sph_transforms_4_MHD.f90:309-426) unless calypso_leg_set_autotune(ctx, 0). */
int calypso_leg_plan_info(calypso_leg_t *ctx, int ncomp, int nvector,
                          int nscalar, int fwd, long long *info);
int calypso_leg_set_autotune(calypso_leg_t *ctx, int on);
/* Number of candidate tile plans for a field set and direction, and use of
 * candidate idx instead of the autotuner's pick (for tests and tuning). */
int calypso_leg_plan_count(calypso_leg_t *ctx, int ncomp, int nvector,
                           int nscalar, int fwd);
int calypso_leg_force_plan(calypso_leg_t *ctx, int ncomp, int nvector,
                           int nscalar, int fwd, int idx);

/* ---- process-wide selector, the shape the Fortran module binds ---------- */
/* One active context, like the reference's module-level state
 * (legendre_transform_select.f90:169-175). */
int calypso_sel_init_legendre_trans(int mode, int l_truncation, int nth,
                                    int nr, const double *radius,
                                    const double *colat, const double *weight,
                                    int ncomp, int nvector, int nscalar);
int calypso_sel_backward_legendre_trans(int ncomp, int nvector, int nscalar,
                                        int n_WR, int n_WS, const double *WR,
                                        double *WS);
int calypso_sel_forward_legendre_trans(int ncomp, int nvector, int nscalar,
                                       int n_WR, int n_WS, const double *WR,
                                       double *WS);
int calypso_sel_finalize_legendre_trans(void);


/* ---- longitude FFT, grid products, rtm <-> rtp exchange ----------------- */
typedef struct calypso_fft calypso_fft_t;

/* FFT context for nphi = 2^a 3^b longitudes (init_MHD_FFT_select,
 * MHD_FFT_selector.F90:86-114). */
int calypso_fft_create(int l_truncation, int nth, int nr, int nphi,
                       calypso_fft_t **out);
int calypso_fft_destroy(calypso_fft_t *fft);

/* rtm spectra (a_m in slot L+m, b_m in slot L-m) -> grid
 * fld_rtp[(k + nr*t + nr*nth*i) + nnod_rtp*nd]: back_MHD_FFT_sel_from_recv
 * (MHD_FFT_selector.F90:219-249) with the rtm->rtp exchange folded in. */
int calypso_fft_backward_device(calypso_fft_t *fft, int ncomp,
                                const double *d_vr_rtm, double *d_fld_rtp,
                                void *stream);
/* grid -> rtm spectra: fwd_MHD_FFT_sel_from_recv (MHD_FFT_selector.F90:
 * 185-215) with the rtp->rtm exchange folded in. */
int calypso_fft_forward_device(calypso_fft_t *fft, int ncomp,
                               const double *d_frc_rtp, double *d_vr_rtm,
                               void *stream);

/* s_cal_nonlinear_sph_MHD (cal_nonlinear_sph_MHD.f90:35-75) on the grid:
 * fld(nnod, 13: u, w, B, J, T) -> frc(nnod, 19: w x u, Coriolis(0),
 * rot Coriolis(0), J x B, u x B, u T, div Coriolis(0)).
 * coef[4] = coef_velo, coef_lor, coef_induct, coef_temp. */
int calypso_nonlinear_products_device(long long nnod, const double *d_fld,
                                      double *d_frc, const double *coef,
                                      void *stream);
/* The same on host arrays fld_rtp(nnod, 13) -> frc_rtp(nnod, 19), copied to
 * the device and back before returning (Fortran:
 * s_cal_nonlinear_sph_MHD_hip). */
int calypso_nonlinear_products(long long nnod, const double *fld_rtp,
                               double *frc_rtp, const double *coef);

/* Fused pencil step: rtm spectra of the 13 backward fields -> inverse FFT ->
 * products -> forward FFT -> rtm spectra of the 19 forward fields, the grid
 * kept on chip (back_MHD_FFT_sel_from_recv + s_cal_nonlinear_sph_MHD +
 * fwd_MHD_FFT_sel_from_recv). */
int calypso_fft_mhd_fused_device(calypso_fft_t *fft, const double *d_bwd_rtm,
                                 double *d_fwd_rtm, const double *coef,
                                 void *stream);

/* Host-buffer FFT selector with one process-wide context, shaped after the
 * reference module MHD_FFT_selector (MHD_FFT_selector.F90):
 *   init_MHD_FFT_select (:86-114): the grid (l_truncation, nth, nr, nphi)
 *     that the reference reads from module variables, and the component
 *     counts; device buffers for ncomp components are allocated here
 *   back_MHD_FFT_sel_from_recv (:219-249): WR(n_WR) = the rtm spectra,
 *     component fastest with stride ncomp_bwd (what the single-domain
 *     rtm -> rtp self-copy leaves in the receive buffer) -> fld_rtp(nnod_rtp,
 *     ncomp_bwd), nnod_rtp = nr nth nphi, node k + nr t + nr nth i
 *   fwd_MHD_FFT_sel_from_recv (:185-215): frc_rtp(nnod_rtp, ncomp_fwd) ->
 *     WS(n_WS), stride ncomp_fwd
 *   finalize_MHD_FFT_select (:118-145)
 * Each call copies its buffers to the device, transforms and copies back
 * before returning. */
int calypso_init_MHD_FFT_select(int l_truncation, int nth, int nr, int nphi,
                                int ncomp, int ncomp_fwd, int ncomp_bwd);
int calypso_back_MHD_FFT_sel_from_recv(int ncomp_bwd, int n_WR,
                                       const double *WR, double *fld_rtp);
int calypso_fwd_MHD_FFT_sel_from_recv(int ncomp_fwd, int n_WS,
                                      const double *frc_rtp, double *WS);
int calypso_finalize_MHD_FFT_select(void);

/* ---- banded LU of the Crank-Nicolson radial matrices (FP64) ------------ */
/* Factor mcomp band matrices band_a(nb, n, mcomp) (nb = 3: ludcmp_3band,
 * m_ludcmp_3band.f90:76-173; nb = 5: ludcmp_band with nband 5,
 * m_ludcmp_band.f90:83-207) into band_lu(2nb-1, n, mcomp) and
 * i_pivot(n, mcomp); ierr(mcomp) = 100 for a singular 3-band matrix.
 * Same pivoting and operation order as the reference (bitwise results). */
int calypso_ludcmp_band_device(int nb, int n, int mcomp, const double *d_band_a,
                               double *d_band_lu, int *d_i_pivot, int *d_ierr,
                               void *stream);
/* Solve in place x(mcomp, n) (mode fastest, a d_rj field):
 * lubksb_3band_mul / lubksb_5band_mul (lubksb_357band_mul.f90:73-195). */
int calypso_lubksb_band_mul_device(int nb, int n, int mcomp,
                                   const double *d_band_lu,
                                   const int *d_i_pivot, double *d_x,
                                   void *stream);
/* The same solves on host arrays with the reference's argument list:
 * lubksb_3band_mul / lubksb_5band_mul(Msmp, Msmp_stack, mcomp, n, band_lu,
 * i_pivot, x) (lubksb_357band_mul.f90:73-195).  Modes Msmp_stack[0]+1 ..
 * Msmp_stack[Msmp] of x(mcomp, n) are solved (the union of the reference's
 * per-thread ranges), the others are left as they are; a pivot outside 1..n
 * is an error.  The device copies are kept for the next call. */
int calypso_lubksb_3band_mul(int msmp, const int *msmp_stack, int mcomp, int n,
                             const double *band_lu, const int *i_pivot, double *x);
int calypso_lubksb_5band_mul(int msmp, const int *msmp_stack, int mcomp, int n,
                             const double *band_lu, const int *i_pivot, double *x);

/* ---- explicit time integration and energy diagnostics (FP64) ----------- */
/* Pointwise update of n d_rj values (one field component):
 *  kind 0 (vorticity, magnetic field): x += dt (coef_exp D + 3/2 F - 1/2 P),
 *         P <- F  (cal_vorticity_eq_adams, cal_vorticity_terms_adams.f90:31-51;
 *         cal_diff_induction_MHD_adams, cal_explicit_terms.f90:22-42)
 *  kind 1 (scalar): x += dt (coef_exp D - 3/2 A - 1/2 P), P <- -A
 *         (scalar_diff_advect_adams, cal_diff_adv_src_explicit.f90:129-154)
 *  scheme 0 Adams-Bashforth, 1 Euler, 2 Euler and start of the history
 *  (cal_expricit_sph_euler, cal_momentum_eq_explicit.f90:82-128). */
int calypso_explicit_update_device(int kind, int scheme, long long n,
                                   double dt, double coef_exp, double *d_x,
                                   const double *d_diffuse,
                                   const double *d_force, double *d_pre,
                                   void *stream);

/* Volume mean square of one field over radial points kr_st..kr_ed (1-based):
 * the numbers of sph_pwr_volume_s.dat (cal_rms_by_sph_spectr.f90:63-154,
 * radial_int_for_sph_spec.f90:66-94, cal_ave_4_rms_vector_sph.f90:47-148).
 * d_f: the field's d_rj columns (pol, dpol/dr, tor or one scalar), each
 * J*nr with the mode fastest; l_of_j[J] degrees of the rj modes (host);
 * radius[nr] (host).  kind 0 vector, 1 vector energy (x 1/2), 2 scalar.
 * out[3] (host): pol, tor, total (vector) or out[0] (scalar). */
int calypso_sph_mean_square(int kind, int nr, int J, const int *l_of_j,
                            const double *radius, int kr_st, int kr_ed,
                            const double *d_f, double *out, void *stream);

/* calypso_sph_mean_square without any device allocation: no counterpart in
 * the reference; for callers that cannot call hipMalloc (diagnostics inside
 * tight serving or time-stepping loops, harnesses that forbid allocation).
 * Same arguments and bitwise the same results.  The per-mode partial sums,
 * the radial grid and the degrees live in one library-owned device array of
 * 7.4 MB per GPU (held until the process ends), so J <= 262144 (L <= 511),
 * nr <= 4096 and degrees 0..2047; larger inputs are rejected with an error.
 * Calls that overlap on one GPU share that array and race. */
int calypso_sph_mean_square_noalloc(int kind, int nr, int J, const int *l_of_j,
                                    const double *radius, int kr_st, int kr_ed,
                                    const double *d_f, double *out, void *stream);

/* ---- seeded synthetic inputs (host only, no GPU) ---------------------- */
/* out[j + J k] (J = (L+1)^2, mode j = l(l+1)+m, k < nr) = amplitude x
 * (l+1)^-power x g(seed, field, k, j) x (envelope ? ((r_k - r_0)(r_{nr-1} -
 * r_k))^2 / ((r_{nr-1} - r_0)/2)^4 : 1) for l >= l_min, 0 below: g is a
 * standard Gaussian from a counter-based generator keyed by (seed, field,
 * k, j) (SplitMix64 output function, Box-Muller), so a value does not depend
 * on the array's size or on the order of generation. */
int calypso_seeded_spectrum(unsigned long long seed, int field, int l_truncation, int nr,
                            const double *radius, double amplitude, double power, int envelope,
                            int l_min, double *out);

/* ---- grid tables of one domain (host only, no GPU) --------------------- */
/* Writes <prefix>.0.rj, .rlm, .rtm and .rtp: the index tables, radial grid,
 * communication tables and groups of a one-domain spherical shell, byte for
 * byte as the reference's grid generator gen_sph_grids writes them for one
 * subdomain (write_spectr_modes_rj_data, write_modes_rlm_data,
 * write_geom_rtm_data, write_geom_rtp_data, sph_modes_grids_data_IO.f90:
 * 128-245; programs/mesh_utilities/main_generate_sph_grids.f90).  radius[nr]
 * from the inner to the outer wall (no inner-core points); nphi = 2 nth and
 * L < nth, as the generator's zonal FFT table requires. */
int calypso_sph_write_grid_files(const char *prefix, int l_truncation, int nth, int nphi,
                                 int nr, const double *radius);
/* Host-only: the reference's ASCII field file of one domain (the restart and
 * spectral-field format): write_step_data (m_time_data_IO.f90:107-119) with
 * (rank, istep, time, dt), then write_field_data (field_data_IO.f90:267-298)
 * of nfld fields with ncomp[i] components named names[i], data(nnod, ntot)
 * column-major as the reference's fld_IO%d_IO (component c at data + c nnod). */
int calypso_sph_write_field_file(const char *path, int rank, int istep, double time,
                                 double dt, long long nnod, int nfld, const int *ncomp,
                                 const char *const *names, const double *data);
/* The local mode order of the reference's one-domain rj (and rlm) arrays:
 * j_global[j] = l(l+1)+m of local mode j, orders m = -L..L, degrees |m|..L
 * inside (idx_gl_1d_rj_j(:,1) after gen_sph_grids); J = (L+1)^2 entries. */
int calypso_sph_rj_mode_order(int l_truncation, int *j_global);

/* ---- the full MHD time step, device resident (FP64 state) ---------------
 * Replaces the body of module SPH_analyzer_MHD: SPH_initialize_MHD and
 * SPH_analyze_MHD(i_step, iflag_finish) (programs/SPH_MHD/
 * SPH_analyzer_MHD.f90:31-202) for a rotating Boussinesq shell with rigid
 * or free-slip walls, insulating magnetic boundaries and fixed-temperature
 * walls around the spherical-shell reference temperature (the model of the
 * reference's Dynamobench_case1).  The coefficients are those module
 * m_physical_property holds after set_material_property
 * (material_property.f90); radius[nr] is the radial grid.  kr_icb and
 * kr_cmb are the 1-based points of the reference's ICB and CMB radial groups
 * (0: the grid's first and last points).  bc_velo_icb / bc_velo_cmb choose
 * each velocity wall as the reference's bc_velocity does (non_slip_sph,
 * free_slip_sph, and at the ICB rot_inner_core; set_bc_flag_sph_velo.f90,
 * set_bc_sph_mhd.f90:47-57); bc_magne_icb the magnetic condition below the
 * shell (insulator, or sph_to_center: set_bc_sph_mhd.f90:147-197).  Points
 * below the ICB come with the rotatable inner core and the field solved to
 * the centre (the model of the reference's Dynamobench_case2: the core
 * turns under the viscous, Lorentz and Coriolis torques,
 * cal_inner_core_rotation.f90; the field continues above the CMB as a
 * potential field); any other combination is rejected by
 * calypso_mhd_create. */
#define CALYPSO_VELO_RIGID          0
#define CALYPSO_VELO_FREE_SLIP      1
#define CALYPSO_VELO_ROT_INNER_CORE 2
#define CALYPSO_MAGNE_INSULATOR     0
#define CALYPSO_MAGNE_TO_CENTER     1
typedef struct calypso_mhd calypso_mhd_t;
typedef struct calypso_mhd_params {
    int l_truncation, nth, nphi, nr;
    int legendre_mode;            /* calypso_legendre_mode_id() value */
    double dt;
    double coef_imp_v, coef_imp_t, coef_imp_b;
    double coef_velo, coef_d_velo, coef_buo, coef_cor, coef_lor;
    double coef_temp, coef_d_temp;
    double coef_magne, coef_d_magne, coef_induct;
    double depth_high_t, depth_low_t, high_temp, low_temp;
    int bc_velo_icb, bc_velo_cmb; /* CALYPSO_VELO_RIGID / _FREE_SLIP / _ROT_INNER_CORE */
    int kr_icb, kr_cmb;           /* 1-based ICB / CMB points; 0: first / last point */
    int bc_magne_icb;             /* CALYPSO_MAGNE_INSULATOR / _TO_CENTER */
    double coef_press;            /* coef_4_press (E^-1 for the benchmarks); only
                                   * calypso_mhd_pressure reads it, 0: no pressure */
} calypso_mhd_params;

int calypso_mhd_create(const calypso_mhd_params *params, const double *radius,
                       void *stream, calypso_mhd_t **out);
int calypso_mhd_destroy(calypso_mhd_t *mhd);
/* The caller's local rj mode order: j_global[j] = l(l+1)+m of its local mode
 * j (the reference's idx_gl_1d_rj_j(:,1); for one domain from gen_sph_grids
 * that is calypso_sph_rj_mode_order's list, orders grouped, not
 * l(l+1)+m).  From then on calypso_mhd_set_field / _get_field and the
 * restart files take and give d_rj in that order.  Default: j = l(l+1)+m. */
int calypso_mhd_set_rj_order(calypso_mhd_t *mhd, int J, const int *j_global);
/* One d_rj component (nr*J doubles, mode j = l(l+1)+m fastest unless
 * calypso_mhd_set_rj_order gave another order), host copy.
 * Components: 0-2 velocity, 3-5 vorticity, 6-8 magnetic field, 9-11 current
 * density, 12 temperature (perturbation once initialised), 13-15 vorticity
 * diffusion, 16-18 magnetic diffusion, 19 thermal diffusion, 20-22 forces,
 * 26-28 induction, 32 heat advection, 34 buoyancy (m_sph_phys_address). */
int calypso_mhd_set_field(calypso_mhd_t *mhd, int comp, const double *host);
int calypso_mhd_get_field(calypso_mhd_t *mhd, int comp, double *host);
/* Fields on the physical grid at an output step: sph_back_trans_snapshot_MHD
 * (sph_transforms_4_MHD.f90:200-224) without the pole values.  nvec vector
 * fields (vec_comp: the first d_rj component of each, pol / dpol/dr / tor)
 * and nscl scalar components (scl_comp), 3 nvec + nscl <= 64, through the
 * context's backward Legendre transform (its precision mode) and FFT.
 * host_rtp(nnod_rtp, 3 nvec + nscl), nnod_rtp = nr nth nphi, node
 * k + nr t + nr nth i; vectors as (r, theta, phi), scalars after.  The
 * temperature component holds the perturbation from the reference profile
 * during a run. */
int calypso_mhd_snapshot_rtp(calypso_mhd_t *mhd, int nvec, const int *vec_comp,
                             int nscl, const int *scl_comp, double *host_rtp);
/* dynamo_benchmark_1 initial fields (set_initial_sph_dynamo.f90:
 * sph_initial_data_control, initial_magne_dynamobench.f90:
 * initial_b_dynamobench_1); isig selects the (m, m) temperature mode. */
int calypso_mhd_set_dynamobench_ic(calypso_mhd_t *mhd, int isig);
/* dynamo_benchmark_2 initial fields (initial_b_dynamobench_2,
 * initial_magne_dynamobench.f90:103-147): the temperature as above, the
 * magnetic field from the first point above the centre to the CMB; needs
 * the inner-core model. */
int calypso_mhd_set_dynamobench2_ic(calypso_mhd_t *mhd, int isig);
/* Seeded random initial fields (the synthetic inputs of the benchmark
 * configurations; the reference reads such fields from a restart file):
 * poloidal/toroidal velocity and magnetic field for l >= 1 and the
 * temperature perturbation for l >= 0, each amplitude (amp_u, amp_b,
 * amp_t) x (l+1)^-2 x a Gaussian of calypso_seeded_spectrum (field ids 0, 2,
 * 6, 8, 12) x the wall envelope; the l = 0 temperature adds the reference
 * profile, so the state holds the full temperature like the benchmark start. */
int calypso_mhd_set_seeded_ic(calypso_mhd_t *mhd, unsigned long long seed, double amp_u,
                              double amp_b, double amp_t);
/* SPH_initialize_MHD after the initial fields: perturbation temperature,
 * set_sph_field_to_start and the first nonlinear terms. */
int calypso_mhd_initialize(calypso_mhd_t *mhd);
/* nsteps calls of SPH_analyze_MHD (Euler on step 1, then Adams-Bashforth). */
int calypso_mhd_step(calypso_mhd_t *mhd, int nsteps);
/* The pressure of the current state as an output step computes it
 * (pressure_4_sph_mhd, lead_fields_4_sph_mhd.f90:70-95, replacing
 * cal_sol_pressure_by_div_v, cal_sol_sph_fluid_crank.f90:136-175): the
 * divergence of the advection, Lorentz force and buoyancy with the walls'
 * radial forces, the Poisson solve of const_radial_mat_4_press_sph
 * (const_r_mat_4_scalar_sph.f90:73-129) and adjust_by_ave_pressure_on_CMB.
 * host: nr*J doubles in the caller's rj order, zero outside the shell.  The
 * matrix is factored on the first call (5 nr J doubles of device memory,
 * kept until calypso_mhd_destroy); params.coef_press must be set. */
int calypso_mhd_pressure(calypso_mhd_t *mhd, double *host);
/* The energy fluxes of the current state as an output step computes them
 * (enegy_fluxes_4_sph_mhd, lead_fields_4_sph_mhd.f90:98-121, replacing
 * s_cal_energy_flux_rtp, cal_energy_flux_rtp.f90:86-150, between
 * sph_back_trans_snapshot_MHD and sph_forward_trans_snapshot_MHD,
 * sph_transforms_4_MHD.f90:200-278): u, B, J, the magnetic induction and the
 * temperature perturbation to the grid, there the scalar fluxes, and their
 * forward transform.  flux_ids[nflux] (1 <= nflux <= 8): 0 Lorentz_work
 * u.(coef_lor J x B), 1 work_against_Lorentz (its negative), 2
 * magnetic_ene_generation B.(curl(u x B)), 3 buoyancy_flux coef_buo T u_r r
 * (m_phys_labels.f90:425-441).  host(nflux, nr, J): spectra in the caller's
 * rj order.  Temporary device buffers of about 13 grids are allocated and
 * freed per call. */
int calypso_mhd_energy_fluxes(calypso_mhd_t *mhd, int nflux, const int *flux_ids,
                              double *host);
/* out[7]: K_ene_pol, K_ene_tor, K_ene, temperature, M_ene_pol, M_ene_tor,
 * M_ene over the shell (sph_pwr_volume_s.dat, output_rms_sph_mhd_control). */
int calypso_mhd_energies(calypso_mhd_t *mhd, double *out);
/* info[4]: i_step, nr, J, number of d_rj components. */
int calypso_mhd_info(calypso_mhd_t *mhd, long long *info);
/* Model time of the current step (time of the start or restart + steps * dt). */
int calypso_mhd_time(calypso_mhd_t *mhd, double *time);
/* Appends the current step's row to the volume mean square file
 * (sph_pwr_volume_s.dat), creating it with its header first
 * (write_sph_vol_ms_file, output_sph_m_square_file.f90:134-156;
 * write_sph_mean_sq_header, m_rms_4_sph_spectr.f90:245-280). */
int calypso_mhd_write_volume_ms(calypso_mhd_t *mhd, const char *path);
/* Writes the ASCII restart file <prefix>.<istep_fld>.0.fst (velocity,
 * temperature, magnetic_field, previous_momentum, previous_induction,
 * previous_heat; sel_write_step_SPH_field_file, field_IO_select.F90:106-122,
 * write_step_field_file, field_file_IO.f90:40-62); its name is copied to
 * name_out[name_len] when given. */
int calypso_mhd_write_restart(calypso_mhd_t *mhd, const char *prefix, int istep_fld,
                              char *name_out, int name_len);
/* Reads a restart file (domain 0) into the state; calypso_mhd_initialize then
 * resumes at its step and time (sel_read_alloc_step_SPH_file and
 * set_sph_restart_from_IO, sph_mhd_rst_IO_control.f90:125-140). */
int calypso_mhd_read_restart(calypso_mhd_t *mhd, const char *path);
/* Execution options: "fused_fft" (1: FFT, products and FFT in one pencil
 * kernel, the default where the pencil fits in LDS; 0: three kernels with
 * the grids in HBM), "legendre_autotune" (0/1). */
int calypso_mhd_set_option(calypso_mhd_t *mhd, const char *name, long long value);
/* ---- scalars diffusing in a shell at rest ---------------------------------
 * Replaces SPH_initialize_MHD / SPH_analyze_MHD (programs/SPH_MHD/
 * SPH_analyzer_MHD.f90:31-202) for a model that evolves only scalars
 * (time_evo_ctl temperature, composition) in a fluid at rest, each wall
 * holding a value or a flux per spherical-harmonic mode from the spectral
 * boundary file (set_bc_sph_scalars.f90, m_sph_boundary_input_data.f90:
 * 147-300; the reference's tests/heterogineous_temp).  FP64, one stream. */
#define CALYPSO_BC_FIXED_VALUE 0
#define CALYPSO_BC_FIXED_FLUX 1
typedef struct calypso_scalar_shell calypso_scalar_shell_t;
typedef struct calypso_scalar_bc {
    double coef_d;        /* diffusion coefficient (coef_d_temp, coef_d_light) */
    double coef_imp;      /* Crank-Nicolson weight (coef_imp_t, coef_imp_c) */
    int bc_icb, bc_cmb;   /* CALYPSO_BC_FIXED_VALUE or CALYPSO_BC_FIXED_FLUX */
    const char *name;     /* column name in the energy file, or NULL */
} calypso_scalar_bc;
/* radius[nr] the radial grid (walls at its ends), 1..8 scalars;
 * val_icb / val_cmb [nscalar][J], J = (L+1)^2, mode l(l+1)+m fastest: the
 * wall value or the flux as the boundary file gives it (the inner flux is
 * negated inside, as inner_sph_bc_scalar_sph does).  Errors: unknown
 * boundary kind, grid not increasing, singular radial matrix. */
int calypso_scalar_shell_create(int l_truncation, int nr, const double *radius, double dt,
                                int nscalar, const calypso_scalar_bc *scalars,
                                const double *val_icb, const double *val_cmb, void *stream,
                                calypso_scalar_shell_t **out);
int calypso_scalar_shell_destroy(calypso_scalar_shell_t *sh);
/* set_noize_scalar_sph (set_initial_sph_dynamo.f90:390-441) on every scalar. */
int calypso_scalar_shell_set_noise(calypso_scalar_shell_t *sh);
/* host[nr*J] into scalar `scalar` (mode fastest). */
int calypso_scalar_shell_set_field(calypso_scalar_shell_t *sh, int scalar, const double *host);
/* field = which * nscalar + scalar, which 0 value, 1 diffusion term,
 * 2 advection term, 3 Adams history; host[nr*J]. */
int calypso_scalar_shell_get_field(calypso_scalar_shell_t *sh, int field, double *host);
/* The diffusion terms of the initial fields (fixed walls take their values). */
int calypso_scalar_shell_initialize(calypso_scalar_shell_t *sh);
/* nsteps time steps (Euler on the first, then Adams-Bashforth). */
int calypso_scalar_shell_step(calypso_scalar_shell_t *sh, int nsteps);
/* out[nscalar]: volume mean square of each scalar over the shell. */
int calypso_scalar_shell_mean_square(calypso_scalar_shell_t *sh, double *out);
/* Appends the current step's row to sph_pwr_volume_s.dat, header first. */
int calypso_scalar_shell_write_volume_ms(calypso_scalar_shell_t *sh, const char *path);

/* Host-only: coupling table of the curl of the Coriolis force for
 * truncation L (init_sum_coriolis_rlm, const_coriolis_sph_rlm.f90:34-67).
 * cor[12*J]: sw11 sw21 sw12 sw22 sw13 tw13 tw23 tw14 tw24 tw11 tw12 g3 per rj
 * mode; idx[3*J]: rj indices of (l-1,m), (l+1,m), (l,-m). */
int calypso_mhd_coriolis_table(int l_truncation, double *cor, int *idx);

#ifdef __cplusplus
}
#endif
#endif
