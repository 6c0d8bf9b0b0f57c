// mhd_internal.hpp -- device-resident MHD time step (rotating Boussinesq
// shell, rigid or free-slip walls, insulating magnetic boundaries,
// fixed-temperature boundaries with the spherical-shell reference
// temperature; or a rotatable conducting inner core below the ICB with the
// magnetic field solved to the centre and points above the CMB).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <string>
#include <vector>

#include "fft_internal.hpp"
#include "leg_internal.hpp"

namespace calypso {

// d_rj components, each nr*J doubles with the mode index fastest
// (the reference's d_rj(inod, comp), inod = j + (k-1) J).
enum MhdField : int {
    kVelo = 0, kVort = 3, kMagne = 6, kCurrent = 9, kTemp = 12,
    kWDiffuse = 13, kBDiffuse = 16, kTDiffuse = 19,
    kForces = 20, kPreMom = 23, kInduction = 26, kPreUxB = 29,
    kHAdvect = 32, kPreHeat = 33, kRotBuoyancy = 34,
    kMAdvect = 35, kLorentz = 38, kVpInduct = 41, kHFlux = 44,
    kRotCoriolis = 47, kRotInertia = 50, kRotLorentz = 53,
    kMhdNumComp = 56
};

// Radial finite differences of a grid (nr >= 3): interior three-point
// weights d1/d2 [nr][3] at offsets -1, 0, +1; wall formulas for a fixed value
// (ffl_*) and a fixed gradient (fdr_*) at the inner (_i) and outer (_c)
// walls; the free-slip formulas of the poloidal (fvp_*) and toroidal (fvt_*)
// velocity, [pt + 1][c] with pt -1 / 0 / +1 the point below / at / above the
// wall and c = 0 value, 1 d/dr, 2 d2/dr2; powers r, 1/r, 1/r^2 of the wall
// radii.
// The walls are the points ki (ICB) and ko (CMB); with points below the ICB
// the first point's lower neighbour is the centre (distance r[0]), ctr1 is
// the centre formula at the first point ([pt + 1][c], pt -1 the centre) and
// fdm1_i the two-point d/dr at the ICB.
struct RadialFdm {
    std::vector<double> d1, d2;
    double ffl_i[3][3], fdr_i[3][3], ffl_c[3][3], fdr_c[3][3];
    double fvp_i[3][3], fvt_i[3][3], fvp_c[3][3], fvt_c[3][3];
    double ctr1[3][3], fdm1_i[2];
    double ri[3], rc[3];
};

// velocity wall conditions (bc_velocity of the reference's control file):
// non_slip_sph, free_slip_sph, rot_inner_core (ICB only)
enum MhdVeloBC : int { kVeloRigid = 0, kVeloFreeSlip = 1, kVeloRotInnerCore = 2 };
// magnetic condition below the first shell point (bc_magnetic_field):
// insulator at the ICB, or sph_to_center (field solved to the centre)
enum MhdMagneBC : int { kMagneInsulator = 0, kMagneToCenter = 1 };
void radial_fdm(const double *r, int nr, RadialFdm *out);
void radial_fdm(const double *r, int nr, int ki, int ko, RadialFdm *out);

struct MhdCoefs {
    double dt = 0;
    double coef_imp_v = 0, coef_imp_t = 0, coef_imp_b = 0;
    double coef_velo = 1, coef_d_velo = 1, coef_buo = 0, coef_cor = 0, coef_lor = 0;
    double coef_temp = 1, coef_d_temp = 1;
    double coef_magne = 1, coef_d_magne = 1, coef_induct = 1;
    double depth_high_t = 0, depth_low_t = 0, high_temp = 0, low_temp = 0;
    int bc_velo_icb = kVeloRigid, bc_velo_cmb = kVeloRigid;
    int kr_icb = 0, kr_cmb = -1;  // 0-based ICB / CMB points (kr_cmb -1: the last)
    int bc_magne_icb = kMagneInsulator;
    double coef_press = 0;        // coef_4_press; 0: the pressure is not asked for
};

// Everything a kernel needs, passed by value.
struct MhdDev {
    int nr = 0, J = 0, ki = 0, ko = 0, L = 0;
    int kb = 0;                                // first point of the magnetic field
    int icore = 0;                             // rotatable inner core, B to the centre
    double fdm1_i[2];                          // two-point d/dr at the ICB
    double rJ = 0.0;                           // 1 / J
    size_t nnod = 0;
    double *d = nullptr;                       // d_rj, kMhdNumComp * nnod
    const double *r = nullptr, *ar1 = nullptr, *ar2 = nullptr;
    const double *d1 = nullptr, *d2 = nullptr; // [nr][3] offsets -1, 0, +1
    const double *g1 = nullptr, *g3 = nullptr; // rj normalisations l, l(l+1)
    const double *t_icb = nullptr, *t_cmb = nullptr;   // [J]
    const double *reft0 = nullptr, *reft1 = nullptr;   // [nr]
    const int *jk1 = nullptr, *jk2 = nullptr, *jl1 = nullptr;
    const double *cor = nullptr;               // [12][J] Coriolis couplings
    const int *rj_of_rlm = nullptr, *rlm_of_rj = nullptr;
    double *sp_b = nullptr, *sp_f = nullptr;   // rlm buffers, 13 / 19 comps
    double ffl_i[3][3], fdr_i[3][3], ffl_c[3][3], fdr_c[3][3];
    double fvp_i[3][3], fvt_i[3][3], fvp_c[3][3], fvt_c[3][3];
    int free_i = 0, free_c = 0;                // free-slip velocity walls
    double ri[3], rc[3];
    double dt, coef_exp_v, coef_exp_t, coef_exp_b;
    double coef_d_velo, coef_d_magne, coef_d_temp, coef_buo, coef_temp, coef_cor;
};

struct MhdContext {
    int L = 0, nth = 0, nphi = 0, nr = 0, J = 0, M = 0;
    int leg_mode = kF64;
    int i_step = 0;
    bool initialized = false;
    int istep0 = 0;               // step and time the run started from (restart)
    double time0 = 0.0;
    int restart_step = -1;        // set by a restart read, taken by the initialisation
    double restart_time = 0.0;
    MhdCoefs co;
    std::vector<double> radius;
    std::vector<double> reft0;    // reference temperature at every point
    LegContext leg;
    FftContext fft;
    MhdDev p;
    hipStream_t stream = nullptr;
    // device allocations
    double *d_rj = nullptr, *d_sp_b = nullptr, *d_sp_f = nullptr;
    double *d_vr_b = nullptr, *d_vr_f = nullptr;
    bool fused_fft = true;        // one pencil kernel: FFT, products, FFT
    bool use_graph = true;        // replay the captured Adams-Bashforth step
    hipStream_t gstream = nullptr;   // capture stream (default priority)
    hipGraphExec_t gexec = nullptr;
    double *d_grid_b = nullptr, *d_grid_f = nullptr;   // grids when not fused
    double *d_consts = nullptr;   // r, ar1, ar2, d1, d2, g1, g3, t_icb, t_cmb, reft0, reft1, cor
    int *d_iconsts = nullptr;     // jk1, jk2, jl1, rj_of_rlm, rlm_of_rj, l_of_j
    int *d_l_of_j = nullptr;
    double *d_radius = nullptr;
    double *d_lu[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // vp vt temp bs bt
    int *d_piv[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    std::vector<int> rj_order;    // caller's local rj mode j -> l(l+1)+m; empty: the same
    double *d_press = nullptr;    // pressure (nr*J), its 3-band LU and pivots, built
    double *d_lu_p = nullptr;     // at the first pressure request
    int *d_piv_p = nullptr;
    double *d_scratch = nullptr;  // energies: full temperature + per-mode sums
    double *d_ms_out = nullptr;
};

// the boundary model the step supports, checked before any device work
int  mhd_check_model(int nr, const double *radius, const MhdCoefs &co, std::string *err);
int  mhd_create(MhdContext *c, int L, int nth, int nphi, int nr, const double *radius,
                const MhdCoefs &co, int leg_mode, hipStream_t s, std::string *err);
void mhd_free(MhdContext *c);
int  mhd_set_dynamobench_ic(MhdContext *c, int isig, std::string *err);
// dynamo_benchmark_2: the same temperature, the magnetic field to the centre
int  mhd_set_dynamobench2_ic(MhdContext *c, int isig, std::string *err);
int  mhd_set_seeded_ic(MhdContext *c, unsigned long long seed, double amp_u, double amp_b,
                       double amp_t, std::string *err);
int  mhd_initialize(MhdContext *c, std::string *err);
int  mhd_steps(MhdContext *c, int nsteps, std::string *err);
int  mhd_energies(MhdContext *c, double *out7, std::string *err);
int  mhd_pressure(MhdContext *c, double *host, std::string *err);
// output-step energy fluxes (up to kMaxEnergyFluxes flux ids) as spectra
constexpr int kMaxEnergyFluxes = 8;
int  mhd_energy_fluxes(MhdContext *c, int nflux, const int *ids, double *host,
                       std::string *err);
int  mhd_field_io(MhdContext *c, int comp, double *host, bool to_host, std::string *err);
// the same with the host array in the caller's rj mode order (mhd_set_rj_order)
int  mhd_field_io_caller(MhdContext *c, int comp, double *host, bool to_host, std::string *err);
int  mhd_set_rj_order(MhdContext *c, int J, const int *j_global, std::string *err);
// backward transform of chosen d_rj fields to the rtp grid (host array)
int  mhd_snapshot_rtp(MhdContext *c, int nvec, const int *vec, int nscl, const int *scl,
                      double *host, std::string *err);
int  mhd_set_option(MhdContext *c, const std::string &name, long long value, std::string *err);
// output files in the reference formats (sph_io.cpp)
double mhd_time(const MhdContext *c);
int  mhd_write_volume_ms(MhdContext *c, const std::string &path, std::string *err);
std::string mhd_restart_name(const std::string &prefix, int istep_fld);
int  mhd_write_restart(MhdContext *c, const std::string &path, std::string *err);
int  mhd_read_restart(MhdContext *c, const std::string &path, std::string *err);

// single-domain grid tables and the reference's grid files (sph_grids.cpp)
void sph_rj_mode_order(int L, int *j_global);
// host: the reference's ASCII field file of one domain (restart / snapshot)
int sph_write_field_file(const std::string &path, int rank, long long istep, double time,
                         double dt, long long nnod, int nfld, const int *ncomp,
                         const char *const *names, const double *const *comp, std::string *err);
int  sph_write_grid_files(const std::string &prefix, int L, int nth, int nphi, int nr,
                          const double *radius, std::string *err);

// closed forms of the Coriolis couplings (exposed for the host tests)
double gaunt_p1(int l, int m);   // K(1,0; l-1,m; l,m)
double elsasser_p1(int l, int m); // E(1,0; l,-m; l,m)
void coriolis_table(int L, double *cor, int *jk1, int *jk2, int *jl1);

}  // namespace calypso
