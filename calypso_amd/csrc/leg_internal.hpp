// leg_internal.hpp -- device-side context of the spherical-harmonic
// transform library (gfx950).  Not part of the public C ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <deque>
#include <string>
#include <vector>

#include "hip_check.hpp"

namespace calypso {

enum Precision : int { kF64 = 1, kF32 = 2, kBF16x3 = 3, kInt8x6 = 4 };

// One (order m, radial tile, row group) unit of Legendre work; the third
// kernel form runs nkt consecutive radial tiles from ktile on.
struct WorkItem {
    int m, ktile, rgroup, nkt;
};

struct LegGeometry {
    int L = 0;      // truncation degree
    int nth = 0;    // Gauss points
    int nr = 0;     // radial points
    int th_h = 0;   // base (southern) hemisphere rows, (nth+1)/2
    int J = 0;      // (L+1)^2
    int M = 0;      // 2L+1
};

struct LegPlan {
    int ncomp = 0, nvec = 0, nscl = 0;
    bool fwd = false;
    int KT = 1;             // radial points per work item
    int rpw = 3;            // row tiles per wave
    int mct1 = 1, mct2 = 1; // column-tile capacity of the kernel instance
    int pair = 0;           // fwd: orders m and L-m in one work item
    int tc = 64;            // fwd: colatitudes per LDS chunk
    int ver = 1;            // kernel form (3: two P tables, GEMM-ready LDS
                            // columns, radial tiles looped per work-group)
    int nw = 4;             // waves per work-group
    int pd = 2;             // K steps the P fragments are fetched ahead
    float tuned_ms = -1.f;  // autotuner timing of the chosen plan
    int n_items = 0;
    WorkItem *d_items = nullptr;
};

struct OzTables;   // int8 split-once form (leg_ozaki.hip)

struct LegContext {
    LegGeometry g;
    int ncomp = 0, nvec = 0, nscl = 0;   // field set given at creation
    int prec = kF64;
    int kstep = 4;          // MFMA K per instruction (4 for f64/f32, 32 bf16)
    int rt_bwd = 0;         // row tiles of 16 over the hemisphere
    int ks_fwd = 0;         // K steps over the hemisphere (fwd)

    double *d_radius = nullptr, *d_colat = nullptr, *d_weight = nullptr;
    int *d_lstack = nullptr;

    void *d_tab_bwd = nullptr, *d_tab_fwd = nullptr;
    long long *d_off_bwd = nullptr;   // [L+2] unit offsets per m
    long long *d_off_fwd = nullptr;   // [2(L+1)+1] unit offsets per (m,par)
    // third-form tables (P, dP only; K steps padded to the prefetch depth)
    void *d_tab_bwd3 = nullptr, *d_tab_fwd3 = nullptr;
    long long *d_off_bwd3 = nullptr, *d_off_fwd3 = nullptr;
    int ks_fwd3 = 0;
    double *d_rsin = nullptr;         // 1/sin(theta) at the hemisphere rows
    double *d_gfac = nullptr;         // g7, g6 normalisations of order m > 0
    size_t tab_bytes = 0;
    OzTables *oz = nullptr;           // int8 form: digit tables

    std::deque<LegPlan> plans;    // per (ncomp, nvec, nscl, direction)
    bool autotune = true;         // time candidate plans on first use

    // staging buffers for the host-pointer entry points
    double *d_rlm = nullptr, *d_rtm = nullptr;
    size_t cap_rlm = 0, cap_rtm = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0.f;
};

// host-side helpers (legendre.hip)
void gauss_points_host(int n, double *point, double *weight, double *colat);
int  leg_context_build(LegContext *c, int L, int nth, int nr,
                       const double *radius, const double *colat,
                       const double *weight, int ncomp, int nvec, int nscl,
                       int prec, std::string *err);
void leg_context_free(LegContext *c);
int  leg_backward_dev(LegContext *c, int ncomp, int nvec, int nscl,
                      const double *d_sp, double *d_vr, hipStream_t s,
                      std::string *err);
int  leg_forward_dev(LegContext *c, int ncomp, int nvec, int nscl,
                     const double *d_vr, double *d_sp, hipStream_t s,
                     std::string *err);
const LegPlan *leg_plan(LegContext *c, int ncomp, int nvec, int nscl, bool fwd,
                        std::string *err);
int  leg_plan_count(LegContext *c, int ncomp, int nvec, int nscl, bool fwd,
                    std::string *err);
int  leg_force_plan(LegContext *c, int ncomp, int nvec, int nscl, bool fwd, int idx,
                    std::string *err);
int  leg_tables_to_host(LegContext *c, double *P, double *dP,
                        std::string *err);

// int8 split-once form (leg_ozaki.hip): digit tables from the FP64 P, dP and
// P m/sin(theta) at the hemisphere rows; plans of kernel form 5
int  oz_build_tables(LegContext *c, const double *rP, const double *rdP, const double *rPM,
                     std::string *err);
void oz_free(LegContext *c);
void oz_plan_candidates(int ncomp, int nv, int ns, bool fwd, std::vector<LegPlan> *out);
int  oz_build_items(LegContext *c, LegPlan *p, bool fwd, std::string *err);
int  oz_launch(LegContext *c, const LegPlan &p, bool fwd, const double *in, double *out,
               hipStream_t s, std::string *err);

}  // namespace calypso
