// pbh_batch.h -- descriptors and launchers shared by the single-problem adjoint
// and the batched (robust-scenario) entry points.
#pragma once

#include "pbh_fwd_core.h"

namespace pbh {

// Adjoint of one problem (pbh_adjoint.hip).
struct AdjArgs {
    FwdArgs f;
    const float* R;        // (nz, ny, nx)
    float* part;           // [nz][L][nsy][nsx] per-slice c P
    double* grad_idd;      // (L, nz) or null
    double* grad_sigma;    // (L, nz) or null
    int split;
    int tab;
    AxisTable tx, ty;
};

// Weighted-quadratic objective of one problem: R = 2 wv (D - Dp), partial sums
// of wv (D - Dp)^2 per workgroup, J = their fixed-order sum.
struct BatchObjItem {
    const float* D;
    const float* Dp;
    const float* wv;       // null: weights 1
    float* R;
    double* part;          // [OBJ_BATCH_BLOCKS]
    double* j_out;
    size_t n;
};

// Depth reduction of one problem's per-slice weight-gradient partials.
struct BatchRedItem {
    const float* part;     // [nz][n]: slice z, layer l's block written only where IDD_l(z) != 0
    const double* idd;     // [L][nz]
    float* grad_w;         // [n]
    size_t n;
    int nz;
    int per_layer;         // nsy * nsx
    // optional: one more workgroup (after the reduction's own) writes j_out[0] = the
    // fixed-order sum of the njpart objective partials jpart (null j_out: none)
    const double* jpart;
    int njpart;
    double* j_out;
};

constexpr int OBJ_BATCH_BLOCKS = 256;   // objective workgroups per item
constexpr int RED_ELEMS = 64;           // dJ/dw elements per depth-reduction workgroup

}  // namespace pbh

// Host-side builders and launchers (pbh_forward.hip, pbh_mixed.hip, pbh_adjoint.hip).
pbh::FwdArgs make_fwd_args(const PbhGrid* g, const PbhLattice* lat, const float* w,
                           const double* idd, const double* sigma, float* dose,
                           int tile_rows = pbh::FT_TY);
int pbh_make_adj_args(const PbhGrid* g, const PbhLattice* lat, const float* w, const double* idd,
                      const double* sigma, const float* R, float* part, double* grad_idd,
                      double* grad_sigma, int split, pbh::AdjArgs* out);
int pbh_adj_default_split(const PbhGrid* g, const PbhLattice* lat);
int pbh_launch_fwd_fp32_batch(const pbh::BatchFwdItem* items, const int* wg0, int n, int total, int scp,
                              bool tab, bool q5, hipStream_t s);
bool fp32_q5(const pbh::FwdPlan& pl);
int pbh_launch_fwd_mixed_batch(const pbh::BatchFwdItem* items, const int* wg0, int n, int total,
                               int precision, int kb, bool tab, bool q5, int rows, unsigned* counters,
                               hipStream_t s);
int pbh_launch_wsplit_batch(const pbh::BatchFwdItem* items, int n, int max_chunks, int precision, hipStream_t s);
int pbh_mixed_slot_mode(const pbh::FwdPlan& pl);
bool pbh_adj_mixed_q5(const pbh::AdjArgs& A);
bool pbh_adj_mixed_fits(const pbh::AdjArgs& A);
int pbh_launch_adj_mixed(const pbh::AdjArgs& A, int precision, int nz, hipStream_t s);
int pbh_launch_adj_mixed_batch(const pbh::AdjArgs* items, const int* wg0, int n, int total, int precision, bool q5,
                               hipStream_t s);
int pbh_launch_combine_batch(const pbh::BatchFwdItem* items, const int* wg0, int n, int total, hipStream_t s);
int pbh_launch_peak_batch(const pbh::BatchPeakItem* items, int n, hipStream_t s);
int pbh_launch_adj_batch(const pbh::AdjArgs* items, const int* wg0, int n, int total, bool tab, bool sig,
                         bool q5, hipStream_t s);
int pbh_launch_obj_batch(const pbh::BatchObjItem* items, int n, hipStream_t s);
int pbh_launch_reduce_batch(const pbh::BatchRedItem* items, const int* wg0, int n, int total, hipStream_t s);
