// pbh_fwd_core.h -- pieces shared by the FP32 and mixed-precision forward kernels:
// launch plan, Gaussian difference tables, and the deterministic combine of
// layer-split partial tiles.
#pragma once

#include "pbh_common.h"

namespace pbh {

constexpr int FT_TY = 32;        // dose tile rows (y) per workgroup
constexpr int FT_TX = 128;       // dose tile columns (x) per workgroup
constexpr int FT_THREADS = 256;  // 4 waves; two workgroups share a CU
constexpr int FT_TAB = 1024;     // entries per axis table

// Regular-grid geometry of one axis: voxel a (in a tile) and spot i (in a chunk)
// are d = C + (P a - Q i) u apart, so the Gaussian of a (layer, depth) takes
// only P (T-1) + Q (K-1) + 1 distinct values over a T x K tile.
struct AxisTable {
    int P, Q;
    double u;
    int half;           // P == 2: entries stored parity-interleaved, odd ones from `half`
};

// LDS slot of table entry t. With P == 2 the lanes of a half-wave read entries
// of one parity at stride 2; storing even and odd entries in separate halves
// makes those reads consecutive (no bank conflicts).
__device__ __forceinline__ int table_slot(const AxisTable& g, int t)
{
    return g.half ? (t & 1) * g.half + (t >> 1) : t;
}

// Host-side launch plan of a forward call.
struct FwdPlan {
    int split;          // workgroups per dose tile sharing its layers (1, 2, 4 or 8)
    int table;          // 1: factors gathered from per-(layer, depth) tables
    AxisTable tx, ty;
    float* part;        // [split][tiles][FT_TY * FT_TX] partial tiles (split > 1)
    unsigned* ctr;      // [tiles] arrival counters, multiples of 8 between calls
};

// Table entry t of an axis (t = Q i - P a + P (T - 1)) for base offset C:
// d = C + (P (T - 1) - t) u.
__device__ __forceinline__ double table_distance(const AxisTable& g, double C, int T, int t)
{
    return C + (double)(g.P * (T - 1) - t) * g.u;
}

// Writes a 32 x 32 block of the dose tile held in the C layout of a 32x32 MFMA
// (row 8 (r/4) + 4 (lane/32) + r%4, column lane%32) at (rm, cn) in the tile.
// With a layer split, every workgroup stores its partial tile, and the last of
// the `split` workgroups of the tile to arrive (atomic counter) adds all
// partials in split order 0..S-1 and writes the dose, so the result does not
// depend on arrival order. Counters are never reset: each arrival adds 8/S
// (S in {2, 4, 8}), so a counter is a multiple of 8 between calls whatever S
// was used, and the last arrival is the one that brings it to a multiple of 8.
__device__ inline void store_block(const FwdArgs& a, const FwdPlan& pl, int tile, int s,
                                   int z, int y0, int x0, int rm, int cn, v16f acc, size_t tiles)
{
    __shared__ int last_flag;
    const int lane = threadIdx.x & 63;
    const int xl = cn + (lane & 31);
    if (pl.split > 1) {
        float* mine = pl.part + ((size_t)s * tiles + tile) * (FT_TY * FT_TX);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = rm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            mine[row * FT_TX + xl] = acc[r];
        }
        __threadfence();
        __syncthreads();
        if (threadIdx.x == 0) {
            const unsigned inc = 8u / (unsigned)pl.split;
            const unsigned old = atomicAdd(&pl.ctr[tile], inc);
            last_flag = ((old + inc) & 7u) == 0u;
        }
        __syncthreads();
        if (!last_flag) return;
        __threadfence();
        v16f sum = {};
        for (int q = 0; q < pl.split; ++q) {
            if (q == s) {
                sum += acc;
            } else {
                const float* p = pl.part + ((size_t)q * tiles + tile) * (FT_TY * FT_TX);
#pragma unroll
                for (int r = 0; r < 16; ++r) {
                    const int row = rm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
                    sum[r] += p[row * FT_TX + xl];
                }
            }
        }
        acc = sum;
    }
    const int xB = x0 + xl;
    if (xB < a.nx) {
        float* base = a.dose + (size_t)z * a.ny * a.nx + xB;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int y = y0 + rm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            if (y < a.ny) base[(size_t)y * a.nx] = acc[r];
        }
    }
}

constexpr int FT_LB = 256;   // layers gathered per pass

// Active layers of one depth slice, in layer order, with their FP64-derived
// scalars: c = IDD / (2 pi s^2), k2 = log2(e) / (2 s^2) (FP32 and FP64) and the
// distance to the layer's Bragg peak (mixed "_peak" forms).
struct LayerList {
    int lay[FT_LB];
    float c[FT_LB];
    float k2[FT_LB];
    double k2d[FT_LB];
    float zpk[FT_LB];
    int wcount[FT_THREADS / 64];
};

// Gathers layers l = lb + t * step (t < FT_LB, l < L) whose IDD at depth z is
// non-zero into ll (stable order, wave ballots + prefix); returns their count.
// Contains two barriers: every thread of the workgroup must call it.
__device__ inline int collect_layers(const FwdArgs& a, int lb, int step, int z, float zc,
                                     const float* r_peak, LayerList& ll)
{
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int l = lb + t * step;
    double iddv = 0.0, sg = 0.0;
    bool act = false;
    if (t < FT_LB && l < a.L) {
        iddv = a.idd[(size_t)l * a.nz + z];
        sg = a.sigma[(size_t)l * a.nz + z];
        act = iddv != 0.0;
    }
    const unsigned long long m = __ballot(act);
    const int before = __popcll(m & ((1ull << lane) - 1ull));
    if (lane == 0) ll.wcount[wave] = __popcll(m);
    __syncthreads();
    int off = 0, total = 0;
#pragma unroll
    for (int w = 0; w < FT_THREADS / 64; ++w) {
        const int c = ll.wcount[w];
        off += w < wave ? c : 0;
        total += c;
    }
    if (act) {
        const int p = off + before;
        const LayerScalars sc = layer_scalars(iddv, sg);
        const double sd = sg > kSigmaFloor ? sg : kSigmaFloor;
        ll.lay[p] = l;
        ll.c[p] = sc.c;
        ll.k2[p] = sc.k2;
        ll.k2d[p] = kLog2e / (2.0 * sd * sd);
        ll.zpk[p] = r_peak ? fabsf(zc - r_peak[l]) : 0.0f;
    }
    __syncthreads();
    return total;
}

// Padded spot-chunk width (a multiple of 4, at most PBH_SPOT_CHUNK) for a lattice.
__host__ __device__ inline int chunk_pad(int nsx, int nsy)
{
    int m = 0;
    const int cx = chunk_count(nsx), cy = chunk_count(nsy);
    for (int c = 0; c < cx; ++c) m = max(m, chunk_begin(nsx, cx, c + 1) - chunk_begin(nsx, cx, c));
    for (int c = 0; c < cy; ++c) m = max(m, chunk_begin(nsy, cy, c + 1) - chunk_begin(nsy, cy, c));
    return (m + 3) & ~3;
}

// ---- batches of independent problems (robust scenarios) ----
// The workgroups of one batched launch are split among its items by a prefix
// table: item k owns workgroups wg0[k] .. wg0[k+1]-1 and runs them as the
// single-problem kernel would run blocks 0 .. wg0[k+1]-wg0[k]-1.

// Largest k with wg0[k] <= blk (wg0 ascending, wg0[0] = 0). blk is uniform in
// the workgroup, so the search runs on scalar loads.
__device__ __forceinline__ int batch_item(const int* __restrict__ wg0, int n, int blk)
{
    int lo = 0, hi = n - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (wg0[mid] <= blk) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

struct BatchFwdItem {
    FwdArgs a;
    FwdPlan pl;
    double oz, dz;
    const float* r_peak;   // per-layer Bragg-peak depth ("_peak" forms) or null
};

struct BatchPeakItem {
    const double* idd;
    int L, nz;
    double oz, dz;
    float* r_peak;
};

}  // namespace pbh

// Host helpers (defined in pbh_forward.hip).
int pbh_plan_forward(const pbh::FwdArgs& a, int kx_pad, int ky_pad, void* ws, size_t ws_bytes,
                     int split_override, pbh::FwdPlan* plan);
size_t pbh_forward_ws_bytes(const pbh::FwdArgs& a, int split);
int pbh_auto_split(const pbh::FwdArgs& a);
