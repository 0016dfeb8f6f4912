// pbh_fwd_core.h -- pieces shared by the FP32 and mixed-precision forward kernels:
// launch plan, Gaussian difference tables, and the deterministic combine of
// layer-split partial tiles.
#pragma once

#include "pbh_common.h"

namespace pbh {

constexpr int FT_TY = 32;        // dose tile rows (y) per workgroup (FP32; the mixed forms also 64)
constexpr int FT_TX = 128;       // dose tile columns (x) per workgroup
constexpr int FT_THREADS = 256;  // 4 waves; two workgroups share a CU
constexpr int FT_TAB = 1024;     // entries per axis table
constexpr int FT_KYP = 64;       // mixed forms: stage A K (spots in y) padded to 2 steps of 32
// mixed forms: 16-bit (hi, lo) weights of one (layer, spot chunk), [hi|lo][i < 48][j < 64]
constexpr int FT_WSPLIT = 2 * PBH_SPOT_CHUNK * FT_KYP;

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
// makes those reads consecutive.
__device__ __forceinline__ int table_slot(const AxisTable& g, int t)
{
    return g.half ? (t & 1) * g.half + (t >> 1) : t;
}

// Start of the odd entries of a parity-interleaved table of n entries: the first
// slot at or after the (n + 1) / 2 even ones that is 14 mod 32. LDS serves a
// wave's read in two 32-lane passes over 32 four-byte banks; where a pass reads
// even and odd entries (the adjoint's gathers, the FP32 forward's y operand),
// this offset puts its 32 dwords on 32 distinct banks.
__host__ __device__ constexpr int table_half(int n)
{
    return (n + 1) / 2 + ((14 - (n + 1) / 2 % 32) % 32 + 32) % 32;
}

// LDS slots a parity-interleaved table of n entries spans.
__host__ __device__ constexpr int table_slots(int n)
{
    return table_half(n) + n / 2 + 1;
}

// Host-side launch plan of a forward call.
//
// Layer chunks: the active layers of a depth slice (IDD != 0, in layer order)
// are cut into chunks of at most `chunk` layers, and chunk c of a dose tile is
// computed by its own workgroup (blocks c * tiles + tile). Chunk 0 writes the
// dose; chunk c >= 1 writes a partial tile, and combine_kernel adds the
// partials to the dose in chunk order, so results are deterministic and no
// workgroup waits for another. Chunks keep every workgroup to a bounded amount
// of work although shallow slices have many more active layers than deep ones.
// One active layer of a depth slice, as the mixed forward uses it: layer index,
// c = IDD / (2 pi s^2), k2 = log2(e) / (2 s^2), window half-width FT_WIN_SIGMAS s (cm).
struct LayerEntry {
    int lay;
    float c, k2, cut;
};

struct FwdPlan {
    int chunk;          // at most this many active layers per workgroup (>= 1)
    int nchunk;         // workgroups per dose tile: ceil(L / chunk)
    int table;          // 1: factors gathered from per-(layer, depth) tables
    AxisTable tx, ty;
    float* part;        // [nchunk - 1][tiles][tile_rows * FT_TX] partial tiles (nchunk > 1)
    unsigned short* wsplit;   // mixed forms: [L][chunk y][chunk x][FT_WSPLIT] split weights
    // FP16 forms: per (layer, chunk y, chunk x) {2^e, b}: the split weights are
    // w 2^-e with max |w 2^-e| in [1, 2), and b bounds their column sums
    // max_i sum_j |w_ji 2^-e|, so |c Gy W| <= c 2^e b for Gaussian factors <= 1
    float2* wpar;
    // mixed forms: the active layers of every slice in layer order, [nz][L] (first
    // nlayers[z] of row z), written by the weight-split launch
    LayerEntry* layers;
    int* nlayers;
};

// Table entry t of an axis (t = Q i - P a + P (T - 1)) for base offset C:
// d = C + (P (T - 1) - t) u.
__device__ __forceinline__ double table_distance(const AxisTable& g, double C, int T, int t)
{
    return C + (double)(g.P * (T - 1) - t) * g.u;
}

// The same distances in FP32 without cancellation: d(t) = (k0 - t + f) u with
// k0 = rint(C / u) + P (T - 1) an exact integer and |f| <= 1/2, so d carries
// only the roundings of (k0 - t + f) and of the product (relative 2^-23) where
// C + (P (T - 1) - t) u in FP32 would lose the digits of C. A factor
// exp2(-d^2 k2) then differs from the FP64-argument value by at most
// ~3e-7 x ln 2 relative (x = d^2 k2 <= 32 inside the window), which is below
// 1.2e-7 of the factor's peak everywhere.
struct TableRow {
    int k0;
    float f, u;
};

__device__ __forceinline__ TableRow table_row(const AxisTable& g, double C, int T)
{
    const double cu = C / g.u;
    const double r = rint(cu);
    return TableRow{(int)r + g.P * (T - 1), (float)(cu - r), (float)g.u};
}

__device__ __forceinline__ float table_distance_f(const TableRow& r, int t)
{
    return ((float)(r.k0 - t) + r.f) * r.u;
}

// A thread's share of one parity-interleaved axis table (P = 2, half set): entries
// t = threadIdx.x + FT_THREADS k (k < N) of the NT entries, slot(t) = slot(threadIdx.x)
// + FT_THREADS / 2 k. The squared distances depend on the tile and spot chunk only,
// so they are formed once per chunk; each layer then costs one product and one
// exp2 per entry. Entries equal the general build's exp2(-(d * d) * k2).
template <int NT>
struct TabShare {
    static constexpr int N = (NT + FT_THREADS - 1) / FT_THREADS;
    float d2[N];
    int s0;     // slot of entry threadIdx.x
    // entry threadIdx.x + FT_THREADS k exists
    __device__ static __forceinline__ bool ok(int k)
    {
        return FT_THREADS * (k + 1) <= NT || (int)threadIdx.x + FT_THREADS * k < NT;
    }
    __device__ __forceinline__ void init(const AxisTable& g, const TableRow& row)
    {
        s0 = table_slot(g, (int)threadIdx.x);
#pragma unroll
        for (int k = 0; k < N; ++k) {
            const float d = table_distance_f(row, (int)threadIdx.x + FT_THREADS * k);
            d2[k] = d * d;
        }
    }
    // FP32 entries
    __device__ __forceinline__ void build_f32(float* tab, float k2) const
    {
        const float nk2 = -k2;
#pragma unroll
        for (int k = 0; k < N; ++k)
            if (ok(k)) tab[s0 + FT_THREADS / 2 * k] = __builtin_amdgcn_exp2f(d2[k] * nk2);
    }
};

// Writes a 32 x 32 block of the dose tile held in the C layout of a 32x32 MFMA
// (row 8 (r/4) + 4 (lane/32) + r%4, column lane%32) at (rm, cn) in the tile:
// chunk 0 to the dose (rows and columns inside the grid), chunk c >= 1 to its
// partial tile (all 32 x 128 entries).
__device__ inline void store_block(const FwdArgs& a, const FwdPlan& pl, int tile, int c,
                                   int z, int y0, int x0, int rm, int cn, v16f acc, size_t tiles)
{
    const int lane = threadIdx.x & 63;
    const int xl = cn + (lane & 31);
    if (c > 0) {
        float* mine = pl.part + ((size_t)(c - 1) * tiles + tile) * ((size_t)a.tile_rows * FT_TX);
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = rm + 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            mine[row * FT_TX + xl] = acc[r];
        }
        return;
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

// Active layers (IDD != 0) of depth slice z, counted by one workgroup.
__device__ inline int count_active(const FwdArgs& a, int z, int* scratch)
{
    int n = 0;
    for (int l = threadIdx.x; l < a.L; l += blockDim.x) n += a.idd[(size_t)l * a.nz + z] != 0.0;
    for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o, 64);
    if ((threadIdx.x & 63) == 0) scratch[threadIdx.x >> 6] = n;
    __syncthreads();
    int t = 0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += scratch[w];
    return t;
}

// Rows of one dose tile (from row y0, column x0, NIT x 1024 voxels) += NP partial tiles
// in chunk order, for a 256-thread workgroup: thread t takes voxels 4 t + 1024 it (it <
// NIT), loads all of them first, then adds and stores. The tile's columns lie inside the
// grid and the dose rows are 16-byte aligned; rows beyond the grid are skipped.
template <int NIT, int NP>
__device__ __forceinline__ void combine_rows(const FwdArgs& a, const float* part0, size_t cstride, int z, int y0,
                                             int x0)
{
    float4 v[NIT], p[NP][NIT];
    int n = 0;
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        const int e = 4 * (int)threadIdx.x + 1024 * it;
        const int y = y0 + e / FT_TX;
        if (y >= a.ny) break;
        n = it + 1;
        v[it] = *(const float4*)(a.dose + ((size_t)z * a.ny + y) * a.nx + x0 + e % FT_TX);
#pragma unroll
        for (int c = 0; c < NP; ++c) p[c][it] = *(const float4*)(part0 + (size_t)c * cstride + e);
    }
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
        if (it >= n) break;
        const int e = 4 * (int)threadIdx.x + 1024 * it;
        float4 r = v[it];
#pragma unroll
        for (int c = 0; c < NP; ++c) {
            r.x += p[c][it].x;
            r.y += p[c][it].y;
            r.z += p[c][it].z;
            r.w += p[c][it].w;
        }
        *(float4*)(a.dose + ((size_t)z * a.ny + y0 + e / FT_TX) * a.nx + x0 + e % FT_TX) = r;
    }
}

// dose tile += partial tiles of chunks 1 .. nchunk(z) - 1, in chunk order.
// One workgroup of 256 threads per dose tile.
__device__ inline void combine_body(const FwdArgs& a, const FwdPlan& pl, int tile)
{
    __shared__ int cnt[4];
    const size_t tiles = (size_t)a.tiles_x * a.tiles_y * a.nz;
    const int tx = tile % a.tiles_x;
    const int ty = (tile / a.tiles_x) % a.tiles_y;
    const int z = tile / (a.tiles_x * a.tiles_y);
    const int nact = pl.nlayers ? pl.nlayers[z] : count_active(a, z, cnt);
    const int nch = nact > pl.chunk ? (nact + pl.chunk - 1) / pl.chunk : 1;
    if (nch <= 1) return;
    const int TR = a.tile_rows;
    const float* part0 = pl.part + (size_t)tile * ((size_t)TR * FT_TX);
    const size_t cstride = tiles * ((size_t)TR * FT_TX);
    const bool vec = (a.nx & 3) == 0 && tx * FT_TX + FT_TX <= a.nx && ((uintptr_t)a.dose & 15) == 0;
    if (vec && blockDim.x == 256 && nch <= 4) {
        // whole 128-column rows inside the grid, four consecutive voxels per thread; the
        // thread's float4s of the dose and of every partial are loaded before any store,
        // so that their latencies overlap (one load, add and store at a time paid one
        // memory latency per four voxels)
        if (TR == 64) {
            if (nch == 2) combine_rows<8, 1>(a, part0, cstride, z, ty * TR, tx * FT_TX);
            else if (nch == 3) combine_rows<8, 2>(a, part0, cstride, z, ty * TR, tx * FT_TX);
            else combine_rows<8, 3>(a, part0, cstride, z, ty * TR, tx * FT_TX);
        } else {
            if (nch == 2) combine_rows<4, 1>(a, part0, cstride, z, ty * TR, tx * FT_TX);
            else if (nch == 3) combine_rows<4, 2>(a, part0, cstride, z, ty * TR, tx * FT_TX);
            else combine_rows<4, 3>(a, part0, cstride, z, ty * TR, tx * FT_TX);
        }
        return;
    }
    if (vec) {
        // whole 128-column rows inside the grid: four consecutive voxels per thread
        for (int e = 4 * threadIdx.x; e < TR * FT_TX; e += 4 * blockDim.x) {
            const int y = ty * TR + e / FT_TX, x = tx * FT_TX + e % FT_TX;
            if (y >= a.ny) break;
            float4* d = (float4*)(a.dose + ((size_t)z * a.ny + y) * a.nx + x);
            float4 v = *d;
            for (int c = 1; c < nch; ++c) {
                const float4 p = *(const float4*)(part0 + (size_t)(c - 1) * cstride + e);
                v.x += p.x;
                v.y += p.y;
                v.z += p.z;
                v.w += p.w;
            }
            *d = v;
        }
        return;
    }
    for (int e = threadIdx.x; e < TR * FT_TX; e += blockDim.x) {
        const int y = ty * TR + e / FT_TX, x = tx * FT_TX + e % FT_TX;
        if (y >= a.ny || x >= a.nx) continue;
        float* d = a.dose + ((size_t)z * a.ny + y) * a.nx + x;
        float v = *d;
        for (int c = 1; c < nch; ++c) v += part0[(size_t)(c - 1) * cstride + e];
        *d = v;
    }
}

// One voxel of the dose objective J = sum wv (D - Dp)^2 (FP64 terms): adds its term
// to acc and returns its residual dJ/dD = 2 wv (D - Dp) as FP32. Shared by the
// objective kernel and the combine that forms the objective of a planning step, so
// that both give the same residual bit for bit.
__device__ __forceinline__ float objective_term(float d, float p, float wf, double& acc)
{
    const double r = (double)d - (double)p;
    const double w = (double)wf;
    acc += w * r * r;
    return (float)(2.0 * w * r);
}

// The objective fused into the layer-chunk combine (one planning step): R and the
// partial sums of J, one per combine workgroup (FT_OBJ_PARTS per dose tile).
struct ObjFuse {
    const float* presc;    // Dp (nz, ny, nx)
    const float* vox_w;    // wv (nz, ny, nx), or null: all 1
    float* R;              // out: 2 wv (D - Dp)
    double* jpart;         // out: tiles * FT_OBJ_PARTS partial sums of J
};
constexpr int FT_OBJ_PARTS = 4;    // row bands of a dose tile, one combine workgroup each

// Row band `band` of dose tile `tile`: the dose (plus the partial tiles of chunks
// 1 .. nchunk(z) - 1 in chunk order, written back when there are any), then R and this
// workgroup's partial sum of J over the band's voxels inside the grid.
__device__ inline void combine_objective_body(const FwdArgs& a, const FwdPlan& pl, const ObjFuse& f, int tile,
                                              int band)
{
    __shared__ int cnt[4];
    __shared__ double red[4];
    const size_t tiles = (size_t)a.tiles_x * a.tiles_y * a.nz;
    const int tx = tile % a.tiles_x;
    const int ty = (tile / a.tiles_x) % a.tiles_y;
    const int z = tile / (a.tiles_x * a.tiles_y);
    const int nact = pl.nlayers ? pl.nlayers[z] : count_active(a, z, cnt);
    const int nch = nact > pl.chunk ? (nact + pl.chunk - 1) / pl.chunk : 1;
    const int TR = a.tile_rows, BR = TR / FT_OBJ_PARTS;
    const float* part0 = nch > 1 ? pl.part + (size_t)tile * ((size_t)TR * FT_TX) : nullptr;
    const size_t cstride = tiles * ((size_t)TR * FT_TX);
    double acc = 0.0;
    const bool vec = (a.nx & 3) == 0 && tx * FT_TX + FT_TX <= a.nx &&
                     (((uintptr_t)a.dose | (uintptr_t)f.presc | (uintptr_t)f.vox_w | (uintptr_t)f.R) & 15) == 0;
    if (vec) {
        // whole 128-column rows inside the grid: four consecutive voxels per thread
        for (int e = BR * band * FT_TX + 4 * threadIdx.x; e < BR * (band + 1) * FT_TX; e += 4 * blockDim.x) {
            const int y = ty * TR + e / FT_TX, x = tx * FT_TX + e % FT_TX;
            if (y >= a.ny) break;
            const size_t i = ((size_t)z * a.ny + y) * a.nx + x;
            float4* d = (float4*)(a.dose + i);
            float4 v = *d;
            if (nch > 1) {
                for (int c = 1; c < nch; ++c) {
                    const float4 p = *(const float4*)(part0 + (size_t)(c - 1) * cstride + e);
                    v.x += p.x;
                    v.y += p.y;
                    v.z += p.z;
                    v.w += p.w;
                }
                *d = v;
            }
            const float4 p = *(const float4*)(f.presc + i);
            const float4 w = f.vox_w ? *(const float4*)(f.vox_w + i) : make_float4(1.0f, 1.0f, 1.0f, 1.0f);
            float4 r;
            r.x = objective_term(v.x, p.x, w.x, acc);
            r.y = objective_term(v.y, p.y, w.y, acc);
            r.z = objective_term(v.z, p.z, w.z, acc);
            r.w = objective_term(v.w, p.w, w.w, acc);
            *(float4*)(f.R + i) = r;
        }
    } else {
        for (int e = BR * band * FT_TX + threadIdx.x; e < BR * (band + 1) * FT_TX; e += blockDim.x) {
            const int y = ty * TR + e / FT_TX, x = tx * FT_TX + e % FT_TX;
            if (y >= a.ny || x >= a.nx) continue;
            const size_t i = ((size_t)z * a.ny + y) * a.nx + x;
            float v = a.dose[i];
            if (nch > 1) {
                for (int c = 1; c < nch; ++c) v += part0[(size_t)(c - 1) * cstride + e];
                a.dose[i] = v;
            }
            f.R[i] = objective_term(v, f.presc[i], f.vox_w ? f.vox_w[i] : 1.0f, acc);
        }
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[w];
        f.jpart[(size_t)tile * FT_OBJ_PARTS + band] = t;
    }
}

constexpr int FT_LB = 64;    // layers gathered per pass

// Window of the contraction: a Gaussian factor exp(-d^2 / (2 s^2)) below
// 2^-32 of its peak (|d| > sqrt(64 ln 2) s = 6.66 s) is dropped. Dropped taps
// change a sum by less than 2^-32 of its largest term, far below FP32 rounding.
constexpr double FT_WIN_SIGMAS = 6.6604;     // sqrt(2 * 32 * ln 2), rounded up

// Active layers of one depth slice, in layer order, with their FP64-derived
// scalars: c = IDD / (2 pi s^2), k2 = log2(e) / (2 s^2) (FP32 and FP64), the
// window half-width cut = FT_WIN_SIGMAS s (cm) and the distance to the layer's
// Bragg peak (mixed "_peak" forms).
struct LayerList {
    int lay[FT_LB];
    float c[FT_LB];
    float k2[FT_LB];
    double k2d[FT_LB];
    float cut[FT_LB];
    float zpk[FT_LB];
    int wcount[FT_THREADS / 64];
};

// Spots of a chunk that lie within `cut` of the coordinate range [lo, hi] (all
// in spot-index units relative to the chunk's first spot), as nstep steps of
// `step` spots from k0. k0 is a multiple of 8 and the window ends within kpad
// (a multiple of 8, >= n); one spot of margin on each side absorbs rounding.
// nstep = 0 when no spot is in reach.
struct SpotWindow { int k0, nstep; };

// First and last spot (of n) within `cut` of [lo, hi], with one spot of margin;
// first > last when none is.
__device__ __forceinline__ void spot_range(float lo, float hi, float cut, int n, int& first, int& last)
{
    const float f = fminf(fmaxf(lo - cut - 1.0f, -1.0f), (float)(n + 1));
    const float l = fmaxf(fminf(hi + cut + 1.0f, (float)n), -2.0f);
    first = max((int)ceilf(f), 0);
    last = min((int)floorf(l), n - 1);
}

__device__ __forceinline__ SpotWindow spot_window(float lo, float hi, float cut, int n, int step, int kpad)
{
    int first, last;
    spot_range(lo, hi, cut, n, first, last);
    SpotWindow w{0, 0};
    if (first > last) return w;
    w.k0 = first & ~7;
    w.nstep = (last + 1 - w.k0 + step - 1) / step;
    if (w.k0 + w.nstep * step > kpad) w.k0 = kpad - w.nstep * step;
    return w;
}


// Gathers layers l = lb + t (t < FT_LB, l < L) whose IDD at depth z is non-zero
// into ll (stable order, wave ballots + prefix); returns their count.
// Contains two barriers: every thread of the workgroup must call it.
__device__ inline int collect_layers(const FwdArgs& a, int lb, int z, float zc,
                                     const float* r_peak, LayerList& ll)
{
    const int t = threadIdx.x;
    const int lane = t & 63, wave = t >> 6;
    const int l = lb + t;
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
        ll.cut[p] = (float)(FT_WIN_SIGMAS * sd);
        ll.zpk[p] = r_peak ? fabsf(zc - r_peak[l]) : 0.0f;
    }
    __syncthreads();
    return total;
}

// Active layers lb .. lb + FT_LB - 1 (in active order) of depth slice z from the
// precomputed lists into ll; returns their count. Contains one barrier: every
// thread of the workgroup must call it.
__device__ inline int copy_layers(const FwdPlan& pl, int L, int z, int lb, float zc, const float* r_peak,
                                  LayerList& ll)
{
    const int n = min(FT_LB, pl.nlayers[z] - lb);
    const int t = threadIdx.x;
    if (t < n) {
        const LayerEntry e = pl.layers[(size_t)z * L + lb + t];
        ll.lay[t] = e.lay;
        ll.c[t] = e.c;
        ll.k2[t] = e.k2;
        ll.cut[t] = e.cut;
        ll.zpk[t] = r_peak ? fabsf(zc - r_peak[e.lay]) : 0.0f;
    }
    __syncthreads();
    return n;
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

// The workgroups of a launch are dispatched to the 8 XCDs in turn (workgroup b to XCD
// b mod 8), and each XCD has its own L2. xcd_unit maps launch index b to the work
// unit (slice z, share s) of a launch of split workgroups per slice so that all shares
// of a slice run on one XCD and read the slice's data through one L2. Round r gives
// the XCDs slices 8 r .. 8 r + 7, in reverse XCD order on odd rounds, so that the work
// per slice, which falls with depth, evens out over the XCDs. The grid is
// xcd_grid(nz, split); indices with z >= nz have no work.
constexpr int kXcds = 8;

__host__ __device__ inline int xcd_grid(int nz, int split) { return kXcds * split * cdiv(nz, kXcds); }

__device__ inline void xcd_unit(int b, int split, int& z, int& s)
{
    const int k = b / kXcds, x = b % kXcds, r = k / split;
    s = k % split;
    z = kXcds * r + ((r & 1) ? kXcds - 1 - x : x);
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
                     int layer_chunk, pbh::FwdPlan* plan);
size_t pbh_forward_ws_bytes(const pbh::FwdArgs& a, int layer_chunk);
int pbh_auto_chunk(const pbh::FwdArgs& a);
// The layer-chunk combine of a forward call; with `fuse`, the combine over every dose
// tile that also forms the objective (combine_objective_body).
int pbh_launch_combine(const pbh::FwdArgs& a, const pbh::FwdPlan& pl, hipStream_t s,
                       const pbh::ObjFuse* fuse = nullptr);
// A forward call (pbh_forward_shaped) whose combine forms the objective (`fuse`, not
// null); *njpart receives the number of J partial sums written.
int pbh_forward_fused(const PbhGrid* g, const PbhLattice* lat, const float* w, const double* idd,
                      const double* sigma, float* dose, int precision, int layer_chunk, int tile_rows,
                      void* workspace, size_t workspace_bytes, hipStream_t s, const pbh::ObjFuse* fuse,
                      int* njpart);
int pbh_device_cus();
