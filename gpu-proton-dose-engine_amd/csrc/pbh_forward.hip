// pbh_forward.hip -- fused forward dose contraction, FP32 MFMA (gfx950), and the
// host-side planning shared with the mixed-precision forms.
//
// For one depth slice z, a dose tile D[y0:y0+32, x0:x0+128] is
//     D = sum_l sum_chunks (c_lz * Gy_lz[:, J] W_l[J, I]) Gx_lz[:, I]^T
// with c_lz = IDD_l(z) / (2 pi s^2). A workgroup (4 waves, two per CU) owns one
// tile and one chunk of the slice's active layers, keeping D in MFMA
// accumulators, so no per-layer intermediate reaches HBM (the reference
// materialises Gx, Gy, A, T and Dl per layer: engines.py:261-274). Chunk 0
// writes the dose; later chunks write partial tiles that combine_kernel adds in
// chunk order.
//
//   stage A  U[32 x 48]  = Gy[32 x J] W[J x I]      v_mfma_f32_16x16x4f32
//            Gy is the A operand, W (LDS) the B operand.
//   stage B  D[32 x 128] += (c U)[32 x I] Gx^T       v_mfma_f32_32x32x2f32
//            wave w: columns 32w..32w+31; (c U) from LDS, Gx the B operand.
//
// Gaussian factors. On regular grids whose voxel and spot pitches have a small
// integer ratio (dx = P u, ds = Q u: 2 mm voxels and 5 mm spots give P = 2,
// Q = 5), voxel-spot distances within a tile take only P (T-1) + Q (K-1) + 1
// values, so each (layer, depth) builds one small table per axis in LDS (FP32
// argument with an exact integer part, one exp2 each) and the MFMA operands are
// gathered from it: ~490 exponentials per 128 x 48 tile instead of ~6000. Other
// grids evaluate every factor directly (FP64 distance, FP32 exponent).
//
// Each layer's contribution is summed in fresh accumulators and added to the
// running dose once (rounding chain per voxel: K steps + layers). Tables, W and
// U are double-buffered in LDS: one barrier per layer.
#include <algorithm>
#include <functional>
#include <vector>

#include "pbh_fwd_core.h"

namespace pbh {

// Per spot chunk, the active layers of the depth slice are processed as a
// software pipeline with one barrier per layer: phase k runs stage B of layer k
// and stage A of layer k+1, and builds the x table of k+1 and the y table and
// W chunk of k+2 (double-buffered by parity).
// Stage A: 16x16 blocks of U (2 row blocks x NCB column blocks); wave w takes
// block (w % 2, w / 2) and, when NCB = 3, block (w % 2, 2) for w < 2.
// Stage B: wave w owns the 32 x 32 dose block at columns 32 w.
//
// blk: the workgroup's index in the problem's own grid (chunk * tiles + tile).
template <bool TAB, int SCP, int QC>
__device__ __forceinline__ void fwd_fp32_body(const FwdArgs& a, const FwdPlan& pl, int blk)
{
    constexpr int KA = SCP / 4;            // stage A K steps
    constexpr int KB = SCP / 2;            // stage B K steps
    constexpr int NCB = (SCP + 15) / 16;   // stage A column blocks
    constexpr int UW = 16 * NCB;           // U / W columns
    constexpr int US = UW + 1;             // odd LDS row stride
    constexpr int NW = (SCP * UW + FT_THREADS - 1) / FT_THREADS;   // W elements per thread

    __shared__ float Ws[2][SCP * US];       // W chunk [j][i]
    __shared__ float Us[2][FT_TY * US];     // c U [y][i]
    __shared__ float Tx[TAB ? 2 : 1][TAB ? FT_TAB : 1];
    __shared__ float Ty[TAB ? 2 : 1][TAB ? FT_TAB : 1];
    __shared__ LayerList ll;

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int tiles = a.tiles_x * a.tiles_y * a.nz;
    const int ch = blk / tiles;                     // layer chunk of this workgroup
    const int tile = blk % tiles;
    const int tx = tile % a.tiles_x;
    const int ty = (tile / a.tiles_x) % a.tiles_y;
    const int z = tile / (a.tiles_x * a.tiles_y);
    const int x0 = tx * FT_TX, y0 = ty * FT_TY;

    const int rbA = wave & 1;                        // stage A row block
    const int bA = 16 * rbA + (lane & 15);           // stage A row in the tile
    const int cbA0 = wave >> 1;                      // first column block
    const bool oneA = cbA0 < NCB;
    const bool twoA = NCB == 3 && wave < 2;          // second block: column block 2
    const int cnB = 32 * wave;                       // stage B columns
    const int aB = cnB + (lane & 31);
    const double yAc = a.oy + (double)(y0 + bA) * a.dy;
    const double xBc = a.ox + (double)(x0 + aB) * a.dx;
    // table entry of this lane's k = 0 operand, and its step per K step
    const int tA0 = pl.ty.Q * (lane >> 4) - pl.ty.P * bA + pl.ty.P * (FT_TY - 1);
    const int tB0 = pl.tx.Q * (lane >> 5) - pl.tx.P * aB + pl.tx.P * (FT_TX - 1);
    const int ntY = pl.ty.P * (FT_TY - 1) + pl.ty.Q * (SCP - 1) + 1;
    const int ntX = pl.tx.P * (FT_TX - 1) + pl.tx.Q * (SCP - 1) + 1;
    // QC > 0 (P = 2, Q = QC on both axes): an operand's entry t0 + 2 Q k keeps the parity
    // of t0, so its parity-interleaved slot is slot(t0) + Q k: one base register plus
    // compile-time offsets per gather
    const int sA0 = TAB && QC ? table_slot(pl.ty, tA0) : 0;
    const int sB0 = TAB && QC ? table_slot(pl.tx, tB0) : 0;

    const int ncx = chunk_count(a.nsx), ncy = chunk_count(a.nsy);

    v16f acc = {};
    float wreg[NW];

    const int lo = ch * pl.chunk, hi = lo + pl.chunk;   // this workgroup's active layers
    int base = 0;                                       // active layers of earlier passes
    for (int lb = 0; lb < a.L && base < hi; lb += FT_LB) {
        const int nact = collect_layers(a, lb, z, 0.0f, nullptr, ll);
        const int m0 = max(0, lo - base), m1 = min(nact, hi - base);
        base += nact;
        if (m1 <= m0) continue;
        for (int cy = 0; cy < ncy; ++cy) {
            const int j0 = chunk_begin(a.nsy, ncy, cy), nj = chunk_begin(a.nsy, ncy, cy + 1) - j0;
            const double Cy = (a.oy + (double)y0 * a.dy) - (a.sy0 + (double)j0 * a.dsy);
            for (int cx = 0; cx < ncx; ++cx) {
                const int i0 = chunk_begin(a.nsx, ncx, cx), ni = chunk_begin(a.nsx, ncx, cx + 1) - i0;
                const double Cx = (a.ox + (double)x0 * a.dx) - (a.sx0 + (double)i0 * a.dsx);
                const TableRow rowx = TAB ? table_row(pl.tx, Cx, FT_TX) : TableRow{0, 0.0f, 0.0f};
                const TableRow rowy = TAB ? table_row(pl.ty, Cy, FT_TY) : TableRow{0, 0.0f, 0.0f};
                // QC > 0: this thread's table entries, squared distances formed once per chunk
                TabShare<QC ? 2 * (FT_TX - 1) + QC * (SCP - 1) + 1 : 1> shx;
                TabShare<QC ? 2 * (FT_TY - 1) + QC * (SCP - 1) + 1 : 1> shy;
                if (TAB && QC) {
                    shx.init(pl.tx, rowx);
                    shy.init(pl.ty, rowy);
                }

                auto load_w = [&](int m) {
                    const float* wl = a.w + (size_t)ll.lay[m] * a.nsy * a.nsx + (size_t)j0 * a.nsx + i0;
#pragma unroll
                    for (int q = 0; q < NW; ++q) {
                        const int e = threadIdx.x + q * FT_THREADS;
                        const int j = e / UW, i = e % UW;
                        wreg[q] = (e < SCP * UW && j < nj && i < ni) ? wl[(size_t)j * a.nsx + i] : 0.0f;
                    }
                };
                auto store_w = [&](int b) {
#pragma unroll
                    for (int q = 0; q < NW; ++q) {
                        const int e = threadIdx.x + q * FT_THREADS;
                        if (e < SCP * UW) Ws[b][(e / UW) * US + (e % UW)] = wreg[q];
                    }
                };
                auto build_tx = [&](int m) {
                    if (!TAB) return;
                    if (QC) {
                        shx.build_f32(Tx[m & 1], ll.k2[m]);
                        return;
                    }
                    const float k2 = ll.k2[m];
                    for (int t = threadIdx.x; t < ntX; t += FT_THREADS) {
                        const float d = table_distance_f(rowx, t);
                        Tx[m & 1][table_slot(pl.tx, t)] = __builtin_amdgcn_exp2f(-(d * d) * k2);
                    }
                };
                auto build_ty = [&](int m) {
                    if (!TAB) return;
                    if (QC) {
                        shy.build_f32(Ty[m & 1], ll.k2[m]);
                        return;
                    }
                    const float k2 = ll.k2[m];
                    for (int t = threadIdx.x; t < ntY; t += FT_THREADS) {
                        const float d = table_distance_f(rowy, t);
                        Ty[m & 1][table_slot(pl.ty, t)] = __builtin_amdgcn_exp2f(-(d * d) * k2);
                    }
                };
                // stage A of layer m: U[m&1] = c Gy W
                auto stage_a = [&](int m) {
                    const int b = m & 1;
                    const float c = ll.c[m], k2 = ll.k2[m];
                    if (!oneA) return;
                    float g[KA];
#pragma unroll
                    for (int kk = 0; kk < KA; ++kk) {
                        if (TAB) {
                            g[kk] = QC ? Ty[b][sA0 + 2 * QC * kk] : Ty[b][table_slot(pl.ty, tA0 + 4 * pl.ty.Q * kk)];
                        } else {
                            const int j = 4 * kk + (lane >> 4);
                            const double d = yAc - (a.sy0 + (double)(j0 + j) * a.dsy);
                            g[kk] = (j < nj) ? gauss2((float)(d * d), k2) : 0.0f;
                        }
                    }
                    v4f u0 = {}, u1 = {};
#pragma unroll
                    for (int kk = 0; kk < KA; ++kk) {
                        const float* wr = &Ws[b][(4 * kk + (lane >> 4)) * US + (lane & 15)];
                        u0 = __builtin_amdgcn_mfma_f32_16x16x4f32(g[kk], wr[16 * cbA0], u0, 0, 0, 0);
                        if (twoA) u1 = __builtin_amdgcn_mfma_f32_16x16x4f32(g[kk], wr[32], u1, 0, 0, 0);
                    }
                    // C layout of 16x16x4: row 4 (lane / 16) + r, column lane % 16
                    const int rbase = 16 * rbA + 4 * (lane >> 4);
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        float* urow = &Us[b][(rbase + r) * US + (lane & 15)];
                        urow[16 * cbA0] = c * u0[r];
                        if (twoA) urow[32] = c * u1[r];
                    }
                };
                // stage B of layer m: acc += (c U) Gx^T; operands are gathered first
                auto stage_b = [&](int m) {
                    const int b = m & 1;
                    const float k2 = ll.k2[m];
                    float gv[KB], av[KB];
#pragma unroll
                    for (int kk = 0; kk < KB; ++kk) {
                        if (TAB) {
                            gv[kk] = QC ? Tx[b][sB0 + QC * kk] : Tx[b][table_slot(pl.tx, tB0 + 2 * pl.tx.Q * kk)];
                        } else {
                            const int i = 2 * kk + (lane >> 5);
                            const double d = xBc - (a.sx0 + (double)(i0 + i) * a.dsx);
                            gv[kk] = (i < ni) ? gauss2((float)(d * d), k2) : 0.0f;
                        }
                        av[kk] = Us[b][(lane & 31) * US + 2 * kk + (lane >> 5)];
                    }
                    v16f p = {};
#pragma unroll
                    for (int kk = 0; kk < KB; ++kk)
                        p = __builtin_amdgcn_mfma_f32_32x32x2f32(av[kk], gv[kk], p, 0, 0, 0);
                    acc += p;
                };

                // prologue: layer m0 staged, U(m0) computed, Tx(m0), W(m0+1), Ty(m0+1) staged
                load_w(m0);
                store_w(m0 & 1);
                build_ty(m0);
                if (m0 + 1 < m1) load_w(m0 + 1);
                __syncthreads();
                stage_a(m0);
                build_tx(m0);
                if (m0 + 1 < m1) {
                    store_w((m0 + 1) & 1);
                    build_ty(m0 + 1);
                }
                __syncthreads();
                for (int m = m0; m < m1; ++m) {
                    if (m + 2 < m1) load_w(m + 2);
                    stage_b(m);
                    if (m + 1 < m1) {
                        stage_a(m + 1);
                        build_tx(m + 1);
                    }
                    if (m + 2 < m1) {
                        store_w(m & 1);
                        build_ty(m + 2);
                    }
                    __syncthreads();
                }
            }
        }
    }
    if (ch > 0 && base <= lo) return;                   // no layers in this chunk: nothing to add
    store_block(a, pl, tile, ch, z, y0, x0, 0, cnB, acc, (size_t)tiles);
}

template <bool TAB, int SCP, int QC>
__global__ __launch_bounds__(FT_THREADS, 2)
void fwd_fp32_kernel(FwdArgs a, FwdPlan pl)
{
    fwd_fp32_body<TAB, SCP, QC>(a, pl, blockIdx.x);
}

// One launch over a batch of problems sharing the template parameters.
template <bool TAB, int SCP, int QC>
__global__ __launch_bounds__(FT_THREADS, 2)
void fwd_fp32_batch_kernel(const BatchFwdItem* __restrict__ items, const int* __restrict__ wg0, int n)
{
    const int k = batch_item(wg0, n, blockIdx.x);
    const FwdArgs a = items[k].a;
    const FwdPlan pl = items[k].pl;
    fwd_fp32_body<TAB, SCP, QC>(a, pl, (int)blockIdx.x - wg0[k]);
}

__global__ __launch_bounds__(256) void combine_kernel(FwdArgs a, FwdPlan pl)
{
    combine_body(a, pl, blockIdx.x);
}

__global__ __launch_bounds__(256) void combine_objective_kernel(FwdArgs a, FwdPlan pl, ObjFuse f)
{
    combine_objective_body(a, pl, f, blockIdx.x / FT_OBJ_PARTS, blockIdx.x % FT_OBJ_PARTS);
}

__global__ __launch_bounds__(256) void combine_batch_kernel(const BatchFwdItem* __restrict__ items,
                                                            const int* __restrict__ wg0, int n)
{
    const int k = batch_item(wg0, n, blockIdx.x);
    const FwdArgs a = items[k].a;
    const FwdPlan pl = items[k].pl;
    combine_body(a, pl, (int)blockIdx.x - wg0[k]);
}

}  // namespace pbh

using namespace pbh;

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
static int validate_problem(const PbhGrid* g, const PbhLattice* lat)
{
    if (!g || !lat) return PBH_ERR_NULL_ARG;
    if (g->nx <= 0 || g->ny <= 0 || g->nz <= 0) return PBH_ERR_BAD_DIM;
    if (g->nx > PBH_MAX_DIM || g->ny > PBH_MAX_DIM || g->nz > PBH_MAX_DIM) return PBH_ERR_UNSUPPORTED;
    if (lat->n_sx <= 0 || lat->n_sy <= 0 || lat->n_layers <= 0) return PBH_ERR_BAD_DIM;
    if (lat->n_sx > PBH_MAX_SPOTS || lat->n_sy > PBH_MAX_SPOTS || lat->n_layers > PBH_MAX_LAYERS)
        return PBH_ERR_UNSUPPORTED;
    if (!(g->dx > 0) || !(g->dy > 0) || !(g->dz > 0)) return PBH_ERR_BAD_DIM;
    if (!(lat->dx_spot > 0) || !(lat->dy_spot > 0)) return PBH_ERR_BAD_DIM;
    return PBH_OK;
}

extern "C" int pbh_validate(const PbhGrid* g, const PbhLattice* lat)
{
    return validate_problem(g, lat);
}

FwdArgs make_fwd_args(const PbhGrid* g, const PbhLattice* lat, const float* w,
                      const double* idd, const double* sigma, float* dose, int tile_rows)
{
    FwdArgs a;
    a.nx = g->nx; a.ny = g->ny; a.nz = g->nz;
    a.nsx = lat->n_sx; a.nsy = lat->n_sy; a.L = lat->n_layers;
    a.dx = g->dx; a.dy = g->dy; a.ox = g->ox; a.oy = g->oy;
    a.dsx = lat->dx_spot; a.dsy = lat->dy_spot;
    a.sx0 = lat->x0 - (lat->n_sx - 1) / 2.0 * lat->dx_spot;
    a.sy0 = lat->y0 - (lat->n_sy - 1) / 2.0 * lat->dy_spot;
    a.w = w; a.idd = idd; a.sigma = sigma; a.dose = dose;
    a.tiles_x = cdiv(g->nx, FT_TX);
    a.tile_rows = tile_rows;
    a.tiles_y = cdiv(g->ny, tile_rows);
    return a;
}

// dv = P u, ds = Q u with small integers and a table that fits FT_TAB entries.
static bool axis_table(double dv, double ds, int T, int K, AxisTable* out)
{
    for (int Q = 1; Q <= 64; ++Q) {
        const double pr = dv * Q / ds;
        const int P = (int)llround(pr);
        if (P < 1 || P > 64) continue;
        if (fabs(pr - P) > 1e-9 * pr) continue;
        const int n = P * (T - 1) + Q * (K - 1) + 1;      // table entries
        if ((P == 2 ? table_slots(n) : n + 1) > FT_TAB) return false;
        out->P = P;
        out->Q = Q;
        out->u = dv / P;
        out->half = (P == 2) ? table_half(n) : 0;
        return true;
    }
    return false;
}

// Compute units of the current device (256 when the query fails), cached per device.
int pbh_device_cus()
{
    static int cus[64];
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) return 256;
    if (cus[dev] == 0) {
        int n = 0;
        if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || n <= 0)
            n = 256;
        cus[dev] = n;
    }
    return cus[dev];
}

static long long tile_count(const FwdArgs& a) { return (long long)a.tiles_x * a.tiles_y * a.nz; }

constexpr int FT_MIN_CHUNK = 4;   // smallest automatic layer chunk

// Layer chunk when the active layers per slice are not known (pbh_forward reads
// none of its device tables on the host): about three workgroups per two
// resident slots (two per CU), so that the chunks of shallow slices (many
// active layers) spread over more workgroups than those of deep slices while
// few partial tiles have to be combined. pbh_forward_choose_chunk, given the
// active-layer counts, replaces this rule.
int pbh_auto_chunk(const FwdArgs& a)
{
    const long long slots = 2LL * pbh_device_cus();
    const long long per_tile = (3 * slots + 2 * tile_count(a) - 1) / (2 * tile_count(a));
    if (per_tile <= 1) return a.L;
    int chunk = (int)((a.L + per_tile - 1) / per_tile);
    return chunk < FT_MIN_CHUNK ? (a.L < FT_MIN_CHUNK ? a.L : FT_MIN_CHUNK) : chunk;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

// Layer chunk from the active-layer count of every depth slice: simulates the
// dispatch of the workgroups in launch order (chunk-major, slices in order)
// onto `slots` resident workgroup slots, each taking the next workgroup when it
// frees, with a workgroup costing its layers plus FT_WG_COST layers of fixed
// work (empty later chunks exit at once), and returns the chunk (at least
// FT_MIN_CHUNK) with the shortest simulated makespan (the larger chunk within
// 2 %: fewer partials).
// On the 128^3 field this picks 17 (bf16x3, 3 slots per CU) and 26 (FP32, 2
// slots per CU); a measured sweep had 18-24 and 26 fastest, 4 % and 14 % ahead
// of two chunks per tile.
constexpr double FT_WG_COST = 2.0;

// Makespan of the simulated dispatch (see model_chunk) for one chunk: workgroups cost
// `layer_cost` per layer plus `fixed` each. Returns the makespan in layers.
static double model_makespan(const FwdArgs& a, const int32_t* active, long long slots, int c,
                             double layer_cost, double fixed, std::vector<double>& heap)
{
    const long long per_z = (long long)a.tiles_x * a.tiles_y;
    const int nch = (a.L + c - 1) / c;
    heap.assign((size_t)slots, 0.0);                              // min-heap of slot finish times
    for (int ch = 0; ch < nch; ++ch) {
        for (int z = 0; z < a.nz; ++z) {
            const int act = active[z] < 0 ? 0 : (active[z] > a.L ? a.L : active[z]);
            const int n = std::min(c, std::max(0, act - ch * c));
            const double w = (ch > 0 && n == 0) ? 0.05 : n * layer_cost + fixed;
            for (long long t = 0; t < per_z; ++t) {
                std::pop_heap(heap.begin(), heap.end(), std::greater<double>());
                heap.back() += w;
                std::push_heap(heap.begin(), heap.end(), std::greater<double>());
            }
        }
    }
    return *std::max_element(heap.begin(), heap.end());
}

// Best chunk and its makespan for one tile height.
static int model_chunk_cost(const FwdArgs& a, const int32_t* active, long long slots, double layer_cost,
                            double fixed, double* best_span)
{
    const long long per_z = (long long)a.tiles_x * a.tiles_y;
    if (slots < 1) slots = 1;
    if (slots > (1 << 20)) slots = 1 << 20;
    int best = a.L;
    double best_t = 0.0;
    std::vector<double> heap;
    for (int c = a.L; c >= std::min(a.L, FT_MIN_CHUNK); --c) {
        const int nch = (a.L + c - 1) / c;
        if ((long long)nch * per_z * a.nz > (1LL << 24)) break;     // keep the model cheap
        const double span = model_makespan(a, active, slots, c, layer_cost, fixed, heap);
        if (c == a.L || span < best_t * 0.98) {
            best = c;
            best_t = span;
        }
    }
    if (best_span) *best_span = best_t;
    return best;
}

static int model_chunk(const FwdArgs& a, const int32_t* active, long long slots)
{
    return model_chunk_cost(a, active, slots, 1.0, FT_WG_COST, nullptr);
}

static int nchunks(const FwdArgs& a, int chunk) { return (a.L + chunk - 1) / chunk; }

// Workspace layout: [layer peaks (mixed "_peak" forms)][split 16-bit weights
// (mixed forms)][partial tiles of layer chunks 1 .. nchunk - 1].
static size_t ws_peaks(const FwdArgs& a) { return align256((size_t)a.L * sizeof(float)); }
static size_t ws_wsplit(const FwdArgs& a)
{
    const size_t chunks = (size_t)a.L * chunk_count(a.nsy) * chunk_count(a.nsx);
    return align256(chunks * FT_WSPLIT * sizeof(unsigned short)) + align256(chunks * sizeof(float2));
}

static size_t ws_layers(const FwdArgs& a)
{
    return align256((size_t)a.nz * sizeof(int)) + align256((size_t)a.nz * a.L * sizeof(LayerEntry));
}

size_t pbh_forward_ws_bytes(const FwdArgs& a, int layer_chunk)
{
    const int chunk = layer_chunk > 0 ? layer_chunk : pbh_auto_chunk(a);
    size_t n = ws_peaks(a) + ws_wsplit(a) + ws_layers(a);
    n += (size_t)(nchunks(a, chunk) - 1) * tile_count(a) * a.tile_rows * FT_TX * sizeof(float);
    return n;
}

float* pbh_forward_ws_peaks(const FwdArgs& a, void* ws)
{
    return (float*)ws;
}

int pbh_plan_forward(const FwdArgs& a, int kx_pad, int ky_pad, void* ws, size_t ws_bytes,
                     int layer_chunk, FwdPlan* plan)
{
    if (layer_chunk < 0) return PBH_ERR_BAD_DIM;
    const int chunk = layer_chunk > 0 ? (layer_chunk < a.L ? layer_chunk : a.L) : pbh_auto_chunk(a);
    plan->chunk = chunk;
    plan->nchunk = nchunks(a, chunk);
    if (tile_count(a) * plan->nchunk > 0x7fffffffLL) return PBH_ERR_UNSUPPORTED;
    plan->table = axis_table(a.dx, a.dsx, FT_TX, kx_pad, &plan->tx) &&
                  axis_table(a.dy, a.dsy, a.tile_rows, ky_pad, &plan->ty);
    if (!plan->table) {
        plan->tx = AxisTable{0, 0, 0.0, 0};
        plan->ty = AxisTable{0, 0, 0.0, 0};
    }
    plan->part = nullptr;
    plan->wsplit = nullptr;
    plan->wpar = nullptr;
    plan->layers = nullptr;
    plan->nlayers = nullptr;
    const size_t need = pbh_forward_ws_bytes(a, chunk);
    if (ws && ws_bytes >= need) {
        char* p = (char*)ws;
        plan->wsplit = (unsigned short*)(p + ws_peaks(a));
        plan->wpar = (float2*)(p + ws_peaks(a) +
                               align256((size_t)a.L * chunk_count(a.nsy) * chunk_count(a.nsx) * FT_WSPLIT *
                                        sizeof(unsigned short)));
        char* lp = p + ws_peaks(a) + ws_wsplit(a);
        plan->nlayers = (int*)lp;
        plan->layers = (LayerEntry*)(lp + align256((size_t)a.nz * sizeof(int)));
        if (plan->nchunk > 1) plan->part = (float*)(lp + ws_layers(a));
    }
    if (plan->nchunk > 1) {
        if (!ws) return PBH_ERR_NULL_ARG;
        if (ws_bytes < need) return PBH_ERR_WORKSPACE;
    }
    return PBH_OK;
}

int pbh_launch_combine(const FwdArgs& a, const FwdPlan& pl, hipStream_t s, const ObjFuse* fuse)
{
    if (fuse) {
        hipLaunchKernelGGL(combine_objective_kernel, dim3((unsigned)(tile_count(a) * FT_OBJ_PARTS)), dim3(256), 0, s,
                           a, pl, *fuse);
        return launch_rc("combine_objective_kernel");
    }
    if (pl.nchunk <= 1) return PBH_OK;
    hipLaunchKernelGGL(combine_kernel, dim3((unsigned)tile_count(a)), dim3(256), 0, s, a, pl);
    return launch_rc("combine_kernel");
}

int pbh_launch_combine_batch(const BatchFwdItem* items, const int* wg0, int n, int total, hipStream_t s)
{
    hipLaunchKernelGGL(combine_batch_kernel, dim3((unsigned)total), dim3(256), 0, s, items, wg0, n);
    return launch_rc("combine_batch_kernel");
}

int pbh_forward_mixed(const FwdArgs& a, double oz, double dz, int precision, int layer_chunk,
                      void* ws, size_t ws_bytes, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1,
                      const ObjFuse* fuse);

extern "C" size_t pbh_forward_workspace_bytes_tuned(const PbhGrid* g, const PbhLattice* lat, int precision,
                                                   int layer_chunk)
{
    if (validate_problem(g, lat) != PBH_OK || layer_chunk < 0) return 0;
    (void)precision;
    const FwdArgs a = make_fwd_args(g, lat, nullptr, nullptr, nullptr, nullptr, FT_TY);
    return pbh_forward_ws_bytes(a, layer_chunk > 0 && layer_chunk < a.L ? layer_chunk : (layer_chunk ? a.L : 0));
}

// Table geometry with P = 2, Q = 5 on both axes (2 mm voxels, 5 mm spots): the FP32
// kernel's compile-time-slot instantiation.
bool fp32_q5(const FwdPlan& pl)
{
    return pl.table && pl.tx.P == 2 && pl.tx.Q == 5 && pl.ty.P == 2 && pl.ty.Q == 5;
}

extern "C" int pbh_forward_choose_chunk(const PbhGrid* g, const PbhLattice* lat, int precision,
                                        const int32_t* active_per_z, int compute_units)
{
    int rc = validate_problem(g, lat);
    if (rc != PBH_OK) return rc;
    if (!active_per_z) return PBH_ERR_NULL_ARG;
    if (precision < PBH_PREC_FP32 || precision > PBH_PREC_FP16X3_PEAK) return PBH_ERR_UNSUPPORTED;
    const FwdArgs a = make_fwd_args(g, lat, nullptr, nullptr, nullptr, nullptr, FT_TY);
    // workgroup slots per CU: 3 for the mixed kernels on the P = 2, Q = 5 table
    // geometry, 2 for the others. (The single-product kernel holds 4, but those four
    // share the CU's issue bandwidth; a sweep measured its best chunk where the
    // 3-slot model puts it: 16-17, against 13 with 4 slots.)
    int per_cu = 2;
    if (precision != PBH_PREC_FP32) {
        FwdPlan pl;
        if (pbh_plan_forward(a, chunk_pad(a.nsx, a.nsy), chunk_pad(a.nsx, a.nsy), nullptr, 0, a.L, &pl) ==
                PBH_OK &&
            pl.table && pl.tx.P == 2 && pl.tx.Q == 5 && pl.ty.P == 2 && pl.ty.Q == 5)
            per_cu = 3;
    }
    const int cus = compute_units > 0 ? compute_units : pbh_device_cus();
    return model_chunk(a, active_per_z, (long long)per_cu * cus);
}

// Tile height and layer chunk from the active layers per slice: the dispatch model of
// each tile height, in units of one 32-row tile-layer. A 64-row workgroup (16-bit
// forms on the P = 2, Q = 5 geometry) costs about as much per layer as a 32-row one
// although it covers twice the rows (operand gathers, weight copies and table builds
// shared by two row blocks; 1.05 per layer and 2.25 fixed, fitted to a sweep of
// chunks 8-40 at both heights on the 128^3 / 41 x 41 x 40-layer field), but only
// two fit per CU for the split forms (LDS) and three for the single-product forms.
constexpr double FT64_LAYER_COST = 1.05, FT64_WG_COST = 2.25;

extern "C" int pbh_forward_choose_shape(const PbhGrid* g, const PbhLattice* lat, int precision,
                                        const int32_t* active_per_z, int compute_units, int32_t* tile_rows)
{
    int rc = validate_problem(g, lat);
    if (rc != PBH_OK) return rc;
    if (!active_per_z || !tile_rows) return PBH_ERR_NULL_ARG;
    if (precision < PBH_PREC_FP32 || precision > PBH_PREC_FP16X3_PEAK) return PBH_ERR_UNSUPPORTED;
    const int cus = compute_units > 0 ? compute_units : pbh_device_cus();
    const int c32 = pbh_forward_choose_chunk(g, lat, precision, active_per_z, cus);
    *tile_rows = FT_TY;
    if (c32 < 0 || precision == PBH_PREC_FP32) return c32;
    const FwdArgs a32 = make_fwd_args(g, lat, nullptr, nullptr, nullptr, nullptr, FT_TY);
    FwdPlan pl;
    if (pbh_plan_forward(a32, chunk_pad(a32.nsx, a32.nsy), chunk_pad(a32.nsx, a32.nsy), nullptr, 0, a32.L, &pl) !=
            PBH_OK ||
        !(pl.table && pl.tx.P == 2 && pl.tx.Q == 5 && pl.ty.P == 2 && pl.ty.Q == 5))
        return c32;
    double t32 = 0.0, t64 = 0.0;
    model_chunk_cost(a32, active_per_z, 3LL * cus, 1.0, FT_WG_COST, &t32);
    const FwdArgs a64 = make_fwd_args(g, lat, nullptr, nullptr, nullptr, nullptr, 64);
    const bool plain = precision == PBH_PREC_BF16 || precision == PBH_PREC_FP16;
    const int c64 = model_chunk_cost(a64, active_per_z, (plain ? 3LL : 2LL) * cus, FT64_LAYER_COST,
                                     FT64_WG_COST, &t64);
    if (t64 < t32) {
        *tile_rows = 64;
        return c64;
    }
    return c32;
}

extern "C" size_t pbh_forward_workspace_bytes(const PbhGrid* g, const PbhLattice* lat, int precision)
{
    return pbh_forward_workspace_bytes_tuned(g, lat, precision, 0);
}

// Dose tile rows of a call: 0 = the default (32); 32 for every form, 64 for the mixed
// forms on the P = 2, Q = 5 table geometry (checked when planned).
static int tile_rows_of(int precision, int tile_rows)
{
    if (tile_rows == 0) return FT_TY;
    if (tile_rows != 32 && tile_rows != 64) return PBH_ERR_BAD_DIM;
    if (tile_rows == 64 && precision == PBH_PREC_FP32) return PBH_ERR_UNSUPPORTED;
    return tile_rows;
}

// pbh_forward_tuned with a tile height, recording ev0 / ev1 (when not null) on the
// stream right before and right after the contraction kernel.
static int forward_impl(const PbhGrid* g, const PbhLattice* lat, const float* w, const double* idd,
                        const double* sigma, float* dose, int precision, int layer_chunk, int tile_rows,
                        void* workspace, size_t workspace_bytes, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1,
                        const ObjFuse* fuse = nullptr)
{
    int rc = validate_problem(g, lat);
    if (rc != PBH_OK) return rc;
    if (!w || !idd || !sigma || !dose) return PBH_ERR_NULL_ARG;
    if (precision < PBH_PREC_FP32 || precision > PBH_PREC_FP16X3_PEAK) return PBH_ERR_UNSUPPORTED;
    if (layer_chunk < 0) return PBH_ERR_BAD_DIM;
    const int rows = tile_rows_of(precision, tile_rows);
    if (rows < 0) return rows;
    const FwdArgs a = make_fwd_args(g, lat, w, idd, sigma, dose, rows);
    if (precision != PBH_PREC_FP32)
        return pbh_forward_mixed(a, g->oz, g->dz, precision, layer_chunk, workspace, workspace_bytes, s, ev0, ev1,
                                 fuse);
    FwdPlan pl;
    rc = pbh_plan_forward(a, chunk_pad(a.nsx, a.nsy), chunk_pad(a.nsx, a.nsy), workspace, workspace_bytes,
                          layer_chunk, &pl);
    if (rc != PBH_OK) return rc;
    pl.layers = nullptr;            // the layer lists are written by the mixed forms' split launch only
    pl.nlayers = nullptr;
    const dim3 grid((unsigned)(tile_count(a) * pl.nchunk));
    const bool q5 = fp32_q5(pl);
    if (ev0 && (rc = hip_rc(hipEventRecord(ev0, s), "hipEventRecord")) != PBH_OK) return rc;
    switch (chunk_pad(a.nsx, a.nsy)) {
#define PBH_FP32_CASE(N) \
    case N: \
        if (q5) hipLaunchKernelGGL((fwd_fp32_kernel<true, N, 5>), grid, dim3(FT_THREADS), 0, s, a, pl); \
        else if (pl.table) hipLaunchKernelGGL((fwd_fp32_kernel<true, N, 0>), grid, dim3(FT_THREADS), 0, s, a, pl); \
        else hipLaunchKernelGGL((fwd_fp32_kernel<false, N, 0>), grid, dim3(FT_THREADS), 0, s, a, pl); \
        break;
    PBH_FP32_CASE(4) PBH_FP32_CASE(8) PBH_FP32_CASE(12) PBH_FP32_CASE(16) PBH_FP32_CASE(20)
    PBH_FP32_CASE(24) PBH_FP32_CASE(28) PBH_FP32_CASE(32) PBH_FP32_CASE(36) PBH_FP32_CASE(40)
    PBH_FP32_CASE(44) PBH_FP32_CASE(48)
#undef PBH_FP32_CASE
    default: return PBH_ERR_UNSUPPORTED;
    }
    if ((rc = launch_rc("fwd_fp32_kernel")) != PBH_OK) return rc;
    if (ev1 && (rc = hip_rc(hipEventRecord(ev1, s), "hipEventRecord")) != PBH_OK) return rc;
    return pbh_launch_combine(a, pl, s, fuse);
}

int pbh_forward_fused(const PbhGrid* g, const PbhLattice* lat, const float* w, const double* idd,
                      const double* sigma, float* dose, int precision, int layer_chunk, int tile_rows,
                      void* workspace, size_t workspace_bytes, hipStream_t s, const ObjFuse* fuse, int* njpart)
{
    if (!fuse || !fuse->presc || !fuse->R || !fuse->jpart || !njpart) return PBH_ERR_NULL_ARG;
    const int rows = tile_rows_of(precision, tile_rows);
    if (rows < 0) return rows;
    if (validate_problem(g, lat) != PBH_OK) return PBH_ERR_BAD_DIM;
    *njpart = (int)tile_count(make_fwd_args(g, lat, nullptr, nullptr, nullptr, nullptr, rows)) * FT_OBJ_PARTS;
    return forward_impl(g, lat, w, idd, sigma, dose, precision, layer_chunk, tile_rows, workspace, workspace_bytes,
                        s, nullptr, nullptr, fuse);
}

extern "C" int pbh_forward_tuned(const PbhGrid* g, const PbhLattice* lat, const float* w,
                                 const double* idd, const double* sigma, float* dose, int precision,
                                 int layer_chunk, void* workspace, size_t workspace_bytes, void* stream)
{
    return forward_impl(g, lat, w, idd, sigma, dose, precision, layer_chunk, 0, workspace, workspace_bytes,
                        (hipStream_t)stream, nullptr, nullptr);
}

extern "C" int pbh_forward_shaped(const PbhGrid* g, const PbhLattice* lat, const float* w,
                                  const double* idd, const double* sigma, float* dose, int precision,
                                  int layer_chunk, int tile_rows, void* workspace, size_t workspace_bytes,
                                  void* stream)
{
    return forward_impl(g, lat, w, idd, sigma, dose, precision, layer_chunk, tile_rows, workspace,
                        workspace_bytes, (hipStream_t)stream, nullptr, nullptr);
}

extern "C" size_t pbh_forward_workspace_bytes_shaped(const PbhGrid* g, const PbhLattice* lat, int precision,
                                                     int layer_chunk, int tile_rows)
{
    if (validate_problem(g, lat) != PBH_OK || layer_chunk < 0) return 0;
    const int rows = tile_rows_of(precision, tile_rows);
    if (rows < 0) return 0;
    const FwdArgs a = make_fwd_args(g, lat, nullptr, nullptr, nullptr, nullptr, rows);
    return pbh_forward_ws_bytes(a, layer_chunk > 0 && layer_chunk < a.L ? layer_chunk : (layer_chunk ? a.L : 0));
}

extern "C" int pbh_forward_profiled(const PbhGrid* g, const PbhLattice* lat, const float* w,
                                    const double* idd, const double* sigma, float* dose, int precision,
                                    int layer_chunk, int tile_rows, void* workspace, size_t workspace_bytes,
                                    void* ev_start, void* ev_end, void* stream)
{
    if (!ev_start || !ev_end) return PBH_ERR_NULL_ARG;
    return forward_impl(g, lat, w, idd, sigma, dose, precision, layer_chunk, tile_rows, workspace,
                        workspace_bytes, (hipStream_t)stream, (hipEvent_t)ev_start, (hipEvent_t)ev_end);
}

extern "C" int pbh_forward(const PbhGrid* g, const PbhLattice* lat, const float* w,
                           const double* idd, const double* sigma, float* dose, int precision,
                           void* workspace, size_t workspace_bytes, void* stream)
{
    return pbh_forward_tuned(g, lat, w, idd, sigma, dose, precision, 0, workspace,
                             workspace_bytes, stream);
}

// Batched FP32 forward: items[0..n) share SCP (their padded chunk width) and the
// table mode; total = wg0[n].
int pbh_launch_fwd_fp32_batch(const BatchFwdItem* items, const int* wg0, int n, int total, int scp,
                              bool tab, bool q5, hipStream_t s)
{
    const dim3 grid((unsigned)total), block(FT_THREADS);
    switch (scp) {
#define PBH_FP32_BATCH_CASE(N) \
    case N: \
        if (q5) hipLaunchKernelGGL((fwd_fp32_batch_kernel<true, N, 5>), grid, block, 0, s, items, wg0, n); \
        else if (tab) hipLaunchKernelGGL((fwd_fp32_batch_kernel<true, N, 0>), grid, block, 0, s, items, wg0, n); \
        else hipLaunchKernelGGL((fwd_fp32_batch_kernel<false, N, 0>), grid, block, 0, s, items, wg0, n); \
        break;
    PBH_FP32_BATCH_CASE(16) PBH_FP32_BATCH_CASE(24) PBH_FP32_BATCH_CASE(32)
    PBH_FP32_BATCH_CASE(40) PBH_FP32_BATCH_CASE(48)
#undef PBH_FP32_BATCH_CASE
    default: return PBH_ERR_UNSUPPORTED;
    }
    return launch_rc("fwd_fp32_batch_kernel");
}
