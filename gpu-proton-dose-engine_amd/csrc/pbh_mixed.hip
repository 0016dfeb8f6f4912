// pbh_mixed.hip -- mixed-precision forward dose: BF16 / FP16 Gaussian factors
// on gfx950 MFMA with FP32 accumulation, optionally as split products.
//
// Same tiling, layer split, Gaussian tables and double buffering as the FP32
// kernel (pbh_forward.hip):
//   stage A  U^T[48 x 32] = W^T[48 x J] Gy^T[J x 32]   v_mfma_f32_16x16x32_{bf16,f16}
//   stage B  D[32 x 128] += (c U) Gx^T                 v_mfma_f32_32x32x16_{bf16,f16}
// Stage A computes U transposed so that each lane's four results are four
// consecutive spots of one row of U, stored with one 8-byte LDS write per
// (hi, lo) half. Factors are formed in FP32 (table: exact-integer-part argument), then rounded
// to the 16-bit type. In the split ("x3") forms every operand a is stored as
// hi = T(a), lo = T(a - hi) and a product is hi*hi + hi*lo + lo*hi (lo*lo
// dropped), which restores ~2^-16 (BF16) or ~2^-21 (FP16) relative accuracy per
// product at three MFMAs instead of one. Table entries hold the (hi, lo) pair
// packed in 32 bits; operands are assembled with v_perm. The weights are split
// once per call by wsplit_kernel into the layout the kernel copies to LDS with
// 16-byte loads and stores. The "_peak" forms use split products only for
// (layer, depth) pairs within 1 cm of the layer's Bragg peak (depth of its IDD
// maximum) and single products elsewhere.
//
// On the 2 mm voxel / 5 mm spot geometry (P = 2, Q = 5) the table slots of a
// lane's operands are a base register plus compile-time offsets (QC = 5); other
// geometries compute the slots at run time (QC = 0).
//
// BF16 shares FP32's exponent range: stage B accumulates every layer's products
// straight into the running dose accumulators. FP16 has a narrow exponent range:
// the weight-split launch scales the weights of every (layer, spot chunk) by a
// power of two (largest magnitude into [1, 2)) and records a bound on their column
// sums. Each workgroup takes the largest bound of its layers, 2^eacc, and keeps the
// dose scaled by 2^-eacc: every layer's U is stored below 2^14 with no reduction
// in the kernel, and its products accumulate straight into the accumulators, as
// for BF16; the tile is scaled back by 2^eacc when it is stored.
#include <algorithm>

#include "pbh_mixed_core.h"

namespace pbh {

constexpr int SCH = PBH_SPOT_CHUNK;       // 48 spots per chunk
constexpr int KYP = FT_KYP;               // stage A K (j) padded to 2 steps of 32
constexpr int WT_STRIDE = KYP + 8;        // W^T rows in halves
constexpr int U_STRIDE = SCH + 8;         // U rows in halves
constexpr int WV = SCH * KYP / 8;         // 16-byte vectors per split half of a weight chunk

// Index of the first maximum of row[0 .. nz) (np.argmax semantics), by one wave:
// lane-strided scans, then a reduction across lanes in which the lower index wins
// ties. Every lane returns the result.
__device__ inline int first_max_wave(const double* row, int nz)
{
    const int lane = threadIdx.x & 63;
    double bv = 0.0;
    int bk = -1;
    for (int k = lane; k < nz; k += 64) {
        const double v = row[k];
        if (bk < 0 || v > bv) { bv = v; bk = k; }
    }
    for (int o = 32; o > 0; o >>= 1) {
        const double ov = __shfl_xor(bv, o, 64);
        const int ok = __shfl_xor(bk, o, 64);
        if (ok >= 0 && (bk < 0 || ov > bv || (ov == bv && ok < bk))) { bv = ov; bk = ok; }
    }
    return bk;
}

// Split 16-bit weights of one (layer, spot chunk): hi and lo at [i][j] (i < 48,
// j < 64), zero outside the chunk. blk = (layer * chunks_y + chunk_y) *
// chunks_x + chunk_x; part (0 .. WS_PARTS-1) selects a slice of the chunk's
// (i, j pair) entries. Each thread writes one pair of consecutive j of hi and of
// lo as 32-bit words, consecutive threads consecutive words.
constexpr int WS_PAIRS = SCH * KYP / 2;                  // 1536 (i, j pair) entries per chunk
constexpr int WS_PARTS = WS_PAIRS / 256;                 // workgroups of 256 threads per chunk

template <typename T>
__device__ __forceinline__ void wsplit_body(const FwdArgs& a, unsigned short* dst, float2* wpar, int blk, int part)
{
    constexpr bool F16 = !__is_same(T, __bf16);
    const int ncx = chunk_count(a.nsx), ncy = chunk_count(a.nsy);
    const int cx = blk % ncx, cy = (blk / ncx) % ncy, l = blk / (ncx * ncy);
    if (l >= a.L) return;
    const int i0 = chunk_begin(a.nsx, ncx, cx), ni = chunk_begin(a.nsx, ncx, cx + 1) - i0;
    const int j0 = chunk_begin(a.nsy, ncy, cy), nj = chunk_begin(a.nsy, ncy, cy + 1) - j0;
    const float* wl = a.w + (size_t)l * a.nsy * a.nsx;
    // this thread's pair (loaded before the FP16 scale is known, so that the loads
    // overlap the reduction below)
    const int p = part * 256 + threadIdx.x;               // pair index: i = p / 32, j = 2 (p % 32) + {0, 1}
    const int i = p / (KYP / 2), j = 2 * (p % (KYP / 2));
    float v0 = (i < ni && j < nj) ? wl[(size_t)(j0 + j) * a.nsx + i0 + i] : 0.0f;
    float v1 = (i < ni && j + 1 < nj) ? wl[(size_t)(j0 + j + 1) * a.nsx + i0 + i] : 0.0f;
    if (F16) {
        // FP16's exponent range: the chunk's weights are scaled by 2^-e so that their
        // largest magnitude lies in [1, 2), and part 0 records 2^e and the column-sum
        // bound. Every wave forms the chunk's maxima by itself (lane i sums column i over
        // the chunk's rows in row order; maxima over the lanes), so every wave of every
        // part has the same values and no barrier is needed.
        const int col = threadIdx.x & 63;
        float cmax = 0.0f, wmax = 0.0f;
        if (col < ni) {
            const float* wc = wl + (size_t)j0 * a.nsx + i0 + col;
#pragma unroll 8
            for (int r = 0; r < nj; ++r) {
                const float v = fabsf(wc[(size_t)r * a.nsx]);
                cmax += v;
                wmax = fmaxf(wmax, v);
            }
        }
        for (int o = 32; o > 0; o >>= 1) {
            cmax = fmaxf(cmax, __shfl_xor(cmax, o, 64));
            wmax = fmaxf(wmax, __shfl_xor(wmax, o, 64));
        }
        const int e = (wmax > 0.0f && wmax <= 3.0e38f) ? ilogbf(wmax) : 0;
        const float sc = ldexpf(1.0f, -e);
        if (part == 0 && threadIdx.x == 0) wpar[blk] = make_float2(ldexpf(1.0f, e), cmax * sc);
        v0 *= sc;
        v1 *= sc;
    }
    unsigned* hi = (unsigned*)(dst + (size_t)blk * FT_WSPLIT);
    unsigned* lo = hi + WS_PAIRS;
    T h0, r0, h1, r1;
    split<T>(v0, h0, r0);
    split<T>(v1, h1, r1);
    hi[p] = (unsigned)bits16(h0) | ((unsigned)bits16(h1) << 16);
    lo[p] = (unsigned)bits16(r0) | ((unsigned)bits16(r1) << 16);
}

// The active layers of every depth slice, in layer order, with their scalars
// (the values collect_layers forms): one wave per slice, its lanes taking 64
// layers at a time (ballot compaction); slices 4 blk + wave with a stride of 4 nblk.
__device__ __forceinline__ void layer_lists_body(const FwdArgs& a, const FwdPlan& pl, int blk, int nblk)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int z = 4 * blk + wave; z < a.nz; z += 4 * nblk) {
        LayerEntry* row = pl.layers + (size_t)z * a.L;
        int n = 0;
        for (int lb = 0; lb < a.L; lb += 64) {
            const int l = lb + lane;
            double iddv = 0.0, sg = 0.0;
            if (l < a.L) {
                iddv = a.idd[(size_t)l * a.nz + z];
                sg = a.sigma[(size_t)l * a.nz + z];
            }
            const bool act = iddv != 0.0;
            const unsigned long long m = __ballot(act);
            if (act) {
                const LayerScalars sc = layer_scalars(iddv, sg);
                const double sd = sg > kSigmaFloor ? sg : kSigmaFloor;
                row[n + __popcll(m & ((1ull << lane) - 1ull))] = LayerEntry{l, sc.c, sc.k2, (float)(FT_WIN_SIGMAS * sd)};
            }
            n += __popcll(m);
        }
        if (lane == 0) pl.nlayers[z] = n;
    }
}

// Grid (chunks, WS_PARTS + 1): rows 0 .. WS_PARTS-1 split the weights, the last
// row writes the layer lists of the slices and, for the "_peak" forms (r_peak not
// null), r_peak[l] = depth of the first IDD maximum of layer l, one wave per layer.
template <typename T>
__global__ __launch_bounds__(256) void wsplit_kernel(FwdArgs a, FwdPlan pl, double oz, double dz, float* r_peak)
{
    if (blockIdx.y != WS_PARTS) {
        wsplit_body<T>(a, pl.wsplit, pl.wpar, blockIdx.x, blockIdx.y);
        return;
    }
    layer_lists_body(a, pl, blockIdx.x, gridDim.x);
    if (!r_peak) return;
    const int wave = threadIdx.x >> 6;
    for (int l = 4 * blockIdx.x + wave; l < a.L; l += 4 * gridDim.x) {
        const int best = first_max_wave(a.idd + (size_t)l * a.nz, a.nz);
        if ((threadIdx.x & 63) == 0) r_peak[l] = (float)(oz + (double)best * dz);
    }
}

template <typename T>
__global__ __launch_bounds__(256) void wsplit_batch_kernel(const BatchFwdItem* __restrict__ items)
{
    const FwdArgs a = items[blockIdx.z].a;
    wsplit_body<T>(a, items[blockIdx.z].pl.wsplit, items[blockIdx.z].pl.wpar, blockIdx.x, blockIdx.y);
}

// 4 waves, one-barrier software pipeline over the active layers of each spot
// chunk, as in the FP32 kernel, on a TY x 128 dose tile (TY = 32 or 64).
// Stage A: 16x16 blocks of U^T (3 spot blocks x TY / 16 row blocks).
//   TY = 32: wave w takes spot block w / 2 of row block w % 2 and, for w < 2,
//            spot block 2 of the same row block;
//   TY = 64: wave w takes row block w and all three spot blocks, so each
//            gathered Gy operand feeds three products.
// Stage B: wave w owns the TY x 32 dose columns 32 w, as TY / 32 blocks of 32 x 32
// that share each gathered Gx operand; KB K steps of 16 spots.
//
// blk: the workgroup's index in the problem's own grid (chunk * tiles + tile).
template <typename T, int MODE, bool TAB, int KB, int QC, bool LISTS, int TY>
__device__ __forceinline__ void fwd_mixed_body(const FwdArgs& a, const FwdPlan& pl, double oz, double dz,
                                               const float* r_peak, int blk)
{
    typedef typename Vec8<T>::type V8;
    constexpr bool F16 = !__is_same(T, __bf16);
    constexpr bool SPLIT = MODE != MODE_PLAIN;
    constexpr int KX = 16 * KB;                 // stage B K (spots in x), padded
    constexpr int NWV = SPLIT ? 2 * WV : WV;    // 16-byte weight vectors per chunk and layer
    constexpr int NV = (NWV + FT_THREADS - 1) / FT_THREADS;
    static_assert(TY == 32 || TY == 64, "tile rows");
    constexpr int NCB = TY == 64 ? 3 : 2;       // stage A spot blocks per wave (TY = 32: the second on w < 2)
    constexpr int NBB = TY / 32;                // stage B row blocks of 32 per wave

    __shared__ __attribute__((aligned(16))) T Wh[2][SCH * WT_STRIDE];      // W^T hi [i][j]
    __shared__ __attribute__((aligned(16))) T Wl[2][SPLIT ? SCH * WT_STRIDE : 8];
    __shared__ __attribute__((aligned(16))) T Uh[2][TY * U_STRIDE];        // c U hi [y][i]
    __shared__ __attribute__((aligned(16))) T Ul[2][SPLIT ? TY * U_STRIDE : 8];
    // table slots: the interleaved span on the P = 2, Q = QC geometry, FT_TAB otherwise
    constexpr int TXN = !TAB ? 1 : QC ? table_slots(2 * (FT_TX - 1) + QC * (KX - 1) + 1) : FT_TAB;
    constexpr int TYN = !TAB ? 1 : QC ? table_slots(2 * (TY - 1) + QC * (KYP - 1) + 1) : FT_TAB;
    // single-product forms on 64-row tiles: operand-pair tables (build_pairs, no
    // unpacking at the gathers); otherwise (hi, lo) packed per entry. (Pair planes for
    // the split forms, hi and lo, were measured 1 us slower: twice the table writes
    // and LDS for 40 fewer VALU instructions per layer and wave.)
    constexpr bool PAIRS = TAB && QC > 0 && TY == 64 && !SPLIT;
    // FP16 split forms on 64-row tiles (two workgroups per CU by LDS, VGPRs to spare):
    // each stage loads all its operands before its products (fp16x3 36.1 -> 34.9 us on
    // the 128^3 / 41 x 41 x 40-layer field; the BF16 split form measured no gain)
    constexpr bool PRE = TAB && QC > 0 && TY == 64 && SPLIT && F16;
    __shared__ unsigned Tx[TAB ? 2 : 1][TXN];
    __shared__ unsigned Ty[TAB ? 2 : 1][TYN];
    __shared__ LayerList ll;
    // contraction windows of the current spot chunk, per active layer: [rb] stage A of
    // row block rb (< TY / 16), [4 + w] stage B of wave w; k0 | nstep << 6
    __shared__ unsigned char win[FT_LB][8];
    // FP16 forms: the largest exponent of each wave's bounds (pass set-up), and the
    // factor c 2^e 2^-eacc of every active layer on the current spot chunk
    __shared__ int emax[F16 ? FT_THREADS / 64 : 1];
    __shared__ float ufac[F16 ? FT_LB : 1];

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int tiles = a.tiles_x * a.tiles_y * a.nz;
    const int ch = blk / tiles;                     // layer chunk of this workgroup
    const int tile = blk % tiles;
    const int tx = tile % a.tiles_x;
    const int ty = (tile / a.tiles_x) % a.tiles_y;
    const int z = tile / (a.tiles_x * a.tiles_y);
    const int x0 = tx * FT_TX, y0 = ty * TY;
    const float zc = (float)(oz + (double)z * dz);

    const int rbA = TY == 64 ? wave : (wave & 1);   // stage A row block (y)
    const int yA = 16 * rbA + (lane & 15);          // stage A row of the tile: B-operand column, output row
    const int gA = lane >> 4;                       // stage A k group (8 values)
    const int cbA0 = TY == 64 ? 0 : wave >> 1;      // TY = 32: stage A spot blocks cbA0 and, w < 2, 2
    const bool twoA = TY == 64 || wave < 2;
    const int cnB = 32 * wave;                      // stage B columns
    const int aB = cnB + (lane & 31);
    const int gB = lane >> 5;                       // stage B k group (8 values)
    const double yAc = a.oy + (double)(y0 + yA) * a.dy;
    const double xBc = a.ox + (double)(x0 + aB) * a.dx;
    TabStream<QC> sA, sB;
    sA.init(pl.ty, 8 * pl.ty.Q * gA - pl.ty.P * yA + pl.ty.P * (TY - 1));
    sB.init(pl.tx, 8 * pl.tx.Q * gB - pl.tx.P * aB + pl.tx.P * (FT_TX - 1));
    const int ntY = pl.ty.P * (TY - 1) + pl.ty.Q * (KYP - 1) + 1;
    const int ntX = pl.tx.P * (FT_TX - 1) + pl.tx.Q * (KX - 1) + 1;

    const int ncx = chunk_count(a.nsx), ncy = chunk_count(a.nsy);
    const double inv_dsx = 1.0 / a.dsx, inv_dsy = 1.0 / a.dsy;

    v16f acc[NBB] = {};
    int eacc = -1000;     // FP16 forms: the accumulators hold dose 2^-eacc (-1000: none yet)
    u4v wreg[NV];

    const int lo = ch * pl.chunk, hi = lo + pl.chunk;   // this workgroup's active layers
    int base = 0;                                       // active layers of earlier passes
    // LISTS: passes over 64 active layers of the slice's precomputed list (written by
    // the weight-split launch); otherwise passes over 64 layers, scanned here
    const int npass = LISTS ? pl.nlayers[z] : a.L;
    const int nq = ncy * ncx;                           // spot chunks
    const size_t wlayer = (size_t)nq * FT_WSPLIT;       // split weights per layer
    for (int lb = 0; lb < npass && base < hi; lb += FT_LB) {
        const int nact = LISTS ? copy_layers(pl, a.L, z, lb, zc, r_peak, ll)
                               : collect_layers(a, lb, z, zc, r_peak, ll);
        const int m0 = max(0, lo - base), m1 = min(nact, hi - base);
        base += nact;
        if (m1 <= m0) continue;
        if (F16) {
            // the accumulators hold dose 2^-eacc: eacc covers the largest bound of this
            // pass's layers on every spot chunk (the accumulated part is rescaled exactly
            // when a later pass needs a larger one). The (layer, chunk) bounds are read
            // in parallel, one per thread, and maximised over the workgroup.
            int e = -1000;
            for (int t = threadIdx.x; t < (m1 - m0) * nq; t += FT_THREADS) {
                const int m = m0 + t / nq;
                const float2 wp = pl.wpar[(size_t)ll.lay[m] * nq + t % nq];
                const float bound = ll.c[m] * wp.x * wp.y;
                if (bound > 0.0f && bound <= 3.0e38f) e = max(e, ilogbf(bound) - 13);
            }
            for (int o = 32; o > 0; o >>= 1) e = max(e, __shfl_xor(e, o, 64));
            if (lane == 0) emax[wave] = e;
            __syncthreads();
            e = max(max(emax[0], emax[1]), max(emax[2], emax[3]));
            if (e > eacc) {
                if (eacc > -1000) {
                    const float r = ldexpf(1.0f, eacc - e);
#pragma unroll
                    for (int rb = 0; rb < NBB; ++rb)
#pragma unroll
                        for (int r16 = 0; r16 < 16; ++r16) acc[rb][r16] *= r;
                }
                eacc = e;
            }
        }
        for (int cy = 0; cy < ncy; ++cy) {
            const int j0 = chunk_begin(a.nsy, ncy, cy), nj = chunk_begin(a.nsy, ncy, cy + 1) - j0;
            const double Cy = (a.oy + (double)y0 * a.dy) - (a.sy0 + (double)j0 * a.dsy);
            for (int cx = 0; cx < ncx; ++cx) {
                const int i0 = chunk_begin(a.nsx, ncx, cx), ni = chunk_begin(a.nsx, ncx, cx + 1) - i0;
                const double Cx = (a.ox + (double)x0 * a.dx) - (a.sx0 + (double)i0 * a.dsx);
                const TableRow rowx = TAB ? table_row(pl.tx, Cx, FT_TX) : TableRow{0, 0.0f, 0.0f};
                const TableRow rowy = TAB ? table_row(pl.ty, Cy, TY) : TableRow{0, 0.0f, 0.0f};
                // QC > 0: this thread's table entries, squared distances formed once per chunk
                TabShare<QC ? 2 * (FT_TX - 1) + QC * (KX - 1) + 1 : 1> shx;
                TabShare<QC ? 2 * (TY - 1) + QC * (KYP - 1) + 1 : 1> shy;
                if (TAB && QC) {
                    shx.init(pl.tx, rowx);
                    shy.init(pl.ty, rowy);
                }
                // PAIRS: slot of entry threadIdx.x - Q, the pair slot of each thread's entries
                const int s1x = PAIRS ? table_slot(pl.tx, (int)threadIdx.x - QC) : 0;
                const int s1y = PAIRS ? table_slot(pl.ty, (int)threadIdx.x - QC) : 0;
                // spots within reach of each wave's rows (stage A, K = y spots) and columns
                // (stage B, K = x spots), in spot units of the chunk; read after the next barrier
                for (int q = threadIdx.x; q < (m1 - m0) * 8; q += FT_THREADS) {
                    const int m = m0 + (q >> 3), w = q & 7;
                    if (w < 4 && w >= TY / 16) continue;
                    SpotWindow sw;
                    if (w < 4) {
                        const float lo = (float)((Cy + (double)(16 * w) * a.dy) * inv_dsy);
                        sw = spot_window(lo, lo + (float)(15.0 * a.dy * inv_dsy),
                                              ll.cut[m] * (float)inv_dsy, nj, 32, KYP);
                    } else {
                        const float lo = (float)((Cx + (double)(32 * (w - 4)) * a.dx) * inv_dsx);
                        sw = spot_window(lo, lo + (float)(31.0 * a.dx * inv_dsx),
                                              ll.cut[m] * (float)inv_dsx, ni, 16, 16 * KB);
                    }
                    win[m][w] = (unsigned char)(sw.k0 | (sw.nstep << 6));
                }
                // FP16 forms: U of layer m on this chunk is stored as c 2^e 2^-eacc U_s, below
                // 2^14 by the weight-split bound c 2^e b of |c Gy W| (U_s = Gy W 2^-e); the
                // factors of the chunk's layers, read after the next barrier (eacc still
                // unset: every bound of the pass is 0, all weights 0)
                if (F16) {
                    for (int m = m0 + (int)threadIdx.x; m < m1; m += FT_THREADS) {
                        const float2 wp = pl.wpar[((size_t)ll.lay[m] * ncy + cy) * ncx + cx];
                        ufac[m] = eacc > -1000 ? ldexpf(ll.c[m] * wp.x, -eacc) : 0.0f;
                    }
                }
                auto window = [&](int m, int w, int& k0) {
                    const int code = __builtin_amdgcn_readfirstlane((int)win[m][w]);
                    k0 = code & 63;
                    return code >> 6;
                };
                const unsigned short* wchunk = pl.wsplit + ((size_t)cy * ncx + cx) * FT_WSPLIT;

                auto is_x3 = [&](int m) {
                    if (MODE == MODE_X3) return true;
                    if (MODE == MODE_PLAIN) return false;
                    return ll.zpk[m] <= 1.0f;
                };
                auto load_w = [&](int m) {
                    const u4v* src = (const u4v*)(wchunk + (size_t)ll.lay[m] * wlayer);
#pragma unroll
                    for (int q = 0; q < NV; ++q) {
                        const int v = threadIdx.x + q * FT_THREADS;
                        if (NWV % FT_THREADS == 0 || v < NWV) wreg[q] = src[v];
                    }
                };
                auto store_w = [&](int b) {
#pragma unroll
                    for (int q = 0; q < NV; ++q) {
                        const int v = threadIdx.x + q * FT_THREADS;
                        if (NWV % FT_THREADS == 0 || v < NWV) {
                            const int e8 = v % WV;
                            T* dst = (v < WV ? Wh[b] : Wl[b]) + (e8 >> 3) * WT_STRIDE + 8 * (e8 & 7);
                            *(u4v*)dst = wreg[q];
                        }
                    }
                };
                auto build_tx = [&](int m) {
                    if (!TAB) return;
                    if (PAIRS) {
                        build_pairs<T, QC>(shx, s1x, Tx[m & 1], ll.k2[m]);
                        return;
                    }
                    if (QC) {
                        build_packed<T>(shx, Tx[m & 1], ll.k2[m]);
                        return;
                    }
                    const float k2 = ll.k2[m];
                    for (int t = threadIdx.x; t < ntX; t += FT_THREADS) {
                        const float d = table_distance_f(rowx, t);
                        Tx[m & 1][table_slot(pl.tx, t)] = pack_split<T>(__builtin_amdgcn_exp2f(-(d * d) * k2));
                    }
                };
                auto build_ty = [&](int m) {
                    if (!TAB) return;
                    if (PAIRS) {
                        build_pairs<T, QC>(shy, s1y, Ty[m & 1], ll.k2[m]);
                        return;
                    }
                    if (QC) {
                        build_packed<T>(shy, Ty[m & 1], ll.k2[m]);
                        return;
                    }
                    const float k2 = ll.k2[m];
                    for (int t = threadIdx.x; t < ntY; t += FT_THREADS) {
                        const float d = table_distance_f(rowy, t);
                        Ty[m & 1][table_slot(pl.ty, t)] = pack_split<T>(__builtin_amdgcn_exp2f(-(d * d) * k2));
                    }
                };
                // stage A of layer m: U[m&1] = c Gy W (computed as W^T Gy^T), split into (hi, lo)
                auto stage_a = [&](int m) {
                    const int b = m & 1;
                    const float c = ll.c[m], k2 = ll.k2[m];
                    const bool x3 = is_x3(m);
                    int k0;                  // K steps k0 + 32 ks, ks < nkA: y spots within reach
                    const int nkA = window(m, rbA, k0);
                    TabStream<QC> sw = sA;
                    sw.shift(pl.ty, k0);
                    // spot blocks of this wave: TY = 32: cbA0 and (w < 2) 2; TY = 64: 0, 1, 2
                    auto cb_of = [&](int k) { return TY == 64 ? k : (k == 0 ? cbA0 : 2); };
                    v4f u[NCB];
#pragma unroll
                    for (int k = 0; k < NCB; ++k) u[k] = v4f{};
                    if (PRE) {
                        // every operand of the layer's products first: the LDS reads of both
                        // K steps issue back to back and the products wait once
                        V8 gh[2], gl[2], wh[2][NCB], wl[2][NCB];
#pragma unroll
                        for (int ks = 0; ks < 2; ++ks) {
                            if (ks < nkA) {
                                unsigned t8[8];
#pragma unroll
                                for (int e = 0; e < 8; ++e) t8[e] = Ty[b][sw.slot(pl.ty, 32 * ks + e)];
                                unpack8(t8, gh[ks], gl[ks]);
                                const int kofs = k0 + 32 * ks + 8 * gA;
#pragma unroll
                                for (int k = 0; k < NCB; ++k) {
                                    const int off = (16 * k + (lane & 15)) * WT_STRIDE + kofs;
                                    wh[ks][k] = *(const V8*)&Wh[b][off];
                                    if (SPLIT) wl[ks][k] = *(const V8*)&Wl[b][off];
                                }
                            }
                        }
#pragma unroll
                        for (int ks = 0; ks < 2; ++ks) {
                            if (ks < nkA) {
#pragma unroll
                                for (int k = 0; k < NCB; ++k) {
                                    u[k] = mfma16(wh[ks][k], gh[ks], u[k]);
                                    if (SPLIT && x3) {
                                        u[k] = mfma16(wh[ks][k], gl[ks], u[k]);
                                        u[k] = mfma16(wl[ks][k], gh[ks], u[k]);
                                    }
                                }
                            }
                        }
                    }
#pragma unroll
                    for (int ks = 0; ks < 2 && !PRE; ++ks) {
                        if (ks < nkA) {
                            V8 gh, gl;      // Gy^T operand: j = k0 + 32 ks + 8 gA + e, y = yA
                            if (PAIRS) {
                                u4v h;
#pragma unroll
                                for (int q = 0; q < 4; ++q) h[q] = Ty[b][sw.slot(pl.ty, 32 * ks + 2 * q)];
                                gh = __builtin_bit_cast(V8, h);
                            } else if (TAB) {
                                unsigned t8[8];
#pragma unroll
                                for (int e = 0; e < 8; ++e) t8[e] = Ty[b][sw.slot(pl.ty, 32 * ks + e)];
                                unpack8(t8, gh, gl);
                            } else {
#pragma unroll
                                for (int e = 0; e < 8; ++e) {
                                    const int j = k0 + 32 * ks + 8 * gA + e;
                                    const double d = yAc - (a.sy0 + (double)(j0 + j) * a.dsy);
                                    const float g = (j < nj) ? gauss2((float)(d * d), k2) : 0.0f;
                                    T h, lo;
                                    split<T>(g, h, lo);
                                    gh[e] = h;
                                    gl[e] = lo;
                                }
                            }
                            const int kofs = k0 + 32 * ks + 8 * gA;
#pragma unroll
                            for (int k = 0; k < NCB; ++k) {
                                if (k > 0 && !twoA) continue;
                                const int off = (16 * cb_of(k) + (lane & 15)) * WT_STRIDE + kofs;
                                const V8 wh = *(const V8*)&Wh[b][off];
                                u[k] = mfma16(wh, gh, u[k]);
                                if (SPLIT && x3) {
                                    u[k] = mfma16(wh, gl, u[k]);
                                    u[k] = mfma16(*(const V8*)&Wl[b][off], gh, u[k]);
                                }
                            }
                        }
                    }
                    // FP16: U is stored as (c 2^e 2^-eU) U_s with the layer's power of two 2^eU,
                    // from the weight-split bound, putting |U| below 2^14 (no reduction needed)
                    const float cs = F16 ? ufac[m] : c;
                    // C layout of U^T (16x16): spots 16 cb + 4 gA + r (r < 4), row yA:
                    // four consecutive spots of U's row yA, one 8-byte store per half
                    auto put = [&](v4f u, int cb) {
                        u2v hv, lv;
#pragma unroll
                        for (int q = 0; q < 2; ++q) {
                            T h0, l0, h1, l1;
                            split<T>(cs * u[2 * q], h0, l0);
                            split<T>(cs * u[2 * q + 1], h1, l1);
                            hv[q] = (unsigned)bits16(h0) | ((unsigned)bits16(h1) << 16);
                            lv[q] = (unsigned)bits16(l0) | ((unsigned)bits16(l1) << 16);
                        }
                        const int off = yA * U_STRIDE + 16 * cb + 4 * gA;
                        *(u2v*)&Uh[b][off] = hv;
                        if (SPLIT) *(u2v*)&Ul[b][off] = lv;
                    };
#pragma unroll
                    for (int k = 0; k < NCB; ++k)
                        if (k == 0 || twoA) put(u[k], cb_of(k));
                };
                // stage B of layer m: acc += (c U) Gx^T
                auto stage_b = [&](int m) {
                    const int b = m & 1;
                    const bool x3 = is_x3(m);
                    const float k2 = ll.k2[m];
                    int k0;                  // K steps k0 + 16 ks, ks < nkB: x spots within reach
                    const int nkB = window(m, 4 + wave, k0);
                    TabStream<QC> sw = sB;
                    sw.shift(pl.tx, k0);
                    V8 bh[KB], bl[KB];
                    if (PRE) {
                        // every operand first (Gx gathers and the c U blocks of both row
                        // blocks), then the products
                        V8 ah[KB][NBB], al[KB][NBB];
#pragma unroll
                        for (int ks = 0; ks < KB; ++ks) {
                            if (ks < nkB) {
                                unsigned t8[8];
#pragma unroll
                                for (int e = 0; e < 8; ++e) t8[e] = Tx[b][sw.slot(pl.tx, 16 * ks + e)];
                                unpack8(t8, bh[ks], bl[ks]);
#pragma unroll
                                for (int rb = 0; rb < NBB; ++rb) {
                                    const int off = (32 * rb + (lane & 31)) * U_STRIDE + k0 + 16 * ks + 8 * gB;
                                    ah[ks][rb] = *(const V8*)&Uh[b][off];
                                    if (SPLIT) al[ks][rb] = *(const V8*)&Ul[b][off];
                                }
                            }
                        }
#pragma unroll
                        for (int ks = 0; ks < KB; ++ks) {
                            if (ks < nkB) {
#pragma unroll
                                for (int rb = 0; rb < NBB; ++rb) {
                                    acc[rb] = mfma32(ah[ks][rb], bh[ks], acc[rb]);
                                    if (SPLIT && x3) {
                                        acc[rb] = mfma32(ah[ks][rb], bl[ks], acc[rb]);
                                        acc[rb] = mfma32(al[ks][rb], bh[ks], acc[rb]);
                                    }
                                }
                            }
                        }
                        return;
                    }
#pragma unroll
                    for (int ks = 0; ks < KB; ++ks) {
                        if (ks >= nkB) break;
                        if (PAIRS) {
                            u4v h;
#pragma unroll
                            for (int q = 0; q < 4; ++q) h[q] = Tx[b][sw.slot(pl.tx, 16 * ks + 2 * q)];
                            bh[ks] = __builtin_bit_cast(V8, h);
                        } else if (TAB) {
                            unsigned t8[8];
#pragma unroll
                            for (int e = 0; e < 8; ++e) t8[e] = Tx[b][sw.slot(pl.tx, 16 * ks + e)];
                            unpack8(t8, bh[ks], bl[ks]);
                        } else {
#pragma unroll
                            for (int e = 0; e < 8; ++e) {
                                const int i = k0 + 16 * ks + 8 * gB + e;
                                const double d = xBc - (a.sx0 + (double)(i0 + i) * a.dsx);
                                const float g = (i < ni) ? gauss2((float)(d * d), k2) : 0.0f;
                                T h, lo;
                                split<T>(g, h, lo);
                                bh[ks][e] = h;
                                bl[ks][e] = lo;
                            }
                        }
                    }
                    // the products accumulate straight into acc (FP16: scaled by 2^-eacc); each
                    // gathered Gx operand feeds the wave's TY / 32 row blocks
#pragma unroll
                    for (int ks = 0; ks < KB; ++ks) {
                        if (ks >= nkB) break;
#pragma unroll
                        for (int rb = 0; rb < NBB; ++rb) {
                            const int off = (32 * rb + (lane & 31)) * U_STRIDE + k0 + 16 * ks + 8 * gB;
                            const V8 ah = *(const V8*)&Uh[b][off];
                            acc[rb] = mfma32(ah, bh[ks], acc[rb]);
                            if (SPLIT && x3) {
                                acc[rb] = mfma32(ah, bl[ks], acc[rb]);
                                acc[rb] = mfma32(*(const V8*)&Ul[b][off], bh[ks], acc[rb]);
                            }
                        }
                    }
                };

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
#pragma unroll
    for (int rb = 0; rb < NBB; ++rb) {
        if (F16 && eacc > -1000) {
            const float r = ldexpf(1.0f, eacc);
#pragma unroll
            for (int r16 = 0; r16 < 16; ++r16) acc[rb][r16] *= r;
        }
        store_block(a, pl, tile, ch, z, y0, x0, 32 * rb, cnB, acc[rb], (size_t)tiles);
    }
}

// Workgroups per CU: on the P = 2, Q = 5 geometry (tables of their interleaved span)
// three for the split forms (LDS near 52 KB) and four for the single-product forms (31 KB,
// VGPRs held to 128); two otherwise.
// 64-row tiles (QC = 5): two workgroups per CU for the split forms (LDS near 66 KB),
// three for the single-product forms.
template <int MODE, int QC, int TY> struct MixedOcc {
    static constexpr int value = !QC ? 2 : TY == 64 ? (MODE == MODE_PLAIN ? 3 : 2) : (MODE == MODE_PLAIN ? 4 : 3);
};

template <typename T, int MODE, bool TAB, int KB, int QC, int TY>
__global__ __launch_bounds__(FT_THREADS, (MixedOcc<MODE, QC, TY>::value))
void fwd_mixed_kernel(FwdArgs a, FwdPlan pl, double oz, double dz, const float* r_peak)
{
    fwd_mixed_body<T, MODE, TAB, KB, QC, true, TY>(a, pl, oz, dz, r_peak, blockIdx.x);
}

// One launch over a batch of problems sharing the template parameters. Its
// workgroups scan their slices' layers (no precomputed lists: with both the item
// lookup and the copy path the kernel needed more than the 168 VGPRs of 3 workgroups
// per CU and spilled).
//
// counters == null: workgroup u of the launch runs work unit u (u < total).
// counters != null (persistent launch, grid = the workgroups that fit at once): each
// workgroup claims units from counters[0] until they run out; the last workgroup to
// leave (counters[1]) sets both counters back to 0 for the next launch. Every unit
// is claimed once, so the results equal the one-workgroup-per-unit launch's.
template <typename T, int MODE, bool TAB, int KB, int QC, int TY>
__global__ __launch_bounds__(FT_THREADS, (MixedOcc<MODE, QC, TY>::value))
void fwd_mixed_batch_kernel(const BatchFwdItem* __restrict__ items, const int* __restrict__ wg0, int n, int total,
                            unsigned* counters)
{
    auto run = [&](int u) {
        const int k = batch_item(wg0, n, u);
        const FwdArgs a = items[k].a;
        const FwdPlan pl = items[k].pl;
        fwd_mixed_body<T, MODE, TAB, KB, QC, false, TY>(a, pl, items[k].oz, items[k].dz, items[k].r_peak,
                                                      u - wg0[k]);
    };
    if (!counters) {
        run((int)blockIdx.x);
        return;
    }
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0) next = (int)atomicAdd(&counters[0], 1u);
        __syncthreads();
        const int u = next;
        __syncthreads();           // every wave has read `next` (and left the last unit)
        if (u >= total) break;
        run(u);
    }
    if (threadIdx.x == 0 && atomicAdd(&counters[1], 1u) == gridDim.x - 1) {
        atomicExch(&counters[0], 0u);
        atomicExch(&counters[1], 0u);
    }
}

// r_peak of every layer of every item: one workgroup per item, one wave per layer
__global__ __launch_bounds__(256) void peak_depth_batch_kernel(const BatchPeakItem* __restrict__ items)
{
    const BatchPeakItem it = items[blockIdx.x];
    for (int l = threadIdx.x >> 6; l < it.L; l += blockDim.x >> 6) {
        const int best = first_max_wave(it.idd + (size_t)l * it.nz, it.nz);
        if ((threadIdx.x & 63) == 0) it.r_peak[l] = (float)(it.oz + (double)best * it.dz);
    }
}

// Template choice shared by the single-problem and batched launches.
struct MixedKey {
    int kb;       // stage B K steps (1..3)
    bool tab;     // table factors
    bool q5;      // P = 2, Q = 5 on both axes: compile-time table slots
};

static MixedKey mixed_key(const FwdPlan& pl, int kb)
{
    MixedKey k{kb, pl.table != 0, false};
    k.q5 = k.tab && pl.tx.P == 2 && pl.tx.Q == 5 && pl.ty.P == 2 && pl.ty.Q == 5;
    return k;
}

// Calls L.template go<T, MODE, TAB, KB, QC>() for the key's instantiation.
template <typename T, int MODE, typename Launch>
static int dispatch_mixed(const MixedKey& k, Launch&& L)
{
    if (!k.tab) return L.template go<T, MODE, false, 3, 0>();
    if (k.q5) {
        switch (k.kb) {
        case 1: return L.template go<T, MODE, true, 1, 5>();
        case 2: return L.template go<T, MODE, true, 2, 5>();
        case 3: return L.template go<T, MODE, true, 3, 5>();
        default: return PBH_ERR_UNSUPPORTED;
        }
    }
    switch (k.kb) {
    case 1: return L.template go<T, MODE, true, 1, 0>();
    case 2: return L.template go<T, MODE, true, 2, 0>();
    case 3: return L.template go<T, MODE, true, 3, 0>();
    default: return PBH_ERR_UNSUPPORTED;
    }
}

template <typename Launch>
static int dispatch_precision(int precision, const MixedKey& k, Launch&& L)
{
    switch (precision) {
    case PBH_PREC_BF16: return dispatch_mixed<__bf16, MODE_PLAIN>(k, L);
    case PBH_PREC_FP16: return dispatch_mixed<_Float16, MODE_PLAIN>(k, L);
    case PBH_PREC_BF16X3: return dispatch_mixed<__bf16, MODE_X3>(k, L);
    case PBH_PREC_FP16X3: return dispatch_mixed<_Float16, MODE_X3>(k, L);
    case PBH_PREC_BF16X3_PEAK: return dispatch_mixed<__bf16, MODE_PEAK>(k, L);
    case PBH_PREC_FP16X3_PEAK: return dispatch_mixed<_Float16, MODE_PEAK>(k, L);
    default: return PBH_ERR_UNSUPPORTED;
    }
}

static bool is_f16(int precision)
{
    return precision == PBH_PREC_FP16 || precision == PBH_PREC_FP16X3 || precision == PBH_PREC_FP16X3_PEAK;
}

struct SingleLaunch {
    const FwdArgs& a;
    const FwdPlan& pl;
    double oz, dz;
    const float* rp;
    hipStream_t s;
    template <typename T, int MODE, bool TAB, int KB, int QC>
    int go() const
    {
        const dim3 grid((unsigned)((long long)a.tiles_x * a.tiles_y * a.nz * pl.nchunk)), block(FT_THREADS);
        if (a.tile_rows == 64) {
            // 64-row tiles on the P = 2, Q = 5 table geometry only
            if (!(TAB && QC == 5)) return PBH_ERR_UNSUPPORTED;
            hipLaunchKernelGGL((fwd_mixed_kernel<T, MODE, TAB, KB, QC, 64>), grid, block, 0, s, a, pl, oz, dz, rp);
        } else {
            hipLaunchKernelGGL((fwd_mixed_kernel<T, MODE, TAB, KB, QC, 32>), grid, block, 0, s, a, pl, oz, dz, rp);
        }
        return launch_rc("fwd_mixed_kernel");
    }
};

struct BatchLaunch {
    const BatchFwdItem* items;
    const int* wg0;
    int n, total, rows;
    unsigned* counters;     // persistent launch: claim counters; null: one workgroup per unit
    hipStream_t s;
    template <typename T, int MODE, bool TAB, int KB, int QC>
    int go() const
    {
        if (rows == 64 && !(TAB && QC == 5)) return PBH_ERR_UNSUPPORTED;
        // a persistent launch has as many workgroups as are resident at once
        const int occ = rows == 64 ? MixedOcc<MODE, QC, 64>::value : MixedOcc<MODE, QC, 32>::value;
        const long long resident = (long long)occ * pbh_device_cus();
        const unsigned grid = counters ? (unsigned)std::min<long long>(total, resident) : (unsigned)total;
        if (rows == 64) {
            if constexpr (TAB && QC == 5)
                hipLaunchKernelGGL((fwd_mixed_batch_kernel<T, MODE, TAB, KB, QC, 64>), dim3(grid), dim3(FT_THREADS),
                                   0, s, items, wg0, n, total, counters);
        } else {
            hipLaunchKernelGGL((fwd_mixed_batch_kernel<T, MODE, TAB, KB, QC, 32>), dim3(grid), dim3(FT_THREADS), 0, s,
                               items, wg0, n, total, counters);
        }
        return launch_rc("fwd_mixed_batch_kernel");
    }
};

}  // namespace pbh

using namespace pbh;

float* pbh_forward_ws_peaks(const FwdArgs& a, void* ws);

int pbh_forward_mixed(const FwdArgs& a, double oz, double dz, int precision, int layer_chunk, void* ws,
                      size_t ws_bytes, hipStream_t s, hipEvent_t ev0, hipEvent_t ev1, const ObjFuse* fuse)
{
    const int kb = (chunk_pad(a.nsx, 1) + 15) / 16;   // stage B K steps of 16 spots
    if (!ws) return PBH_ERR_NULL_ARG;
    if (ws_bytes < pbh_forward_ws_bytes(a, a.L)) return PBH_ERR_WORKSPACE;
    FwdPlan pl;
    int rc = pbh_plan_forward(a, 16 * kb, KYP, ws, ws_bytes, layer_chunk, &pl);
    if (rc != PBH_OK) return rc;
    if (a.tile_rows == 64 && !mixed_key(pl, kb).q5) return PBH_ERR_UNSUPPORTED;   // before any launch
    float* r_peak = nullptr;
    if (precision == PBH_PREC_BF16X3_PEAK || precision == PBH_PREC_FP16X3_PEAK) {
        r_peak = pbh_forward_ws_peaks(a, ws);   // written by the weight-split launch
    }
    if (!pl.wsplit || !pl.layers) return PBH_ERR_WORKSPACE;
    const dim3 wgrid((unsigned)(a.L * chunk_count(a.nsy) * chunk_count(a.nsx)), WS_PARTS + 1);
    if (is_f16(precision))
        hipLaunchKernelGGL(wsplit_kernel<_Float16>, wgrid, dim3(256), 0, s, a, pl, oz, dz, r_peak);
    else
        hipLaunchKernelGGL(wsplit_kernel<__bf16>, wgrid, dim3(256), 0, s, a, pl, oz, dz, r_peak);
    if ((rc = launch_rc("wsplit_kernel")) != PBH_OK) return rc;
    if (ev0 && (rc = hip_rc(hipEventRecord(ev0, s), "hipEventRecord")) != PBH_OK) return rc;
    if ((rc = dispatch_precision(precision, mixed_key(pl, kb), SingleLaunch{a, pl, oz, dz, r_peak, s})) != PBH_OK)
        return rc;
    if (ev1 && (rc = hip_rc(hipEventRecord(ev1, s), "hipEventRecord")) != PBH_OK) return rc;
    return pbh_launch_combine(a, pl, s, fuse);
}

int pbh_launch_peak_batch(const BatchPeakItem* items, int n, hipStream_t s)
{
    hipLaunchKernelGGL(peak_depth_batch_kernel, dim3((unsigned)n), dim3(256), 0, s, items);
    return launch_rc("peak_depth_batch_kernel");
}

// Split weights of every item of a batch (grid: max chunks x items).
int pbh_launch_wsplit_batch(const BatchFwdItem* items, int n, int max_chunks, int precision, hipStream_t s)
{
    const dim3 grid((unsigned)max_chunks, WS_PARTS, (unsigned)n);
    if (is_f16(precision))
        hipLaunchKernelGGL(wsplit_batch_kernel<_Float16>, grid, dim3(256), 0, s, items);
    else
        hipLaunchKernelGGL(wsplit_batch_kernel<__bf16>, grid, dim3(256), 0, s, items);
    return launch_rc("wsplit_batch_kernel");
}

// Batched mixed-precision forward: items[0..n) share kb (stage B K steps), the
// table mode and the slot mode (q5); total = wg0[n].
int pbh_launch_fwd_mixed_batch(const BatchFwdItem* items, const int* wg0, int n, int total, int precision,
                               int kb, bool tab, bool q5, int rows, unsigned* counters, hipStream_t s)
{
    if (kb < 1 || kb > 3) return PBH_ERR_UNSUPPORTED;
    return dispatch_precision(precision, MixedKey{kb, tab, q5 && tab},
                              BatchLaunch{items, wg0, n, total, rows, counters, s});
}

// Template key of a planned item, for grouping a batch: bit 0 = table, bit 1 = q5.
int pbh_mixed_slot_mode(const FwdPlan& pl)
{
    const MixedKey k = mixed_key(pl, 1);
    return (k.tab ? 1 : 0) | (k.q5 ? 2 : 0);
}
