// pbh_mixed.hip -- mixed-precision forward dose: BF16 / FP16 Gaussian factors
// on gfx950 MFMA with FP32 accumulation, optionally as split products.
//
// Same tiling, layer split, Gaussian tables and double buffering as the FP32
// kernel (pbh_forward.hip):
//   stage A  U[64 x 48]  = Gy[64 x J] W[J x I]      v_mfma_f32_16x16x32_{bf16,f16}
//   stage B  D[64 x 128] += (c U) Gx^T              v_mfma_f32_32x32x16_{bf16,f16}
// Factors are formed in FP32 (table: FP64 argument), then rounded to the 16-bit
// type. In the split ("x3") forms every operand a is stored as hi = T(a),
// lo = T(a - hi) and a product is hi*hi + hi*lo + lo*hi (lo*lo dropped), which
// restores ~2^-16 (BF16) or ~2^-21 (FP16) relative accuracy per product at
// three MFMAs instead of one. Table entries hold the (hi, lo) pair packed in 32
// bits; operands are assembled with v_perm. The "_peak" forms use split products
// only for (layer, depth) pairs within 1 cm of the layer's Bragg peak (depth of
// its IDD maximum) and single products elsewhere.
//
// Each layer's contribution is summed in fresh FP32 accumulators and added to
// the running dose once. FP16 has a narrow exponent range, so U is scaled per
// wave by a power of two (largest magnitude into [1, 2)) and the inverse scale
// is applied to the layer's partial when it joins the running dose. BF16
// shares FP32's exponent range and needs no scaling.
#include "pbh_fwd_core.h"

namespace pbh {

constexpr int SCH = PBH_SPOT_CHUNK;       // 48 spots per chunk
constexpr int KYP = 64;                   // stage A K (j) padded to 2 steps of 32
constexpr int WT_STRIDE = KYP + 8;        // W^T rows in halves
constexpr int U_STRIDE = SCH + 8;         // U rows in halves

enum { MODE_PLAIN = 0, MODE_X3 = 1, MODE_PEAK = 2 };

typedef unsigned u4v __attribute__((ext_vector_type(4)));

template <typename T> struct Vec8;
template <> struct Vec8<__bf16> { typedef v8bf type; };
template <> struct Vec8<_Float16> { typedef v8h type; };

__device__ __forceinline__ v4f mfma16(v8bf a, v8bf b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
__device__ __forceinline__ v4f mfma16(v8h a, v8h b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
__device__ __forceinline__ v16f mfma32(v8bf a, v8bf b, v16f c) { return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0); }
__device__ __forceinline__ v16f mfma32(v8h a, v8h b, v16f c) { return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0); }

template <typename T>
__device__ __forceinline__ void split(float x, T& hi, T& lo)
{
    hi = (T)x;
    lo = (T)(x - (float)hi);
}

template <typename T>
__device__ __forceinline__ unsigned short bits16(T v) { return __builtin_bit_cast(unsigned short, v); }

// (hi, lo) of x packed as hi | lo << 16
template <typename T>
__device__ __forceinline__ unsigned pack_split(float x)
{
    T h, l;
    split<T>(x, h, l);
    return (unsigned)bits16(h) | ((unsigned)bits16(l) << 16);
}

// Eight packed entries -> hi and lo operand vectors.
template <typename V8>
__device__ __forceinline__ void unpack8(const unsigned* t, V8& hi, V8& lo)
{
    u4v h, l;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        h[q] = __builtin_amdgcn_perm(t[2 * q + 1], t[2 * q], 0x05040100u);
        l[q] = __builtin_amdgcn_perm(t[2 * q + 1], t[2 * q], 0x07060302u);
    }
    hi = __builtin_bit_cast(V8, h);
    lo = __builtin_bit_cast(V8, l);
}

// r_peak[l] = depth of the first IDD maximum of layer l (np.argmax semantics).
__global__ void peak_depth_f32_kernel(const double* idd, int L, int nz, double oz, double dz, float* r_peak)
{
    const int l = blockIdx.x * blockDim.x + threadIdx.x;
    if (l >= L) return;
    const double* row = idd + (size_t)l * nz;
    int best = 0;
    for (int k = 1; k < nz; ++k)
        if (row[k] > row[best]) best = k;
    r_peak[l] = (float)(oz + (double)best * dz);
}

// 4 waves (two workgroups per CU), one-barrier software pipeline over the
// active layers of each spot chunk, as in the FP32 kernel. Stage A: 16x16 blocks
// of U (2 row blocks x 3 column blocks); wave w takes block (w % 2, w / 2) and,
// for w < 2, block (w % 2, 2). Stage B: wave w owns the 32 x 32 dose block at
// columns 32 w; KB K steps of 16 spots.
//
// blk: the workgroup's index in the problem's own grid (tile * split + share).
template <typename T, int MODE, bool TAB, int KB>
__device__ __forceinline__ void fwd_mixed_body(const FwdArgs& a, const FwdPlan& pl, double oz, double dz,
                                               const float* r_peak, int blk)
{
    typedef typename Vec8<T>::type V8;
    constexpr bool F16 = !__is_same(T, __bf16);
    constexpr bool SPLIT = MODE != MODE_PLAIN;
    constexpr int KX = 16 * KB;                 // stage B K (spots in x), padded
    constexpr int NW = (SCH * KYP + FT_THREADS - 1) / FT_THREADS;   // W elements per thread

    __shared__ __attribute__((aligned(16))) T Wh[2][SCH * WT_STRIDE];      // W^T hi [i][j]
    __shared__ __attribute__((aligned(16))) T Wl[2][SPLIT ? SCH * WT_STRIDE : 8];
    __shared__ __attribute__((aligned(16))) T Uh[2][FT_TY * U_STRIDE];     // c U hi [y][i]
    __shared__ __attribute__((aligned(16))) T Ul[2][SPLIT ? FT_TY * U_STRIDE : 8];
    __shared__ unsigned Tx[TAB ? 2 : 1][TAB ? FT_TAB : 1];
    __shared__ unsigned Ty[TAB ? 2 : 1][TAB ? FT_TAB : 1];
    __shared__ float uscale[2][2];
    __shared__ float bmax[2][4];
    __shared__ LayerList ll;

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int s = blk % pl.split;
    const int tile = blk / pl.split;
    const int tx = tile % a.tiles_x;
    const int ty = (tile / a.tiles_x) % a.tiles_y;
    const int z = tile / (a.tiles_x * a.tiles_y);
    const int x0 = tx * FT_TX, y0 = ty * FT_TY;
    const float zc = (float)(oz + (double)z * dz);

    const int rbA = wave & 1;                       // stage A row block
    const int bA = 16 * rbA + (lane & 15);          // stage A row in the tile
    const int gA = lane >> 4;                       // stage A k group (8 values)
    const int cbA0 = wave >> 1;                     // stage A column blocks: cbA0 and, w < 2, 2
    const bool twoA = wave < 2;
    const int cnB = 32 * wave;                      // stage B columns
    const int aB = cnB + (lane & 31);
    const int gB = lane >> 5;                       // stage B k group (8 values)
    const double yAc = a.oy + (double)(y0 + bA) * a.dy;
    const double xBc = a.ox + (double)(x0 + aB) * a.dx;
    const int tA0 = 8 * pl.ty.Q * gA - pl.ty.P * bA + pl.ty.P * (FT_TY - 1);
    const int tB0 = 8 * pl.tx.Q * gB - pl.tx.P * aB + pl.tx.P * (FT_TX - 1);
    const int ntY = pl.ty.P * (FT_TY - 1) + pl.ty.Q * (KYP - 1) + 1;
    const int ntX = pl.tx.P * (FT_TX - 1) + pl.tx.Q * (KX - 1) + 1;

    const int ncx = chunk_count(a.nsx), ncy = chunk_count(a.nsy);

    v16f acc = {};
    float wreg[NW];

    for (int lb = s; lb < a.L; lb += FT_LB * pl.split) {
        const int nact = collect_layers(a, lb, pl.split, z, zc, r_peak, ll);
        if (nact == 0) continue;
        for (int cy = 0; cy < ncy; ++cy) {
            const int j0 = chunk_begin(a.nsy, ncy, cy), nj = chunk_begin(a.nsy, ncy, cy + 1) - j0;
            const int nkA = TAB ? 2 : ((nj + 31) >> 5);
            const double Cy = (a.oy + (double)y0 * a.dy) - (a.sy0 + (double)j0 * a.dsy);
            for (int cx = 0; cx < ncx; ++cx) {
                const int i0 = chunk_begin(a.nsx, ncx, cx), ni = chunk_begin(a.nsx, ncx, cx + 1) - i0;
                const double Cx = (a.ox + (double)x0 * a.dx) - (a.sx0 + (double)i0 * a.dsx);

                auto is_x3 = [&](int m) {
                    if (MODE == MODE_X3) return true;
                    if (MODE == MODE_PLAIN) return false;
                    return ll.zpk[m] <= 1.0f;
                };
                auto load_w = [&](int m) {
                    const float* wl = a.w + (size_t)ll.lay[m] * a.nsy * a.nsx + (size_t)j0 * a.nsx + i0;
#pragma unroll
                    for (int q = 0; q < NW; ++q) {
                        const int e = threadIdx.x + q * FT_THREADS;
                        const int i = e / KYP, j = e % KYP;
                        wreg[q] = (e < SCH * KYP && i < ni && j < nj) ? wl[(size_t)j * a.nsx + i] : 0.0f;
                    }
                };
                auto store_w = [&](int b) {
#pragma unroll
                    for (int q = 0; q < NW; ++q) {
                        const int e = threadIdx.x + q * FT_THREADS;
                        if (e < SCH * KYP) {
                            T h, lo;
                            split<T>(wreg[q], h, lo);
                            const int off = (e / KYP) * WT_STRIDE + (e % KYP);
                            Wh[b][off] = h;
                            if (SPLIT) Wl[b][off] = lo;
                        }
                    }
                };
                auto build_tx = [&](int m) {
                    if (!TAB) return;
                    const double k2d = ll.k2d[m];
                    for (int t = threadIdx.x; t < ntX; t += FT_THREADS) {
                        const double d = table_distance(pl.tx, Cx, FT_TX, t);
                        Tx[m & 1][table_slot(pl.tx, t)] =
                            pack_split<T>(__builtin_amdgcn_exp2f((float)(-(d * d) * k2d)));
                    }
                };
                auto build_ty = [&](int m) {
                    if (!TAB) return;
                    const double k2d = ll.k2d[m];
                    for (int t = threadIdx.x; t < ntY; t += FT_THREADS) {
                        const double d = table_distance(pl.ty, Cy, FT_TY, t);
                        Ty[m & 1][table_slot(pl.ty, t)] =
                            pack_split<T>(__builtin_amdgcn_exp2f((float)(-(d * d) * k2d)));
                    }
                };
                // stage A of layer m: U[m&1] = c Gy W, split into (hi, lo)
                auto stage_a = [&](int m) {
                    const int b = m & 1;
                    const float c = ll.c[m], k2 = ll.k2[m];
                    const bool x3 = is_x3(m);
                    v4f u0 = {}, u1 = {};
#pragma unroll
                    for (int ks = 0; ks < 2; ++ks) {
                        if (ks < nkA) {
                            V8 ah, al;
                            if (TAB) {
                                unsigned t8[8];
#pragma unroll
                                for (int e = 0; e < 8; ++e)
                                    t8[e] = Ty[b][table_slot(pl.ty, tA0 + pl.ty.Q * (32 * ks + e))];
                                unpack8(t8, ah, al);
                            } else {
#pragma unroll
                                for (int e = 0; e < 8; ++e) {
                                    const int j = 32 * ks + 8 * gA + e;
                                    const double d = yAc - (a.sy0 + (double)(j0 + j) * a.dsy);
                                    const float g = (j < nj) ? gauss2((float)(d * d), k2) : 0.0f;
                                    T h, lo;
                                    split<T>(g, h, lo);
                                    ah[e] = h;
                                    al[e] = lo;
                                }
                            }
                            const int kofs = 32 * ks + 8 * gA;
                            const int off0 = (16 * cbA0 + (lane & 15)) * WT_STRIDE + kofs;
                            const int off1 = (32 + (lane & 15)) * WT_STRIDE + kofs;
                            const V8 b0h = *(const V8*)&Wh[b][off0];
                            u0 = mfma16(ah, b0h, u0);
                            if (SPLIT && x3) {
                                u0 = mfma16(ah, *(const V8*)&Wl[b][off0], u0);
                                u0 = mfma16(al, b0h, u0);
                            }
                            if (twoA) {
                                const V8 b1h = *(const V8*)&Wh[b][off1];
                                u1 = mfma16(ah, b1h, u1);
                                if (SPLIT && x3) {
                                    u1 = mfma16(ah, *(const V8*)&Wl[b][off1], u1);
                                    u1 = mfma16(al, b1h, u1);
                                }
                            }
                        }
                    }
                    float inv_scale = 1.0f;
                    if (F16) {
                        // row block rbA is shared by waves rbA and rbA + 2: combine maxima in LDS
                        float mx = 0.0f;
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            mx = fmaxf(mx, fabsf(c * u0[r]));
                            if (twoA) mx = fmaxf(mx, fabsf(c * u1[r]));
                        }
                        for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
                        if (lane == 0) bmax[b][wave] = mx;
                        __syncthreads();
                        mx = fmaxf(bmax[b][rbA], bmax[b][rbA + 2]);
                        float scale = 1.0f;
                        if (mx > 0.0f && mx < 3.0e38f) {
                            const int ex = ilogbf(mx);            // 2^ex <= mx < 2^(ex+1)
                            scale = ldexpf(1.0f, ex);
                            inv_scale = ldexpf(1.0f, -ex);
                        }
                        if (lane == 0 && wave < 2) uscale[b][wave] = scale;
                    }
                    const int rbase = 16 * rbA + 4 * gA;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        const int row = (rbase + r) * U_STRIDE + (lane & 15);
                        T h, lo;
                        split<T>(c * u0[r] * inv_scale, h, lo);
                        Uh[b][row + 16 * cbA0] = h;
                        if (SPLIT) Ul[b][row + 16 * cbA0] = lo;
                        if (twoA) {
                            split<T>(c * u1[r] * inv_scale, h, lo);
                            Uh[b][row + 32] = h;
                            if (SPLIT) Ul[b][row + 32] = lo;
                        }
                    }
                };
                // stage B of layer m: acc += (c U) Gx^T
                auto stage_b = [&](int m) {
                    const int b = m & 1;
                    const bool x3 = is_x3(m);
                    const float k2 = ll.k2[m];
                    V8 bh[KB], bl[KB];
#pragma unroll
                    for (int ks = 0; ks < KB; ++ks) {
                        if (TAB) {
                            unsigned t8[8];
#pragma unroll
                            for (int e = 0; e < 8; ++e)
                                t8[e] = Tx[b][table_slot(pl.tx, tB0 + pl.tx.Q * (16 * ks + e))];
                            unpack8(t8, bh[ks], bl[ks]);
                        } else {
#pragma unroll
                            for (int e = 0; e < 8; ++e) {
                                const int i = 16 * ks + 8 * gB + e;
                                const double d = xBc - (a.sx0 + (double)(i0 + i) * a.dsx);
                                const float g = (i < ni) ? gauss2((float)(d * d), k2) : 0.0f;
                                T h, lo;
                                split<T>(g, h, lo);
                                bh[ks][e] = h;
                                bl[ks][e] = lo;
                            }
                        }
                    }
                    v16f p = {};
#pragma unroll
                    for (int ks = 0; ks < KB; ++ks) {
                        const int off = (lane & 31) * U_STRIDE + 16 * ks + 8 * gB;
                        const V8 ah = *(const V8*)&Uh[b][off];
                        p = mfma32(ah, bh[ks], p);
                        if (SPLIT && x3) {
                            p = mfma32(ah, bl[ks], p);
                            p = mfma32(*(const V8*)&Ul[b][off], bh[ks], p);
                        }
                    }
                    if (F16) {
                        // rows 0-15 / 16-31 came from stage-A row blocks 0 / 1
                        const float s0 = uscale[b][0], s1 = uscale[b][1];
#pragma unroll
                        for (int r = 0; r < 16; ++r) acc[r] += (r < 8 ? s0 : s1) * p[r];
                    } else {
                        acc += p;
                    }
                };

                load_w(0);
                store_w(0);
                build_ty(0);
                if (nact > 1) load_w(1);
                __syncthreads();
                stage_a(0);
                build_tx(0);
                if (nact > 1) {
                    store_w(1);
                    build_ty(1);
                }
                __syncthreads();
                for (int m = 0; m < nact; ++m) {
                    if (m + 2 < nact) load_w(m + 2);
                    stage_b(m);
                    if (m + 1 < nact) {
                        stage_a(m + 1);
                        build_tx(m + 1);
                    }
                    if (m + 2 < nact) {
                        store_w(m & 1);
                        build_ty(m + 2);
                    }
                    __syncthreads();
                }
            }
        }
    }
    store_block(a, pl, tile, s, z, y0, x0, 0, cnB, acc, (size_t)a.tiles_x * a.tiles_y * a.nz);
}

template <typename T, int MODE, bool TAB, int KB>
__global__ __launch_bounds__(FT_THREADS, 2)
void fwd_mixed_kernel(FwdArgs a, FwdPlan pl, double oz, double dz, const float* r_peak)
{
    fwd_mixed_body<T, MODE, TAB, KB>(a, pl, oz, dz, r_peak, blockIdx.x);
}

// One launch over a batch of problems sharing the template parameters.
template <typename T, int MODE, bool TAB, int KB>
__global__ __launch_bounds__(FT_THREADS, 2)
void fwd_mixed_batch_kernel(const BatchFwdItem* __restrict__ items, const int* __restrict__ wg0, int n)
{
    const int k = batch_item(wg0, n, blockIdx.x);
    const FwdArgs a = items[k].a;
    const FwdPlan pl = items[k].pl;
    fwd_mixed_body<T, MODE, TAB, KB>(a, pl, items[k].oz, items[k].dz, items[k].r_peak,
                                     (int)blockIdx.x - wg0[k]);
}

// r_peak of every layer of every item (one workgroup per item)
__global__ void peak_depth_batch_kernel(const BatchPeakItem* __restrict__ items)
{
    const BatchPeakItem it = items[blockIdx.x];
    for (int l = threadIdx.x; l < it.L; l += blockDim.x) {
        const double* row = it.idd + (size_t)l * it.nz;
        int best = 0;
        for (int k = 1; k < it.nz; ++k)
            if (row[k] > row[best]) best = k;
        it.r_peak[l] = (float)(it.oz + (double)best * it.dz);
    }
}

template <typename T, int MODE>
static void launch_mixed_batch(const BatchFwdItem* items, const int* wg0, int n, dim3 grid, int kb,
                               bool tab, hipStream_t s)
{
    const dim3 block(FT_THREADS);
    if (!tab) {
        hipLaunchKernelGGL((fwd_mixed_batch_kernel<T, MODE, false, 3>), grid, block, 0, s, items, wg0, n);
        return;
    }
    switch (kb) {
    case 1: hipLaunchKernelGGL((fwd_mixed_batch_kernel<T, MODE, true, 1>), grid, block, 0, s, items, wg0, n); break;
    case 2: hipLaunchKernelGGL((fwd_mixed_batch_kernel<T, MODE, true, 2>), grid, block, 0, s, items, wg0, n); break;
    default: hipLaunchKernelGGL((fwd_mixed_batch_kernel<T, MODE, true, 3>), grid, block, 0, s, items, wg0, n); break;
    }
}

template <typename T, int MODE>
static int launch_mixed(const FwdArgs& a, const FwdPlan& pl, int kb, double oz, double dz,
                        const float* rp, hipStream_t s)
{
    const dim3 grid((unsigned)((long long)a.tiles_x * a.tiles_y * a.nz * pl.split)), block(FT_THREADS);
    if (!pl.table) {
        hipLaunchKernelGGL((fwd_mixed_kernel<T, MODE, false, 3>), grid, block, 0, s, a, pl, oz, dz, rp);
        return PBH_OK;
    }
    switch (kb) {
    case 1: hipLaunchKernelGGL((fwd_mixed_kernel<T, MODE, true, 1>), grid, block, 0, s, a, pl, oz, dz, rp); break;
    case 2: hipLaunchKernelGGL((fwd_mixed_kernel<T, MODE, true, 2>), grid, block, 0, s, a, pl, oz, dz, rp); break;
    case 3: hipLaunchKernelGGL((fwd_mixed_kernel<T, MODE, true, 3>), grid, block, 0, s, a, pl, oz, dz, rp); break;
    default: return PBH_ERR_UNSUPPORTED;
    }
    return PBH_OK;
}

}  // namespace pbh

using namespace pbh;

float* pbh_forward_ws_peaks(const FwdArgs& a, void* ws);

int pbh_forward_mixed(const FwdArgs& a, double oz, double dz, int precision, int split, void* ws,
                      size_t ws_bytes, hipStream_t s)
{
    const int kb = (chunk_pad(a.nsx, 1) + 15) / 16;   // stage B K steps of 16 spots
    FwdPlan pl;
    int rc = pbh_plan_forward(a, 16 * kb, KYP, ws, ws_bytes, split, &pl);
    if (rc != PBH_OK) return rc;
    float* r_peak = nullptr;
    if (precision == PBH_PREC_BF16X3_PEAK || precision == PBH_PREC_FP16X3_PEAK) {
        if (!ws) return PBH_ERR_NULL_ARG;
        if (ws_bytes < pbh_forward_ws_bytes(a, 1)) return PBH_ERR_WORKSPACE;
        r_peak = pbh_forward_ws_peaks(a, ws);
        hipLaunchKernelGGL(peak_depth_f32_kernel, dim3(cdiv(a.L, 64)), dim3(64), 0, s,
                           a.idd, a.L, a.nz, oz, dz, r_peak);
        if ((rc = launch_rc("peak_depth_f32_kernel")) != PBH_OK) return rc;
    }
    switch (precision) {
    case PBH_PREC_BF16: rc = launch_mixed<__bf16, MODE_PLAIN>(a, pl, kb, oz, dz, r_peak, s); break;
    case PBH_PREC_FP16: rc = launch_mixed<_Float16, MODE_PLAIN>(a, pl, kb, oz, dz, r_peak, s); break;
    case PBH_PREC_BF16X3: rc = launch_mixed<__bf16, MODE_X3>(a, pl, kb, oz, dz, r_peak, s); break;
    case PBH_PREC_FP16X3: rc = launch_mixed<_Float16, MODE_X3>(a, pl, kb, oz, dz, r_peak, s); break;
    case PBH_PREC_BF16X3_PEAK: rc = launch_mixed<__bf16, MODE_PEAK>(a, pl, kb, oz, dz, r_peak, s); break;
    case PBH_PREC_FP16X3_PEAK: rc = launch_mixed<_Float16, MODE_PEAK>(a, pl, kb, oz, dz, r_peak, s); break;
    default: return PBH_ERR_UNSUPPORTED;
    }
    if (rc != PBH_OK) return rc;
    return launch_rc("fwd_mixed_kernel");
}

int pbh_launch_peak_batch(const BatchPeakItem* items, int n, hipStream_t s)
{
    hipLaunchKernelGGL(peak_depth_batch_kernel, dim3((unsigned)n), dim3(64), 0, s, items);
    return launch_rc("peak_depth_batch_kernel");
}

// Batched mixed-precision forward: items[0..n) share kb (stage B K steps) and
// the table mode; total = wg0[n].
int pbh_launch_fwd_mixed_batch(const BatchFwdItem* items, const int* wg0, int n, int total, int precision,
                               int kb, bool tab, hipStream_t s)
{
    if (kb < 1 || kb > 3) return PBH_ERR_UNSUPPORTED;
    const dim3 grid((unsigned)total);
    switch (precision) {
    case PBH_PREC_BF16: launch_mixed_batch<__bf16, MODE_PLAIN>(items, wg0, n, grid, kb, tab, s); break;
    case PBH_PREC_FP16: launch_mixed_batch<_Float16, MODE_PLAIN>(items, wg0, n, grid, kb, tab, s); break;
    case PBH_PREC_BF16X3: launch_mixed_batch<__bf16, MODE_X3>(items, wg0, n, grid, kb, tab, s); break;
    case PBH_PREC_FP16X3: launch_mixed_batch<_Float16, MODE_X3>(items, wg0, n, grid, kb, tab, s); break;
    case PBH_PREC_BF16X3_PEAK: launch_mixed_batch<__bf16, MODE_PEAK>(items, wg0, n, grid, kb, tab, s); break;
    case PBH_PREC_FP16X3_PEAK: launch_mixed_batch<_Float16, MODE_PEAK>(items, wg0, n, grid, kb, tab, s); break;
    default: return PBH_ERR_UNSUPPORTED;
    }
    return launch_rc("fwd_mixed_batch_kernel");
}
