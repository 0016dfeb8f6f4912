// pbh_mixed.hip -- mixed-precision forward dose: BF16 / FP16 Gaussian factors
// on gfx950 MFMA with FP32 accumulation, optionally as split products.
//
// Same tiling as the FP32 kernel (one workgroup = 64 rows x 128 columns of one
// depth slice, 4 waves, layers and spot chunks looped inside):
//   stage A  U[64 x 48]  = Gy[64 x J] W[J x I]      v_mfma_f32_16x16x32_{bf16,f16}
//   stage B  D[64 x 128] += (c U) Gx^T              v_mfma_f32_32x32x16_{bf16,f16}
// Factors are formed in FP32 (exp2 of an FP32 argument), then rounded to the
// 16-bit type. In the split ("x3") form every operand a is stored as
// hi = T(a), lo = T(a - hi) and a product is hi*hi + hi*lo + lo*hi (lo*lo is
// dropped), which restores ~2^-16 (BF16) or ~2^-21 (FP16) relative accuracy per
// product at three MFMAs instead of one. The "_peak" forms use the split
// products only for (layer, depth) pairs within 1 cm of the layer's Bragg peak
// (the depth of its IDD maximum) and single products elsewhere.
//
// Each layer's contribution is summed in fresh FP32 accumulators and added to
// the running dose once. FP16 has a narrow exponent range, so U is scaled per
// wave by a power of two (its largest magnitude is brought into [1, 2)) and the
// inverse scale is applied to the layer's partial when it joins the running
// dose. BF16 shares FP32's exponent range and needs no scaling.
#include "pbh_common.h"

namespace pbh {

constexpr int MX_TY = 64, MX_TX = 128, MX_THREADS = 256;
constexpr int SCH = PBH_SPOT_CHUNK;       // 48 spots per chunk
constexpr int WT_STRIDE = 64 + 8;         // W^T rows: k (j) padded to 64, +8 halves of pad
constexpr int U_STRIDE = SCH + 8;         // U rows: k (i) padded to 48, +8 halves of pad

enum { MODE_PLAIN = 0, MODE_X3 = 1, MODE_PEAK = 2 };

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

template <typename T, int MODE>
__global__ __launch_bounds__(MX_THREADS)
void fwd_mixed_kernel(FwdArgs a, double oz, double dz, const float* r_peak)
{
    typedef typename Vec8<T>::type V8;
    constexpr bool F16 = sizeof(T) == 2 && !__is_same(T, __bf16);
    constexpr bool SPLIT = MODE != MODE_PLAIN;

    __shared__ __attribute__((aligned(16))) T Wh[SCH * WT_STRIDE];      // W^T hi [i][j]
    __shared__ __attribute__((aligned(16))) T Wl[SPLIT ? SCH * WT_STRIDE : 8];
    __shared__ __attribute__((aligned(16))) T Uh[MX_TY * U_STRIDE];     // c U hi [y][i]
    __shared__ __attribute__((aligned(16))) T Ul[SPLIT ? MX_TY * U_STRIDE : 8];
    __shared__ float uscale[4];

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    int b = blockIdx.x;
    const int tx = b % a.tiles_x; b /= a.tiles_x;
    const int ty = b % a.tiles_y; b /= a.tiles_y;
    const int z = b;
    const int x0 = tx * MX_TX, y0 = ty * MX_TY;
    const float zc = (float)(oz + (double)z * dz);

    const int yA = y0 + 16 * wave + (lane & 15);      // stage A row of this lane
    const int gA = lane >> 4;                         // stage A k group (8 values)
    const int xB = x0 + 32 * wave + (lane & 31);      // stage B column of this lane
    const int gB = lane >> 5;                         // stage B k group (8 values)
    const double yAc = a.oy + (double)yA * a.dy;
    const double xBc = a.ox + (double)xB * a.dx;

    v16f acc0 = {}, acc1 = {};

    const int ncx = chunk_count(a.nsx), ncy = chunk_count(a.nsy);
    for (int cy = 0; cy < ncy; ++cy) {
        const int j0 = chunk_begin(a.nsy, ncy, cy), nj = chunk_begin(a.nsy, ncy, cy + 1) - j0;
        const int nkA = (nj + 31) >> 5;               // K steps of 32
        float dyA[2][8];
#pragma unroll
        for (int ks = 0; ks < 2; ++ks)
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const int j = 32 * ks + 8 * gA + e;
                const double d = yAc - (a.sy0 + (double)(j0 + j) * a.dsy);
                dyA[ks][e] = (j < nj && yA < a.ny) ? (float)(d * d) : -1.0f;
            }
        for (int cx = 0; cx < ncx; ++cx) {
            const int i0 = chunk_begin(a.nsx, ncx, cx), ni = chunk_begin(a.nsx, ncx, cx + 1) - i0;
            const int nkB = (ni + 15) >> 4;           // K steps of 16
            float dxB[3][8];
#pragma unroll
            for (int ks = 0; ks < 3; ++ks)
#pragma unroll
                for (int e = 0; e < 8; ++e) {
                    const int i = 16 * ks + 8 * gB + e;
                    const double d = xBc - (a.sx0 + (double)(i0 + i) * a.dsx);
                    dxB[ks][e] = (i < ni) ? (float)(d * d) : -1.0f;
                }
            for (int l = 0; l < a.L; ++l) {
                const double iddv = a.idd[(size_t)l * a.nz + z];
                if (iddv == 0.0) continue;
                const LayerScalars sc = layer_scalars(iddv, a.sigma[(size_t)l * a.nz + z]);
                bool x3 = MODE == MODE_X3;
                if (MODE == MODE_PEAK) x3 = fabsf(zc - r_peak[l]) <= 1.0f;

                // W chunk -> LDS as W^T (hi, lo), zero padded
                const float* wl = a.w + (size_t)l * a.nsy * a.nsx;
                for (int e = threadIdx.x; e < SCH * 64; e += MX_THREADS) {
                    const int i = e >> 6, j = e & 63;
                    const float v = (i < ni && j < nj) ? wl[(size_t)(j0 + j) * a.nsx + i0 + i] : 0.0f;
                    T h, lo;
                    split<T>(v, h, lo);
                    Wh[i * WT_STRIDE + j] = h;
                    if (SPLIT) Wl[i * WT_STRIDE + j] = lo;
                }
                __syncthreads();

                // stage A: 16 rows x 48 columns per wave
                v4f u[3] = {};
#pragma unroll
                for (int ks = 0; ks < 2; ++ks) {
                    if (ks < nkA) {
                        V8 ah, al;
#pragma unroll
                        for (int e = 0; e < 8; ++e) {
                            const float g = dyA[ks][e] >= 0.0f ? gauss2(dyA[ks][e], sc.k2) : 0.0f;
                            T h, lo;
                            split<T>(g, h, lo);
                            ah[e] = h;
                            al[e] = lo;
                        }
#pragma unroll
                        for (int nb = 0; nb < 3; ++nb) {
                            const int off = (16 * nb + (lane & 15)) * WT_STRIDE + 32 * ks + 8 * gA;
                            const V8 bh = *(const V8*)&Wh[off];
                            u[nb] = mfma16(ah, bh, u[nb]);
                            if (SPLIT && x3) {
                                const V8 bl = *(const V8*)&Wl[off];
                                u[nb] = mfma16(ah, bl, u[nb]);
                                u[nb] = mfma16(al, bh, u[nb]);
                            }
                        }
                    }
                }
                // c U -> LDS (hi, lo); FP16: scale the wave's rows into [1, 2)
                float inv_scale = 1.0f, scale = 1.0f;
                if (F16) {
                    float m = 0.0f;
#pragma unroll
                    for (int nb = 0; nb < 3; ++nb)
#pragma unroll
                        for (int r = 0; r < 4; ++r) m = fmaxf(m, fabsf(sc.c * u[nb][r]));
                    for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
                    if (m > 0.0f && m < 3.0e38f) {
                        const int ex = ilogbf(m);            // 2^ex <= m < 2^(ex+1)
                        scale = ldexpf(1.0f, ex);
                        inv_scale = ldexpf(1.0f, -ex);
                    }
                    if (lane == 0) uscale[wave] = scale;
                }
                {
                    const int rbase = 16 * wave + 4 * gA;
#pragma unroll
                    for (int nb = 0; nb < 3; ++nb)
#pragma unroll
                        for (int r = 0; r < 4; ++r) {
                            T h, lo;
                            split<T>(sc.c * u[nb][r] * inv_scale, h, lo);
                            const int off = (rbase + r) * U_STRIDE + 16 * nb + (lane & 15);
                            Uh[off] = h;
                            if (SPLIT) Ul[off] = lo;
                        }
                }
                __syncthreads();

                // stage B: 64 rows x 32 columns per wave
                v16f p0 = {}, p1 = {};
#pragma unroll
                for (int ks = 0; ks < 3; ++ks) {
                    if (ks < nkB) {
                        V8 bh, bl;
#pragma unroll
                        for (int e = 0; e < 8; ++e) {
                            const float g = dxB[ks][e] >= 0.0f ? gauss2(dxB[ks][e], sc.k2) : 0.0f;
                            T h, lo;
                            split<T>(g, h, lo);
                            bh[e] = h;
                            bl[e] = lo;
                        }
                        const int k = 16 * ks + 8 * gB;
                        const int off0 = (lane & 31) * U_STRIDE + k;
                        const int off1 = (32 + (lane & 31)) * U_STRIDE + k;
                        const V8 a0h = *(const V8*)&Uh[off0];
                        const V8 a1h = *(const V8*)&Uh[off1];
                        p0 = mfma32(a0h, bh, p0);
                        p1 = mfma32(a1h, bh, p1);
                        if (SPLIT && x3) {
                            const V8 a0l = *(const V8*)&Ul[off0];
                            const V8 a1l = *(const V8*)&Ul[off1];
                            p0 = mfma32(a0h, bl, p0);
                            p0 = mfma32(a0l, bh, p0);
                            p1 = mfma32(a1h, bl, p1);
                            p1 = mfma32(a1l, bh, p1);
                        }
                    }
                }
                // the layer's partial joins the accumulator once (short rounding chain);
                // FP16: rows 0-15 / 16-31 of p0 came from stage-A waves 0 / 1, of p1 from 2 / 3
                if (F16) {
                    const float s0 = uscale[0], s1 = uscale[1], s2 = uscale[2], s3 = uscale[3];
#pragma unroll
                    for (int r = 0; r < 16; ++r) {
                        acc0[r] += (r < 8 ? s0 : s1) * p0[r];
                        acc1[r] += (r < 8 ? s2 : s3) * p1[r];
                    }
                } else {
                    acc0 += p0;
                    acc1 += p1;
                }
                __syncthreads();
            }
        }
    }

    if (xB < a.nx) {
        float* base = a.dose + (size_t)z * a.ny * a.nx + xB;
#pragma unroll
        for (int r = 0; r < 16; ++r) {
            const int row = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
            const int y = y0 + row;
            if (y < a.ny) base[(size_t)y * a.nx] = acc0[r];
            if (y + 32 < a.ny) base[(size_t)(y + 32) * a.nx] = acc1[r];
        }
    }
}

}  // namespace pbh

using namespace pbh;

size_t pbh_forward_mixed_ws(const PbhGrid*, const PbhLattice* lat, int precision)
{
    if (precision == PBH_PREC_BF16X3_PEAK || precision == PBH_PREC_FP16X3_PEAK)
        return (size_t)lat->n_layers * sizeof(float);
    return 0;
}

int pbh_forward_mixed(const FwdArgs& a, double oz, double dz, int precision, void* ws,
                      size_t ws_bytes, hipStream_t s)
{
    const long long blocks = (long long)a.tiles_x * a.tiles_y * a.nz;
    if (blocks > 0x7fffffffLL) return PBH_ERR_UNSUPPORTED;
    const dim3 grid((unsigned)blocks), block(MX_THREADS);
    float* r_peak = nullptr;
    if (precision == PBH_PREC_BF16X3_PEAK || precision == PBH_PREC_FP16X3_PEAK) {
        if (!ws) return PBH_ERR_NULL_ARG;
        if (ws_bytes < (size_t)a.L * sizeof(float)) return PBH_ERR_WORKSPACE;
        r_peak = (float*)ws;
        hipLaunchKernelGGL(peak_depth_f32_kernel, dim3(cdiv(a.L, 64)), dim3(64), 0, s,
                           a.idd, a.L, a.nz, oz, dz, r_peak);
        int rc = launch_rc("peak_depth_f32_kernel");
        if (rc != PBH_OK) return rc;
    }
    switch (precision) {
    case PBH_PREC_BF16:
        hipLaunchKernelGGL((fwd_mixed_kernel<__bf16, MODE_PLAIN>), grid, block, 0, s, a, oz, dz, r_peak); break;
    case PBH_PREC_FP16:
        hipLaunchKernelGGL((fwd_mixed_kernel<_Float16, MODE_PLAIN>), grid, block, 0, s, a, oz, dz, r_peak); break;
    case PBH_PREC_BF16X3:
        hipLaunchKernelGGL((fwd_mixed_kernel<__bf16, MODE_X3>), grid, block, 0, s, a, oz, dz, r_peak); break;
    case PBH_PREC_FP16X3:
        hipLaunchKernelGGL((fwd_mixed_kernel<_Float16, MODE_X3>), grid, block, 0, s, a, oz, dz, r_peak); break;
    case PBH_PREC_BF16X3_PEAK:
        hipLaunchKernelGGL((fwd_mixed_kernel<__bf16, MODE_PEAK>), grid, block, 0, s, a, oz, dz, r_peak); break;
    case PBH_PREC_FP16X3_PEAK:
        hipLaunchKernelGGL((fwd_mixed_kernel<_Float16, MODE_PEAK>), grid, block, 0, s, a, oz, dz, r_peak); break;
    default:
        return PBH_ERR_UNSUPPORTED;
    }
    return launch_rc("fwd_mixed_kernel");
}
