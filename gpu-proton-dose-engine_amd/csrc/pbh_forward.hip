// pbh_forward.hip -- fused forward dose contraction, FP32 MFMA (gfx950).
//
// For one depth slice z, the dose tile D[y0:y0+64, x0:x0+128] is
//     D = sum_l sum_chunks  (c_lz * Gy_lz[:, J] W_l[J, I]) Gx_lz[:, I]^T
// with c_lz = IDD_l(z) / (2 pi s^2). Each workgroup (4 waves) owns one tile and
// loops over layers and spot chunks, keeping D in MFMA accumulators, so no
// per-layer intermediate ever reaches HBM (the reference materialises Gx, Gy,
// A, T and Dl per layer: engines.py:261-274).
//
//   stage A  U[64 x 48]  = Gy[64 x J] * W[J x I]     v_mfma_f32_16x16x4f32
//            wave w computes rows 16w..16w+15; Gy is built in registers
//            (A operand), W is staged in LDS.
//   stage B  D[64 x 128] += (c U)[64 x I] * Gx^T      v_mfma_f32_32x32x2f32
//            wave w computes columns 32w..32w+31; Gx is built in registers
//            (B operand), c*U is read from LDS.
// Squared voxel-spot distances are evaluated in FP64 once per chunk and kept
// in registers; per layer only d2 * k2 and one exp2 remain per factor.
#include "pbh_common.h"

namespace pbh {

constexpr int FWD_TY = 64;
constexpr int FWD_TX = 128;
constexpr int FWD_THREADS = 256;
constexpr int SC = PBH_SPOT_CHUNK;   // 48
constexpr int LDS_STRIDE = SC + 1;   // odd stride: spreads rows over banks


__global__ __launch_bounds__(FWD_THREADS)
void fwd_fp32_kernel(FwdArgs a)
{
    __shared__ float Ws[SC * LDS_STRIDE];        // W chunk [j][i]
    __shared__ float Us[FWD_TY * LDS_STRIDE];    // c * U   [y][i]

    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    int b = blockIdx.x;
    const int tx = b % a.tiles_x; b /= a.tiles_x;
    const int ty = b % a.tiles_y; b /= a.tiles_y;
    const int z = b;
    const int x0 = tx * FWD_TX, y0 = ty * FWD_TY;

    // stage-A lane geometry: row yA (A operand row = lane % 16), k = 4 kk + lane / 16
    const int yA = y0 + 16 * wave + (lane & 15);
    // stage-B lane geometry: column xB (B operand column = lane % 32), k = 2 kk + lane / 32
    const int xB = x0 + 32 * wave + (lane & 31);
    const double yAc = a.oy + (double)yA * a.dy;
    const double xBc = a.ox + (double)xB * a.dx;

    v16f acc0 = {}, acc1 = {};

    const int ncx = chunk_count(a.nsx), ncy = chunk_count(a.nsy);
    for (int cy = 0; cy < ncy; ++cy) {
        const int j0 = chunk_begin(a.nsy, ncy, cy), nj = chunk_begin(a.nsy, ncy, cy + 1) - j0;
        const int nkA = (nj + 3) >> 2;   // K steps of 4 in stage A
        // d2 for Gy: k = 4 kk + lane/16, kk < 12 (48 / 4)
        float dyA[SC / 4];
#pragma unroll
        for (int kk = 0; kk < SC / 4; ++kk) {
            const int j = 4 * kk + (lane >> 4);
            const double d = yAc - (a.sy0 + (double)(j0 + j) * a.dsy);
            dyA[kk] = (j < nj && yA < a.ny) ? (float)(d * d) : -1.0f;   // -1: padded (factor 0)
        }
        for (int cx = 0; cx < ncx; ++cx) {
            const int i0 = chunk_begin(a.nsx, ncx, cx), ni = chunk_begin(a.nsx, ncx, cx + 1) - i0;
            const int nkB = (ni + 1) >> 1;   // K steps of 2 in stage B
            float dxB[SC / 2];
#pragma unroll
            for (int kk = 0; kk < SC / 2; ++kk) {
                const int i = 2 * kk + (lane >> 5);
                const double d = xBc - (a.sx0 + (double)(i0 + i) * a.dsx);
                dxB[kk] = (i < ni) ? (float)(d * d) : -1.0f;
            }
            for (int l = 0; l < a.L; ++l) {
                const double iddv = a.idd[(size_t)l * a.nz + z];
                if (iddv == 0.0) continue;   // layer deposits nothing at this depth (uniform branch)
                const LayerScalars sc = layer_scalars(iddv, a.sigma[(size_t)l * a.nz + z]);

                // stage W chunk into LDS (zero padded to 48 x 48)
                const float* wl = a.w + (size_t)l * a.nsy * a.nsx;
                for (int e = threadIdx.x; e < SC * SC; e += FWD_THREADS) {
                    const int j = e / SC, i = e % SC;
                    Ws[j * LDS_STRIDE + i] = (j < nj && i < ni) ? wl[(size_t)(j0 + j) * a.nsx + i0 + i] : 0.0f;
                }
                __syncthreads();

                // stage A: U[16 rows of this wave][48] = Gy * W
                v4f u0 = {}, u1 = {}, u2 = {};
#pragma unroll
                for (int kk = 0; kk < SC / 4; ++kk) {
                    if (kk < nkA) {
                        const float g = dyA[kk] >= 0.0f ? gauss2(dyA[kk], sc.k2) : 0.0f;
                        const int k = 4 * kk + (lane >> 4);
                        const float* wr = &Ws[k * LDS_STRIDE + (lane & 15)];
                        u0 = __builtin_amdgcn_mfma_f32_16x16x4f32(g, wr[0], u0, 0, 0, 0);
                        u1 = __builtin_amdgcn_mfma_f32_16x16x4f32(g, wr[16], u1, 0, 0, 0);
                        u2 = __builtin_amdgcn_mfma_f32_16x16x4f32(g, wr[32], u2, 0, 0, 0);
                    }
                }
                // C layout of 16x16x4: row = 4 (lane / 16) + r, column = lane % 16
                {
                    const int rbase = 16 * wave + 4 * (lane >> 4);
                    const int col = lane & 15;
#pragma unroll
                    for (int r = 0; r < 4; ++r) {
                        float* urow = &Us[(rbase + r) * LDS_STRIDE + col];
                        urow[0] = sc.c * u0[r];
                        urow[16] = sc.c * u1[r];
                        urow[32] = sc.c * u2[r];
                    }
                }
                __syncthreads();

                // stage B: D[64 x 32 columns of this wave] += (cU) Gx^T. The layer's
                // partial is summed in fresh accumulators and added to D once, which
                // keeps the FP32 rounding chain per voxel at (K steps + layers).
                v16f p0 = {}, p1 = {};
#pragma unroll
                for (int kk = 0; kk < SC / 2; ++kk) {
                    if (kk < nkB) {
                        const float g = dxB[kk] >= 0.0f ? gauss2(dxB[kk], sc.k2) : 0.0f;
                        const int k = 2 * kk + (lane >> 5);
                        const float a0 = Us[(lane & 31) * LDS_STRIDE + k];
                        const float a1 = Us[(32 + (lane & 31)) * LDS_STRIDE + k];
                        p0 = __builtin_amdgcn_mfma_f32_32x32x2f32(a0, g, p0, 0, 0, 0);
                        p1 = __builtin_amdgcn_mfma_f32_32x32x2f32(a1, g, p1, 0, 0, 0);
                    }
                }
                acc0 += p0;
                acc1 += p1;
                __syncthreads();
            }
        }
    }

    // C layout of 32x32: row = 8 (r / 4) + 4 (lane / 32) + r % 4, column = lane % 32
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
                      const double* idd, const double* sigma, float* dose)
{
    FwdArgs a;
    a.nx = g->nx; a.ny = g->ny; a.nz = g->nz;
    a.nsx = lat->n_sx; a.nsy = lat->n_sy; a.L = lat->n_layers;
    a.dx = g->dx; a.dy = g->dy; a.ox = g->ox; a.oy = g->oy;
    a.dsx = lat->dx_spot; a.dsy = lat->dy_spot;
    a.sx0 = lat->x0 - (lat->n_sx - 1) / 2.0 * lat->dx_spot;
    a.sy0 = lat->y0 - (lat->n_sy - 1) / 2.0 * lat->dy_spot;
    a.w = w; a.idd = idd; a.sigma = sigma; a.dose = dose;
    a.tiles_x = cdiv(g->nx, FWD_TX);
    a.tiles_y = cdiv(g->ny, FWD_TY);
    return a;
}

int pbh_forward_mixed(const FwdArgs& a, double oz, double dz, int precision, void* ws, size_t ws_bytes,
                      hipStream_t s);
size_t pbh_forward_mixed_ws(const PbhGrid* g, const PbhLattice* lat, int precision);

extern "C" size_t pbh_forward_workspace_bytes(const PbhGrid* g, const PbhLattice* lat, int precision)
{
    if (!g || !lat) return 0;
    if (precision == PBH_PREC_FP32) return 0;
    return pbh_forward_mixed_ws(g, lat, precision);
}

extern "C" int pbh_forward(const PbhGrid* g, const PbhLattice* lat, const float* w,
                           const double* idd, const double* sigma, float* dose, int precision,
                           void* workspace, size_t workspace_bytes, void* stream)
{
    int rc = validate_problem(g, lat);
    if (rc != PBH_OK) return rc;
    if (!w || !idd || !sigma || !dose) return PBH_ERR_NULL_ARG;
    if (precision < PBH_PREC_FP32 || precision > PBH_PREC_FP16X3_PEAK) return PBH_ERR_UNSUPPORTED;
    const FwdArgs a = make_fwd_args(g, lat, w, idd, sigma, dose);
    hipStream_t s = (hipStream_t)stream;
    if (precision != PBH_PREC_FP32)
        return pbh_forward_mixed(a, g->oz, g->dz, precision, workspace, workspace_bytes, s);
    const long long blocks = (long long)a.tiles_x * a.tiles_y * g->nz;
    if (blocks > 0x7fffffffLL) return PBH_ERR_UNSUPPORTED;
    hipLaunchKernelGGL(fwd_fp32_kernel, dim3((unsigned)blocks), dim3(FWD_THREADS), 0, s, a);
    return launch_rc("fwd_fp32_kernel");
}
