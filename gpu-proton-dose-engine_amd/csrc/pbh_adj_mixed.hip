// pbh_adj_mixed.hip -- spot-weight adjoint in BF16 on gfx950 MFMA (plain or
// split "x3" products), FP32 accumulation.
//
// For a dose cotangent R and each (layer l, depth z):
//   Q   = R_z Gx_lz          (64 x 48 per tile: rows y, spots i; K = 128 voxels x)
//   P  += Gy_lz^T Q          (48 x 48: spots j, i; K = 64 voxels y)
//   dJ/dW_l = sum_z c_lz P_lz,  c_lz = IDD_l(z) / (2 pi s^2)
// One workgroup (4 waves) per depth slice and share of its layers, as in the
// FP32 adjoint (pbh_adjoint.hip). Per dose tile (64 x 128) the wave's 16 rows of
// R are loaded once into registers and split into BF16 (hi, lo). Per layer:
//   stage 1  Q = R Gx      v_mfma_f32_16x16x32_bf16, A = R (registers), B = Gx
//                          gathered from the packed (hi, lo) x table
//   Q^T is stored split in LDS, four consecutive rows per 8-byte write
//   stage 2  P = Gy^T Q    v_mfma_f32_16x16x32_bf16, A = Gy^T from the y table,
//                          B = Q^T rows (16-byte reads)
// and c P joins the per-slice FP32 partial that adj_reduce_kernel sums over z
// in FP64 (deterministic). Split products: hi*hi + hi*lo + lo*hi of the
// (hi, lo) parts of both operands; plain: hi*hi. Layers with IDD = 0 at the
// slice contribute nothing and are skipped. Gradients with respect to the
// look-up tables (dJ/dIDD, dJ/dsigma) use the FP32 adjoint.
#include "pbh_batch.h"
#include "pbh_mixed_core.h"

namespace pbh {

constexpr int AM_TY = 64, AM_TX = 128, AM_SC = PBH_SPOT_CHUNK;   // 64 x 128 tile, 48-spot chunk
constexpr int AM_QS = AM_TY + 8;                                  // Q^T row stride (halves)

// Table geometry of an axis for tile width T and chunk width K (as axis_table in
// the forward): entry t = Q i - P a + P (T - 1) for voxel a, spot i.
template <int QC, int T, int K>
struct AmTab {
    static constexpr int N = QC ? 2 * ((2 * (T - 1) + QC * (K - 1) + 2) / 2) + 2 : FT_TAB;
};

template <int MODE, int QC>
__device__ __forceinline__ void adj_mixed_body(const AdjArgs& A, int blk)
{
    typedef __bf16 T;
    typedef v8bf V8;
    constexpr bool SPLIT = MODE != MODE_PLAIN;
    constexpr int TXN = AmTab<QC, AM_TX, AM_SC>::N, TYN = AmTab<QC, AM_TY, AM_SC>::N;
    const FwdArgs& a = A.f;

    __shared__ unsigned Tx[TXN], Ty[TYN];
    __shared__ __attribute__((aligned(16))) T Qh[AM_SC * AM_QS];
    __shared__ __attribute__((aligned(16))) T Ql[SPLIT ? AM_SC * AM_QS : 8];

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, c16 = lane & 15;
    const int s = blk % A.split, z = blk / A.split;
    const int tiles_x = cdiv(a.nx, AM_TX), tiles_y = cdiv(a.ny, AM_TY);
    const int ncx = chunk_count(a.nsx), ncy = chunk_count(a.nsy);
    const float* Rz = A.R + (size_t)z * a.ny * a.nx;
    const AxisTable& gx = A.tx;
    const AxisTable& gy = A.ty;
    // stage 1 B operand (Gx): x = 32 ks + 8 g + e, i = 16 nb + c16:
    //   t = Q i - P x + P (AM_TX - 1) = t0 + 16 Q nb - 32 P ks - P e
    TabStream<QC> s1, s2;
    s1.init(gx, gx.Q * c16 - 8 * gx.P * g + gx.P * (AM_TX - 1));
    // stage 2 A operand (Gy^T): y = 32 ks + 8 g + e, j = 16 mb + c16
    s2.init(gy, gy.Q * c16 - 8 * gy.P * g + gy.P * (AM_TY - 1));
    const int ntX = gx.P * (AM_TX - 1) + gx.Q * (AM_SC - 1) + 1;
    const int ntY = gy.P * (AM_TY - 1) + gy.Q * (AM_SC - 1) + 1;

    auto slot1 = [&](int nb, int ks, int e) {
        if (QC > 0) return s1.at(16 * QC * nb - 64 * ks - 2 * e);     // P = 2
        return table_slot(gx, s1.bE + 16 * gx.Q * nb - 32 * gx.P * ks - gx.P * e);
    };
    auto slot2 = [&](int mb, int ks, int e) {
        if (QC > 0) return s2.at(16 * QC * mb - 64 * ks - 2 * e);
        return table_slot(gy, s2.bE + 16 * gy.Q * mb - 32 * gy.P * ks - gy.P * e);
    };

    for (int cy = 0; cy < ncy; ++cy) {
        const int j0 = chunk_begin(a.nsy, ncy, cy), nj = chunk_begin(a.nsy, ncy, cy + 1) - j0;
        for (int cx = 0; cx < ncx; ++cx) {
            const int i0 = chunk_begin(a.nsx, ncx, cx), ni = chunk_begin(a.nsx, ncx, cx + 1) - i0;
            for (int tyi = 0; tyi < tiles_y; ++tyi) {
                for (int txi = 0; txi < tiles_x; ++txi) {
                    const int y0 = tyi * AM_TY, x0 = txi * AM_TX;
                    const bool first_tile = tyi == 0 && txi == 0;
                    // this wave's 16 rows of R (x = 32 ks + 8 g + e), split into (hi, lo)
                    V8 rh[4], rl[4];
                    {
                        const int y = y0 + 16 * wave + c16;
#pragma unroll
                        for (int ks = 0; ks < 4; ++ks)
#pragma unroll
                            for (int e = 0; e < 8; ++e) {
                                const int x = x0 + 32 * ks + 8 * g + e;
                                const float v = (y < a.ny && x < a.nx) ? Rz[(size_t)y * a.nx + x] : 0.0f;
                                T h, l;
                                split<T>(v, h, l);
                                rh[ks][e] = h;
                                rl[ks][e] = l;
                            }
                    }
                    const double Cx = (a.ox + (double)x0 * a.dx) - (a.sx0 + (double)i0 * a.dsx);
                    const double Cy = (a.oy + (double)y0 * a.dy) - (a.sy0 + (double)j0 * a.dsy);
                    for (int l = s; l < a.L; l += A.split) {
                        const double iddv = a.idd[(size_t)l * a.nz + z];
                        float* pz = A.part + ((size_t)z * a.L + l) * a.nsy * a.nsx;
                        if (iddv == 0.0) {      // uniform in the workgroup: no dose, zero partial
                            if (first_tile)
                                for (int e = threadIdx.x; e < ni * nj; e += FT_THREADS)
                                    pz[(size_t)(j0 + e / ni) * a.nsx + i0 + e % ni] = 0.0f;
                            continue;
                        }
                        const double sg = a.sigma[(size_t)l * a.nz + z];
                        const double sd = sg > kSigmaFloor ? sg : kSigmaFloor;
                        const double k2d = kLog2e / (2.0 * sd * sd);
                        const double cn = iddv / (2.0 * kPi * sd * sd);
                        __syncthreads();        // previous layer's readers of the tables and Q are done
                        for (int t = threadIdx.x; t < ntX; t += FT_THREADS) {
                            const double d = table_distance(gx, Cx, AM_TX, t);
                            Tx[table_slot(gx, t)] = pack_split<T>(__builtin_amdgcn_exp2f((float)(-(d * d) * k2d)));
                        }
                        for (int t = threadIdx.x; t < ntY; t += FT_THREADS) {
                            const double d = table_distance(gy, Cy, AM_TY, t);
                            Ty[table_slot(gy, t)] = pack_split<T>(__builtin_amdgcn_exp2f((float)(-(d * d) * k2d)));
                        }
                        __syncthreads();
                        // stage 1: Q[y][i] for this wave's 16 rows and the three 16-spot blocks
#pragma unroll 1
                        for (int nb = 0; nb < 3; ++nb) {
                            v4f q = {};
#pragma unroll
                            for (int ks = 0; ks < 4; ++ks) {
                                unsigned t8[8];
#pragma unroll
                                for (int e = 0; e < 8; ++e) t8[e] = Tx[slot1(nb, ks, e)];
                                V8 bh, bl;
                                unpack8(t8, bh, bl);
                                q = mfma16(rh[ks], bh, q);
                                if (SPLIT) {
                                    q = mfma16(rh[ks], bl, q);
                                    q = mfma16(rl[ks], bh, q);
                                }
                            }
                            // C layout: rows y = 16 wave + 4 g + r, column i = 16 nb + c16
                            u2v hv, lv;
#pragma unroll
                            for (int h2 = 0; h2 < 2; ++h2) {
                                T a0, b0, a1, b1;
                                split<T>(q[2 * h2], a0, b0);
                                split<T>(q[2 * h2 + 1], a1, b1);
                                hv[h2] = (unsigned)bits16(a0) | ((unsigned)bits16(a1) << 16);
                                lv[h2] = (unsigned)bits16(b0) | ((unsigned)bits16(b1) << 16);
                            }
                            const int off = (16 * nb + c16) * AM_QS + 16 * wave + 4 * g;
                            *(u2v*)&Qh[off] = hv;
                            if (SPLIT) *(u2v*)&Ql[off] = lv;
                        }
                        __syncthreads();
                        // stage 2: P[j][i] blocks (mb, nb) = wave + 4 qq, qq = 0..2, of the 3 x 3
#pragma unroll 1
                        for (int qq = 0; qq < 3; ++qq) {
                            const int b2 = wave + 4 * qq;
                            if (b2 >= 9) break;
                            const int mb = b2 / 3, nb = b2 % 3;
                            v4f p = {};
#pragma unroll
                            for (int ks = 0; ks < 2; ++ks) {
                                unsigned t8[8];
#pragma unroll
                                for (int e = 0; e < 8; ++e) t8[e] = Ty[slot2(mb, ks, e)];
                                V8 ah, al;
                                unpack8(t8, ah, al);
                                const int off = (16 * nb + c16) * AM_QS + 32 * ks + 8 * g;
                                const V8 bh = *(const V8*)&Qh[off];
                                p = mfma16(ah, bh, p);
                                if (SPLIT) {
                                    p = mfma16(ah, *(const V8*)&Ql[off], p);
                                    p = mfma16(al, bh, p);
                                }
                            }
                            // C layout: j = 16 mb + 4 g + r, i = 16 nb + c16
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                const int j = 16 * mb + 4 * g + r, i = 16 * nb + c16;
                                if (j < nj && i < ni) {
                                    const size_t off = (size_t)(j0 + j) * a.nsx + i0 + i;
                                    const float cp = (float)(cn * (double)p[r]);
                                    pz[off] = first_tile ? cp : pz[off] + cp;
                                }
                            }
                        }
                    }
                }
            }
        }
    }
}

template <int MODE, int QC>
__global__ __launch_bounds__(FT_THREADS, 2)
void adj_mixed_kernel(AdjArgs A)
{
    adj_mixed_body<MODE, QC>(A, blockIdx.x);
}

template <int MODE, int QC>
__global__ __launch_bounds__(FT_THREADS, 2)
void adj_mixed_batch_kernel(const AdjArgs* __restrict__ items, const int* __restrict__ wg0, int n)
{
    const int k = batch_item(wg0, n, blockIdx.x);
    const AdjArgs A = items[k];
    adj_mixed_body<MODE, QC>(A, (int)blockIdx.x - wg0[k]);
}

}  // namespace pbh

using namespace pbh;

// Table geometry of the mixed adjoint for one problem: AdjArgs.tx / ty with the
// 64 x 128 tile and 48-spot chunks; q5 when P = 2, Q = 5 on both axes.
bool pbh_adj_mixed_q5(const AdjArgs& A)
{
    return A.tab && A.tx.P == 2 && A.tx.Q == 5 && A.ty.P == 2 && A.ty.Q == 5;
}

// Launches the mixed adjoint (dJ/dw only) for one problem; A.tab must be set
// (the mixed adjoint gathers every factor from tables).
int pbh_launch_adj_mixed(const AdjArgs& A, int precision, int nz, hipStream_t s)
{
    const dim3 grid((unsigned)(nz * A.split)), block(FT_THREADS);
    const bool x3 = precision != PBH_PREC_BF16;
    const bool q5 = pbh_adj_mixed_q5(A);
    if (x3) {
        if (q5) hipLaunchKernelGGL((adj_mixed_kernel<MODE_X3, 5>), grid, block, 0, s, A);
        else hipLaunchKernelGGL((adj_mixed_kernel<MODE_X3, 0>), grid, block, 0, s, A);
    } else {
        if (q5) hipLaunchKernelGGL((adj_mixed_kernel<MODE_PLAIN, 5>), grid, block, 0, s, A);
        else hipLaunchKernelGGL((adj_mixed_kernel<MODE_PLAIN, 0>), grid, block, 0, s, A);
    }
    return launch_rc("adj_mixed_kernel");
}

int pbh_launch_adj_mixed_batch(const AdjArgs* items, const int* wg0, int n, int total, int precision, bool q5,
                               hipStream_t s)
{
    const dim3 grid((unsigned)total), block(FT_THREADS);
    const bool x3 = precision != PBH_PREC_BF16;
    if (x3) {
        if (q5) hipLaunchKernelGGL((adj_mixed_batch_kernel<MODE_X3, 5>), grid, block, 0, s, items, wg0, n);
        else hipLaunchKernelGGL((adj_mixed_batch_kernel<MODE_X3, 0>), grid, block, 0, s, items, wg0, n);
    } else {
        if (q5) hipLaunchKernelGGL((adj_mixed_batch_kernel<MODE_PLAIN, 5>), grid, block, 0, s, items, wg0, n);
        else hipLaunchKernelGGL((adj_mixed_batch_kernel<MODE_PLAIN, 0>), grid, block, 0, s, items, wg0, n);
    }
    return launch_rc("adj_mixed_batch_kernel");
}
