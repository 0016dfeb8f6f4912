// pbh_adjoint.hip -- adjoint of the forward dose map, FP32 MFMA (gfx950), plus the
// weighted-quadratic objective and the range-shift chain rule.
//
// Forward: D_z = sum_l c_lz Gy_lz W_l Gx_lz^T, c_lz = IDD_l(z) / (2 pi s^2),
// G[k][i] = exp(-(v_k - s_i)^2 / (2 s^2)), s = max(sigma_l(z), 1e-9).
// For a dose cotangent R (nz, ny, nx) and P_lz = Gy_lz^T R_z Gx_lz (nsy x nsx):
//   dJ/dW_l       = sum_z c_lz P_lz
//   dJ/dIDD_l(z)  = <W_l, P_lz> / (2 pi s^2)
//   dJ/dsigma_l(z)= c_lz [ -(2/s) <W_l, P_lz> + <W_l, Gy'^T R Gx + Gy^T R Gx'> ]  (sigma > 1e-9)
//   with G' = dG/ds = G (v - s_i)^2 / s^3.
//
// Kernel 1 (one workgroup per depth slice and share of layers): the slice of R is
// staged in LDS tile by tile (64 x 128, once per tile), then per layer
//   Q  = R Gx, Q' = R Gx'     (M = 64 rows, N = spots, K = 128 columns; v_mfma_f32_16x16x4f32)
//   P += Gy^T Q, P2 += Gy'^T Q + Gy^T Q'   (M = N = spots, K = 64 rows)
// with the Gaussians gathered from per-(layer, depth) difference tables as in the
// forward kernels (direct evaluation on grids without a small pitch ratio). The
// per-slice products c P go to a workspace; <W, P> and <W, P2> are reduced in
// the workgroup in a fixed order.
// Kernel 2 sums the per-slice FP32 partials over z in FP64 in increasing z: the
// weight gradient is deterministic (no atomics).
#include <algorithm>
#include <functional>
#include <vector>

#include "pbh_batch.h"

namespace pbh {

constexpr int AJ_TY = 64, AJ_TX = 128, AJ_THREADS = 256;
constexpr int AJ_SC = PBH_SPOT_CHUNK;            // 48
constexpr int AJ_RS = AJ_TX + 2;                 // R tile row stride: lanes (c, g) hit banks 2c + g
constexpr int AJ_QS = AJ_SC + 1;                 // Q row stride

// Slot of table entry t for axes of AJ_TX / AJ_TY voxels (same layout rules as the forward).
__device__ __forceinline__ float tab_get(const float* T, const AxisTable& g, int t) { return T[table_slot(g, t)]; }

constexpr int AJ_MAX_SPLIT = 16;  // workgroups per depth slice, at most
constexpr int AJ_LB = 64;   // layers of one workgroup handled per pass over the slice tiles

// blk: the workgroup's index in the problem's own grid (slice * split + share).
template <bool TAB, bool SIG, int QC>
__device__ __forceinline__ void adj_fp32_body(const AdjArgs& A, int blk)
{
    const FwdArgs& a = A.f;
    __shared__ float Rs[AJ_TY * AJ_RS];
    __shared__ float Qs[AJ_TY * AJ_QS];
    __shared__ float Q2s[SIG ? AJ_TY * AJ_QS : 1];
    __shared__ float Tx[TAB ? FT_TAB : 1], Tx2[(TAB && SIG) ? FT_TAB : 1];
    __shared__ float Ty[TAB ? FT_TAB : 1], Ty2[(TAB && SIG) ? FT_TAB : 1];
    __shared__ double red[2][AJ_THREADS / 64];
    __shared__ double wacc[2][AJ_LB];          // <W, P>, <W, P2> per layer of the pass

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int s = blk % A.split;
    const int z = blk / A.split;
    const int g = lane >> 4, c16 = lane & 15;
    // QC > 0 (P = 2, Q = QC on both axes): slots of the lane's k = 0, block-0 entries
    const int sgx = TAB && QC ? table_slot(A.tx, A.tx.Q * c16 - A.tx.P * g + A.tx.P * (AJ_TX - 1)) : 0;
    const int sgy = TAB && QC ? table_slot(A.ty, A.ty.Q * c16 - A.ty.P * g + A.ty.P * (AJ_TY - 1)) : 0;
    const int tiles_x = cdiv(a.nx, AJ_TX), tiles_y = cdiv(a.ny, AJ_TY);
    const int ncx = chunk_count(a.nsx), ncy = chunk_count(a.nsy);
    const float* Rz = A.R + (size_t)z * a.ny * a.nx;
    const bool lut_grads = A.grad_idd || (SIG && A.grad_sigma);

    for (int lb = s; lb < a.L; lb += AJ_LB * A.split) {
        const int nl = min(AJ_LB, (a.L - lb + A.split - 1) / A.split);
        if (threadIdx.x < AJ_LB) {
            wacc[0][threadIdx.x] = 0.0;
            wacc[1][threadIdx.x] = 0.0;
        }
        for (int cy = 0; cy < ncy; ++cy) {
            const int j0 = chunk_begin(a.nsy, ncy, cy), nj = chunk_begin(a.nsy, ncy, cy + 1) - j0;
            for (int cx = 0; cx < ncx; ++cx) {
                const int i0 = chunk_begin(a.nsx, ncx, cx), ni = chunk_begin(a.nsx, ncx, cx + 1) - i0;
                for (int tyi = 0; tyi < tiles_y; ++tyi) {
                    for (int txi = 0; txi < tiles_x; ++txi) {
                        const int y0 = tyi * AJ_TY, x0 = txi * AJ_TX;
                        const bool first_tile = tyi == 0 && txi == 0;
                        __syncthreads();   // previous readers of Rs are done
                        for (int e = threadIdx.x; e < AJ_TY * AJ_TX; e += AJ_THREADS) {
                            const int yy = e / AJ_TX, xx = e % AJ_TX;
                            const int y = y0 + yy, x = x0 + xx;
                            Rs[yy * AJ_RS + xx] = (y < a.ny && x < a.nx) ? Rz[(size_t)y * a.nx + x] : 0.0f;
                        }
                        const double Cx = (a.ox + (double)x0 * a.dx) - (a.sx0 + (double)i0 * a.dsx);
                        const double Cy = (a.oy + (double)y0 * a.dy) - (a.sy0 + (double)j0 * a.dsy);
                        // QC > 0: this thread's table entries, squared distances formed once per tile
                        TabShare<QC ? 2 * (AJ_TX - 1) + QC * (AJ_SC - 1) + 1 : 1> shx;
                        TabShare<QC ? 2 * (AJ_TY - 1) + QC * (AJ_SC - 1) + 1 : 1> shy;
                        if (TAB && QC) {
                            shx.init(A.tx, table_row(A.tx, Cx, AJ_TX));
                            shy.init(A.ty, table_row(A.ty, Cy, AJ_TY));
                        }

                        for (int q = 0; q < nl; ++q) {
                            const int l = lb + q * A.split;
                            const double iddv = a.idd[(size_t)l * a.nz + z];
                            // no dose from this layer at this depth: its dJ/dw partial is zero
                            // and the depth reduction skips it (uniform in the workgroup)
                            if (iddv == 0.0 && !lut_grads) continue;
                            const double sg = a.sigma[(size_t)l * a.nz + z];
                            const double sd = sg > kSigmaFloor ? sg : kSigmaFloor;
                            const double k2d = kLog2e / (2.0 * sd * sd);
                            const float k2 = (float)k2d;
                            const float is3 = (float)(1.0 / (sd * sd * sd));   // dG/ds = G d^2 / s^3
                            const double cn = iddv / (2.0 * kPi * sd * sd);
                            if (TAB && QC) {
                                // G and, for dJ/dsigma, G d^2 / s^3 from the per-tile squared distances
                                auto build = [&](const auto& sh, float* t1, float* t2) {
#pragma unroll
                                    for (int k = 0; k < sh.N; ++k) {
                                        if (!sh.ok(k)) continue;
                                        const int sl = sh.s0 + AJ_THREADS / 2 * k;
                                        const float gv = __builtin_amdgcn_exp2f(sh.d2[k] * -k2);
                                        t1[sl] = gv;
                                        if (SIG) t2[sl] = gv * sh.d2[k] * is3;
                                    }
                                };
                                build(shx, Tx, Tx2);
                                build(shy, Ty, Ty2);
                            } else if (TAB) {
                                const int ntX = A.tx.P * (AJ_TX - 1) + A.tx.Q * (AJ_SC - 1) + 1;
                                const int ntY = A.ty.P * (AJ_TY - 1) + A.ty.Q * (AJ_SC - 1) + 1;
                                const TableRow rx = table_row(A.tx, Cx, AJ_TX), ry = table_row(A.ty, Cy, AJ_TY);
                                for (int t = threadIdx.x; t < ntX; t += AJ_THREADS) {
                                    const float d = table_distance_f(rx, t);
                                    const float gv = __builtin_amdgcn_exp2f(-(d * d) * k2);
                                    Tx[table_slot(A.tx, t)] = gv;
                                    if (SIG) Tx2[table_slot(A.tx, t)] = gv * (d * d) * is3;
                                }
                                for (int t = threadIdx.x; t < ntY; t += AJ_THREADS) {
                                    const float d = table_distance_f(ry, t);
                                    const float gv = __builtin_amdgcn_exp2f(-(d * d) * k2);
                                    Ty[table_slot(A.ty, t)] = gv;
                                    if (SIG) Ty2[table_slot(A.ty, t)] = gv * (d * d) * is3;
                                }
                            }
                            __syncthreads();

                            // Gaussian of voxel column xx (tile) and spot i (chunk): table or direct
                            auto gx = [&](int xx, int i, float* dg) -> float {
                                if (TAB && QC) {
                                    // P = 2: with xx = 4 k + g and i = 16 nb + c16 the entry keeps
                                    // the parity of its k = nb = 0 value; slot = base + 8 Q nb - 4 k
                                    const int sl = sgx + (i - c16) / 2 * QC - (xx - g);
                                    if (SIG) *dg = Tx2[sl];
                                    return Tx[sl];
                                }
                                if (TAB) {
                                    const int t = A.tx.Q * i - A.tx.P * xx + A.tx.P * (AJ_TX - 1);
                                    if (SIG) *dg = tab_get(Tx2, A.tx, t);
                                    return tab_get(Tx, A.tx, t);
                                }
                                const double d = (a.ox + (double)(x0 + xx) * a.dx) - (a.sx0 + (double)(i0 + i) * a.dsx);
                                const float d2 = (float)(d * d);
                                const float gv = i < ni ? gauss2(d2, k2) : 0.0f;
                                if (SIG) *dg = gv * d2 * is3;
                                return gv;
                            };
                            auto gy = [&](int yy, int j, float* dg) -> float {
                                if (TAB && QC) {
                                    const int sl = sgy + (j - c16) / 2 * QC - (yy - g);
                                    if (SIG) *dg = Ty2[sl];
                                    return Ty[sl];
                                }
                                if (TAB) {
                                    const int t = A.ty.Q * j - A.ty.P * yy + A.ty.P * (AJ_TY - 1);
                                    if (SIG) *dg = tab_get(Ty2, A.ty, t);
                                    return tab_get(Ty, A.ty, t);
                                }
                                const double d = (a.oy + (double)(y0 + yy) * a.dy) - (a.sy0 + (double)(j0 + j) * a.dsy);
                                const float d2 = (float)(d * d);
                                const float gv = j < nj ? gauss2(d2, k2) : 0.0f;
                                if (SIG) *dg = gv * d2 * is3;
                                return gv;
                            };

                            // Windows: K steps (4 voxels each) within FT_WIN_SIGMAS s of a spot
                            // block; a factor beyond is below 2^-32 of its peak. Returns the steps
                            // [lo, hi) of voxels 0 .. T-1 in reach of block spots b0 .. b0 + 15 of
                            // n, for voxel v at C0 + v dv and spot k at k ds (uniform).
                            const float cut = (float)(FT_WIN_SIGMAS * sd);
                            auto steps = [&](int b0, int n, double C0, double dv, double ds, int T, int& lo) {
                                lo = 0;
                                if (b0 >= n) return 0;
                                const float s0 = (float)(((double)b0 * ds - C0) / dv);
                                const float s1 = (float)(((double)min(b0 + 15, n - 1) * ds - C0) / dv);
                                int first, last;
                                spot_range(s0, s1, cut * (float)(1.0 / dv), T, first, last);
                                if (first > last) return 0;
                                lo = __builtin_amdgcn_readfirstlane(first >> 2);
                                return __builtin_amdgcn_readfirstlane((last >> 2) + 1);
                            };
                            // stage 1: Q[y][i] = sum_x R[y][x] Gx[x][i]; wave w owns rows 16w..16w+15
                            {
                                v4f q1[3] = {}, q2[3] = {};
#pragma unroll
                                for (int nb = 0; nb < 3; ++nb) {
                                    int klo;
                                    const int khi = steps(16 * nb, ni, Cx, a.dx, a.dsx, AJ_TX, klo);
#pragma unroll 4
                                    for (int kk = klo; kk < khi; ++kk) {
                                        const int xx = 4 * kk + g;
                                        const float av = Rs[(16 * wave + c16) * AJ_RS + xx];
                                        float dg = 0.0f;
                                        const float bv = gx(xx, 16 * nb + c16, &dg);
                                        q1[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, q1[nb], 0, 0, 0);
                                        if (SIG) q2[nb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, dg, q2[nb], 0, 0, 0);
                                    }
                                }
#pragma unroll
                                for (int nb = 0; nb < 3; ++nb)
#pragma unroll
                                    for (int r = 0; r < 4; ++r) {
                                        const int off = (16 * wave + 4 * g + r) * AJ_QS + 16 * nb + c16;
                                        Qs[off] = q1[nb][r];
                                        if (SIG) Q2s[off] = q2[nb][r];
                                    }
                            }
                            __syncthreads();

                            // stage 2: P[j][i] = sum_y Gy[y][j] Q[y][i] over this tile (A = Gy^T, B = Q);
                            // the tile's c P joins the per-slice partial (first tile stores)
                            const float* wl = a.w + (size_t)l * a.nsy * a.nsx;
                            float* pz = A.part + ((size_t)z * a.L + l) * a.nsy * a.nsx;
                            double wp = 0.0, wp2 = 0.0;
#pragma unroll
                            for (int qq = 0; qq < 3; ++qq) {
                                const int blk = wave + 4 * qq;
                                if (blk < 9) {
                                    const int mb = blk / 3, nb = blk % 3;
                                    v4f p = {}, p2 = {};
                                    int klo;
                                    const int khi = steps(16 * mb, nj, Cy, a.dy, a.dsy, AJ_TY, klo);
#pragma unroll 4
                                    for (int kk = klo; kk < khi; ++kk) {
                                        const int yy = 4 * kk + g;
                                        float dgy = 0.0f;
                                        const float av = gy(yy, 16 * mb + c16, &dgy);
                                        const float bv = Qs[yy * AJ_QS + 16 * nb + c16];
                                        p = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, p, 0, 0, 0);
                                        if (SIG) {
                                            p2 = __builtin_amdgcn_mfma_f32_16x16x4f32(dgy, bv, p2, 0, 0, 0);
                                            p2 = __builtin_amdgcn_mfma_f32_16x16x4f32(av, Q2s[yy * AJ_QS + 16 * nb + c16], p2, 0, 0, 0);
                                        }
                                    }
#pragma unroll
                                    for (int r = 0; r < 4; ++r) {
                                        const int j = 16 * mb + 4 * g + r, i = 16 * nb + c16;
                                        if (j < nj && i < ni) {
                                            const size_t off = (size_t)(j0 + j) * a.nsx + i0 + i;
                                            const float cp = (float)(cn * (double)p[r]);
                                            pz[off] = first_tile ? cp : pz[off] + cp;
                                            if (lut_grads) {
                                                const double wv = (double)wl[off];
                                                wp += wv * (double)p[r];
                                                if (SIG) wp2 += wv * (double)p2[r];
                                            }
                                        }
                                    }
                                }
                            }
                            if (lut_grads) {   // fixed-order workgroup sum into the pass accumulators
                                for (int o = 32; o > 0; o >>= 1) {
                                    wp += __shfl_xor(wp, o, 64);
                                    wp2 += __shfl_xor(wp2, o, 64);
                                }
                                if (lane == 0) {
                                    red[0][wave] = wp;
                                    red[1][wave] = wp2;
                                }
                            }
                            __syncthreads();
                            if (lut_grads && threadIdx.x == 0) {
                                double t0 = 0.0, t1 = 0.0;
                                for (int w = 0; w < AJ_THREADS / 64; ++w) {
                                    t0 += red[0][w];
                                    t1 += red[1][w];
                                }
                                wacc[0][q] += t0;
                                wacc[1][q] += t1;
                            }
                        }
                    }
                }
            }
        }
        __syncthreads();
        if (lut_grads && threadIdx.x < nl) {
            const int q = threadIdx.x, l = lb + q * A.split;
            const double iddv = a.idd[(size_t)l * a.nz + z];
            const double sg = a.sigma[(size_t)l * a.nz + z];
            const double sd = sg > kSigmaFloor ? sg : kSigmaFloor;
            const double norm = 1.0 / (2.0 * kPi * sd * sd);
            if (A.grad_idd) A.grad_idd[(size_t)l * a.nz + z] = wacc[0][q] * norm;
            if (SIG && A.grad_sigma)
                A.grad_sigma[(size_t)l * a.nz + z] =
                    sg > kSigmaFloor ? iddv * norm * (-(2.0 / sd) * wacc[0][q] + wacc[1][q]) : 0.0;
        }
        __syncthreads();
    }
}

template <bool TAB, bool SIG, int QC>
__global__ __launch_bounds__(AJ_THREADS, 2)
void adj_fp32_kernel(AdjArgs A)
{
    int z, sh;
    xcd_unit((int)blockIdx.x, A.split, z, sh);
    if (z >= A.f.nz) return;
    adj_fp32_body<TAB, SIG, QC>(A, z * A.split + sh);
}

// One launch over a batch of problems sharing the template parameters.
template <bool TAB, bool SIG, int QC>
__global__ __launch_bounds__(AJ_THREADS, 2)
void adj_fp32_batch_kernel(const AdjArgs* __restrict__ items, const int* __restrict__ wg0, int n)
{
    const int k = batch_item(wg0, n, blockIdx.x);
    const AdjArgs A = items[k];
    int z, sh;     // item ranges start at multiples of 8: launch index b and b - wg0[k] share an XCD
    xcd_unit((int)blockIdx.x - wg0[k], A.split, z, sh);
    if (z >= A.f.nz) return;
    adj_fp32_body<TAB, SIG, QC>(A, z * A.split + sh);
}

// grad_w[e] = sum_z part[z][e] in FP64, increasing z (8 loads in flight per thread)
constexpr int OBJ_THREADS = 256;

// Workgroups of the depth reduction for one problem: whole 64-element blocks per layer.
__host__ __device__ inline int reduce_blocks(size_t n, int per_layer)
{
    return (int)(n / (size_t)per_layer) * ((per_layer + RED_ELEMS - 1) / RED_ELEMS);
}

// out[0] = sum of n partials: each thread sums a fixed stride, then a fixed tree
__device__ __forceinline__ void sum_partials_body(const double* partial, int n, double* out)
{
    __shared__ double red[OBJ_THREADS];
    double t = 0.0;
    for (int i = threadIdx.x; i < n; i += OBJ_THREADS) t += partial[i];
    red[threadIdx.x] = t;
    __syncthreads();
    for (int w = OBJ_THREADS / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

// grad_w[e] = sum_z part[z][e] in FP64 over the slices where element e's layer
// has IDD != 0: elsewhere its c P partial is exactly zero, so the adjoint kernels
// neither write nor read it. Workgroup blk takes 64 elements of one layer l
// (blocks of a layer are consecutive). Per pass of 256 slices the workgroup
// lists the active ones in increasing z (ballot compaction), and wave w adds the
// partials of list entries w, w + 4, ... (coalesced rows, four loads in flight);
// the four wave sums are then added in wave order, so the result does not depend
// on scheduling.
__device__ __forceinline__ void adj_reduce_body(const BatchRedItem& it, int blk)
{
    if (it.j_out && blk == reduce_blocks(it.n, it.per_layer)) {     // the objective's J (uniform)
        sum_partials_body(it.jpart, it.njpart, it.j_out);
        return;
    }
    __shared__ double ws[4][RED_ELEMS];
    __shared__ int zl[256];
    __shared__ int wcnt[4];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int bpl = (it.per_layer + RED_ELEMS - 1) / RED_ELEMS;
    const int l = blk / bpl;
    const int el = (blk - l * bpl) * RED_ELEMS + lane;
    const bool own = el < it.per_layer;
    const size_t e = (size_t)l * it.per_layer + el;
    const double* idd = it.idd + (size_t)l * it.nz;
    double acc = 0.0;
    for (int z0 = 0; z0 < it.nz; z0 += 256) {
        const int z = z0 + (int)threadIdx.x;
        const bool act = z < it.nz && idd[z] != 0.0;
        const unsigned long long m = __ballot(act);
        if (lane == 0) wcnt[wave] = __popcll(m);
        __syncthreads();
        int off = 0, cnt = 0;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            off += w < wave ? wcnt[w] : 0;
            cnt += wcnt[w];
        }
        if (act) zl[off + __popcll(m & ((1ull << lane) - 1ull))] = z;
        __syncthreads();
        if (own) {
            int q = wave;
            for (; q + 12 < cnt; q += 16) {
                float v[4];
#pragma unroll
                for (int k = 0; k < 4; ++k) v[k] = it.part[(size_t)zl[q + 4 * k] * it.n + e];
#pragma unroll
                for (int k = 0; k < 4; ++k) acc += (double)v[k];
            }
            for (; q < cnt; q += 4) acc += (double)it.part[(size_t)zl[q] * it.n + e];
        }
        __syncthreads();
    }
    ws[wave][lane] = acc;
    __syncthreads();
    if (wave == 0 && own)
        it.grad_w[e] = (float)(((ws[0][lane] + ws[1][lane]) + ws[2][lane]) + ws[3][lane]);
}

__global__ __launch_bounds__(256) void adj_reduce_kernel(BatchRedItem it)
{
    adj_reduce_body(it, blockIdx.x);
}

__global__ __launch_bounds__(256) void adj_reduce_batch_kernel(const BatchRedItem* __restrict__ items,
                                                               const int* __restrict__ wg0, int n)
{
    const int k = batch_item(wg0, n, blockIdx.x);
    adj_reduce_body(items[k], (int)blockIdx.x - wg0[k]);
}

// ---- weighted quadratic objective: R = 2 wv (D - Dp), J = sum wv (D - Dp)^2 ----

// Workgroup blk of nblk: grid-stride loop over the voxels, partial[blk] = its sum.
__device__ __forceinline__ void objective_body(const float* D, const float* Dp, const float* wv, size_t n,
                                               float* R, double* partial, int blk, int nblk)
{
    __shared__ double red[OBJ_THREADS / 64];
    double acc = 0.0;
    auto one = [&](float d, float p, float wf, float* r_out) {
        const float r = objective_term(d, p, wf, acc);
        if (r_out) *r_out = r;
    };
    const bool vec = (((uintptr_t)D | (uintptr_t)Dp | (uintptr_t)wv | (uintptr_t)R) & 15) == 0;
    size_t done = 0;
    if (vec) {
        // four voxels per thread and step (16-byte loads and stores)
        const size_t n4 = n / 4;
        for (size_t e = (size_t)blk * OBJ_THREADS + threadIdx.x; e < n4; e += (size_t)nblk * OBJ_THREADS) {
            const float4 d = ((const float4*)D)[e], p = ((const float4*)Dp)[e];
            const float4 w = wv ? ((const float4*)wv)[e] : make_float4(1.0f, 1.0f, 1.0f, 1.0f);
            float4 r;
            one(d.x, p.x, w.x, &r.x);
            one(d.y, p.y, w.y, &r.y);
            one(d.z, p.z, w.z, &r.z);
            one(d.w, p.w, w.w, &r.w);
            if (R) ((float4*)R)[e] = r;
        }
        done = 4 * n4;
    }
    for (size_t e = done + (size_t)blk * OBJ_THREADS + threadIdx.x; e < n; e += (size_t)nblk * OBJ_THREADS)
        one(D[e], Dp[e], wv ? wv[e] : 1.0f, R ? R + e : nullptr);
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < OBJ_THREADS / 64; ++w) t += red[w];
        partial[blk] = t;
    }
}

__global__ __launch_bounds__(OBJ_THREADS) void objective_kernel(const float* D, const float* Dp, const float* wv,
                                                                size_t n, float* R, double* partial)
{
    objective_body(D, Dp, wv, n, R, partial, blockIdx.x, gridDim.x);
}

__global__ __launch_bounds__(OBJ_THREADS) void sum_partials_kernel(const double* partial, int n, double* out)
{
    sum_partials_body(partial, n, out);
}

// Batched objective: OBJ_BATCH_BLOCKS workgroups per item.
__global__ __launch_bounds__(OBJ_THREADS) void objective_batch_kernel(const BatchObjItem* __restrict__ items)
{
    const int k = blockIdx.x / OBJ_BATCH_BLOCKS;
    const BatchObjItem it = items[k];
    objective_body(it.D, it.Dp, it.wv, it.n, it.R, it.part, blockIdx.x % OBJ_BATCH_BLOCKS, OBJ_BATCH_BLOCKS);
}

// J of every item (one workgroup per item)
__global__ __launch_bounds__(OBJ_THREADS) void sum_partials_batch_kernel(const BatchObjItem* __restrict__ items)
{
    const BatchObjItem it = items[blockIdx.x];
    sum_partials_body(it.part, OBJ_BATCH_BLOCKS, it.j_out);
}

// out[l] = sum_z (g_idd[l,z] didd[l,z] + g_sig[l,z] dsig[l,z]) (one wave per layer, fixed order)
__global__ void lut_chain_kernel(const double* g_idd, const double* g_sig, const double* didd,
                                 const double* dsig, int L, int nz, double* out)
{
    const int l = blockIdx.x;
    const int lane = threadIdx.x;
    double acc = 0.0;
    for (int k = lane; k < nz; k += 64) {
        const size_t e = (size_t)l * nz + k;
        acc += g_idd[e] * didd[e] + g_sig[e] * dsig[e];
    }
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
    if (lane == 0) out[l] = acc;
}

}  // namespace pbh

using namespace pbh;

extern "C" int pbh_validate(const PbhGrid* g, const PbhLattice* lat);

// Layer split of the adjoint from the active-layer count of every depth slice:
// simulates the dispatch of the workgroups (slice-major, split s of slice z is
// block z split + s) onto `slots` resident slots, each taking the next workgroup
// when it frees. Workgroup s of slice z takes the active layers whose index is s
// modulo split; it costs its layers plus AJ_WG_COST layers of fixed work (R tiles,
// windows), or AJ_WG_EMPTY when it has none. Returns the split (1 to AJ_MAX_SPLIT,
// at least min_split) with the shortest makespan, the smaller on ties.
// On the 128^3 / 41 x 41 x 40-layer field it picks 7 (BF16x3, 3 slots per CU),
// 10 (BF16, 4 slots) and 5 (FP32, 2 slots); a measured sweep of splits 2-16 had
// the same bests, 7 %, 5 % and 12 % ahead of a split of 8.
constexpr double AJ_WG_COST = 2.5, AJ_WG_EMPTY = 0.3;

static bool adj_axis(double dv, double ds, int T, int K, AxisTable* out);

static int model_split(int nz, int L, const int32_t* active, long long slots, int min_split)
{
    if (slots < 1) slots = 1;
    if (slots > (1 << 20)) slots = 1 << 20;
    int best = std::max(1, min_split);
    double best_t = -1.0;
    std::vector<double> heap;
    for (int sp = std::max(1, min_split); sp <= AJ_MAX_SPLIT; ++sp) {
        if ((long long)nz * sp > (1LL << 24)) break;                  // keep the model cheap
        heap.assign((size_t)slots, 0.0);                                // min-heap of slot finish times
        for (int z = 0; z < nz; ++z) {
            const int act = active[z] < 0 ? 0 : (active[z] > L ? L : active[z]);
            for (int s = 0; s < sp; ++s) {
                const int n = act > s ? (act - s + sp - 1) / sp : 0;
                const double c = n > 0 ? n + AJ_WG_COST : AJ_WG_EMPTY;
                std::pop_heap(heap.begin(), heap.end(), std::greater<double>());
                heap.back() += c;
                std::push_heap(heap.begin(), heap.end(), std::greater<double>());
            }
        }
        const double t = *std::max_element(heap.begin(), heap.end());
        if (best_t < 0.0 || t < best_t) {
            best_t = t;
            best = sp;
        }
    }
    return best;
}

// The adjoints with 16-bit operands (adj_mixed_kernel): BF16 or FP16, single or split
// products.
static bool adj_16bit(int precision)
{
    return precision == PBH_PREC_BF16 || precision == PBH_PREC_BF16X3 || precision == PBH_PREC_FP16 ||
           precision == PBH_PREC_FP16X3;
}

extern "C" int pbh_adjoint_choose_split(const PbhGrid* g, const PbhLattice* lat, int precision,
                                        const int32_t* active_per_z, int compute_units)
{
    int rc = pbh_validate(g, lat);
    if (rc != PBH_OK) return rc;
    if (!active_per_z) return PBH_ERR_NULL_ARG;
    const bool mixed = adj_16bit(precision);
    if (precision != PBH_PREC_FP32 && !mixed) return PBH_ERR_UNSUPPORTED;
    // resident workgroups per CU: the FP32 kernel 2; the 16-bit kernels on the P = 2, Q = 5
    // table geometry 3 (split products: 156 VGPRs BF16, 168 FP16), 4 (BF16 single products,
    // 124 VGPRs) or 2 (FP16 single products, 185 VGPRs), else 2
    int per_cu = 2;
    AxisTable tx, ty;
    if (mixed && adj_axis(g->dx, lat->dx_spot, AJ_TX, AJ_SC, &tx) && adj_axis(g->dy, lat->dy_spot, AJ_TY, AJ_SC, &ty) &&
        tx.P == 2 && tx.Q == 5 && ty.P == 2 && ty.Q == 5)
        per_cu = precision == PBH_PREC_BF16 ? 4 : precision == PBH_PREC_FP16 ? 2 : 3;
    // the mixed kernel keeps at most AM_MAXK (128) layers of a slice per workgroup
    const int min_split = mixed ? (lat->n_layers + 127) / 128 : 1;
    if (min_split > AJ_MAX_SPLIT) return PBH_ERR_UNSUPPORTED;
    const int cus = compute_units > 0 ? compute_units : pbh_device_cus();
    return model_split(g->nz, lat->n_layers, active_per_z, (long long)per_cu * cus, min_split);
}

int pbh_adj_default_split(const PbhGrid* g, const PbhLattice* lat)
{
    int s = 1;
    while (s < 16 && (long long)g->nz * s < 1024 && s * 2 <= lat->n_layers) s *= 2;
    return s;
}

extern "C" size_t pbh_adjoint_workspace_bytes(const PbhGrid* g, const PbhLattice* lat, int precision)
{
    (void)precision;
    if (pbh_validate(g, lat) != PBH_OK) return 0;
    return (size_t)g->nz * lat->n_layers * lat->n_sy * lat->n_sx * sizeof(float);
}

static bool adj_axis(double dv, double ds, int T, int K, AxisTable* out)
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

// Adjoint arguments of one problem; split 0 = default (split in 1, 2, 4, 8, 16, at most L).
int pbh_make_adj_args(const PbhGrid* g, const PbhLattice* lat, const float* w, const double* idd,
                      const double* sigma, const float* R, float* part, double* grad_idd,
                      double* grad_sigma, int split, AdjArgs* out)
{
    if (split == 0) split = pbh_adj_default_split(g, lat);
    if (split < 1 || split > AJ_MAX_SPLIT) return PBH_ERR_UNSUPPORTED;
    if ((long long)kXcds * split * cdiv(g->nz, kXcds) > 0x7fffffffLL) return PBH_ERR_UNSUPPORTED;
    AdjArgs& A = *out;
    A.f = make_fwd_args(g, lat, w, idd, sigma, nullptr);
    A.R = R;
    A.part = part;
    A.grad_idd = grad_idd;
    A.grad_sigma = grad_sigma;
    A.split = split;
    A.tab = adj_axis(g->dx, lat->dx_spot, AJ_TX, AJ_SC, &A.tx) && adj_axis(g->dy, lat->dy_spot, AJ_TY, AJ_SC, &A.ty);
    if (!A.tab) {
        A.tx = AxisTable{0, 0, 0.0, 0};
        A.ty = AxisTable{0, 0, 0.0, 0};
    }
    return PBH_OK;
}

// Adjoint arguments of one problem with every check that precedes a launch: the
// split, and for the BF16 kernels the table geometry and the layers per workgroup.
static int prepare_adjoint(const PbhGrid* g, const PbhLattice* lat, const float* w, const double* idd,
                           const double* sigma, const float* R, float* part, double* grad_idd,
                           double* grad_sigma, int precision, int split, AdjArgs* A)
{
    int rc = pbh_make_adj_args(g, lat, w, idd, sigma, R, part, grad_idd, grad_sigma, split, A);
    if (rc != PBH_OK) return rc;
    const bool mixed = adj_16bit(precision);
    // the BF16 adjoint gathers every factor from tables and keeps at most 128 layers per workgroup
    if (mixed && (!A->tab || !pbh_adj_mixed_fits(*A))) return PBH_ERR_UNSUPPORTED;
    return PBH_OK;
}

// The adjoint kernels and the depth reduction of one problem (arguments from
// prepare_adjoint); with j_out, the reduction's launch also sums the njpart
// objective partials.
static int adjoint_launch(const PbhGrid* g, const PbhLattice* lat, const AdjArgs& A, const double* idd,
                          float* grad_w, double* grad_sigma, int precision, const double* jpart, int njpart,
                          double* j_out, hipStream_t s)
{
    const bool mixed = adj_16bit(precision);
    int rc;
    const size_t n = (size_t)lat->n_layers * lat->n_sy * lat->n_sx;
    if (mixed) {
        if ((rc = pbh_launch_adj_mixed(A, precision, g->nz, s)) != PBH_OK) return rc;
    } else {
        const dim3 grid((unsigned)xcd_grid(g->nz, A.split)), block(AJ_THREADS);
        const bool sig = grad_sigma != nullptr;
        if (pbh_adj_mixed_q5(A)) {
            if (sig) hipLaunchKernelGGL((adj_fp32_kernel<true, true, 5>), grid, block, 0, s, A);
            else hipLaunchKernelGGL((adj_fp32_kernel<true, false, 5>), grid, block, 0, s, A);
        } else if (A.tab) {
            if (sig) hipLaunchKernelGGL((adj_fp32_kernel<true, true, 0>), grid, block, 0, s, A);
            else hipLaunchKernelGGL((adj_fp32_kernel<true, false, 0>), grid, block, 0, s, A);
        } else {
            if (sig) hipLaunchKernelGGL((adj_fp32_kernel<false, true, 0>), grid, block, 0, s, A);
            else hipLaunchKernelGGL((adj_fp32_kernel<false, false, 0>), grid, block, 0, s, A);
        }
        if ((rc = launch_rc("adj_fp32_kernel")) != PBH_OK) return rc;
    }
    const int per_layer = lat->n_sy * lat->n_sx;
    const int blocks = reduce_blocks(n, per_layer) + (j_out ? 1 : 0);
    hipLaunchKernelGGL(adj_reduce_kernel, dim3((unsigned)blocks), dim3(256), 0, s,
                       BatchRedItem{A.part, idd, grad_w, n, g->nz, per_layer, jpart, njpart, j_out});
    return launch_rc("adj_reduce_kernel");
}

extern "C" int pbh_adjoint_tuned(const PbhGrid* g, const PbhLattice* lat, const float* w, const double* idd,
                                 const double* sigma, const float* R, float* grad_w, double* grad_idd,
                                 double* grad_sigma, int precision, int split, void* workspace,
                                 size_t workspace_bytes, void* stream)
{
    int rc = pbh_validate(g, lat);
    if (rc != PBH_OK) return rc;
    if (!w || !idd || !sigma || !R || !grad_w || !workspace) return PBH_ERR_NULL_ARG;
    const bool mixed = adj_16bit(precision);
    if (precision != PBH_PREC_FP32 && !mixed) return PBH_ERR_UNSUPPORTED;
    if (mixed && (grad_idd || grad_sigma)) return PBH_ERR_UNSUPPORTED;   // table gradients: FP32 only
    if (split < 0) return PBH_ERR_BAD_DIM;
    if (workspace_bytes < pbh_adjoint_workspace_bytes(g, lat, precision)) return PBH_ERR_WORKSPACE;
    AdjArgs A;
    if ((rc = prepare_adjoint(g, lat, w, idd, sigma, R, (float*)workspace, grad_idd, grad_sigma, precision, split,
                              &A)) != PBH_OK)
        return rc;
    return adjoint_launch(g, lat, A, idd, grad_w, grad_sigma, precision, nullptr, 0, nullptr, (hipStream_t)stream);
}

extern "C" int pbh_adjoint(const PbhGrid* g, const PbhLattice* lat, const float* w, const double* idd,
                           const double* sigma, const float* R, float* grad_w, double* grad_idd,
                           double* grad_sigma, int precision, void* workspace, size_t workspace_bytes,
                           void* stream)
{
    return pbh_adjoint_tuned(g, lat, w, idd, sigma, R, grad_w, grad_idd, grad_sigma, precision, 0, workspace,
                             workspace_bytes, stream);
}

static constexpr int OBJ_BLOCKS = 1024;

extern "C" size_t pbh_objective_workspace_bytes(const PbhGrid* g)
{
    (void)g;
    return OBJ_BLOCKS * sizeof(double);
}

extern "C" int pbh_objective(const PbhGrid* g, const float* dose, const float* presc, const float* vox_w,
                             float* R, double* j_out, void* workspace, size_t workspace_bytes, void* stream)
{
    if (!g || !dose || !presc || !j_out || !workspace) return PBH_ERR_NULL_ARG;
    if (g->nx <= 0 || g->ny <= 0 || g->nz <= 0) return PBH_ERR_BAD_DIM;
    if (workspace_bytes < pbh_objective_workspace_bytes(g)) return PBH_ERR_WORKSPACE;
    hipStream_t s = (hipStream_t)stream;
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    double* part = (double*)workspace;
    hipLaunchKernelGGL(objective_kernel, dim3(OBJ_BLOCKS), dim3(OBJ_THREADS), 0, s, dose, presc, vox_w, n, R, part);
    int rc = launch_rc("objective_kernel");
    if (rc != PBH_OK) return rc;
    hipLaunchKernelGGL(sum_partials_kernel, dim3(1), dim3(OBJ_THREADS), 0, s, (const double*)part, OBJ_BLOCKS, j_out);
    return launch_rc("sum_partials_kernel");
}

// Workspace of pbh_objective_adjoint: the adjoint's per-slice partials, then R
// (nz ny nx FP32), then the OBJ_BLOCKS partial sums of J.
static size_t align_up(size_t x) { return (x + 255) & ~(size_t)255; }

extern "C" size_t pbh_objective_adjoint_workspace_bytes(const PbhGrid* g, const PbhLattice* lat, int precision)
{
    const size_t part = pbh_adjoint_workspace_bytes(g, lat, precision);
    if (part == 0) return 0;
    return align_up(part) + align_up((size_t)g->nx * g->ny * g->nz * sizeof(float)) + OBJ_BLOCKS * sizeof(double);
}

extern "C" int pbh_objective_adjoint(const PbhGrid* g, const PbhLattice* lat, const float* w, const double* idd,
                                     const double* sigma, const float* dose, const float* presc,
                                     const float* vox_w, float* grad_w, double* j_out, int precision, int split,
                                     void* workspace, size_t workspace_bytes, void* stream)
{
    int rc = pbh_validate(g, lat);
    if (rc != PBH_OK) return rc;
    if (!w || !idd || !sigma || !dose || !presc || !grad_w || !j_out || !workspace) return PBH_ERR_NULL_ARG;
    const bool mixed = adj_16bit(precision);
    if (precision != PBH_PREC_FP32 && !mixed) return PBH_ERR_UNSUPPORTED;
    if (split < 0) return PBH_ERR_BAD_DIM;
    if (workspace_bytes < pbh_objective_adjoint_workspace_bytes(g, lat, precision)) return PBH_ERR_WORKSPACE;
    hipStream_t s = (hipStream_t)stream;
    char* ws = (char*)workspace;
    float* part = (float*)ws;
    float* R = (float*)(ws + align_up(pbh_adjoint_workspace_bytes(g, lat, precision)));
    const size_t nv = (size_t)g->nx * g->ny * g->nz;
    double* jpart = (double*)((char*)R + align_up(nv * sizeof(float)));
    AdjArgs A;
    if ((rc = prepare_adjoint(g, lat, w, idd, sigma, R, part, nullptr, nullptr, precision, split, &A)) != PBH_OK)
        return rc;
    hipLaunchKernelGGL(objective_kernel, dim3(OBJ_BLOCKS), dim3(OBJ_THREADS), 0, s, dose, presc, vox_w, nv, R, jpart);
    if ((rc = launch_rc("objective_kernel")) != PBH_OK) return rc;
    return adjoint_launch(g, lat, A, idd, grad_w, nullptr, precision, jpart, OBJ_BLOCKS, j_out, s);
}

// Workspace of pbh_forward_objective_adjoint: the forward's, then the objective +
// adjoint's, then the J partial sums of the fused combine (FT_OBJ_PARTS per dose tile).
extern "C" size_t pbh_forward_objective_adjoint_workspace_bytes(const PbhGrid* g, const PbhLattice* lat,
                                                               int fwd_precision, int layer_chunk, int tile_rows,
                                                               int adj_precision)
{
    const size_t f = pbh_forward_workspace_bytes_shaped(g, lat, fwd_precision, layer_chunk, tile_rows);
    const size_t o = pbh_objective_adjoint_workspace_bytes(g, lat, adj_precision);
    if (f == 0 || o == 0) return 0;
    const long long rows = tile_rows ? tile_rows : FT_TY;
    const long long tiles = ((g->nx + FT_TX - 1) / FT_TX) * ((g->ny + rows - 1) / rows) * (long long)g->nz;
    return align_up(f) + align_up(o) + (size_t)tiles * FT_OBJ_PARTS * sizeof(double);
}

extern "C" int pbh_forward_objective_adjoint(const PbhGrid* g, const PbhLattice* lat, const float* w,
                                             const double* idd, const double* sigma, float* dose,
                                             const float* presc, const float* vox_w, float* grad_w,
                                             double* j_out, int fwd_precision, int layer_chunk, int tile_rows,
                                             int adj_precision, int split, void* workspace,
                                             size_t workspace_bytes, void* stream)
{
    int rc = pbh_validate(g, lat);
    if (rc != PBH_OK) return rc;
    if (!w || !idd || !sigma || !dose || !presc || !grad_w || !j_out || !workspace) return PBH_ERR_NULL_ARG;
    const bool mixed = adj_16bit(adj_precision);
    if (adj_precision != PBH_PREC_FP32 && !mixed) return PBH_ERR_UNSUPPORTED;
    if (split < 0 || layer_chunk < 0) return PBH_ERR_BAD_DIM;
    const size_t need = pbh_forward_objective_adjoint_workspace_bytes(g, lat, fwd_precision, layer_chunk,
                                                                      tile_rows, adj_precision);
    if (need == 0) return PBH_ERR_UNSUPPORTED;
    if (workspace_bytes < need) return PBH_ERR_WORKSPACE;
    hipStream_t s = (hipStream_t)stream;
    char* ws = (char*)workspace;
    const size_t fbytes = pbh_forward_workspace_bytes_shaped(g, lat, fwd_precision, layer_chunk, tile_rows);
    char* ows = ws + align_up(fbytes);
    float* part = (float*)ows;
    float* R = (float*)(ows + align_up(pbh_adjoint_workspace_bytes(g, lat, adj_precision)));
    double* jpart = (double*)(ows + align_up(pbh_objective_adjoint_workspace_bytes(g, lat, adj_precision)));
    AdjArgs A;
    if ((rc = prepare_adjoint(g, lat, w, idd, sigma, R, part, nullptr, nullptr, adj_precision, split, &A)) != PBH_OK)
        return rc;
    const ObjFuse fuse{presc, vox_w, R, jpart};
    int njpart = 0;
    if ((rc = pbh_forward_fused(g, lat, w, idd, sigma, dose, fwd_precision, layer_chunk, tile_rows, ws, fbytes, s,
                                &fuse, &njpart)) != PBH_OK)
        return rc;
    return adjoint_launch(g, lat, A, idd, grad_w, nullptr, adj_precision, jpart, njpart, j_out, s);
}

extern "C" int pbh_lut_chain(const double* grad_idd, const double* grad_sigma, const double* didd,
                             const double* dsigma, int n_layers, int nz, double* out, void* stream)
{
    if (!grad_idd || !grad_sigma || !didd || !dsigma || !out) return PBH_ERR_NULL_ARG;
    if (n_layers <= 0 || nz <= 0) return PBH_ERR_BAD_DIM;
    hipLaunchKernelGGL(lut_chain_kernel, dim3(n_layers), dim3(64), 0, (hipStream_t)stream, grad_idd, grad_sigma,
                       didd, dsigma, n_layers, nz, out);
    return launch_rc("lut_chain_kernel");
}

// ---- batched launches (robust-scenario batches) ----
int pbh_launch_adj_batch(const AdjArgs* items, const int* wg0, int n, int total, bool tab, bool sig, bool q5,
                         hipStream_t s)
{
    const dim3 grid((unsigned)total), block(AJ_THREADS);
    if (tab && q5) {
        if (sig) hipLaunchKernelGGL((adj_fp32_batch_kernel<true, true, 5>), grid, block, 0, s, items, wg0, n);
        else hipLaunchKernelGGL((adj_fp32_batch_kernel<true, false, 5>), grid, block, 0, s, items, wg0, n);
    } else if (tab) {
        if (sig) hipLaunchKernelGGL((adj_fp32_batch_kernel<true, true, 0>), grid, block, 0, s, items, wg0, n);
        else hipLaunchKernelGGL((adj_fp32_batch_kernel<true, false, 0>), grid, block, 0, s, items, wg0, n);
    } else {
        if (sig) hipLaunchKernelGGL((adj_fp32_batch_kernel<false, true, 0>), grid, block, 0, s, items, wg0, n);
        else hipLaunchKernelGGL((adj_fp32_batch_kernel<false, false, 0>), grid, block, 0, s, items, wg0, n);
    }
    return launch_rc("adj_fp32_batch_kernel");
}

int pbh_launch_obj_batch(const BatchObjItem* items, int n, hipStream_t s)
{
    hipLaunchKernelGGL(objective_batch_kernel, dim3((unsigned)(n * OBJ_BATCH_BLOCKS)), dim3(OBJ_THREADS), 0, s, items);
    int rc = launch_rc("objective_batch_kernel");
    if (rc != PBH_OK) return rc;
    hipLaunchKernelGGL(sum_partials_batch_kernel, dim3((unsigned)n), dim3(OBJ_THREADS), 0, s, items);
    return launch_rc("sum_partials_batch_kernel");
}

int pbh_launch_reduce_batch(const BatchRedItem* items, const int* wg0, int n, int total, hipStream_t s)
{
    hipLaunchKernelGGL(adj_reduce_batch_kernel, dim3((unsigned)total), dim3(256), 0, s, items, wg0, n);
    return launch_rc("adj_reduce_batch_kernel");
}
