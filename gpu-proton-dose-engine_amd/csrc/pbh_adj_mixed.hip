// pbh_adj_mixed.hip -- spot-weight adjoint in BF16 on gfx950 MFMA (plain or
// split "x3" products), FP32 accumulation.
//
// For a dose cotangent R and each (layer l, depth z):
//   Q   = R_z Gx_lz          (64 x 48 per tile: rows y, spots i; K = 128 voxels x)
//   P  += Gy_lz^T Q          (48 x 48: spots j, i; K = 64 voxels y)
//   dJ/dW_l = sum_z c_lz P_lz,  c_lz = IDD_l(z) / (2 pi s^2)
// Workgroup (z, c) of the `split` workgroups of depth slice z takes the active
// layers (IDD != 0) of the slice whose active index is c modulo split (the
// partials of inactive layers are zero and never written: the depth reduction
// skips them): the workgroups of a slice share its active layers evenly (at most
// ceil(active / split) each), and those left without one skip the slice's R
// tiles entirely. The slice is processed in groups of two 64 x 128 tiles
// (y): the wave's 16 rows of R of both are loaded once into registers and split
// into BF16 (hi, lo). Per layer (x table once per group, y table per tile):
//   stage 1  Q = R Gx      v_mfma_f32_16x16x32_bf16, A = R (registers), B = Gx
//                          gathered from the packed (hi, lo) x table, once per
//                          (spot block, K step) for both tiles of the group
//   Q^T is stored split in LDS, four consecutive rows per 8-byte write
//   stage 2  P = Gy^T Q    v_mfma_f32_16x16x32_bf16, A = Gy^T from the y table,
//                          B = Q^T rows (16-byte reads)
// with P summed over the group's tiles in registers, and c P joins the
// per-slice FP32 partial (written by the first group, added by later ones) that
// adj_reduce_kernel sums over z
// in FP64 (deterministic). Split products: hi*hi + hi*lo + lo*hi of the
// (hi, lo) parts of both operands; plain: hi*hi. Layers with IDD = 0 at the
// slice contribute nothing and are skipped. Gradients with respect to the
// look-up tables (dJ/dIDD, dJ/dsigma) use the FP32 adjoint.
#include "pbh_batch.h"
#include "pbh_mixed_core.h"

namespace pbh {

constexpr int AM_TY = 64, AM_TX = 128, AM_SC = PBH_SPOT_CHUNK;   // 64 x 128 tile, 48-spot chunk
constexpr int AM_QS = AM_TY + 8;                                  // Q^T row stride (halves)
constexpr int AM_MAXK = 128;                                      // active layers per workgroup, at most
constexpr int AM_NT = 2;                                          // 64-row tiles per register group

// Table geometry of an axis for tile width T and chunk width K (as axis_table in
// the forward): entry t = Q i - P a + P (T - 1) for voxel a, spot i.
template <int QC, int T, int K>
struct AmTab {
    static constexpr int N = QC ? table_slots(2 * (T - 1) + QC * (K - 1) + 1) : FT_TAB;
};

// T = __bf16 or _Float16. FP16 has five exponent bits, so each wave scales its rows of
// a tile group's R by the power of two 2^-e_w that puts their largest |R| in [1, 2)
// before the split (then |Q| <= 2 x 128 voxels). Stage 1 brings every wave's Q rows to
// the group's common scale 2^-e (e the largest e_w) before Q's split, and the group's
// c P partial is scaled back by 2^e. All scalings are exact powers of two, so the
// gradient does not depend on the scale of R.
template <typename T, int MODE, int QC>
__device__ __forceinline__ void adj_mixed_body(const AdjArgs& A, int blk)
{
    typedef typename Vec8<T>::type V8;
    constexpr bool F16 = !__is_same(T, __bf16);
    constexpr bool SPLIT = MODE != MODE_PLAIN;
    constexpr int TXN = AmTab<QC, AM_TX, AM_SC>::N, TYN = AmTab<QC, AM_TY, AM_SC>::N;
    const FwdArgs& a = A.f;

    // tables of two layers (built one layer ahead, during stage 2); y tables per tile
    __shared__ unsigned Tx[2][TXN], Ty[2][AM_NT][TYN];
    __shared__ __attribute__((aligned(16))) T Qh[AM_NT][AM_SC * AM_QS];
    __shared__ __attribute__((aligned(16))) T Ql[AM_NT][SPLIT ? AM_SC * AM_QS : 8];

    __shared__ int lay[AM_MAXK];        // this workgroup's active layers, in layer order,
    __shared__ double lk2[AM_MAXK];     //   their log2(e) / (2 s^2)
    __shared__ double lcn[AM_MAXK];     //   and c = IDD / (2 pi s^2), in layer order
    __shared__ float lcut[AM_MAXK];     //   and window half-width FT_WIN_SIGMAS s (cm)
    // K-step windows per layer of the current tile group (double-buffered by group):
    // [0..2] stage 1, x voxels within reach of spot block nb (steps of 32 voxels);
    // [3 + 3 u + mb] stage 2, rows of tile u within reach of spot block mb (steps of 32
    // rows); each lo | hi << 4, steps lo .. hi - 1
    __shared__ unsigned char awin[2][AM_MAXK][9];
    __shared__ int wcnt[FT_THREADS / 64];
    __shared__ int rexp[F16 ? FT_THREADS / 64 : 1];      // FP16: e_w of each wave's R rows

    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int g = lane >> 4, c16 = lane & 15;
    const int s = blk % A.split, z = blk / A.split;

    const int tiles_x = cdiv(a.nx, AM_TX), tiles_y = cdiv(a.ny, AM_TY);
    const int ncx = chunk_count(a.nsx), ncy = chunk_count(a.nsy);
    const float* Rz = A.R + (size_t)z * a.ny * a.nx;
    const AxisTable& gx = A.tx;
    const AxisTable& gy = A.ty;
    // stage 1 B operand (Gx): x = 32 ks + 8 g + e, i = 16 nb + c16:
    //   t = Q i - P x + P (AM_TX - 1) = t0 + 16 Q nb - 32 P ks - P e.
    // The streams start at their smallest entry (ks = 3 or 1, e = 7) so that
    // every offset is non-negative and becomes an LDS immediate.
    constexpr int D1 = 32 * 2 * 3 + 2 * 7, D2 = 32 * 2 * 1 + 2 * 7;   // P = 2 (QC > 0)
    TabStream<QC> s1, s2;
    s1.init(gx, gx.Q * c16 - 8 * gx.P * g + gx.P * (AM_TX - 1) - (QC ? D1 : 0));
    // stage 2 A operand (Gy^T): y = 32 ks + 8 g + e, j = 16 mb + c16
    s2.init(gy, gy.Q * c16 - 8 * gy.P * g + gy.P * (AM_TY - 1) - (QC ? D2 : 0));
    const int ntX = gx.P * (AM_TX - 1) + gx.Q * (AM_SC - 1) + 1;
    const int ntY = gy.P * (AM_TY - 1) + gy.Q * (AM_SC - 1) + 1;

    auto slot1 = [&](int nb, int ks, int e) {
        if (QC > 0) return s1.at(16 * QC * nb + D1 - 64 * ks - 2 * e);     // P = 2
        return table_slot(gx, s1.bE + 16 * gx.Q * nb - 32 * gx.P * ks - gx.P * e);
    };
    auto slot2 = [&](int mb, int ks, int e) {
        if (QC > 0) return s2.at(16 * QC * mb + D2 - 64 * ks - 2 * e);
        return table_slot(gy, s2.bE + 16 * gy.Q * mb - 32 * gy.P * ks - gy.P * e);
    };

    // Inactive layers (IDD = 0 at z) have a zero partial, which the depth reduction skips:
    // nothing is written for them.
    // active layers of the slice (stable order, ballots over passes of 256); keep
    // those with active index = s (mod split), at most ceil(L / split) <= AM_MAXK
    int nmine = 0, base = 0;
    for (int lb = 0; lb < a.L; lb += FT_THREADS) {
        const int l = lb + threadIdx.x;
        const bool act = l < a.L && a.idd[(size_t)l * a.nz + z] != 0.0;
        const unsigned long long m = __ballot(act);
        if (lane == 0) wcnt[wave] = __popcll(m);
        __syncthreads();
        int off = base, tot = 0;
        for (int w = 0; w < FT_THREADS / 64; ++w) {
            off += w < wave ? wcnt[w] : 0;
            tot += wcnt[w];
        }
        const int idx = off + __popcll(m & ((1ull << lane) - 1ull));
        if (act && idx % A.split == s) {
            const double sg = a.sigma[(size_t)l * a.nz + z];
            const double sd = sg > kSigmaFloor ? sg : kSigmaFloor;
            lk2[idx / A.split] = kLog2e / (2.0 * sd * sd);
            lcn[idx / A.split] = a.idd[(size_t)l * a.nz + z] / (2.0 * kPi * sd * sd);
            lcut[idx / A.split] = (float)(FT_WIN_SIGMAS * sd);
            lay[idx / A.split] = l;
        }
        base += tot;
        __syncthreads();
    }
    nmine = base > s ? (base - s + A.split - 1) / A.split : 0;
    if (nmine <= 0) return;             // uniform: nothing for this workgroup

    for (int cy = 0; cy < ncy; ++cy) {
        const int j0 = chunk_begin(a.nsy, ncy, cy), nj = chunk_begin(a.nsy, ncy, cy + 1) - j0;
        for (int cx = 0; cx < ncx; ++cx) {
            const int i0 = chunk_begin(a.nsx, ncx, cx), ni = chunk_begin(a.nsx, ncx, cx + 1) - i0;
            for (int txi = 0; txi < tiles_x; ++txi) {
                const int x0 = txi * AM_TX;
                const bool interior_x = x0 + AM_TX <= a.nx && (a.nx & 3) == 0 && ((uintptr_t)A.R & 15) == 0;
                const double Cx = (a.ox + (double)x0 * a.dx) - (a.sx0 + (double)i0 * a.dsx);
                // groups of AM_NT consecutive 64-row tiles: R of the group in registers, P of a
                // layer accumulated over the group and joined to the partial once
                for (int tg = 0; tg < tiles_y; tg += AM_NT) {
                    const int nt = min(AM_NT, tiles_y - tg);
                    const bool first_group = txi == 0 && tg == 0;
                    // windows of this group's layers (read after the first layer's barrier;
                    // the other buffer may still be read by the previous group's stage 2)
                    const int par = (tg / AM_NT) & 1;
                    for (int e = threadIdx.x; e < nmine * 9; e += FT_THREADS) {
                        const int q = e / 9, w = e % 9;
                        int first = 1, last = 0;
                        if (w < 3) {
                            if (16 * w < ni) {
                                const float s0 = (float)((16.0 * w * a.dsx - Cx) / a.dx);
                                const float s1 = (float)(((double)min(16 * w + 15, ni - 1) * a.dsx - Cx) / a.dx);
                                spot_range(s0, s1, lcut[q] * (float)(1.0 / a.dx), AM_TX, first, last);
                            }
                        } else {
                            const int u = (w - 3) / 3, mb = (w - 3) % 3;
                            if (u < nt && 16 * mb < nj) {
                                const double Cy = (a.oy + (double)((tg + u) * AM_TY) * a.dy) -
                                                  (a.sy0 + (double)j0 * a.dsy);
                                const float s0 = (float)((16.0 * mb * a.dsy - Cy) / a.dy);
                                const float s1 = (float)(((double)min(16 * mb + 15, nj - 1) * a.dsy - Cy) / a.dy);
                                spot_range(s0, s1, lcut[q] * (float)(1.0 / a.dy), AM_TY, first, last);
                            }
                        }
                        awin[par][q][w] = first > last ? 0 : (unsigned char)((first >> 5) | ((last / 32 + 1) << 4));
                    }
                    V8 rh[AM_NT][4], rl[AM_NT][4];
                    // R of tile u, this wave's row: columns 32 ks + 8 g + e
                    auto load_r = [&](int u, float (&rv)[4][8]) {
                        const int y = (tg + u) * AM_TY + 16 * wave + c16;
                        if (u < nt && interior_x && y < a.ny) {
                            // whole 128-column row segment inside the grid, 16-byte aligned
                            const float4* src = (const float4*)(Rz + (size_t)y * a.nx + x0 + 8 * g);
#pragma unroll
                            for (int ks = 0; ks < 4; ++ks) {
                                const float4 v0 = src[8 * ks], v1 = src[8 * ks + 1];
                                rv[ks][0] = v0.x; rv[ks][1] = v0.y; rv[ks][2] = v0.z; rv[ks][3] = v0.w;
                                rv[ks][4] = v1.x; rv[ks][5] = v1.y; rv[ks][6] = v1.z; rv[ks][7] = v1.w;
                            }
                        } else {
#pragma unroll
                            for (int ks = 0; ks < 4; ++ks)
#pragma unroll
                                for (int e = 0; e < 8; ++e) {
                                    const int x = x0 + 32 * ks + 8 * g + e;
                                    rv[ks][e] = (u < nt && y < a.ny && x < a.nx) ? Rz[(size_t)y * a.nx + x] : 0.0f;
                                }
                        }
                    };
                    auto split_r = [&](int u, const float (&rv)[4][8]) {
#pragma unroll
                        for (int ks = 0; ks < 4; ++ks)
#pragma unroll
                            for (int e = 0; e < 8; ++e) {
                                T h, l;
                                split<T>(rv[ks][e], h, l);
                                rh[u][ks][e] = h;
                                rl[u][ks][e] = l;
                            }
                    };
                    int eR = 0;     // FP16: this wave's e_w
                    if (!F16) {
#pragma unroll
                        for (int u = 0; u < AM_NT; ++u) {
                            float rv[4][8];
                            load_r(u, rv);
                            split_r(u, rv);
                        }
                    } else {
                        float rv[AM_NT][4][8];
                        float m = 0.0f;
#pragma unroll
                        for (int u = 0; u < AM_NT; ++u) {
                            load_r(u, rv[u]);
#pragma unroll
                            for (int ks = 0; ks < 4; ++ks)
#pragma unroll
                                for (int e = 0; e < 8; ++e) m = fmaxf(m, fabsf(rv[u][ks][e]));
                        }
                        for (int o = 32; o > 0; o >>= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
                        eR = (m > 0.0f && m <= 3.0e38f) ? ilogbf(m) : -1000;
                        if (lane == 0) rexp[wave] = eR;     // read after the tables' barrier below
                        const float sc = eR > -1000 ? ldexpf(1.0f, -eR) : 0.0f;
#pragma unroll
                        for (int u = 0; u < AM_NT; ++u) {
#pragma unroll
                            for (int ks = 0; ks < 4; ++ks)
#pragma unroll
                                for (int e = 0; e < 8; ++e) rv[u][ks][e] *= sc;
                            split_r(u, rv[u]);
                        }
                    }
                    const TableRow rowx = table_row(gx, Cx, AM_TX);
                    TableRow rowy[AM_NT];
#pragma unroll
                    for (int u = 0; u < AM_NT; ++u)
                        rowy[u] = table_row(gy, (a.oy + (double)((tg + u) * AM_TY) * a.dy) -
                                                    (a.sy0 + (double)j0 * a.dsy), AM_TY);
                    // QC > 0: this thread's table entries, squared distances formed once per group
                    TabShare<QC ? 2 * (AM_TX - 1) + QC * (AM_SC - 1) + 1 : 1> shx;
                    TabShare<QC ? 2 * (AM_TY - 1) + QC * (AM_SC - 1) + 1 : 1> shy[AM_NT];
                    if (QC) {
                        shx.init(gx, rowx);
#pragma unroll
                        for (int u = 0; u < AM_NT; ++u) shy[u].init(gy, rowy[u]);
                    }
                    // the tables of layer q into buffer b
                    auto build = [&](int q, int b) {
                        const float k2 = (float)lk2[q];
                        if (QC) {
                            build_packed<T>(shx, Tx[b], k2);
#pragma unroll
                            for (int u = 0; u < AM_NT; ++u)
                                if (u < nt) build_packed<T>(shy[u], Ty[b][u], k2);
                        } else {
                            for (int t = threadIdx.x; t < ntX; t += FT_THREADS) {
                                const float d = table_distance_f(rowx, t);
                                Tx[b][table_slot(gx, t)] = pack_split<T>(__builtin_amdgcn_exp2f(-(d * d) * k2));
                            }
#pragma unroll
                            for (int u = 0; u < AM_NT; ++u) {
                                if (u >= nt) break;
                                for (int t = threadIdx.x; t < ntY; t += FT_THREADS) {
                                    const float d = table_distance_f(rowy[u], t);
                                    Ty[b][u][table_slot(gy, t)] =
                                        pack_split<T>(__builtin_amdgcn_exp2f(-(d * d) * k2));
                                }
                            }
                        }
                    };
                    __syncthreads();            // the previous group's readers of buffer 0 are done
                    build(0, 0);
                    __syncthreads();
                    // FP16: the group's scale e and this wave's factor 2^(e_w - e) for its Q rows
                    int eG = 0;
                    float qsc = 1.0f;
                    if (F16) {
                        eG = max(max(rexp[0], rexp[1]), max(rexp[2], rexp[3]));
                        qsc = eR > -1000 ? ldexpf(1.0f, eR - eG) : 0.0f;
                        if (eG == -1000) eG = 0;             // the whole group's R is zero
                    }
                    // per layer q: stage 1 (Tx[b] -> Q), barrier, stage 2 (Ty[b], Q) with the
                    // tables of layer q + 1 built into the other buffer, barrier
                    for (int q = 0; q < nmine; ++q) {
                        const int bt = q & 1;
                        // stage 1 of both tiles of the group: Q[y][i], this wave's 16 rows of each,
                        // 3 spot blocks (unrolled: their product chains interleave); the Gx operand
                        // of a (spot block, K step) is gathered once and used for both tiles
#pragma unroll
                        for (int nb = 0; nb < 3; ++nb) {
                            const int wc = __builtin_amdgcn_readfirstlane((int)awin[par][q][nb]);
                            const int k_lo = wc & 15, k_hi = wc >> 4;
                            v4f acc1[AM_NT];
#pragma unroll
                            for (int u = 0; u < AM_NT; ++u) acc1[u] = v4f{};
#pragma unroll
                            for (int ks = 0; ks < 4; ++ks) {
                                if (ks < k_lo || ks >= k_hi) continue;
                                unsigned t8[8];
#pragma unroll
                                for (int e = 0; e < 8; ++e) t8[e] = Tx[bt][slot1(nb, ks, e)];
                                V8 bh, bl;
                                unpack8(t8, bh, bl);
#pragma unroll
                                for (int u = 0; u < AM_NT; ++u) {
                                    if (u >= nt) break;
                                    acc1[u] = mfma16(rh[u][ks], bh, acc1[u]);
                                    if (SPLIT) {
                                        acc1[u] = mfma16(rh[u][ks], bl, acc1[u]);
                                        acc1[u] = mfma16(rl[u][ks], bh, acc1[u]);
                                    }
                                }
                            }
                            // C layout: rows y = 16 wave + 4 g + r, column i = 16 nb + c16
#pragma unroll
                            for (int u = 0; u < AM_NT; ++u) {
                                if (u >= nt) break;
                                u2v hv, lv;
#pragma unroll
                                for (int h2 = 0; h2 < 2; ++h2) {
                                    T a0, b0, a1, b1;
                                    split<T>(F16 ? acc1[u][2 * h2] * qsc : acc1[u][2 * h2], a0, b0);
                                    split<T>(F16 ? acc1[u][2 * h2 + 1] * qsc : acc1[u][2 * h2 + 1], a1, b1);
                                    hv[h2] = (unsigned)bits16(a0) | ((unsigned)bits16(a1) << 16);
                                    lv[h2] = (unsigned)bits16(b0) | ((unsigned)bits16(b1) << 16);
                                }
                                const int off = (16 * nb + c16) * AM_QS + 16 * wave + 4 * g;
                                *(u2v*)&Qh[u][off] = hv;
                                if (SPLIT) *(u2v*)&Ql[u][off] = lv;
                            }
                        }
                        __syncthreads();
                        // stage 2: P[j][i] blocks (mb, nb) = wave + 4 qq of the 3 x 3, summed over
                        // the group's tiles, joined to the slice's partial once
                        float* pz = A.part + ((size_t)z * a.L + lay[q]) * a.nsy * a.nsx;
                        const double cn = F16 ? ldexp(lcn[q], eG) : lcn[q];     // FP16: R's scale undone
#pragma unroll 1
                        for (int qq = 0; qq < 3; ++qq) {
                            const int b2 = wave + 4 * qq;
                            if (b2 >= 9) break;
                            const int mb = b2 / 3, nb = b2 % 3;
                            v4f p = {};
                            for (int u = 0; u < nt; ++u) {
                                const int wc = __builtin_amdgcn_readfirstlane((int)awin[par][q][3 + 3 * u + mb]);
                                const int k_lo = wc & 15, k_hi = wc >> 4;
#pragma unroll
                                for (int ks = 0; ks < 2; ++ks) {
                                    if (ks < k_lo || ks >= k_hi) continue;
                                    unsigned t8[8];
#pragma unroll
                                    for (int e = 0; e < 8; ++e) t8[e] = Ty[bt][u][slot2(mb, ks, e)];
                                    V8 ah, al;
                                    unpack8(t8, ah, al);
                                    const int off = (16 * nb + c16) * AM_QS + 32 * ks + 8 * g;
                                    const V8 bh = *(const V8*)&Qh[u][off];
                                    p = mfma16(ah, bh, p);
                                    if (SPLIT) {
                                        p = mfma16(ah, *(const V8*)&Ql[u][off], p);
                                        p = mfma16(al, bh, p);
                                    }
                                }
                            }
                            // C layout: j = 16 mb + 4 g + r, i = 16 nb + c16
#pragma unroll
                            for (int r = 0; r < 4; ++r) {
                                const int j = 16 * mb + 4 * g + r, i = 16 * nb + c16;
                                if (j < nj && i < ni) {
                                    const size_t off = (size_t)(j0 + j) * a.nsx + i0 + i;
                                    const float cp = (float)(cn * (double)p[r]);
                                    pz[off] = first_group ? cp : pz[off] + cp;
                                }
                            }
                        }
                        if (q + 1 < nmine) build(q + 1, bt ^ 1);
                        __syncthreads();        // stage 2 has read Q and Ty[bt]; the next tables are built
                    }
                }
            }
        }
    }
}

// FP16 split products on the tabulated geometry: held to the registers of 3 waves per
// SIMD, which the BF16 kernel has by itself
template <typename T, int MODE, int QC>
__global__ __launch_bounds__(FT_THREADS, (!__is_same(T, __bf16) && MODE == MODE_X3 && QC > 0 ? 3 : 2))
void adj_mixed_kernel(AdjArgs A)
{
    int z, sh;
    xcd_unit((int)blockIdx.x, A.split, z, sh);
    if (z >= A.f.nz) return;
    adj_mixed_body<T, MODE, QC>(A, z * A.split + sh);
}

template <typename T, int MODE, int QC>
__global__ __launch_bounds__(FT_THREADS, (!__is_same(T, __bf16) && MODE == MODE_X3 && QC > 0 ? 3 : 2))
void adj_mixed_batch_kernel(const AdjArgs* __restrict__ items, const int* __restrict__ wg0, int n)
{
    const int k = batch_item(wg0, n, blockIdx.x);
    const AdjArgs A = items[k];
    int z, sh;     // item ranges start at multiples of 8: launch index b and b - wg0[k] share an XCD
    xcd_unit((int)blockIdx.x - wg0[k], A.split, z, sh);
    if (z >= A.f.nz) return;
    adj_mixed_body<T, MODE, QC>(A, z * A.split + sh);
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
// Layers per workgroup of a split: the mixed adjoint keeps at most AM_MAXK.
static bool split_fits(const AdjArgs& A)
{
    return (A.f.L + A.split - 1) / A.split <= AM_MAXK;
}

int pbh_launch_adj_mixed(const AdjArgs& A, int precision, int nz, hipStream_t s)
{
    if (!split_fits(A)) return PBH_ERR_UNSUPPORTED;
    const dim3 grid((unsigned)xcd_grid(nz, A.split)), block(FT_THREADS);
    const bool q5 = pbh_adj_mixed_q5(A);
    switch (precision) {
    case PBH_PREC_BF16X3:
        if (q5) hipLaunchKernelGGL((adj_mixed_kernel<__bf16, MODE_X3, 5>), grid, block, 0, s, A);
        else hipLaunchKernelGGL((adj_mixed_kernel<__bf16, MODE_X3, 0>), grid, block, 0, s, A);
        break;
    case PBH_PREC_BF16:
        if (q5) hipLaunchKernelGGL((adj_mixed_kernel<__bf16, MODE_PLAIN, 5>), grid, block, 0, s, A);
        else hipLaunchKernelGGL((adj_mixed_kernel<__bf16, MODE_PLAIN, 0>), grid, block, 0, s, A);
        break;
    case PBH_PREC_FP16X3:
        if (q5) hipLaunchKernelGGL((adj_mixed_kernel<_Float16, MODE_X3, 5>), grid, block, 0, s, A);
        else hipLaunchKernelGGL((adj_mixed_kernel<_Float16, MODE_X3, 0>), grid, block, 0, s, A);
        break;
    case PBH_PREC_FP16:
        if (q5) hipLaunchKernelGGL((adj_mixed_kernel<_Float16, MODE_PLAIN, 5>), grid, block, 0, s, A);
        else hipLaunchKernelGGL((adj_mixed_kernel<_Float16, MODE_PLAIN, 0>), grid, block, 0, s, A);
        break;
    default:
        return PBH_ERR_UNSUPPORTED;
    }
    return launch_rc("adj_mixed_kernel");
}

bool pbh_adj_mixed_fits(const AdjArgs& A)
{
    return split_fits(A);
}

int pbh_launch_adj_mixed_batch(const AdjArgs* items, const int* wg0, int n, int total, int precision, bool q5,
                               hipStream_t s)
{
    const dim3 grid((unsigned)total), block(FT_THREADS);
    switch (precision) {
    case PBH_PREC_BF16X3:
        if (q5) hipLaunchKernelGGL((adj_mixed_batch_kernel<__bf16, MODE_X3, 5>), grid, block, 0, s, items, wg0, n);
        else hipLaunchKernelGGL((adj_mixed_batch_kernel<__bf16, MODE_X3, 0>), grid, block, 0, s, items, wg0, n);
        break;
    case PBH_PREC_BF16:
        if (q5) hipLaunchKernelGGL((adj_mixed_batch_kernel<__bf16, MODE_PLAIN, 5>), grid, block, 0, s, items, wg0, n);
        else hipLaunchKernelGGL((adj_mixed_batch_kernel<__bf16, MODE_PLAIN, 0>), grid, block, 0, s, items, wg0, n);
        break;
    case PBH_PREC_FP16X3:
        if (q5) hipLaunchKernelGGL((adj_mixed_batch_kernel<_Float16, MODE_X3, 5>), grid, block, 0, s, items, wg0, n);
        else hipLaunchKernelGGL((adj_mixed_batch_kernel<_Float16, MODE_X3, 0>), grid, block, 0, s, items, wg0, n);
        break;
    case PBH_PREC_FP16:
        if (q5) hipLaunchKernelGGL((adj_mixed_batch_kernel<_Float16, MODE_PLAIN, 5>), grid, block, 0, s, items, wg0, n);
        else hipLaunchKernelGGL((adj_mixed_batch_kernel<_Float16, MODE_PLAIN, 0>), grid, block, 0, s, items, wg0, n);
        break;
    default:
        return PBH_ERR_UNSUPPORTED;
    }
    return launch_rc("adj_mixed_batch_kernel");
}
