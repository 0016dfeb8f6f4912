// pbh_fp64.hip -- forward dose in FP64: the per-layer-exact model with FP64
// weights, tables and dose (what the reference's torch_separable computes with
// dtype=float64 on a GPU, engines.py:208-283).
//
// For each depth slice z and active layer l (IDD_l(z) != 0), with s = max(sigma, 1e-9):
//   Gy[y][j] = exp(-(y - sy_j)^2 / (2 s^2)),  Gx[x][i] likewise,
//   U[y][i]  = c sum_j Gy[y][j] W_l[j][i],     c = IDD_l(z) / (2 pi s^2),
//   D[y][x] += sum_i Gx[x][i] U[y][i]
// (the continuous normalisation 1 / (sqrt(2 pi) s) of each factor folded into c).
// One workgroup of 256 threads per (slice, band of FP_TY rows, FP_TX columns), all
// active layers in layer order, spot chunks of at most FP_SC x FP_SC. Sixteen rows
// per workgroup give the 128^3 grid 1,024 workgroups, four waves per SIMD. Every
// voxel sums its layers, chunks and spots in a fixed order, so results are
// deterministic. No workspace, no layer chunks.
//
// fwd_fp64_mfma_kernel, where the voxel and spot pitches are exact multiples P u and
// Q u of one length u (2 mm voxels and 5 mm spots: u = 1 mm, P = 2, Q = 5): the
// distance from voxel a of the tile to spot i of the chunk is C + (P a - Q i) u, so
// the factors of a (tile, chunk, layer) take only P (T - 1) + Q (K - 1) + 1 values
// (570 for 128 columns and 64 spots instead of 8,192). They are evaluated once into
// LDS and gathered at entry Q i - P a + P (T - 1) as MFMA operands:
//   stage A  U^T blocks = v_mfma_f64_16x16x4_f64 over K = 4 y spots: A = Gy (16 rows
//            x 4 spots, from the y table), B = W (4 spots x 16 x spots, global
//            memory); wave w forms the 16 x 16 block of x spots 16 w .. 16 w + 15
//   c U goes to LDS spot-major
//   stage B  D += U Gx^T over K = 4 x spots: A = U (16 rows x 4 spots), B = Gx^T
//            (4 spots x 16 columns, from the x table); wave w owns columns
//            32 w .. 32 w + 31, two accumulator blocks sharing each A operand.
// (On MI355X the FP64 matrix and vector peaks are equal, 78.6 TFLOP/s: the matrix
// form pays because each operand read from LDS feeds 16 products, where the vector
// form read one U value per product.)
//
// fwd_fp64_kernel, any other pitch ratio, on the vector ALUs with every factor
// evaluated where it is used:
//   - the y factors of the band and chunk go to LDS (FP_TY x FP_SC);
//   - thread t forms the FP_TY * FP_SC / 256 entries U[t / 64 + 4 q][t % 64] of U,
//     summing W over the y chunks: each W value is read from global memory once
//     (consecutive threads, consecutive spots i) and used for all its rows;
//   - c U goes to LDS transposed (spot-major, rows contiguous), and thread t (column
//     x = t % FP_TX, rows FP_RPT (t / FP_TX) + r) evaluates Gx[x][i] once per spot and
//     adds it times U[y][i] to its FP_RPT voxels, reading the U values in pairs.
#include "pbh_fwd_core.h"

namespace pbh {

constexpr int FP_TY = 16;          // rows per workgroup (16: four waves per SIMD at 128^3)
constexpr int FP_TX = 128;         // columns per workgroup
constexpr int FP_SC = 64;          // spots per chunk and axis
constexpr int FP_THREADS = 256;
constexpr int FP_RPT = FP_TY * FP_TX / FP_THREADS;       // rows per thread in the accumulation (8)
constexpr int FP_UPT = FP_TY * FP_SC / FP_THREADS;       // U entries per thread (4)
constexpr int FP_TAB = 1024;       // factor table entries per axis, at most
// MFMA kernel: stage-A K steps whose weights load ahead of the table build (8: 168
// registers, three waves per SIMD, 0.19-0.21 ms at 128^3 against 0.21-0.23 ms for 1 at
// four waves; 16 take two waves per SIMD)
constexpr int FP_WPRE = 8;
static_assert(FP_RPT == 8 && FP_UPT == 4, "tile shape");
static_assert(FP_TY == 16 && FP_TX == 128 && FP_SC == 64, "MFMA tiling: one 16-row block, four waves");

typedef double v4d __attribute__((ext_vector_type(4)));

// One axis of the factor tables: voxel pitch P u, spot pitch Q u.
struct Fp64Axis {
    int P, Q;
    double u;
};

struct Fp64Args {
    int nx, ny, nz, nsx, nsy, L;
    double dx, dy, ox, oy;
    double dsx, dsy, sx0, sy0;     // spot i centre = sx0 + i dsx
    const double* w;               // (L, nsy, nsx)
    const double* idd;             // (L, nz)
    const double* sigma;           // (L, nz)
    double* dose;                  // (nz, ny, nx)
    int tiles_x, tiles_y;
    Fp64Axis tx, ty;               // factor tables (fwd_fp64_mfma_kernel)
};

__global__ __launch_bounds__(FP_THREADS) void fwd_fp64_kernel(Fp64Args a)
{
    __shared__ double Gy[FP_TY][FP_SC + 1];
    __shared__ __align__(16) double Ut[FP_SC][FP_TY + 2];    // Ut[i][y]: c U, spot-major

    const int t = threadIdx.x;
    const int tx = blockIdx.x % a.tiles_x;
    const int ty = (blockIdx.x / a.tiles_x) % a.tiles_y;
    const int z = blockIdx.x / (a.tiles_x * a.tiles_y);
    const int x0 = tx * FP_TX, y0 = ty * FP_TY;
    const int xl = t % FP_TX, rb = FP_RPT * (t / FP_TX);      // accumulation: column, first row
    const double xc = a.ox + (double)(x0 + xl) * a.dx;

    double acc[FP_RPT];
#pragma unroll
    for (int r = 0; r < FP_RPT; ++r) acc[r] = 0.0;

    const int ncx = (a.nsx + FP_SC - 1) / FP_SC, ncy = (a.nsy + FP_SC - 1) / FP_SC;
    for (int l = 0; l < a.L; ++l) {
        const double iddv = a.idd[(size_t)l * a.nz + z];
        if (iddv == 0.0) continue;                            // uniform: no dose from this layer
        const double sg = a.sigma[(size_t)l * a.nz + z];
        const double s = sg > kSigmaFloor ? sg : kSigmaFloor;
        const double k = 1.0 / (2.0 * s * s);
        const double c = iddv / (2.0 * kPi * s * s);
        const double* wl = a.w + (size_t)l * a.nsy * a.nsx;
        for (int cx = 0; cx < ncx; ++cx) {
            const int i0 = cx * FP_SC, ni = min(FP_SC, a.nsx - i0);
            double u[FP_UPT];
#pragma unroll
            for (int q = 0; q < FP_UPT; ++q) u[q] = 0.0;
            for (int cy = 0; cy < ncy; ++cy) {
                const int j0 = cy * FP_SC, nj = min(FP_SC, a.nsy - j0);
                __syncthreads();                              // earlier readers of Gy are done
                for (int e = t; e < FP_TY * FP_SC; e += FP_THREADS) {
                    const int r = e / FP_SC, j = e % FP_SC;
                    const double d = (a.oy + (double)(y0 + r) * a.dy) - (a.sy0 + (double)(j0 + j) * a.dsy);
                    Gy[r][j] = j < nj ? exp(-(d * d) * k) : 0.0;
                }
                __syncthreads();
                const int i = t % FP_SC, r0 = t / FP_SC;          // rows r0 + 4 q
                if (i < ni) {
                    const double* wc = wl + (size_t)j0 * a.nsx + i0 + i;
                    for (int j = 0; j < nj; ++j) {
                        const double wv = wc[(size_t)j * a.nsx];
#pragma unroll
                        for (int q = 0; q < FP_UPT; ++q) u[q] = fma(Gy[r0 + 4 * q][j], wv, u[q]);
                    }
                }
            }
            __syncthreads();                                  // earlier readers of U are done
#pragma unroll
            for (int q = 0; q < FP_UPT; ++q) Ut[t % FP_SC][t / FP_SC + 4 * q] = c * u[q];
            __syncthreads();
            for (int i = 0; i < ni; ++i) {
                const double d = xc - (a.sx0 + (double)(i0 + i) * a.dsx);
                const double g = exp(-(d * d) * k);
                const double2* ur = reinterpret_cast<const double2*>(&Ut[i][rb]);
#pragma unroll
                for (int r = 0; r < FP_RPT / 2; ++r) {
                    const double2 v = ur[r];
                    acc[2 * r] = fma(g, v.x, acc[2 * r]);
                    acc[2 * r + 1] = fma(g, v.y, acc[2 * r + 1]);
                }
            }
        }
    }
    if (x0 + xl >= a.nx) return;
    double* base = a.dose + (size_t)z * a.ny * a.nx + x0 + xl;
#pragma unroll
    for (int r = 0; r < FP_RPT; ++r) {
        const int y = y0 + rb + r;
        if (y < a.ny) base[(size_t)y * a.nx] = acc[r];
    }
}


__global__ __launch_bounds__(FP_THREADS) void fwd_fp64_mfma_kernel(Fp64Args a)
{
    __shared__ double Ty[FP_TAB], Tx[FP_TAB];             // factors by table entry
    __shared__ double Ut[FP_SC][FP_TY + 1];               // Ut[i][y]: c U, spot-major

    const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
    const int r16 = lane & 15, k4 = lane >> 4;            // operand row / column, K index
    const int tx = blockIdx.x % a.tiles_x;
    const int ty = (blockIdx.x / a.tiles_x) % a.tiles_y;
    const int z = blockIdx.x / (a.tiles_x * a.tiles_y);
    const int x0 = tx * FP_TX, y0 = ty * FP_TY;
    const int Px = a.tx.P, Qx = a.tx.Q, Py = a.ty.P, Qy = a.ty.Q;
    // stage A: row y = r16 of the A operand at entry Q j + eA; stage B: column
    // x = 32 wave + 16 nb + r16 of the B operand at entry Q i + eB - 16 nb P
    const int eA = Py * (FP_TY - 1 - r16);
    const int eB = Px * (FP_TX - 1 - 32 * wave - r16);

    v4d acc[2] = {};
    const int ncx = (a.nsx + FP_SC - 1) / FP_SC, ncy = (a.nsy + FP_SC - 1) / FP_SC;
    for (int l = 0; l < a.L; ++l) {
        const double iddv = a.idd[(size_t)l * a.nz + z];
        if (iddv == 0.0) continue;                            // uniform: no dose from this layer
        const double sg = a.sigma[(size_t)l * a.nz + z];
        const double s = sg > kSigmaFloor ? sg : kSigmaFloor;
        const double k = 1.0 / (2.0 * s * s);
        const double c = iddv / (2.0 * kPi * s * s);
        const double* wl = a.w + (size_t)l * a.nsy * a.nsx;
        for (int cx = 0; cx < ncx; ++cx) {
            const int i0 = cx * FP_SC, ni = min(FP_SC, a.nsx - i0), nip = (ni + 3) & ~3;
            const int iA = 16 * wave + r16;                   // stage A: x spot of the B column
            v4d u = {};
            for (int cy = 0; cy < ncy; ++cy) {
                const int j0 = cy * FP_SC, nj = min(FP_SC, a.nsy - j0), njp = (nj + 3) & ~3;
                // the B operands (weights) of the first FP_WPRE K steps, loaded ahead of the
                // table build; later steps load theirs in the loop
                const bool col = 16 * wave < ni && iA < ni;
                const double* wc = wl + (size_t)j0 * a.nsx + i0 + iA;
                double wb[FP_WPRE];
#pragma unroll
                for (int ks = 0; ks < FP_WPRE; ++ks) {
                    const int j = 4 * ks + k4;
                    wb[ks] = (col && j < nj) ? wc[(size_t)j * a.nsx] : 0.0;
                }
                __syncthreads();                              // earlier readers of Ty are done
                const int n = Py * (FP_TY - 1) + Qy * (njp - 1) + 1;
                const double C = (a.oy + (double)y0 * a.dy) - (a.sy0 + (double)j0 * a.dsy);
                for (int e = t; e < n; e += FP_THREADS) {
                    const double d = C + (double)(Py * (FP_TY - 1) - e) * a.ty.u;
                    Ty[e] = exp(-(d * d) * k);
                }
                __syncthreads();
                if (16 * wave < ni) {                         // wave-uniform
#pragma unroll
                    for (int ks = 0; ks < FP_WPRE; ++ks)
                        if (4 * ks < njp)
                            u = __builtin_amdgcn_mfma_f64_16x16x4f64(Ty[Qy * (4 * ks + k4) + eA], wb[ks], u, 0, 0, 0);
                    for (int ks = FP_WPRE; ks < njp / 4; ++ks) {
                        const int j = 4 * ks + k4;
                        const double w = (col && j < nj) ? wc[(size_t)j * a.nsx] : 0.0;
                        u = __builtin_amdgcn_mfma_f64_16x16x4f64(Ty[Qy * j + eA], w, u, 0, 0, 0);
                    }
                }
            }
            // u[r]: U[y = 4 k4 + r][i = iA] (zero for i >= ni: padding of both stages)
            __syncthreads();                                  // earlier readers of Ut and Tx are done
#pragma unroll
            for (int r = 0; r < 4; ++r) Ut[iA][4 * k4 + r] = c * u[r];
            const int n = Px * (FP_TX - 1) + Qx * (nip - 1) + 1;
            const double C = (a.ox + (double)x0 * a.dx) - (a.sx0 + (double)i0 * a.dsx);
            for (int e = t; e < n; e += FP_THREADS) {
                const double d = C + (double)(Px * (FP_TX - 1) - e) * a.tx.u;
                Tx[e] = exp(-(d * d) * k);
            }
            __syncthreads();
            for (int ks = 0; ks < nip / 4; ++ks) {
                const int i = 4 * ks + k4;
                const double ua = Ut[i][r16];
                const int e = Qx * i + eB;
                acc[0] = __builtin_amdgcn_mfma_f64_16x16x4f64(ua, Tx[e], acc[0], 0, 0, 0);
                acc[1] = __builtin_amdgcn_mfma_f64_16x16x4f64(ua, Tx[e - 16 * Px], acc[1], 0, 0, 0);
            }
        }
    }
    // acc[nb][r]: D[y = 4 k4 + r][x = 32 wave + 16 nb + r16]
#pragma unroll
    for (int nb = 0; nb < 2; ++nb) {
        const int x = x0 + 32 * wave + 16 * nb + r16;
        if (x >= a.nx) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int y = y0 + 4 * k4 + r;
            if (y < a.ny) a.dose[((size_t)z * a.ny + y) * a.nx + x] = acc[nb][r];
        }
    }
}
}  // namespace pbh

using namespace pbh;

extern "C" int pbh_validate(const PbhGrid* g, const PbhLattice* lat);

// Voxel pitch dv = P u and spot pitch ds = Q u for small integers, to 1e-15 relative
// (decimal pitches such as 0.2 and 0.5 cm pass; a pitch ratio off by more would move
// the spots by more than FP64 resolves), with at most FP_TAB table entries for tile
// width T and chunk width K.
static bool fp64_axis(double dv, double ds, int T, int K, Fp64Axis* out)
{
    for (int Q = 1; Q <= 64; ++Q) {
        const double pr = dv * Q / ds;
        const int P = (int)llround(pr);
        if (P < 1 || P > 64 || fabs(pr - P) > 1e-15 * pr) continue;
        if (P * (T - 1) + Q * (K - 1) + 1 > FP_TAB) return false;
        *out = Fp64Axis{P, Q, dv / P};
        return true;
    }
    return false;
}

extern "C" int pbh_forward_fp64(const PbhGrid* g, const PbhLattice* lat, const double* w, const double* idd,
                                const double* sigma, double* dose, void* stream)
{
    int rc = pbh_validate(g, lat);
    if (rc != PBH_OK) return rc;
    if (!w || !idd || !sigma || !dose) return PBH_ERR_NULL_ARG;
    Fp64Args a;
    a.nx = g->nx; a.ny = g->ny; a.nz = g->nz;
    a.nsx = lat->n_sx; a.nsy = lat->n_sy; a.L = lat->n_layers;
    a.dx = g->dx; a.dy = g->dy; a.ox = g->ox; a.oy = g->oy;
    a.dsx = lat->dx_spot; a.dsy = lat->dy_spot;
    a.sx0 = lat->x0 - (lat->n_sx - 1) / 2.0 * lat->dx_spot;
    a.sy0 = lat->y0 - (lat->n_sy - 1) / 2.0 * lat->dy_spot;
    a.w = w; a.idd = idd; a.sigma = sigma; a.dose = dose;
    a.tiles_x = cdiv(g->nx, FP_TX);
    a.tiles_y = cdiv(g->ny, FP_TY);
    const long long blocks = (long long)a.tiles_x * a.tiles_y * a.nz;
    if (blocks > 0x7fffffffLL) return PBH_ERR_UNSUPPORTED;
    a.tx = a.ty = Fp64Axis{0, 0, 0.0};
    const bool tab = fp64_axis(a.dx, a.dsx, FP_TX, FP_SC, &a.tx) && fp64_axis(a.dy, a.dsy, FP_TY, FP_SC, &a.ty);
    if (tab) {
        hipLaunchKernelGGL(fwd_fp64_mfma_kernel, dim3((unsigned)blocks), dim3(FP_THREADS), 0, (hipStream_t)stream, a);
        return launch_rc("fwd_fp64_mfma_kernel");
    }
    hipLaunchKernelGGL(fwd_fp64_kernel, dim3((unsigned)blocks), dim3(FP_THREADS), 0, (hipStream_t)stream, a);
    return launch_rc("fwd_fp64_kernel");
}
