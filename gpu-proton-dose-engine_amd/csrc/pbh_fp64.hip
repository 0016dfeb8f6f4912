// pbh_fp64.hip -- forward dose in FP64: the per-layer-exact model with FP64
// weights, tables and dose, FP64 FMA on the vector ALUs (what the reference's
// torch_separable computes with dtype=float64 on a GPU, engines.py:208-283).
//
// For each depth slice z and active layer l (IDD_l(z) != 0), with s = max(sigma, 1e-9):
//   Gy[y][j] = exp(-(y - sy_j)^2 / (2 s^2)),  Gx[x][i] likewise,
//   U[y][i]  = c sum_j Gy[y][j] W_l[j][i],     c = IDD_l(z) / (2 pi s^2),
//   D[y][x] += sum_i Gx[x][i] U[y][i]
// (the continuous normalisation 1 / (sqrt(2 pi) s) of each factor folded into c).
//
// One workgroup of 256 threads per (slice, band of FP_TY rows, FP_TX columns), all
// active layers in layer order, spot chunks of at most FP_SC x FP_SC:
//   - the y factors of the band and chunk go to LDS (FP_TY x FP_SC);
//   - thread t forms the FP_TY * FP_SC / 256 entries U[t / 64 + 4 q][t % 64] of U,
//     summing W over the y chunks: each W value is read from global memory once
//     (consecutive threads, consecutive spots i) and used for all its rows;
//   - c U goes to LDS transposed (spot-major, rows contiguous), and thread t (column
//     x = t % FP_TX, rows FP_RPT (t / FP_TX) + r) evaluates Gx[x][i] once per spot and
//     adds it times U[y][i] to its FP_RPT voxels, reading the U values in pairs.
// Sixteen rows per workgroup: an exp serves 8 voxels, and the 128^3 grid gets 1,024
// workgroups, four waves per SIMD (32 rows: 16 voxels per exp but two waves per SIMD,
// 1.36x slower; 8 rows: 1.5x slower).
// Every voxel sums its layers, chunks and spots in a fixed order, so results are
// deterministic. No workspace, no layer chunks.
#include "pbh_fwd_core.h"

namespace pbh {

constexpr int FP_TY = 16;          // rows per workgroup (16: four waves per SIMD at 128^3)
constexpr int FP_TX = 128;         // columns per workgroup
constexpr int FP_SC = 64;          // spots per chunk and axis
constexpr int FP_THREADS = 256;
constexpr int FP_RPT = FP_TY * FP_TX / FP_THREADS;       // rows per thread in the accumulation (8)
constexpr int FP_UPT = FP_TY * FP_SC / FP_THREADS;       // U entries per thread (4)
static_assert(FP_RPT == 8 && FP_UPT == 4, "tile shape");

struct Fp64Args {
    int nx, ny, nz, nsx, nsy, L;
    double dx, dy, ox, oy;
    double dsx, dsy, sx0, sy0;     // spot i centre = sx0 + i dsx
    const double* w;               // (L, nsy, nsx)
    const double* idd;             // (L, nz)
    const double* sigma;           // (L, nz)
    double* dose;                  // (nz, ny, nx)
    int tiles_x, tiles_y;
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

}  // namespace pbh

using namespace pbh;

extern "C" int pbh_validate(const PbhGrid* g, const PbhLattice* lat);

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
    hipLaunchKernelGGL(fwd_fp64_kernel, dim3((unsigned)blocks), dim3(FP_THREADS), 0, (hipStream_t)stream, a);
    return launch_rc("fwd_fp64_kernel");
}
