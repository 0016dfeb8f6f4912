// pbh_gamma.hip -- gamma index (gamma.py:41-160) on the GPU, FP64 arithmetic.
//
// For every voxel p of the reference dose:
//   gamma^2(p) = min over integer voxel offsets o with |o| <= DTA of
//                (|o| / DTA)^2 + ((eval[p + o] - ref[p]) / crit)^2
// with eval taken as 0 outside the grid (zero-filled shifts, gamma.py:63-75),
// crit = dose% / 100 * ref_max (global) or dose% / 100 * max(ref[p], 1e-12)
// (local), and offsets per axis limited to min(floor(DTA / d + 1e-9), 6)
// voxels (gamma.py:41-60). Voxels with ref >= threshold% / 100 * ref_max are
// evaluated (if none are, voxels with ref > 0). Pass: gamma <= 1.
//
// HBM-bound stencil: one thread per voxel (x fastest, coalesced), the offset
// list in LDS, neighbour reads served by L1/L2. Statistics are reduced per
// workgroup in a fixed order and summed by one workgroup in a fixed order, so
// every result is deterministic. The arithmetic follows the reference's
// operation order in FP64 with contraction off, so gamma values and pass
// decisions equal an FP64 evaluation of gamma.py on the same inputs.
#include "pbh_common.h"

#pragma clang fp contract(off)

namespace pbh {

constexpr int GM_THREADS = 256;
constexpr int GM_BLOCKS = 2048;
constexpr int GM_MAX_VOX = 6;
constexpr int GM_MAX_OFFS = (2 * GM_MAX_VOX + 1) * (2 * GM_MAX_VOX + 1) * (2 * GM_MAX_VOX + 1);

struct GammaPart {
    double max_t, sum_t, max_p, sum_p;                 // threshold mask / ref > 0 mask
    unsigned long long n_t, pass_t, n_p, pass_p;
};

struct GammaArgs {
    int nx, ny, nz;
    double dx, dy, dz;
    double dta;            // cm
    double dose_frac;      // dose% / 100
    double thr_frac;       // threshold% / 100
    double ref_max;        // > 0: given; otherwise read from *ref_max_dev
    const double* ref_max_dev;
    int local;
    int na, nb, nc;        // offset ranges in z, y, x
    double* gamma_out;     // optional map
    GammaPart* part;       // [GM_BLOCKS]
};

template <typename T>
__device__ __forceinline__ double ld(const T* p, size_t i) { return (double)p[i]; }

// per-workgroup maximum of ref (fixed-order tree); part[blk]
template <typename TR>
__global__ __launch_bounds__(GM_THREADS) void gamma_refmax_kernel(const TR* ref, size_t n, double* part)
{
    __shared__ double red[GM_THREADS];
    double m = -INFINITY;
    for (size_t e = (size_t)blockIdx.x * GM_THREADS + threadIdx.x; e < n; e += (size_t)GM_BLOCKS * GM_THREADS)
        m = fmax(m, ld(ref, e));
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = GM_THREADS / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ __launch_bounds__(GM_THREADS) void gamma_refmax_final_kernel(const double* part, double* out)
{
    __shared__ double red[GM_THREADS];
    double m = -INFINITY;
    for (int i = threadIdx.x; i < GM_BLOCKS; i += GM_THREADS) m = fmax(m, part[i]);
    red[threadIdx.x] = m;
    __syncthreads();
    for (int w = GM_THREADS / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = red[0];
}

template <typename TR, typename TE>
__global__ __launch_bounds__(GM_THREADS) void gamma_kernel(const TR* __restrict__ ref, const TE* __restrict__ ev,
                                                           GammaArgs g)
{
    __shared__ short oz[GM_MAX_OFFS], oy[GM_MAX_OFFS], ox[GM_MAX_OFFS];
    __shared__ double odr2[GM_MAX_OFFS];
    __shared__ int n_off;
    __shared__ double rd[4][GM_THREADS / 64];
    __shared__ unsigned long long rc[4][GM_THREADS / 64];

    // offsets inside the DTA sphere (gamma.py:50-59); their order does not
    // matter: the minimum is exact
    if (threadIdx.x == 0) n_off = 0;
    __syncthreads();
    const int ca = 2 * g.na + 1, cb = 2 * g.nb + 1, cc = 2 * g.nc + 1;
    for (int t = threadIdx.x; t < ca * cb * cc; t += GM_THREADS) {
        const int a = t / (cb * cc) - g.na, b = (t / cc) % cb - g.nb, c = t % cc - g.nc;
        const double az = a * g.dz, by = b * g.dy, cx = c * g.dx;
        const double d = sqrt(az * az + by * by + cx * cx);
        if (d <= g.dta + 1e-12) {
            const int slot = atomicAdd(&n_off, 1);
            oz[slot] = (short)a;
            oy[slot] = (short)b;
            ox[slot] = (short)c;
            const double dr = d / g.dta;
            odr2[slot] = dr * dr;
        }
    }
    __syncthreads();
    const int no = n_off;
    const double rmax = g.ref_max > 0.0 ? g.ref_max : g.ref_max_dev[0];
    const double crit_g = g.dose_frac * rmax;
    const double thr = g.thr_frac * rmax;

    double max_t = -INFINITY, sum_t = 0.0, max_p = -INFINITY, sum_p = 0.0;
    unsigned long long n_t = 0, pass_t = 0, n_p = 0, pass_p = 0;
    const size_t plane = (size_t)g.nx * g.ny, n = plane * g.nz;
    for (size_t e = (size_t)blockIdx.x * GM_THREADS + threadIdx.x; e < n; e += (size_t)GM_BLOCKS * GM_THREADS) {
        const double r = ld(ref, e);
        const bool in_t = r >= thr, in_p = r > 0.0;
        if (!(in_t || in_p || g.gamma_out)) continue;
        const int k = (int)(e / plane), j = (int)((e / g.nx) % g.ny), i = (int)(e % g.nx);
        const double crit = g.local ? g.dose_frac * fmax(r, 1e-12) : crit_g;
        double g2 = INFINITY;
        for (int q = 0; q < no; ++q) {
            const int kk = k + oz[q], jj = j + oy[q], ii = i + ox[q];
            double v = 0.0;
            if (kk >= 0 && kk < g.nz && jj >= 0 && jj < g.ny && ii >= 0 && ii < g.nx)
                v = ld(ev, (size_t)kk * plane + (size_t)jj * g.nx + ii);
            const double dd = (v - r) / crit;
            g2 = fmin(g2, odr2[q] + dd * dd);
        }
        const double gm = sqrt(g2);
        if (g.gamma_out) g.gamma_out[e] = gm;
        const bool pass = gm <= 1.0;
        if (in_t) {
            ++n_t;
            pass_t += pass;
            max_t = fmax(max_t, gm);
            sum_t += gm;
        }
        if (in_p) {
            ++n_p;
            pass_p += pass;
            max_p = fmax(max_p, gm);
            sum_p += gm;
        }
    }
    // fixed-order workgroup reduction
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int o = 32; o > 0; o >>= 1) {
        max_t = fmax(max_t, __shfl_xor(max_t, o, 64));
        sum_t += __shfl_xor(sum_t, o, 64);
        max_p = fmax(max_p, __shfl_xor(max_p, o, 64));
        sum_p += __shfl_xor(sum_p, o, 64);
        n_t += __shfl_xor(n_t, o, 64);
        pass_t += __shfl_xor(pass_t, o, 64);
        n_p += __shfl_xor(n_p, o, 64);
        pass_p += __shfl_xor(pass_p, o, 64);
    }
    if (lane == 0) {
        rd[0][wave] = max_t; rd[1][wave] = sum_t; rd[2][wave] = max_p; rd[3][wave] = sum_p;
        rc[0][wave] = n_t; rc[1][wave] = pass_t; rc[2][wave] = n_p; rc[3][wave] = pass_p;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        GammaPart p{-INFINITY, 0.0, -INFINITY, 0.0, 0, 0, 0, 0};
        for (int w = 0; w < GM_THREADS / 64; ++w) {
            p.max_t = fmax(p.max_t, rd[0][w]); p.sum_t += rd[1][w];
            p.max_p = fmax(p.max_p, rd[2][w]); p.sum_p += rd[3][w];
            p.n_t += rc[0][w]; p.pass_t += rc[1][w]; p.n_p += rc[2][w]; p.pass_p += rc[3][w];
        }
        g.part[blockIdx.x] = p;
    }
}

// counts = {n_eval, n_pass, n_offsets}, stats = {gamma_max, gamma_mean, ref_max}
__global__ __launch_bounds__(GM_THREADS) void gamma_final_kernel(const GammaPart* part, GammaArgs g, int n_offsets,
                                                                 unsigned long long* counts, double* stats)
{
    __shared__ double sd[4][GM_THREADS];
    __shared__ unsigned long long sc[4][GM_THREADS];
    GammaPart p{-INFINITY, 0.0, -INFINITY, 0.0, 0, 0, 0, 0};
    for (int b = threadIdx.x; b < GM_BLOCKS; b += GM_THREADS) {
        const GammaPart q = part[b];
        p.max_t = fmax(p.max_t, q.max_t); p.sum_t += q.sum_t;
        p.max_p = fmax(p.max_p, q.max_p); p.sum_p += q.sum_p;
        p.n_t += q.n_t; p.pass_t += q.pass_t; p.n_p += q.n_p; p.pass_p += q.pass_p;
    }
    const int t = threadIdx.x;
    sd[0][t] = p.max_t; sd[1][t] = p.sum_t; sd[2][t] = p.max_p; sd[3][t] = p.sum_p;
    sc[0][t] = p.n_t; sc[1][t] = p.pass_t; sc[2][t] = p.n_p; sc[3][t] = p.pass_p;
    __syncthreads();
    for (int w = GM_THREADS / 2; w > 0; w >>= 1) {
        if (t < w) {
            sd[0][t] = fmax(sd[0][t], sd[0][t + w]); sd[1][t] += sd[1][t + w];
            sd[2][t] = fmax(sd[2][t], sd[2][t + w]); sd[3][t] += sd[3][t + w];
            sc[0][t] += sc[0][t + w]; sc[1][t] += sc[1][t + w];
            sc[2][t] += sc[2][t + w]; sc[3][t] += sc[3][t + w];
        }
        __syncthreads();
    }
    if (t == 0) {
        const bool thr = sc[0][0] > 0;      // gamma.py:139-141: fall back to ref > 0
        const unsigned long long ne = thr ? sc[0][0] : sc[2][0];
        counts[0] = ne;
        counts[1] = thr ? sc[1][0] : sc[3][0];
        counts[2] = (unsigned long long)n_offsets;
        stats[0] = ne ? (thr ? sd[0][0] : sd[2][0]) : NAN;
        stats[1] = ne ? (thr ? sd[1][0] : sd[3][0]) / (double)ne : NAN;
        stats[2] = g.ref_max > 0.0 ? g.ref_max : g.ref_max_dev[0];
    }
}

template <typename TR, typename TE>
static int launch_gamma(const void* ref, const void* ev, GammaArgs g, size_t n, double* refmax_part, int n_offsets,
                        unsigned long long* counts, double* stats, hipStream_t s)
{
    int rc;
    if (!(g.ref_max > 0.0)) {
        hipLaunchKernelGGL((gamma_refmax_kernel<TR>), dim3(GM_BLOCKS), dim3(GM_THREADS), 0, s, (const TR*)ref, n,
                           refmax_part);
        if ((rc = launch_rc("gamma_refmax_kernel")) != PBH_OK) return rc;
        hipLaunchKernelGGL(gamma_refmax_final_kernel, dim3(1), dim3(GM_THREADS), 0, s, (const double*)refmax_part,
                           refmax_part + GM_BLOCKS);
        if ((rc = launch_rc("gamma_refmax_final_kernel")) != PBH_OK) return rc;
        g.ref_max_dev = refmax_part + GM_BLOCKS;
    }
    hipLaunchKernelGGL((gamma_kernel<TR, TE>), dim3(GM_BLOCKS), dim3(GM_THREADS), 0, s, (const TR*)ref, (const TE*)ev, g);
    if ((rc = launch_rc("gamma_kernel")) != PBH_OK) return rc;
    hipLaunchKernelGGL(gamma_final_kernel, dim3(1), dim3(GM_THREADS), 0, s, (const GammaPart*)g.part, g, n_offsets,
                       counts, stats);
    return launch_rc("gamma_final_kernel");
}

}  // namespace pbh

using namespace pbh;

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

extern "C" size_t pbh_gamma_workspace_bytes(const PbhGrid* g)
{
    if (!g || g->nx <= 0 || g->ny <= 0 || g->nz <= 0) return 0;
    return align256(GM_BLOCKS * sizeof(GammaPart)) + (GM_BLOCKS + 1) * sizeof(double);
}

// Offsets per axis (gamma.py:47-50).
static int axis_reach(double dta, double d)
{
    const int n = (int)floor(dta / d + 1e-9);
    return n < GM_MAX_VOX ? (n < 0 ? 0 : n) : GM_MAX_VOX;
}

extern "C" int pbh_gamma_count_offsets(const PbhGrid* g, double dta_mm)
{
    if (!g) return PBH_ERR_NULL_ARG;
    if (!(g->dx > 0) || !(g->dy > 0) || !(g->dz > 0) || !(dta_mm > 0)) return PBH_ERR_BAD_DIM;
    const double dta = dta_mm / 10.0;
    const int na = axis_reach(dta, g->dz), nb = axis_reach(dta, g->dy), nc = axis_reach(dta, g->dx);
    int cnt = 0;
    for (int a = -na; a <= na; ++a)
        for (int b = -nb; b <= nb; ++b)
            for (int c = -nc; c <= nc; ++c) {
                const double az = a * g->dz, by = b * g->dy, cx = c * g->dx;
                if (sqrt(az * az + by * by + cx * cx) <= dta + 1e-12) ++cnt;
            }
    return cnt;
}

extern "C" int pbh_gamma_index(const PbhGrid* g, const void* ref, int ref_dtype, const void* eval, int eval_dtype,
                               double dose_pct, double dta_mm, double threshold_pct, double ref_max, int local,
                               double* gamma_out, unsigned long long* counts, double* stats, void* workspace,
                               size_t workspace_bytes, void* stream)
{
    if (!g || !ref || !eval || !counts || !stats || !workspace) return PBH_ERR_NULL_ARG;
    if (g->nx <= 0 || g->ny <= 0 || g->nz <= 0) return PBH_ERR_BAD_DIM;
    if (!(g->dx > 0) || !(g->dy > 0) || !(g->dz > 0)) return PBH_ERR_BAD_DIM;
    if (!(dose_pct > 0) || !(dta_mm > 0) || !(threshold_pct >= 0)) return PBH_ERR_BAD_DIM;
    if ((ref_dtype != PBH_DTYPE_F32 && ref_dtype != PBH_DTYPE_F64) ||
        (eval_dtype != PBH_DTYPE_F32 && eval_dtype != PBH_DTYPE_F64))
        return PBH_ERR_UNSUPPORTED;
    if (workspace_bytes < pbh_gamma_workspace_bytes(g)) return PBH_ERR_WORKSPACE;
    GammaArgs a;
    a.nx = g->nx; a.ny = g->ny; a.nz = g->nz;
    a.dx = g->dx; a.dy = g->dy; a.dz = g->dz;
    a.dta = dta_mm / 10.0;
    a.dose_frac = dose_pct / 100.0;
    a.thr_frac = threshold_pct / 100.0;
    a.ref_max = ref_max > 0.0 ? ref_max : 0.0;
    a.ref_max_dev = nullptr;
    a.local = local != 0;
    a.na = axis_reach(a.dta, g->dz);
    a.nb = axis_reach(a.dta, g->dy);
    a.nc = axis_reach(a.dta, g->dx);
    a.gamma_out = gamma_out;
    a.part = (GammaPart*)workspace;
    double* refmax_part = (double*)((char*)workspace + align256(GM_BLOCKS * sizeof(GammaPart)));
    const int n_off = pbh_gamma_count_offsets(g, dta_mm);
    const size_t n = (size_t)g->nx * g->ny * g->nz;
    hipStream_t s = (hipStream_t)stream;
    const bool r64 = ref_dtype == PBH_DTYPE_F64, e64 = eval_dtype == PBH_DTYPE_F64;
    if (r64 && e64) return launch_gamma<double, double>(ref, eval, a, n, refmax_part, n_off, counts, stats, s);
    if (r64) return launch_gamma<double, float>(ref, eval, a, n, refmax_part, n_off, counts, stats, s);
    if (e64) return launch_gamma<float, double>(ref, eval, a, n, refmax_part, n_off, counts, stats, s);
    return launch_gamma<float, float>(ref, eval, a, n, refmax_part, n_off, counts, stats, s);
}
