// pbh_lut.hip -- FP64 physics look-up tables on the device: CSDA range table,
// per-layer pristine IDD and Fermi-Eyges lateral sigma, and their exact
// derivatives with respect to the layer energies.
//
// Restates physics.py:72-139 (stopping power, range table, interpolation),
// :154-174 and :310-322 (range straggling), :185-286 (scattering power,
// Fermi-Eyges moments) and :328-387 (pristine IDD), in FP64 throughout: the
// moment sums cancel strongly (B = z^2 S0 - 2 z S1 + S2 loses ~8 digits), so
// B is evaluated as the reference's explicit double sum, term by term.
//
// FP contraction is disabled in this file so that every product and sum rounds
// as NumPy's does; transcendental functions (log, exp, pow, hypot) are the
// device library's and may differ from the host's by an ulp.
//
// The IDD's range-spread convolution is a zero-padded real FFT in the
// reference (nfft = 16384 >= 2 n, multiplier exp(-2 (pi f sigma)^2)). By
// Poisson summation that product is exactly the linear convolution with the
// sampled Gaussian h / (sqrt(2 pi) sigma) exp(-(m h)^2 / (2 sigma^2)) up to
// terms below exp(-pi^2 sigma^2 / (2 h^2)) (zero in FP64 for sigma >> h), so
// the kernel sums that convolution directly, only at the fine-grid points the
// final interpolation reads.
//
// A single depth z0 (nz = 1) follows physics.py:247-249: the moment integrals
// run over [0, z0] (or [0, 1e-6] when z0 <= 0) and sigma is read at the end of
// that grid; the IDD is read at z0.
//
// Tangent tables (pbh_build_lut_tangent): every kernel is a template over its
// scalar type. With pbh::Dual (value, derivative) the same kernels give the
// tables (bit-identical to the double instantiation) and their derivatives
// along a direction dE0 of the layer energies: through R0 = R(E0), the Bohr
// integral on linspace(0.5, E0, 4000), the moving fine grid and bin edges of
// the CSDA binning (top = R0 + 8 sigma_R + 0.2), the sampled Gaussian, the
// convolution, the interpolation onto z, and the scattering power inside the
// B double sum. Interpolation brackets, bin membership and window sizes are
// those of the current energies, so the derivative is that of the discrete
// algorithm wherever it is differentiable.
//
// Depth-dependent range straggling (straggling_depth = 1) replaces the single
// sigma_R by a spread per CSDA source bin: the dose binned at depth z' is
// spread with s(z') = hypot(sigma_B(z'), sigma_E), where sigma_B(z')^2 =
// C(E0) - C(E(R0 - z')) and C is the cumulative trapezoid of the Bohr
// integrand on the reference's 4000-point energy grid (physics.py:168-174).
// The reference documents this as its open model defect (README.md:180-184):
// its kernel is 6x too wide at 2 mm depth. Each bin keeps unit weight: the
// sampled Gaussian is normalised by the continuous factor h / (sqrt(2 pi) s)
// when s >= 2h (the discrete sum differs by < 1e-30) and by its explicit
// discrete sum below that; s = 0 leaves the bin in place.
#pragma clang fp contract(off)

#include "pbh_common.h"
#include "pbh_dual.h"

namespace pbh {

// ---- constants (physics.py:42-65, :151) ----
constexpr double MP = 938.27208816;
constexpr double ME = 0.51099895000;
constexpr double KBB = 0.307075;
constexpr double ZA = 10.0 / 18.01528;
constexpr double IMEV = 75.0 * 1e-6;
constexpr double RHO = 1.0;
constexpr double X0W = 36.08;
constexpr double EHL = 13.6;
constexpr double ELO = 0.5, EHI = 400.0;
constexpr double PR = 1.77;
constexpr double BOHR = KBB * ZA * RHO * ME;

constexpr int N_TABLE = 3797;      // 800 + 899 + 899 + 1199 (shared segment ends removed)
constexpr int LUT_THREADS = 256;
constexpr int STRAG_N = 4000;      // Bohr integral grid, linspace(0.5, E0, 4000) (physics.py:168)

// ---- kinematics and stopping power (physics.py:72-100) ----
template <class T> __device__ inline T beta2_of(T E, T* gamma_out)
{
    const T g = 1.0 + E / MP;
    T b2 = 1.0 - 1.0 / (g * g);
    b2 = clamp_hi(clamp_lo(b2, 1e-14), 1.0);
    if (gamma_out) *gamma_out = g;
    return b2;
}

template <class T> __device__ inline T stopping_power(T E)
{
    T g;
    const T b2 = beta2_of(E, &g);
    const double r = ME / MP;
    const T tmax = 2.0 * ME * b2 * (g * g) / (1.0 + 2.0 * g * r + r * r);
    T arg = 2.0 * ME * b2 * (g * g) * tmax / (IMEV * IMEV);
    arg = clamp_lo(arg, 1.0 + 1e-12);
    return KBB * ZA * RHO / b2 * (0.5 * r_log(arg) - b2);
}

// Local Highland scattering power with L = 1 (ln L = 0). physics.py:202-217
template <class T> __device__ inline T scattering_power(T E)
{
    T g;
    const T b2 = beta2_of(E, &g);
    const T beta = r_sqrt(b2);
    T p2 = g * g - 1.0;
    p2 = clamp_lo(p2, 0.0);
    const T pc = r_sqrt(p2) * MP;
    T bp = beta * pc;
    bp = clamp_lo(bp, 1e-9);
    const T q = EHL / bp;
    const T base = q * q / X0W;
    const double lnL = 0.0;
    return base * (1.0 + 0.038 * lnL) * (1.076 + 0.038 * lnL);
}

// ---- NumPy linspace(start, stop, num)[i] ----
template <class T> __device__ inline T lin_at(double start, T stop, int num, int i)
{
    if (i == num - 1) return stop;
    const T step = (stop - start) / (double)(num - 1);
    return (double)i * step + start;
}

// Energy grid of the range table: index -> energy. physics.py:118-123
__device__ inline double table_energy(int k)
{
    if (k < 800) return lin_at(ELO, 2.0, 800, k);
    k -= 800;
    if (k < 899) return lin_at(2.0, 20.0, 900, k + 1);
    k -= 899;
    if (k < 899) return lin_at(20.0, 60.0, 900, k + 1);
    k -= 899;
    return lin_at(60.0, EHI, 1200, k + 1);
}

// NumPy interp for sorted xp with the default end values (fp[0], fp[n-1]).
// Finds j with xp[j] <= x < xp[j+1]; x == xp[j] returns fp[j] exactly. The
// derivative is the bracket's slope inside the table and 0 outside it.
template <class T> __device__ inline T interp_sorted(const double* xp, const double* fp, int n, T xx)
{
    const double x = val(xx);
    if (x > xp[n - 1]) return cst<T>(fp[n - 1]);
    if (x < xp[0]) return cst<T>(fp[0]);
    int lo = 0, hi = n - 1;   // invariant xp[lo] <= x, and x < xp[hi] or hi == n-1
    if (x >= xp[n - 1]) return cst<T>(fp[n - 1]);
    while (hi - lo > 1) {
        const int mid = (lo + hi) >> 1;
        if (xp[mid] <= x) lo = mid; else hi = mid;
    }
    const double slope = (fp[lo + 1] - fp[lo]) / (xp[lo + 1] - xp[lo]);
    if (xp[lo] == x) return make<T>(fp[lo], slope * der(xx));
    return slope * (xx - xp[lo]) + fp[lo];
}

template <class T> __device__ inline T csda_range(const double* E, const double* R, T e)
{
    return interp_sorted(E, R, N_TABLE, e);
}

template <class T> __device__ inline T energy_from_range(const double* E, const double* R, T r)
{
    r = clamp_hi(clamp_lo(r, 0.0), R[N_TABLE - 1]);
    return interp_sorted(R, E, N_TABLE, r);
}

// Block-wide sum (LUT_THREADS threads).
template <class T> __device__ inline T block_sum(T v, T* red)
{
    for (int o = 32; o > 0; o >>= 1) v += shfl_down(v, o);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) red[wv] = v;
    __syncthreads();
    T t = cst<T>(0.0);
    if (threadIdx.x == 0) {
        for (int i = 0; i < LUT_THREADS / 64; ++i) t += red[i];
        red[0] = t;
    }
    __syncthreads();
    t = red[0];
    __syncthreads();
    return t;
}

template <class T> __device__ inline T wave_sum(T v)
{
    for (int o = 32; o > 0; o >>= 1) v += shfl_xor(v, o);
    return v;
}

// ---- range table: E grid, S, cumulative trapezoid of 1/S ----
__global__ __launch_bounds__(LUT_THREADS) void range_table_kernel(double* E, double* R)
{
    __shared__ double seg[N_TABLE];
    for (int k = threadIdx.x; k < N_TABLE; k += LUT_THREADS) E[k] = table_energy(k);
    __syncthreads();
    for (int k = threadIdx.x; k < N_TABLE - 1; k += LUT_THREADS) {
        const double e0 = table_energy(k), e1 = table_energy(k + 1);
        const double i0 = 1.0 / stopping_power(e0), i1 = 1.0 / stopping_power(e1);
        seg[k] = 0.5 * (i1 + i0) * (e1 - e0);
    }
    __syncthreads();
    if (threadIdx.x == 0) {   // sequential, as np.cumsum
        // segments read 16 at a time ahead of the dependent additions (same order);
        // the chain of FP64 additions, not the reads, sets this loop's time
        constexpr int U = 16;
        double acc = 0.0;
        R[0] = 0.0;
        int k = 0;
        for (; k + U <= N_TABLE - 1; k += U) {
            double v[U];
#pragma unroll
            for (int u = 0; u < U; ++u) v[u] = seg[k + u];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                acc += v[u];
                R[k + u + 1] = acc;
            }
        }
        for (; k < N_TABLE - 1; ++k) { acc += seg[k]; R[k + 1] = acc; }
    }
}

// out = interp(x); dout (if given) = its derivative in x (the bracket's slope, 0 outside).
__global__ void interp_kernel(const double* xp, const double* fp, int nt, const double* x,
                              double* out, double* dout, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    if (dout) {
        const Dual r = interp_sorted(xp, fp, nt, Dual{x[i], 1.0});
        out[i] = r.v;
        dout[i] = r.d;
    } else {
        out[i] = interp_sorted(xp, fp, nt, x[i]);
    }
}

// ---- per-layer arguments; T = double (tables) or Dual (tables + derivatives) ----
template <class T> struct LayerLutArgs {
    const double* E;      // range table energies
    const double* R;      // range table ranges
    const double* energies;
    const double* denergies;   // direction of the derivative (Dual only)
    int L, nz;
    double z0, dz;
    double fe_lo, fe_hi;  // Fermi-Eyges depth grid [z[0], z[-1]] ([0, z0] for one depth)
    int fe_single;        // 1: one output depth, read at fe_hi (physics.py:247-249)
    PbhBeamModel m;
    int n_fine;           // IDD fine grid
    int nf;               // Fermi-Eyges fine grid
    T* R0;                // [L]
    T* sR;                // [L]
    T* dcsda;             // [L][n_fine]
    T* gk;                // [L][n_fine]
    T* tsc;               // [L][nf] scattering power on the fine grid
    double* hf;           // [nf]
    double* idd;          // [L][nz]
    double* sigma;        // [L][nz]
    double* didd;         // [L][nz] derivatives (Dual only)
    double* dsigma;       // [L][nz]
    T* cum;               // [L][STRAG_N] cumulative Bohr integral (depth-dependent mode)
    T* amp;               // [L][n_fine] per-source Gaussian amplitude (depth-dependent mode;
                          // gk then holds 1 / (2 s^2), or -1 for s = 0)
};

template <class T> __device__ inline T load_E0(const LayerLutArgs<T>& a, int l)
{
    return make<T>(a.energies[l], is_dual<T> ? a.denergies[l] : 0.0);
}

template <class T> __device__ inline void store_out(double* v, double* d, size_t i, T x)
{
    v[i] = val(x);
    if constexpr (is_dual<T>) d[i] = x.d;
}

template <class T> __device__ inline T bohr_integrand(T e)
{
    T b = beta2_of(e, (T*)nullptr);
    b = clamp_lo(b, 1e-12);
    const T S = stopping_power(e);
    return BOHR / (b * (S * S * S));
}

template <class T>
__global__ __launch_bounds__(LUT_THREADS) void layer_scalars_kernel(LayerLutArgs<T> a)
{
    __shared__ T red[LUT_THREADS / 64];
    const int l = blockIdx.x;
    const T E0 = load_E0(a, l);
    const T R0 = csda_range(a.E, a.R, E0);
    T s_strag = cst<T>(0.0);
    if (a.m.straggling == 0) {
        // sigma^2 = trapezoid of BOHR / (beta^2 S^3) on linspace(0.5, E0, 4000). physics.py:168-174
        constexpr int n = 4000;
        T part = cst<T>(0.0);
        const T last = lin_at(ELO, E0, n, n - 1), first = lin_at(ELO, E0, n, 0);
        if (val(last) > val(first)) {
            // a contiguous run of cells per thread: each integrand value is evaluated
            // once and serves the two cells it bounds
            constexpr int PER = (n - 1 + LUT_THREADS - 1) / LUT_THREADS;
            const int c0 = threadIdx.x * PER, c1 = c0 + PER < n - 1 ? c0 + PER : n - 1;
            if (c0 < c1) {
                T e0 = lin_at(ELO, E0, n, c0), f0 = bohr_integrand(e0);
                for (int i = c0; i < c1; ++i) {
                    const T e1 = lin_at(ELO, E0, n, i + 1), f1 = bohr_integrand(e1);
                    part += (e1 - e0) * (f1 + f0) / 2.0;
                    e0 = e1;
                    f0 = f1;
                }
            }
        }
        s_strag = r_sqrt(block_sum(part, red));
    } else if (a.m.straggling == 1) {
        s_strag = a.m.icru_coef * r_pow(R0, a.m.icru_exp);
    }
    if (threadIdx.x == 0) {
        const T s_espr = PR * a.m.energy_spread * R0;
        a.R0[l] = R0;
        a.sR[l] = r_hypot(s_strag, s_espr);
    }
}

// ---- depth-dependent straggling, step 1: cumulative trapezoid of the Bohr
// integrand, C[0] = 0, C[k+1] = C[k] + (e_{k+1} - e_k)(f_{k+1} + f_k) / 2.
// Each thread sums a contiguous run of cells; the run totals are scanned in
// order by one thread, so the result is deterministic. ----
template <class T>
__global__ __launch_bounds__(LUT_THREADS) void strag_cum_kernel(LayerLutArgs<T> a)
{
    constexpr int PER = (STRAG_N - 1 + LUT_THREADS - 1) / LUT_THREADS;
    __shared__ T run_tot[LUT_THREADS];
    const int l = blockIdx.x;
    const T E0 = load_E0(a, l);
    T* C = a.cum + (size_t)l * STRAG_N;
    const bool live = val(lin_at(ELO, E0, STRAG_N, STRAG_N - 1)) > val(lin_at(ELO, E0, STRAG_N, 0));
    const int c0 = threadIdx.x * PER;
    T loc[PER];
    T run = cst<T>(0.0);
    T e_prev = lin_at(ELO, E0, STRAG_N, c0 < STRAG_N ? c0 : STRAG_N - 1);
    T f_prev = live ? bohr_integrand(e_prev) : cst<T>(0.0);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int k = c0 + q;
        if (live && k < STRAG_N - 1) {
            const T e1 = lin_at(ELO, E0, STRAG_N, k + 1);
            const T f1 = bohr_integrand(e1);
            run += (e1 - e_prev) * (f1 + f_prev) / 2.0;
            e_prev = e1;
            f_prev = f1;
        }
        loc[q] = run;
    }
    run_tot[threadIdx.x] = run;
    __syncthreads();
    if (threadIdx.x == 0) {
        T acc = cst<T>(0.0);
        for (int t = 0; t < LUT_THREADS; ++t) {
            const T v = run_tot[t];
            run_tot[t] = acc;
            acc += v;
        }
    }
    __syncthreads();
    const T off = run_tot[threadIdx.x];
    if (threadIdx.x == 0) C[0] = cst<T>(0.0);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
        const int k = c0 + q;
        if (k < STRAG_N - 1) C[k + 1] = off + loc[q];
    }
}

// Spread of the CSDA bin at depth zf (depth-dependent mode): hypot of the Bohr
// straggling accumulated up to zf and the energy-spread term.
template <class T>
__device__ inline T source_spread(const LayerLutArgs<T>& a, const T* C, T E0, T R0, T zf, T s_espr)
{
    T s2 = cst<T>(0.0);
    if (a.m.straggling == 0 && val(lin_at(ELO, E0, STRAG_N, STRAG_N - 1)) > val(lin_at(ELO, E0, STRAG_N, 0))) {
        T er = val(zf) < val(R0) ? energy_from_range(a.E, a.R, R0 - zf) : cst<T>(ELO);
        er = val(er) < ELO ? cst<T>(ELO) : (val(er) > val(E0) ? E0 : er);
        // np.interp(er, linspace(0.5, E0, 4000), C)
        int k = (int)((val(er) - ELO) / (val(E0) - ELO) * (double)(STRAG_N - 1));
        k = k < 0 ? 0 : (k > STRAG_N - 2 ? STRAG_N - 2 : k);
        while (k > 0 && val(lin_at(ELO, E0, STRAG_N, k)) > val(er)) --k;
        while (k < STRAG_N - 2 && val(lin_at(ELO, E0, STRAG_N, k + 1)) <= val(er)) ++k;
        const T ek = lin_at(ELO, E0, STRAG_N, k), ek1 = lin_at(ELO, E0, STRAG_N, k + 1);
        const T ce = (C[k + 1] - C[k]) / (ek1 - ek) * (er - ek) + C[k];
        s2 = C[STRAG_N - 1] - ce;
        s2 = clamp_lo(s2, 0.0);
    }
    return r_hypot(r_sqrt(s2), s_espr);
}

// ---- IDD, step 1: energy-conserving CSDA binning and the Gaussian kernel ----
template <class T>
__global__ __launch_bounds__(LUT_THREADS) void idd_prep_kernel(LayerLutArgs<T> a)
{
    const int l = blockIdx.x;
    const int n = a.n_fine;
    const T R0 = a.R0[l], sR = a.sR[l];
    const T top = R0 + 8.0 * sR + 0.2;
    const T h = lin_at(0.0, top, n, 1) - lin_at(0.0, top, n, 0);
    T* dc = a.dcsda + (size_t)l * n;
    T* gk = a.gk + (size_t)l * n;
    for (int i = blockIdx.y * LUT_THREADS + threadIdx.x; i < n; i += gridDim.y * LUT_THREADS) {
        // bin edges zb = linspace(0, top, n + 1); Eb = E(R0 - zb) inside the range, else 0
        const T zb0 = lin_at(0.0, top, n + 1, i), zb1 = lin_at(0.0, top, n + 1, i + 1);
        const T eb0 = val(zb0) < val(R0) ? energy_from_range(a.E, a.R, clamp_lo(R0 - zb0, 0.0)) : cst<T>(0.0);
        const T eb1 = val(zb1) < val(R0) ? energy_from_range(a.E, a.R, clamp_lo(R0 - zb1, 0.0)) : cst<T>(0.0);
        const T zf = lin_at(0.0, top, n, i);
        const T flu = r_exp(-zf / a.m.attenuation_length);
        dc[i] = flu * (eb0 - eb1) / h;
        if (a.m.straggling_depth) {
            const T sp = source_spread(a, a.cum + (size_t)l * STRAG_N, load_E0(a, l), R0, zf,
                                       PR * a.m.energy_spread * R0);
            T amp = cst<T>(1.0), c = cst<T>(-1.0);
            if (val(sp) > 0.0) {
                c = 1.0 / (2.0 * sp * sp);
                if (val(sp) >= 2.0 * val(h)) {
                    amp = h / (sqrt(2.0 * kPi) * sp);
                } else {
                    const int M = (int)ceil(12.0 * val(sp) / val(h));
                    T norm = cst<T>(1.0);
                    for (int m = 1; m <= M; ++m) {
                        const T t = (double)m * h;
                        norm += 2.0 * r_exp(-(t * t) * c);
                    }
                    amp = 1.0 / norm;
                }
            }
            a.amp[(size_t)l * n + i] = amp;
            gk[i] = c;
            continue;
        }
        T g;
        if (val(sR) > 0.0) {
            const T t = (double)i * h;
            g = h / (sqrt(2.0 * kPi) * sR) * r_exp(-(t * t) / (2.0 * sR * sR));
        } else {
            g = cst<T>((i == 0) ? 1.0 : 0.0);
        }
        gk[i] = g;
    }
}

// Convolved fine-grid dose at index j: sum_i dcsda[i] g[|j - i|] (one wave).
// Terms with |j - i| > M (M covering 10 sR: the Gaussian below e^-50 of its peak)
// are dropped. The loop starts at a multiple of 64, so every lane keeps its residue
// class and the kept terms are added in the same order as over the whole range.
template <class T> __device__ inline T conv_at(const T* dc, const T* gk, int n, int j, int M)
{
    const int lo = (j - M > 0 ? j - M : 0) & ~63, hi = j + M + 1 < n ? j + M + 1 : n;
    T s = cst<T>(0.0);
    for (int i = lo + (threadIdx.x & 63); i < hi; i += 64) {
        const int m = j > i ? j - i : i - j;
        s += dc[i] * gk[m];
    }
    return wave_sum(s);
}

// Depth-dependent mode: sum_i dcsda[i] amp[i] exp(-((j - i) h)^2 c[i]) over the
// sources within W bins of j (W covers 12 of the largest spreads). One wave.
template <class T>
__device__ inline T conv_depth_at(const T* dc, const T* amp, const T* cc, int n, int j, int W, T h)
{
    const int i0 = j - W > 0 ? j - W : 0;
    const int i1 = j + W + 1 < n ? j + W + 1 : n;
    T s = cst<T>(0.0);
    for (int i = i0 + (threadIdx.x & 63); i < i1; i += 64) {
        const T c = cc[i];
        T w;
        if (val(c) < 0.0) {
            w = cst<T>((i == j) ? 1.0 : 0.0);
        } else {
            const T t = (double)(j - i) * h;
            w = amp[i] * r_exp(-(t * t) * c);
        }
        s += dc[i] * w;
    }
    return wave_sum(s);
}

constexpr int Z_PER_BLOCK = 4;        // one output depth per wave

// ---- IDD, step 2: evaluate the convolution at the bracketing fine points and
// interpolate onto z (left = right = 0). One wave per depth. ----
template <class T>
__global__ __launch_bounds__(LUT_THREADS) void idd_eval_kernel(LayerLutArgs<T> a)
{
    const int l = blockIdx.x;
    const int n = a.n_fine;
    const T R0 = a.R0[l], sR = a.sR[l];
    const T top = R0 + 8.0 * sR + 0.2;
    const T h = lin_at(0.0, top, n, 1) - lin_at(0.0, top, n, 0);
    const T* dc = a.dcsda + (size_t)l * n;
    const T* gk = a.gk + (size_t)l * n;
    const bool depth = a.m.straggling_depth != 0;
    const T* amp = depth ? a.amp + (size_t)l * n : nullptr;
    const int W = (int)ceil(12.0 * val(sR) / val(h)) + 2;
    const int M = val(sR) > 0.0 ? (int)ceil(10.0 * val(sR) / val(h)) + 1 : 0;
    auto conv = [&](int j) { return depth ? conv_depth_at(dc, amp, gk, n, j, W, h) : conv_at(dc, gk, n, j, M); };
    const int wv = threadIdx.x >> 6;
    for (int q = wv; q < Z_PER_BLOCK; q += LUT_THREADS / 64) {
        const int k = blockIdx.y * Z_PER_BLOCK + q;
        if (k >= a.nz) break;
        const double z = a.z0 + (double)k * a.dz;
        T out;
        if (z > val(top) || z < 0.0) {
            out = cst<T>(0.0);
        } else {
            int j = (int)(z / val(h));
            if (j > n - 1) j = n - 1;
            while (j > 0 && val(lin_at(0.0, top, n, j)) > z) --j;
            while (j < n - 1 && val(lin_at(0.0, top, n, j + 1)) <= z) ++j;
            const T zj = lin_at(0.0, top, n, j);
            // with derivatives the bracket formula also at zj == z: its value is d0
            // exactly, and its derivative carries the motion of the fine point
            if (j == n - 1 || (!is_dual<T> && val(zj) == z)) {
                out = conv(j);
            } else {
                const T d0 = conv(j), d1 = conv(j + 1);
                const T slope = (d1 - d0) / (lin_at(0.0, top, n, j + 1) - zj);
                out = slope * (z - zj) + d0;
            }
        }
        if ((threadIdx.x & 63) == 0) store_out(a.idd, a.didd, (size_t)l * a.nz + k, out);
    }
}

// ---- Fermi-Eyges, step 1: scattering power and np.gradient weights on the
// fine grid zf = linspace(z[0], z[-1], nf). physics.py:255-264 ----
template <class T>
__global__ __launch_bounds__(LUT_THREADS) void fe_prep_kernel(LayerLutArgs<T> a)
{
    const int l = blockIdx.x;
    const int nf = a.nf;
    const double zlo = a.fe_lo, zhi = a.fe_hi;
    const T R0 = a.R0[l];
    for (int i = threadIdx.x; i < nf; i += LUT_THREADS) {
        const double zf = lin_at(zlo, zhi, nf, i);
        T t = cst<T>(0.0);
        if (zf < val(R0)) {
            const T rem = clamp_lo(R0 - zf, 0.0);
            T e = energy_from_range(a.E, a.R, rem);
            e = clamp_lo(e, 1e-3);
            t = scattering_power(e);
        }
        a.tsc[(size_t)l * nf + i] = t;
        if (l == 0) {
            double hv;
            if (i == 0) hv = lin_at(zlo, zhi, nf, 1) - lin_at(zlo, zhi, nf, 0);
            else if (i == nf - 1) hv = lin_at(zlo, zhi, nf, nf - 1) - lin_at(zlo, zhi, nf, nf - 2);
            else hv = (lin_at(zlo, zhi, nf, i + 1) - lin_at(zlo, zhi, nf, i - 1)) / 2.0;
            a.hf[i] = hv;
        }
    }
}

// B at fine index j: sum over i with zf_j - zf_i > 0 of (dz^2 T_i) h_i (one wave).
template <class T>
__device__ inline T bsum_at(const T* ts, const double* hf, double zlo, double zhi, int nf, int j)
{
    const double zj = lin_at(zlo, zhi, nf, j);
    T s = cst<T>(0.0);
    for (int i = (threadIdx.x & 63); i < nf; i += 64) {
        const double d = zj - lin_at(zlo, zhi, nf, i);
        if (d > 0.0) s += d * d * ts[i] * hf[i];
    }
    return wave_sum(s);
}

// ---- Fermi-Eyges, step 2: B interpolated onto z, sigma = sqrt(var). ----
template <class T>
__global__ __launch_bounds__(LUT_THREADS) void fe_eval_kernel(LayerLutArgs<T> a)
{
    const int l = blockIdx.x;
    const int nf = a.nf;
    const double zlo = a.fe_lo, zhi = a.fe_hi;
    const T* ts = a.tsc + (size_t)l * nf;
    const int wv = threadIdx.x >> 6;
    for (int q = wv; q < Z_PER_BLOCK; q += LUT_THREADS / 64) {
        const int k = blockIdx.y * Z_PER_BLOCK + q;
        if (k >= a.nz) break;
        const double z = a.fe_single ? a.fe_hi : a.z0 + (double)k * a.dz;
        T B;
        if (z >= zhi) {
            B = bsum_at(ts, a.hf, zlo, zhi, nf, nf - 1);
        } else if (z <= zlo) {
            B = bsum_at(ts, a.hf, zlo, zhi, nf, 0);
        } else {
            const double step = (zhi - zlo) / (double)(nf - 1);
            int j = (int)((z - zlo) / step);
            if (j > nf - 1) j = nf - 1;
            while (j > 0 && lin_at(zlo, zhi, nf, j) > z) --j;
            while (j < nf - 1 && lin_at(zlo, zhi, nf, j + 1) <= z) ++j;
            const double zj = lin_at(zlo, zhi, nf, j);
            if (j == nf - 1 || zj == z) {
                B = bsum_at(ts, a.hf, zlo, zhi, nf, j);
            } else {
                const T b0 = bsum_at(ts, a.hf, zlo, zhi, nf, j);
                const T b1 = bsum_at(ts, a.hf, zlo, zhi, nf, j + 1);
                const T slope = (b1 - b0) / (lin_at(zlo, zhi, nf, j + 1) - zj);
                B = slope * (z - zj) + b0;
            }
        }
        const double sx = a.m.sigma_x0, st = a.m.sigma_theta0, rho = a.m.rho0;
        T var = sx * sx + 2.0 * rho * sx * st * z + st * st * (z * z) + B;
        var = clamp_lo(var, 0.0);
        if ((threadIdx.x & 63) == 0) store_out(a.sigma, a.dsigma, (size_t)l * a.nz + k, r_sqrt(var));
    }
}

// Per-layer summary: R0, sigma_R, depth and value of the first IDD maximum (np.argmax).
template <class T>
__global__ __launch_bounds__(64) void layer_info_kernel(const double* idd, const T* R0, const T* sR,
                                                        int L, int nz, double z0, double dz, double* info)
{
    // one wave per layer: np.argmax (first maximum) over lane-strided depths, then
    // across lanes, the lower index winning ties
    const int l = blockIdx.x, lane = threadIdx.x;
    if (l >= L) return;
    const double* row = idd + (size_t)l * nz;
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
    if (lane == 0) {
        info[4 * l + 0] = val(R0[l]);
        info[4 * l + 1] = val(sR[l]);
        info[4 * l + 2] = z0 + (double)bk * dz;
        info[4 * l + 3] = bv;
    }
}

// Depth grid of the moment integrals: [z[0], z[-1]], or [0, z0] (z0 > 0) and
// [0, 1e-6] for a single depth (physics.py:247-249).
struct FeGrid {
    double lo, hi;
    int n_user;   // points of the grid the reference builds (2 for a single depth)
    int single;
};

static FeGrid fe_grid(int nz, double z0, double dz)
{
    FeGrid f;
    if (nz == 1) {
        f.lo = 0.0;
        f.hi = z0 > 0.0 ? z0 : 1e-6;
        f.n_user = 2;
        f.single = 1;
    } else {
        f.lo = z0;
        f.hi = z0 + (double)(nz - 1) * dz;
        f.n_user = nz;
        f.single = 0;
    }
    return f;
}

// Fine-grid size of the moment integrals (physics.py:256-258), evaluated on the
// host exactly as the reference does: span = z[-1] - z[0].
static int fe_fine_count(const FeGrid& f, int n_sub, double max_step)
{
    const double span = f.hi - f.lo;
    const double ms = max_step > 1e-6 ? max_step : 1e-6;
    const int need = span > 0 ? (int)ceil(span / ms) + 1 : 2;
    int nf = (f.n_user - 1) * n_sub + 1;
    if (need > nf) nf = need;
    if (nf < 2) nf = 2;
    return nf;
}

struct LutLayout {
    size_t table, scal, dcsda, gk, T, hf, cum, amp, total;
    int nf;
};

// Workspace of one build; `real` = bytes per table entry (8 for the tables,
// 16 with derivatives).
static LutLayout lut_layout(const PbhBeamModel* m, int nz, double z0, double dz, int L, size_t real)
{
    LutLayout o;
    o.nf = fe_fine_count(fe_grid(nz, z0, dz), m->fe_n_sub, m->fe_max_step);
    size_t off = 0;
    auto take = [&](size_t bytes) { size_t p = off; off += (bytes + 255) & ~(size_t)255; return p; };
    o.table = take(2 * (size_t)N_TABLE * 8);
    o.scal = take(2 * (size_t)L * real);
    o.dcsda = take((size_t)L * m->n_fine_idd * real);
    o.gk = take((size_t)L * m->n_fine_idd * real);
    o.T = take((size_t)L * o.nf * real);
    o.hf = take((size_t)o.nf * 8);
    const bool depth = m->straggling_depth != 0;
    o.cum = take(depth ? (size_t)L * STRAG_N * real : 0);
    o.amp = take(depth ? (size_t)L * m->n_fine_idd * real : 0);
    o.total = off;
    return o;
}

static int check_model(const PbhBeamModel* m)
{
    if (!m) return PBH_ERR_NULL_ARG;
    if (m->n_fine_idd < 2 || m->n_fine_idd > (1 << 20)) return PBH_ERR_BAD_DIM;
    if (m->fe_n_sub < 1 || !(m->fe_max_step > 0)) return PBH_ERR_BAD_DIM;
    if (m->straggling < 0 || m->straggling > 2) return PBH_ERR_UNSUPPORTED;
    if (m->straggling_depth < 0 || m->straggling_depth > 1) return PBH_ERR_UNSUPPORTED;
    if (m->straggling_depth && m->straggling == 1) return PBH_ERR_UNSUPPORTED;
    if (!(m->attenuation_length > 0)) return PBH_ERR_BAD_DIM;
    return PBH_OK;
}

static int check_grid(int nz, double z0, double dz, int n_layers)
{
    if (nz < 1 || nz > PBH_MAX_DIM || n_layers < 1 || n_layers > (1 << 20)) return PBH_ERR_BAD_DIM;
    if (!(dz > 0) || !(z0 == z0)) return PBH_ERR_BAD_DIM;
    return PBH_OK;
}

// Size in bytes of the workspace, or 0 for invalid input.
static size_t lut_bytes(const PbhBeamModel* m, int nz, double z0, double dz, int n_layers, size_t real)
{
    if (check_model(m) != PBH_OK || check_grid(nz, z0, dz, n_layers) != PBH_OK) return 0;
    const LutLayout lay = lut_layout(m, nz, z0, dz, n_layers, real);
    if (lay.nf > (1 << 22)) return 0;
    return lay.total;
}

// The whole build: [range table] -> layer scalars -> [Bohr cumulative] -> IDD
// binning -> IDD evaluation -> scattering power -> sigma -> [layer summary].
template <class T>
static int build_lut(const PbhBeamModel* m, int nz, double z0, double dz, int n_layers,
                     const double* energies, const double* denergies,
                     const double* range_E, const double* range_R,
                     double* idd, double* sigma, double* didd, double* dsigma, double* layer_info,
                     void* workspace, size_t workspace_bytes, hipStream_t s)
{
    int rc = check_model(m);
    if (rc != PBH_OK) return rc;
    if (!energies || !idd || !sigma || !workspace) return PBH_ERR_NULL_ARG;
    if (is_dual<T> && (!denergies || !didd || !dsigma)) return PBH_ERR_NULL_ARG;
    if ((range_E == nullptr) != (range_R == nullptr)) return PBH_ERR_NULL_ARG;
    if ((rc = check_grid(nz, z0, dz, n_layers)) != PBH_OK) return rc;
    const LutLayout lay = lut_layout(m, nz, z0, dz, n_layers, sizeof(T));
    if (lay.nf > (1 << 22)) return PBH_ERR_UNSUPPORTED;
    if (workspace_bytes < lay.total) return PBH_ERR_WORKSPACE;
    char* ws = (char*)workspace;
    const FeGrid fg = fe_grid(nz, z0, dz);
    LayerLutArgs<T> a;
    if (range_E) {
        a.E = range_E;
        a.R = range_R;
    } else {
        double* tE = (double*)(ws + lay.table);
        hipLaunchKernelGGL(range_table_kernel, dim3(1), dim3(LUT_THREADS), 0, s, tE, tE + N_TABLE);
        if ((rc = launch_rc("range_table_kernel")) != PBH_OK) return rc;
        a.E = tE;
        a.R = tE + N_TABLE;
    }
    a.energies = energies;
    a.denergies = denergies;
    a.L = n_layers; a.nz = nz; a.z0 = z0; a.dz = dz;
    a.fe_lo = fg.lo; a.fe_hi = fg.hi; a.fe_single = fg.single;
    a.m = *m;
    a.n_fine = m->n_fine_idd;
    a.nf = lay.nf;
    a.R0 = (T*)(ws + lay.scal);
    a.sR = a.R0 + n_layers;
    a.dcsda = (T*)(ws + lay.dcsda);
    a.gk = (T*)(ws + lay.gk);
    a.tsc = (T*)(ws + lay.T);
    a.hf = (double*)(ws + lay.hf);
    a.idd = idd; a.sigma = sigma;
    a.didd = didd; a.dsigma = dsigma;
    a.cum = (T*)(ws + lay.cum);
    a.amp = (T*)(ws + lay.amp);

    hipLaunchKernelGGL(layer_scalars_kernel<T>, dim3(n_layers), dim3(LUT_THREADS), 0, s, a);
    if ((rc = launch_rc("layer_scalars_kernel")) != PBH_OK) return rc;
    if (m->straggling_depth && m->straggling == 0) {
        hipLaunchKernelGGL(strag_cum_kernel<T>, dim3(n_layers), dim3(LUT_THREADS), 0, s, a);
        if ((rc = launch_rc("strag_cum_kernel")) != PBH_OK) return rc;
    }
    // one workgroup per layer and 256 fine bins: every bin is independent
    hipLaunchKernelGGL(idd_prep_kernel<T>, dim3(n_layers, cdiv(a.n_fine, LUT_THREADS)), dim3(LUT_THREADS), 0, s, a);
    if ((rc = launch_rc("idd_prep_kernel")) != PBH_OK) return rc;
    const dim3 eval_grid(n_layers, cdiv(nz, Z_PER_BLOCK));   // layers in x: up to 2^20 of them
    hipLaunchKernelGGL(idd_eval_kernel<T>, eval_grid, dim3(LUT_THREADS), 0, s, a);
    if ((rc = launch_rc("idd_eval_kernel")) != PBH_OK) return rc;
    hipLaunchKernelGGL(fe_prep_kernel<T>, dim3(n_layers), dim3(LUT_THREADS), 0, s, a);
    if ((rc = launch_rc("fe_prep_kernel")) != PBH_OK) return rc;
    hipLaunchKernelGGL(fe_eval_kernel<T>, eval_grid, dim3(LUT_THREADS), 0, s, a);
    if ((rc = launch_rc("fe_eval_kernel")) != PBH_OK) return rc;
    if (layer_info) {
        hipLaunchKernelGGL(layer_info_kernel<T>, dim3(n_layers), dim3(64), 0, s,
                           (const double*)idd, (const T*)a.R0, (const T*)a.sR,
                           n_layers, nz, z0, dz, layer_info);
        if ((rc = launch_rc("layer_info_kernel")) != PBH_OK) return rc;
    }
    return PBH_OK;
}

}  // namespace pbh

using namespace pbh;

extern "C" int pbh_range_table_size(void) { return N_TABLE; }

extern "C" int pbh_range_table(double* E, double* R, void* stream)
{
    if (!E || !R) return PBH_ERR_NULL_ARG;
    hipLaunchKernelGGL(range_table_kernel, dim3(1), dim3(LUT_THREADS), 0, (hipStream_t)stream, E, R);
    return launch_rc("range_table_kernel");
}

extern "C" int pbh_interp_table(const double* xp, const double* fp, int n_table,
                                const double* x, double* out, int n, void* stream)
{
    if (!xp || !fp || !x || !out) return PBH_ERR_NULL_ARG;
    if (n_table < 2 || n < 0) return PBH_ERR_BAD_DIM;
    if (n == 0) return PBH_OK;
    hipLaunchKernelGGL(interp_kernel, dim3(cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream,
                       xp, fp, n_table, x, out, (double*)nullptr, n);
    return launch_rc("interp_kernel");
}

extern "C" int pbh_interp_table_slope(const double* xp, const double* fp, int n_table,
                                      const double* x, double* out, double* dout, int n, void* stream)
{
    if (!xp || !fp || !x || !out || !dout) return PBH_ERR_NULL_ARG;
    if (n_table < 2 || n < 0) return PBH_ERR_BAD_DIM;
    if (n == 0) return PBH_OK;
    hipLaunchKernelGGL(interp_kernel, dim3(cdiv(n, 256)), dim3(256), 0, (hipStream_t)stream,
                       xp, fp, n_table, x, out, dout, n);
    return launch_rc("interp_kernel");
}

extern "C" size_t pbh_lut_workspace_bytes(const PbhBeamModel* m, int nz, double z0, double dz, int n_layers)
{
    return lut_bytes(m, nz, z0, dz, n_layers, sizeof(double));
}

extern "C" size_t pbh_lut_tangent_workspace_bytes(const PbhBeamModel* m, int nz, double z0, double dz,
                                                  int n_layers)
{
    return lut_bytes(m, nz, z0, dz, n_layers, sizeof(Dual));
}

extern "C" int pbh_build_lut(const PbhBeamModel* m, int nz, double z0, double dz, int n_layers,
                             const double* energies, const double* range_E, const double* range_R,
                             double* idd, double* sigma, double* layer_info,
                             void* workspace, size_t workspace_bytes, void* stream)
{
    return build_lut<double>(m, nz, z0, dz, n_layers, energies, nullptr, range_E, range_R, idd, sigma,
                             nullptr, nullptr, layer_info, workspace, workspace_bytes, (hipStream_t)stream);
}

extern "C" int pbh_build_lut_tangent(const PbhBeamModel* m, int nz, double z0, double dz, int n_layers,
                                     const double* energies, const double* denergies,
                                     const double* range_E, const double* range_R,
                                     double* idd, double* sigma, double* didd, double* dsigma,
                                     double* layer_info, void* workspace, size_t workspace_bytes,
                                     void* stream)
{
    return build_lut<Dual>(m, nz, z0, dz, n_layers, energies, denergies, range_E, range_R, idd, sigma,
                           didd, dsigma, layer_info, workspace, workspace_bytes, (hipStream_t)stream);
}
