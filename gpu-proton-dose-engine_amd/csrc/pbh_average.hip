// pbh_average.hip -- the single-sigma approximation of the forward
// (torch_separable / separable_numpy with per_layer=False, engines.py:184-191,
// :243-254): one Gaussian width per depth slice, sigma_bar(z) = sum_l wl_l
// sigma_l(z) with wl_l the layer's share of the total weight (engines.py:75-80),
// and one amplitude map per slice, A(z) = sum_l IDD_l(z) W_l.
//
// The map is written as an ordinary problem of nz "layers" for the forward
// kernels: layer k has weights A(z_k), IDD 1 at slice k and 0 elsewhere, and
// sigma_bar on every slice, so that slice z receives exactly Gy A(z) Gx^T with
// both factors of width sigma_bar(z). Sums run in FP64 in layer order.
#pragma clang fp contract(off)

#include "pbh_common.h"

namespace pbh {

constexpr int AVG_THREADS = 256;

// One workgroup, one wave per layer: wl[l] = sum of layer l's weights
// (lane-strided partial sums, then a fixed butterfly).
__global__ __launch_bounds__(AVG_THREADS) void layer_sum_kernel(const double* w, int L, int nspot, double* wl)
{
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int l = wave; l < L; l += AVG_THREADS / 64) {
        double s = 0.0;
        for (int e = lane; e < nspot; e += 64) s += w[(size_t)l * nspot + e];
        for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
        if (lane == 0) wl[l] = s;
    }
}

// sigma_bar(z) into row 0 of sigma_out: sum_l sigma_l(z) (wl_l / max(sum wl, 1e-30)).
__global__ void sigma_bar_kernel(const double* sigma, const double* wl, int L, int nz, double* sigma_out)
{
    const int z = blockIdx.x * blockDim.x + threadIdx.x;
    if (z >= nz) return;
    double t = 0.0;
    for (int l = 0; l < L; ++l) t += wl[l];
    t = t > 1e-30 ? t : 1e-30;
    double sb = 0.0;
    for (int l = 0; l < L; ++l) sb += sigma[(size_t)l * nz + z] * (wl[l] / t);
    sigma_out[z] = sb;
}

// Rows 1..nz-1 of sigma_out copy row 0; idd_out is the identity (nz x nz).
__global__ void expand_kernel(int nz, double* idd_out, double* sigma_out)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nz * nz) return;
    const int k = (int)(i / nz), z = (int)(i % nz);
    idd_out[i] = (k == z) ? 1.0 : 0.0;
    if (k > 0) sigma_out[i] = sigma_out[z];
}

// A(z)[e] = sum_l IDD_l(z) w_l[e] in FP64, stored as T (nz x nspot).
template <typename T>
__global__ void amplitude_kernel(const double* w, const double* idd, int L, int nz, int nspot, T* amp)
{
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (size_t)nz * nspot) return;
    const int z = (int)(i / nspot), e = (int)(i % nspot);
    double s = 0.0;
    for (int l = 0; l < L; ++l) s += idd[(size_t)l * nz + z] * w[(size_t)l * nspot + e];
    amp[i] = (T)s;
}

}  // namespace pbh

using namespace pbh;

template <typename T>
static int average_sigma(const double* w, const double* idd, const double* sigma, int n_layers, int nz,
                         int nsy, int nsx, T* w_avg, double* idd_avg, double* sigma_avg,
                         double* layer_sums, void* stream)
{
    if (!w || !idd || !sigma || !w_avg || !idd_avg || !sigma_avg || !layer_sums) return PBH_ERR_NULL_ARG;
    if (n_layers < 1 || n_layers > PBH_MAX_LAYERS || nz < 1 || nz > PBH_MAX_DIM || nsy < 1 || nsx < 1 ||
        nsy > PBH_MAX_SPOTS || nsx > PBH_MAX_SPOTS)
        return PBH_ERR_BAD_DIM;
    const long long nspot = (long long)nsy * nsx;
    if ((long long)nz * nspot > 0x7fffffffLL || (long long)n_layers * nspot > 0x7fffffffLL) return PBH_ERR_UNSUPPORTED;
    hipStream_t s = (hipStream_t)stream;
    hipLaunchKernelGGL(layer_sum_kernel, dim3(1), dim3(AVG_THREADS), 0, s, w, n_layers, (int)nspot, layer_sums);
    int rc = launch_rc("layer_sum_kernel");
    if (rc != PBH_OK) return rc;
    hipLaunchKernelGGL(sigma_bar_kernel, dim3((nz + 255) / 256), dim3(256), 0, s, sigma, layer_sums, n_layers, nz,
                       sigma_avg);
    if ((rc = launch_rc("sigma_bar_kernel")) != PBH_OK) return rc;
    const long long nn = (long long)nz * nz;
    hipLaunchKernelGGL(expand_kernel, dim3((unsigned)((nn + 255) / 256)), dim3(256), 0, s, nz, idd_avg, sigma_avg);
    if ((rc = launch_rc("expand_kernel")) != PBH_OK) return rc;
    const long long n = (long long)nz * nspot;
    hipLaunchKernelGGL(amplitude_kernel<T>, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, w, idd, n_layers, nz,
                       (int)nspot, w_avg);
    return launch_rc("amplitude_kernel");
}

extern "C" int pbh_average_sigma(const double* w, const double* idd, const double* sigma, int n_layers, int nz,
                                 int nsy, int nsx, float* w_avg, double* idd_avg, double* sigma_avg,
                                 double* layer_sums, void* stream)
{
    return average_sigma(w, idd, sigma, n_layers, nz, nsy, nsx, w_avg, idd_avg, sigma_avg, layer_sums, stream);
}

extern "C" int pbh_average_sigma_fp64(const double* w, const double* idd, const double* sigma, int n_layers, int nz,
                                      int nsy, int nsx, double* w_avg, double* idd_avg, double* sigma_avg,
                                      double* layer_sums, void* stream)
{
    return average_sigma(w, idd, sigma, n_layers, nz, nsy, nsx, w_avg, idd_avg, sigma_avg, layer_sums, stream);
}
