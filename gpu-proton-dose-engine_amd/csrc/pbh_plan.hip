// pbh_plan.hip -- spot-weight update of the inverse-planning loop.
//
// One projected-gradient step of one field from the dJ/dw of its scenarios:
//   g      = sum_s grads[s]  (summed in scenario order)     mode PBH_ROBUST_SUM
//          = grads[s*], s* = first argmax_s J[s]             mode PBH_ROBUST_WORST
//   w_new  = min(max(w_old - step * g, w_min), w_max)
// Elementwise (HBM-bound); FP32 with contraction off, so every element rounds
// exactly as "t = step * g; w_old - t" in FP32 does.
#include "pbh_common.h"

#pragma clang fp contract(off)

namespace pbh {

__global__ __launch_bounds__(256) void weight_step_kernel(const float* w_old, float* w_new,
                                                          const float* const* grads, const double* J,
                                                          int n_scen, size_t n, int mode, float step,
                                                          float w_min, float w_max)
{
    int pick = 0;
    if (mode == PBH_ROBUST_WORST) {
        double best = J[0];
        for (int k = 1; k < n_scen; ++k)
            if (J[k] > best) {
                best = J[k];
                pick = k;
            }
    }
    for (size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x) {
        float g;
        if (mode == PBH_ROBUST_WORST) {
            g = grads[pick][e];
        } else {
            g = grads[0][e];
            for (int k = 1; k < n_scen; ++k) g = g + grads[k][e];
        }
        const float t = step * g;
        float v = w_old[e] - t;
        v = v < w_min ? w_min : v;
        v = v > w_max ? w_max : v;
        w_new[e] = v;
    }
}

}  // namespace pbh

using namespace pbh;

extern "C" int pbh_weight_step(const float* w_old, float* w_new, const float* const* grads, const double* J,
                               int n_scen, size_t n, int mode, float step, float w_min, float w_max, void* stream)
{
    if (!w_old || !w_new || !grads || (mode == PBH_ROBUST_WORST && !J)) return PBH_ERR_NULL_ARG;
    if (n_scen <= 0 || n == 0) return PBH_ERR_BAD_DIM;
    if (mode != PBH_ROBUST_SUM && mode != PBH_ROBUST_WORST) return PBH_ERR_UNSUPPORTED;
    if (!(w_min <= w_max)) return PBH_ERR_BAD_DIM;
    const unsigned blocks = (unsigned)((n + 255) / 256 < 4096 ? (n + 255) / 256 : 4096);
    hipLaunchKernelGGL(weight_step_kernel, dim3(blocks), dim3(256), 0, (hipStream_t)stream, w_old, w_new, grads, J,
                       n_scen, n, mode, step, w_min, w_max);
    return launch_rc("weight_step_kernel");
}
