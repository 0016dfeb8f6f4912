// pbh_noalloc.hip -- the forward dose without a workspace argument, for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops).
//
// pbh_forward_noalloc runs pbh_forward_tuned's kernels with their workspace in
// `scratch`, a fixed file-scope device array of PBH_NOALLOC_SCRATCH_BYTES that
// each GPU holds until the process ends. In host code the array's name is only a
// stand-in: its device address is resolved once per device (hipGetSymbolAddress,
// with that device current) and kept in g_scratch; only that address is passed
// to the kernels. The forward workspace needs no initialisation (every split
// weight, layer list and partial tile is written before it is read in the same
// call), so nothing is cleared between calls. All calls on one GPU share its
// array: calls that overlap on one GPU (two streams, two threads) race.
#include <mutex>

#include "pbh_common.h"

static __device__ __align__(256) unsigned char scratch[PBH_NOALLOC_SCRATCH_BYTES];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_mutex;
void* g_scratch[kMaxDevices];     // device address of `scratch` per device, once resolved

// Device address of the current device's scratch array, or nullptr on failure.
void* scratch_ptr(int* rc)
{
    int dev = 0;
    if ((*rc = pbh::hip_rc(hipGetDevice(&dev), "hipGetDevice")) != PBH_OK) return nullptr;
    if (dev < 0 || dev >= kMaxDevices) {
        *rc = PBH_ERR_UNSUPPORTED;
        return nullptr;
    }
    std::lock_guard<std::mutex> lock(g_mutex);
    if (!g_scratch[dev]) {
        void* p = nullptr;
        *rc = pbh::hip_rc(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
        if (*rc != PBH_OK) return nullptr;
        g_scratch[dev] = p;
    }
    *rc = PBH_OK;
    return g_scratch[dev];
}

}  // namespace

extern "C" int pbh_forward_noalloc(const PbhGrid* g, const PbhLattice* lat, const float* w,
                                   const double* idd, const double* sigma, float* dose, int precision,
                                   int layer_chunk, void* stream)
{
    int rc = pbh_validate(g, lat);
    if (rc != PBH_OK) return rc;
    if (!w || !idd || !sigma || !dose) return PBH_ERR_NULL_ARG;
    if (precision < PBH_PREC_FP32 || precision > PBH_PREC_FP16X3_PEAK) return PBH_ERR_UNSUPPORTED;
    if (layer_chunk < 0) return PBH_ERR_BAD_DIM;
    const size_t need = pbh_forward_workspace_bytes_tuned(g, lat, precision, layer_chunk);
    if (need == 0) return PBH_ERR_BAD_DIM;
    if (need > (size_t)PBH_NOALLOC_SCRATCH_BYTES) {
        fprintf(stderr, "[pbdose_hip] pbh_forward_noalloc: this problem needs %zu bytes of scratch, "
                        "more than the fixed %zu; use pbh_forward_tuned with a workspace, or a larger "
                        "layer chunk\n", need, (size_t)PBH_NOALLOC_SCRATCH_BYTES);
        return PBH_ERR_SCRATCH;
    }
    void* ws = scratch_ptr(&rc);
    if (!ws) return rc;
    return pbh_forward_tuned(g, lat, w, idd, sigma, dose, precision, layer_chunk, ws,
                             (size_t)PBH_NOALLOC_SCRATCH_BYTES, stream);
}
