// pbh_common.h -- definitions shared by the gfx950 kernels of libpbdose_hip.so.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#include "pbdose_hip.h"

typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v16f __attribute__((ext_vector_type(16)));
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

#define PBH_WAVE 64

// Library limits (checked by pbh_validate).
#define PBH_MAX_DIM 4096        // voxels per axis
#define PBH_MAX_SPOTS 4096      // spots per lattice axis
#define PBH_MAX_LAYERS 65536
#define PBH_SPOT_CHUNK 48       // spots per axis handled by one pass of the contraction

namespace pbh {

constexpr double kPi = 3.14159265358979323846;
constexpr double kLog2e = 1.4426950408889634074;
constexpr double kSigmaFloor = 1e-9;   // engines.py:57, :266

__host__ __device__ inline int cdiv(int a, int b) { return (a + b - 1) / b; }

// Translates a HIP status into a library code, reporting it once on stderr.
static inline int hip_rc(hipError_t e, const char* what)
{
    if (e == hipSuccess) return PBH_OK;
    fprintf(stderr, "[pbdose_hip] %s: %s\n", what, hipGetErrorString(e));
    return PBH_ERR_HIP;
}

static inline int launch_rc(const char* name)
{
    return hip_rc(hipGetLastError(), name);
}

// Split of the lattice axis into chunks of at most PBH_SPOT_CHUNK spots:
// chunk c covers [c * n / nc, (c+1) * n / nc).
__host__ __device__ inline int chunk_count(int n) { return (n + PBH_SPOT_CHUNK - 1) / PBH_SPOT_CHUNK; }
__host__ __device__ inline int chunk_begin(int n, int nc, int c) { return (int)(((long long)c * n) / nc); }

// Per-(layer, depth) scalars of the contraction, evaluated in FP64:
//   c  = IDD / (2 pi s^2)            (both 1/(sqrt(2 pi) s) norms and the IDD)
//   k2 = log2(e) / (2 s^2)            (Gaussian exponent in base 2)
struct LayerScalars { float c, k2; };

__device__ inline LayerScalars layer_scalars(double idd, double sigma)
{
    const double s = sigma > kSigmaFloor ? sigma : kSigmaFloor;
    const double s2 = s * s;
    LayerScalars r;
    r.c = (float)(idd / (2.0 * kPi * s2));
    r.k2 = (float)(kLog2e / (2.0 * s2));
    return r;
}

// exp(-d2 / (2 s^2)) given k2 = log2(e) / (2 s^2): one rounding in the
// product, then the hardware base-2 exponential (underflow flushes to 0).
__device__ __forceinline__ float gauss2(float d2, float k2)
{
    return __builtin_amdgcn_exp2f(-d2 * k2);
}

// Arguments of the forward kernels (geometry in FP64, device pointers).
struct FwdArgs {
    int nx, ny, nz, nsx, nsy, L;
    double dx, dy, ox, oy;           // voxel axes (x, y)
    double dsx, dsy, sx0, sy0;       // spot i centre = sx0 + i * dsx (sx0 = x0 - (nsx-1)/2 dsx)
    const float* w;                  // (L, nsy, nsx)
    const double* idd;               // (L, nz)
    const double* sigma;             // (L, nz)
    float* dose;                     // (nz, ny, nx)
    int tiles_x, tiles_y;
    int tile_rows;                   // dose tile rows (y): 32, or 64 for the mixed forms
};

}  // namespace pbh
