// pbh_mixed_core.h -- 16-bit operand helpers shared by the mixed-precision
// forward and adjoint kernels: MFMA wrappers, (hi, lo) splitting and packing,
// and compile-time table-slot streams.
#pragma once

#include "pbh_fwd_core.h"

namespace pbh {

enum { MODE_PLAIN = 0, MODE_X3 = 1, MODE_PEAK = 2 };

typedef unsigned u4v __attribute__((ext_vector_type(4)));
typedef unsigned u2v __attribute__((ext_vector_type(2)));

template <typename T> struct Vec8;
template <> struct Vec8<__bf16> { typedef v8bf type; };
template <> struct Vec8<_Float16> { typedef v8h type; };

__device__ __forceinline__ v4f mfma16(v8bf a, v8bf b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0); }
__device__ __forceinline__ v4f mfma16(v8h a, v8h b, v4f c) { return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0); }
__device__ __forceinline__ v16f mfma32(v8bf a, v8bf b, v16f c) { return __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, c, 0, 0, 0); }
__device__ __forceinline__ v16f mfma32(v8h a, v8h b, v16f c) { return __builtin_amdgcn_mfma_f32_32x32x16_f16(a, b, c, 0, 0, 0); }

// Table slots of one lane's operand stream: entries t0 + Q k for k = 0, 1, ...
// QC > 0: Q = QC at compile time and the table parity-interleaved (P == 2):
// slot(t0 + c) = (c even ? bE : bO) + (c >> 1), offsets known at compile time.
// QC == 0: slots computed from the run-time geometry.
template <int QC>
struct TabStream {
    int bE, bO;
    __device__ __forceinline__ void init(const AxisTable& g, int t0)
    {
        if (QC > 0) {
            const int p0 = t0 & 1, h0 = t0 >> 1;
            bE = (p0 ? g.half : 0) + h0;
            bO = (p0 ? 0 : g.half) + h0 + p0;
        } else {
            bE = t0;
            bO = 0;
        }
    }
    // Moves the stream k0 spots on (k0 a multiple of 8): slot(k) becomes slot(k0 + k).
    __device__ __forceinline__ void shift(const AxisTable& g, int k0)
    {
        if (QC > 0) {
            const int d = (QC * k0) >> 1;
            bE += d;
            bO += d;
        } else {
            bE += g.Q * k0;
        }
    }
    // QC > 0: slot of entry t0 + c for any integer offset c (compile-time after unrolling)
    __device__ __forceinline__ int at(int c) const { return ((c & 1) ? bO : bE) + (c >> 1); }
    __device__ __forceinline__ int slot(const AxisTable& g, int k) const
    {
        if (QC > 0) {
            const int c = QC * k;
            return ((c & 1) ? bO : bE) + (c >> 1);
        }
        return table_slot(g, bE + g.Q * k);
    }
};

template <typename T>
__device__ __forceinline__ void split(float x, T& hi, T& lo)
{
    hi = (T)x;
    lo = (T)(x - (float)hi);
}

template <typename T>
__device__ __forceinline__ unsigned short bits16(T v) { return __builtin_bit_cast(unsigned short, v); }

// (hi, lo) of x packed as hi | lo << 16
template <typename T>
__device__ __forceinline__ unsigned pack_split(float x)
{
    T h, l;
    split<T>(x, h, l);
    return (unsigned)bits16(h) | ((unsigned)bits16(l) << 16);
}

// pack_split of two values, each half converted as one pair: w0 = pack_split(x0),
// w1 = pack_split(x1), bit for bit.
template <typename T>
__device__ __forceinline__ void pack_split2(float x0, float x1, unsigned& w0, unsigned& w1)
{
    typedef T V2 __attribute__((ext_vector_type(2)));
    typedef float F2 __attribute__((ext_vector_type(2)));
    const V2 h = __builtin_convertvector((F2){x0, x1}, V2);
    const V2 l = __builtin_convertvector((F2){x0 - (float)h[0], x1 - (float)h[1]}, V2);
    const unsigned hb = __builtin_bit_cast(unsigned, h), lb = __builtin_bit_cast(unsigned, l);
    w0 = __builtin_amdgcn_perm(lb, hb, 0x05040100u);
    w1 = __builtin_amdgcn_perm(lb, hb, 0x07060302u);
}

// Packed (hi, lo) entries of a thread's share of a table: entry t gets
// pack_split(exp2(-(d * d) * k2)), two entries per pair conversion.
template <typename T, int NT>
__device__ __forceinline__ void build_packed(const TabShare<NT>& sh, unsigned* tab, float k2)
{
    const float nk2 = -k2;
    unsigned* dst = tab + sh.s0;
#pragma unroll
    for (int k = 0; k < TabShare<NT>::N; k += 2) {
        if (k + 1 < TabShare<NT>::N) {
            unsigned w0, w1;
            pack_split2<T>(__builtin_amdgcn_exp2f(sh.d2[k] * nk2), __builtin_amdgcn_exp2f(sh.d2[k + 1] * nk2), w0, w1);
            if (TabShare<NT>::ok(k)) dst[FT_THREADS / 2 * k] = w0;
            if (TabShare<NT>::ok(k + 1)) dst[FT_THREADS / 2 * (k + 1)] = w1;
        } else if (TabShare<NT>::ok(k)) {
            dst[FT_THREADS / 2 * k] = pack_split<T>(__builtin_amdgcn_exp2f(sh.d2[k] * nk2));
        }
    }
}

// Operand-pair tables (single-product forms): entry t of a thread's share is
// written as the first half of slot(t) and the second half of slot(t - Q), so that
// slot(t) holds (hi(t), hi(t + Q)). A lane's eight operand entries t0 + Q e (e < 8)
// are then the four 32-bit slots of t0 + 2 Q q, read straight into the MFMA operand
// with no unpacking. s1 = slot of entry threadIdx.x - Q (negative for the first Q
// threads: their first entry has no pair slot). Entry values equal build_packed's
// hi halves.
template <typename T, int QC, int NT>
__device__ __forceinline__ void build_pairs(const TabShare<NT>& sh, int s1, unsigned* A, float k2)
{
    typedef T V2 __attribute__((ext_vector_type(2)));
    typedef float F2 __attribute__((ext_vector_type(2)));
    const float nk2 = -k2;
    unsigned short* a16 = (unsigned short*)A;
    auto put = [&](int k, T h) {
        if (!TabShare<NT>::ok(k)) return;
        a16[2 * (sh.s0 + FT_THREADS / 2 * k)] = bits16(h);
        if (k > 0 || (int)threadIdx.x >= QC) a16[2 * (s1 + FT_THREADS / 2 * k) + 1] = bits16(h);
    };
#pragma unroll
    for (int k = 0; k < TabShare<NT>::N; k += 2) {
        const float x0 = __builtin_amdgcn_exp2f(sh.d2[k] * nk2);
        const float x1 = k + 1 < TabShare<NT>::N ? __builtin_amdgcn_exp2f(sh.d2[k + 1] * nk2) : 0.0f;
        const V2 h = __builtin_convertvector((F2){x0, x1}, V2);
        put(k, h[0]);
        if (k + 1 < TabShare<NT>::N) put(k + 1, h[1]);
    }
}

// Eight packed entries -> hi and lo operand vectors.
template <typename V8>
__device__ __forceinline__ void unpack8(const unsigned* t, V8& hi, V8& lo)
{
    u4v h, l;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        h[q] = __builtin_amdgcn_perm(t[2 * q + 1], t[2 * q], 0x05040100u);
        l[q] = __builtin_amdgcn_perm(t[2 * q + 1], t[2 * q], 0x07060302u);
    }
    hi = __builtin_bit_cast(V8, h);
    lo = __builtin_bit_cast(V8, l);
}

}  // namespace pbh
