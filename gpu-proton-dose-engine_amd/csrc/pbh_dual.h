// pbh_dual.h -- forward-mode differentiation for the FP64 look-up-table kernels.
//
// A Dual carries a value v and its derivative d along one input direction. The
// look-up-table kernels are templates over their scalar type: instantiated with
// double they are the plain tables; instantiated with Dual they return the same
// values (every value operation is the double operation, in the same order) and,
// alongside, the exact derivative of the discrete algorithm (interpolation
// brackets, bin edges and grid sizes fixed at their current values).
#pragma once

#include <hip/hip_runtime.h>
#include <type_traits>

namespace pbh {

struct Dual {
    double v, d;
};

template <class T> constexpr bool is_dual = std::is_same<T, Dual>::value;

__device__ __forceinline__ double val(double x) { return x; }
__device__ __forceinline__ double val(Dual x) { return x.v; }
__device__ __forceinline__ double der(double) { return 0.0; }
__device__ __forceinline__ double der(Dual x) { return x.d; }

// a constant (zero derivative) of scalar type T
template <class T> __device__ __forceinline__ T cst(double v)
{
    if constexpr (is_dual<T>) return Dual{v, 0.0};
    else return v;
}

// value v with derivative d (d dropped for double)
template <class T> __device__ __forceinline__ T make(double v, double d)
{
    if constexpr (is_dual<T>) return Dual{v, d};
    else return v;
}

__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ Dual operator+(Dual a, double b) { return {a.v + b, a.d}; }
__device__ __forceinline__ Dual operator+(double a, Dual b) { return {a + b.v, b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ Dual operator-(Dual a, double b) { return {a.v - b, a.d}; }
__device__ __forceinline__ Dual operator-(double a, Dual b) { return {a - b.v, -b.d}; }
__device__ __forceinline__ Dual operator-(Dual a) { return {-a.v, -a.d}; }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ Dual operator*(Dual a, double b) { return {a.v * b, a.d * b}; }
__device__ __forceinline__ Dual operator*(double a, Dual b) { return {a * b.v, a * b.d}; }
__device__ __forceinline__ Dual operator/(Dual a, Dual b)
{
    const double q = a.v / b.v;
    return {q, (a.d - q * b.d) / b.v};
}
__device__ __forceinline__ Dual operator/(Dual a, double b) { return {a.v / b, a.d / b}; }
__device__ __forceinline__ Dual operator/(double a, Dual b)
{
    const double q = a / b.v;
    return {q, -q * b.d / b.v};
}
__device__ __forceinline__ Dual& operator+=(Dual& a, Dual b) { a = a + b; return a; }
__device__ __forceinline__ Dual& operator-=(Dual& a, Dual b) { a = a - b; return a; }

__device__ __forceinline__ double r_exp(double x) { return exp(x); }
__device__ __forceinline__ Dual r_exp(Dual x) { const double e = exp(x.v); return {e, e * x.d}; }
__device__ __forceinline__ double r_log(double x) { return log(x); }
__device__ __forceinline__ Dual r_log(Dual x) { return {log(x.v), x.d / x.v}; }
__device__ __forceinline__ double r_sqrt(double x) { return sqrt(x); }
__device__ __forceinline__ Dual r_sqrt(Dual x)
{
    const double s = sqrt(x.v);
    return {s, s > 0.0 ? x.d / (2.0 * s) : 0.0};
}
__device__ __forceinline__ double r_hypot(double a, double b) { return hypot(a, b); }
__device__ __forceinline__ Dual r_hypot(Dual a, Dual b)
{
    const double h = hypot(a.v, b.v);
    return {h, h > 0.0 ? (a.v * a.d + b.v * b.d) / h : 0.0};
}
__device__ __forceinline__ double r_pow(double a, double e) { return pow(a, e); }
__device__ __forceinline__ Dual r_pow(Dual a, double e)
{
    const double p = pow(a.v, e);
    return {p, a.v > 0.0 ? e * p / a.v * a.d : 0.0};
}

// min/max against a constant bound (derivative 0 where the bound is taken)
template <class T> __device__ __forceinline__ T clamp_lo(T x, double lo) { return val(x) < lo ? cst<T>(lo) : x; }
template <class T> __device__ __forceinline__ T clamp_hi(T x, double hi) { return val(x) > hi ? cst<T>(hi) : x; }

__device__ __forceinline__ double shfl_down(double v, int o) { return __shfl_down(v, o, 64); }
__device__ __forceinline__ Dual shfl_down(Dual v, int o) { return {__shfl_down(v.v, o, 64), __shfl_down(v.d, o, 64)}; }
__device__ __forceinline__ double shfl_xor(double v, int o) { return __shfl_xor(v, o, 64); }
__device__ __forceinline__ Dual shfl_xor(Dual v, int o) { return {__shfl_xor(v.v, o, 64), __shfl_xor(v.d, o, 64)}; }

}  // namespace pbh
