// Forward-mode dual numbers with K tangent directions, for exact Jacobians
// of the pointwise physics inside the adjoint kernels.  Every operation
// propagates the derivative of the operation actually evaluated, so a branch
// (entropy fix, HLLC wave selection, AUSM splitting, floors) contributes the
// derivative of the taken side.
#pragma once
#include <hip/hip_runtime.h>

namespace nz {

// the plain-double versions, visible next to the dual overloads below
using ::fabs;
using ::fmax;
using ::fmin;
using ::pow;
using ::sqrt;

template <int K>
struct Dual {
    double v;
    double d[K];
    __device__ __forceinline__ Dual() : v(0.0) {
        for (int k = 0; k < K; ++k) d[k] = 0.0;
    }
    __device__ __forceinline__ Dual(double x) : v(x) {  // NOLINT: implicit from constants
        for (int k = 0; k < K; ++k) d[k] = 0.0;
    }
    __device__ __forceinline__ static Dual seed(double x, int k) {
        Dual r(x);
        r.d[k] = 1.0;
        return r;
    }
};

// x with tangent direction k among the K carried (none if k is outside [0, K))
template <int K>
__device__ __forceinline__ Dual<K> seeded(double x, int k) {
    Dual<K> r(x);
#pragma unroll
    for (int j = 0; j < K; ++j) r.d[j] = j == k ? 1.0 : 0.0;
    return r;
}

template <int K>
__device__ __forceinline__ Dual<K> operator+(const Dual<K>& a, const Dual<K>& b) {
    Dual<K> r;
    r.v = a.v + b.v;
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(const Dual<K>& a, const Dual<K>& b) {
    Dual<K> r;
    r.v = a.v - b.v;
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator-(const Dual<K>& a) {
    Dual<K> r;
    r.v = -a.v;
    for (int k = 0; k < K; ++k) r.d[k] = -a.d[k];
    return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator*(const Dual<K>& a, const Dual<K>& b) {
    Dual<K> r;
    r.v = a.v * b.v;
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator/(const Dual<K>& a, const Dual<K>& b) {
    Dual<K> r;
    r.v = a.v / b.v;
    const double ib = 1.0 / b.v;
    for (int k = 0; k < K; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}
// mixed forms with plain doubles
template <int K>
__device__ __forceinline__ Dual<K> operator+(const Dual<K>& a, double b) { return a + Dual<K>(b); }
template <int K>
__device__ __forceinline__ Dual<K> operator+(double a, const Dual<K>& b) { return Dual<K>(a) + b; }
template <int K>
__device__ __forceinline__ Dual<K> operator-(const Dual<K>& a, double b) { return a - Dual<K>(b); }
template <int K>
__device__ __forceinline__ Dual<K> operator-(double a, const Dual<K>& b) { return Dual<K>(a) - b; }
template <int K>
__device__ __forceinline__ Dual<K> operator*(const Dual<K>& a, double b) {
    Dual<K> r;
    r.v = a.v * b;
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * b;
    return r;
}
template <int K>
__device__ __forceinline__ Dual<K> operator*(double a, const Dual<K>& b) { return b * a; }
template <int K>
__device__ __forceinline__ Dual<K> operator/(const Dual<K>& a, double b) { return a / Dual<K>(b); }
template <int K>
__device__ __forceinline__ Dual<K> operator/(double a, const Dual<K>& b) { return Dual<K>(a) / b; }
template <int K>
__device__ __forceinline__ Dual<K>& operator+=(Dual<K>& a, const Dual<K>& b) { return a = a + b; }
template <int K>
__device__ __forceinline__ Dual<K>& operator-=(Dual<K>& a, const Dual<K>& b) { return a = a - b; }

#define NZ_DUAL_CMP(op)                                                                        \
    template <int K>                                                                           \
    __device__ __forceinline__ bool operator op(const Dual<K>& a, const Dual<K>& b) {          \
        return a.v op b.v;                                                                     \
    }                                                                                          \
    template <int K>                                                                           \
    __device__ __forceinline__ bool operator op(const Dual<K>& a, double b) { return a.v op b; } \
    template <int K>                                                                           \
    __device__ __forceinline__ bool operator op(double a, const Dual<K>& b) { return a op b.v; }
NZ_DUAL_CMP(<)
NZ_DUAL_CMP(>)
NZ_DUAL_CMP(<=)
NZ_DUAL_CMP(>=)
#undef NZ_DUAL_CMP

template <int K>
__device__ __forceinline__ Dual<K> sqrt(const Dual<K>& a) {
    Dual<K> r;
    r.v = ::sqrt(a.v);
    const double h = 0.5 / r.v;
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * h;
    return r;
}
template <int K>
__device__ __forceinline__ Dual<K> fabs(const Dual<K>& a) {
    return a.v < 0.0 ? -a : a;
}
// max/min: the derivative of the selected argument (first on ties, as fmax)
template <int K>
__device__ __forceinline__ Dual<K> fmax(const Dual<K>& a, const Dual<K>& b) { return a.v >= b.v ? a : b; }
template <int K>
__device__ __forceinline__ Dual<K> fmin(const Dual<K>& a, const Dual<K>& b) { return a.v <= b.v ? a : b; }
template <int K>
__device__ __forceinline__ Dual<K> fmax(const Dual<K>& a, double b) { return a.v >= b ? a : Dual<K>(b); }
template <int K>
__device__ __forceinline__ Dual<K> fmin(const Dual<K>& a, double b) { return a.v <= b ? a : Dual<K>(b); }
template <int K>
__device__ __forceinline__ Dual<K> fmax(double a, const Dual<K>& b) { return fmax(b, a); }
template <int K>
__device__ __forceinline__ Dual<K> fmin(double a, const Dual<K>& b) { return fmin(b, a); }
template <int K>
__device__ __forceinline__ Dual<K> pow(const Dual<K>& a, double e) {
    Dual<K> r;
    r.v = ::pow(a.v, e);
    const double s = e * ::pow(a.v, e - 1.0);
    for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * s;
    return r;
}

__device__ __forceinline__ double value(double x) { return x; }
template <int K>
__device__ __forceinline__ double value(const Dual<K>& x) { return x.v; }

}  // namespace nz
