// swe_dual.hpp — forward-mode dual numbers for the adjoint kernels.
//
// Dual<N> carries a value and N partial derivatives.  The adjoint kernels
// evaluate a face flux or a cell reconstruction with duals seeded on its
// inputs to get the local Jacobian at the forward state, then apply its
// transpose.  Values follow IEEE double arithmetic exactly as the plain
// kernels do (same operations, same order), so every branch the forward
// step took is taken again; derivatives follow the taken branch, as
// forward-mode differentiation of the reference would.
#pragma once

#include <hip/hip_runtime.h>

namespace sfv {

template <int N>
struct Dual {
    double v;
    double d[N];
    __host__ __device__ Dual() : v(0.0) {
        for (int k = 0; k < N; ++k) d[k] = 0.0;
    }
    __host__ __device__ Dual(double x) : v(x) {   // a constant
        for (int k = 0; k < N; ++k) d[k] = 0.0;
    }
    __host__ __device__ static Dual var(double x, int k) {
        Dual r(x);
        r.d[k] = 1.0;
        return r;
    }
};

template <class T> struct DualTraits { static constexpr int N = 0; };
template <int N> struct DualTraits<Dual<N>> { static constexpr int N_ = N; };

__host__ __device__ inline double val(double x) { return x; }
template <int N> __host__ __device__ inline double val(const Dual<N>& x) { return x.v; }

template <int N>
__host__ __device__ inline Dual<N> operator+(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v + b.v);
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] + b.d[k];
    return r;
}
template <int N>
__host__ __device__ inline Dual<N> operator-(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v - b.v);
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] - b.d[k];
    return r;
}
template <int N>
__host__ __device__ inline Dual<N> operator-(const Dual<N>& a) {
    Dual<N> r(-a.v);
    for (int k = 0; k < N; ++k) r.d[k] = -a.d[k];
    return r;
}
template <int N>
__host__ __device__ inline Dual<N> operator*(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v * b.v);
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
    return r;
}
template <int N>
__host__ __device__ inline Dual<N> operator/(const Dual<N>& a, const Dual<N>& b) {
    Dual<N> r(a.v / b.v);
    const double ib = 1.0 / b.v;
    for (int k = 0; k < N; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) * ib;
    return r;
}
template <int N> __host__ __device__ inline Dual<N> operator+(const Dual<N>& a, double b) { return a + Dual<N>(b); }
template <int N> __host__ __device__ inline Dual<N> operator+(double a, const Dual<N>& b) { return Dual<N>(a) + b; }
template <int N> __host__ __device__ inline Dual<N> operator-(const Dual<N>& a, double b) { return a - Dual<N>(b); }
template <int N> __host__ __device__ inline Dual<N> operator-(double a, const Dual<N>& b) { return Dual<N>(a) - b; }
template <int N>
__host__ __device__ inline Dual<N> operator*(const Dual<N>& a, double b) {
    Dual<N> r(a.v * b);
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * b;
    return r;
}
template <int N> __host__ __device__ inline Dual<N> operator*(double a, const Dual<N>& b) {
    Dual<N> r(a * b.v);
    for (int k = 0; k < N; ++k) r.d[k] = a * b.d[k];
    return r;
}
template <int N> __host__ __device__ inline Dual<N> operator/(const Dual<N>& a, double b) { return a / Dual<N>(b); }
template <int N> __host__ __device__ inline Dual<N> operator/(double a, const Dual<N>& b) { return Dual<N>(a) / b; }

template <int N> __host__ __device__ inline bool operator<(const Dual<N>& a, const Dual<N>& b) { return a.v < b.v; }
template <int N> __host__ __device__ inline bool operator>(const Dual<N>& a, const Dual<N>& b) { return a.v > b.v; }
template <int N> __host__ __device__ inline bool operator<(const Dual<N>& a, double b) { return a.v < b; }
template <int N> __host__ __device__ inline bool operator>(const Dual<N>& a, double b) { return a.v > b; }
template <int N> __host__ __device__ inline bool operator<(double a, const Dual<N>& b) { return a < b.v; }
template <int N> __host__ __device__ inline bool operator>(double a, const Dual<N>& b) { return a > b.v; }
template <int N> __host__ __device__ inline bool operator==(const Dual<N>& a, double b) { return a.v == b; }

template <int N>
__host__ __device__ inline Dual<N> dsqrt(const Dual<N>& a) {
    Dual<N> r(sqrt(a.v));
    const double s = r.v > 0.0 ? 0.5 / r.v : 0.0;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * s;
    return r;
}
__host__ __device__ inline double dsqrt(double a) { return sqrt(a); }

template <int N>
__host__ __device__ inline Dual<N> dcbrt(const Dual<N>& a) {
    Dual<N> r(cbrt(a.v));
    const double s = a.v != 0.0 ? r.v / (3.0 * a.v) : 0.0;
    for (int k = 0; k < N; ++k) r.d[k] = a.d[k] * s;
    return r;
}
__host__ __device__ inline double dcbrt(double a) { return cbrt(a); }

__host__ __device__ inline double dabs(double a) { return __builtin_fabs(a); }

__host__ __device__ inline bool is_nan(double a) { return a != a; }
template <int N> __host__ __device__ inline bool is_nan(const Dual<N>& a) { return a.v != a.v; }

// y if py, else x, component by component (a select of whole operands
// through references would be lowered to a pointer select and put both
// operands in scratch memory).
__host__ __device__ inline double dpick(bool py, double x, double y) { return py ? y : x; }
template <int N>
__host__ __device__ inline Dual<N> dpick(bool py, const Dual<N>& x, const Dual<N>& y) {
    Dual<N> r(py ? y.v : x.v);
    for (int k = 0; k < N; ++k) r.d[k] = py ? y.d[k] : x.d[k];
    return r;
}

template <int N>
__host__ __device__ inline Dual<N> dabs(const Dual<N>& a) { return dpick(__builtin_signbit(a.v), a, -a); }

// Julia max/min (NaN-propagating, -0.0 < +0.0) choosing one operand whole.
template <class T>
__host__ __device__ inline T djl_max(const T& x, const T& y) {
    const double a = val(x), b = val(y);
    const bool py = !(a != a) && ((b != b) || b > a || (!(a > b) && __builtin_signbit(a)));
    return dpick(py, x, y);
}
template <class T>
__host__ __device__ inline T djl_min(const T& x, const T& y) {
    const double a = val(x), b = val(y);
    const bool py = !(a != a) && ((b != b) || b < a || (!(a < b) && !__builtin_signbit(a)));
    return dpick(py, x, y);
}

}  // namespace sfv
