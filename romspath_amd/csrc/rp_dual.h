// rp_dual.h -- forward-mode dual numbers for the step's local Jacobians.
//
// Dual<K> carries a value and its derivatives along K seed directions.  The
// value part of every operation is the double operation itself, in the same
// order, so a step evaluated in Dual<K> reproduces the double step's values
// bit for bit (the file is built with -ffp-contract=off) while the derivative
// parts give the exact Jacobian of that piecewise-smooth map.  Comparisons,
// floor and integer conversion act on the value (branches and cell indices
// are locally constant).
#pragma once

#include <hip/hip_runtime.h>

namespace rp {

// the double overloads stay visible next to the Dual ones below
using ::fabs;
using ::sqrt;

template <int K>
struct Dual {
    double v;
    double d[K];
    __device__ __forceinline__ Dual() {}
    __device__ __forceinline__ Dual(double x) : v(x)
    {
#pragma unroll
        for (int i = 0; i < K; i++) d[i] = 0.0;
    }
    __device__ __forceinline__ static Dual seed(double x, int dir)
    {
        Dual r(x);
        r.d[dir] = 1.0;
        return r;
    }
    __device__ __forceinline__ Dual& operator+=(const Dual& b) { return *this = *this + b; }
    __device__ __forceinline__ Dual& operator-=(const Dual& b) { return *this = *this - b; }
    __device__ __forceinline__ Dual& operator*=(const Dual& b) { return *this = *this * b; }
    __device__ __forceinline__ Dual& operator/=(const Dual& b) { return *this = *this / b; }

    __device__ __forceinline__ friend Dual operator+(const Dual& a, const Dual& b)
    {
        Dual r;
        r.v = a.v + b.v;
#pragma unroll
        for (int i = 0; i < K; i++) r.d[i] = a.d[i] + b.d[i];
        return r;
    }
    __device__ __forceinline__ friend Dual operator-(const Dual& a, const Dual& b)
    {
        Dual r;
        r.v = a.v - b.v;
#pragma unroll
        for (int i = 0; i < K; i++) r.d[i] = a.d[i] - b.d[i];
        return r;
    }
    __device__ __forceinline__ friend Dual operator-(const Dual& a)
    {
        Dual r;
        r.v = -a.v;
#pragma unroll
        for (int i = 0; i < K; i++) r.d[i] = -a.d[i];
        return r;
    }
    __device__ __forceinline__ friend Dual operator*(const Dual& a, const Dual& b)
    {
        Dual r;
        r.v = a.v * b.v;
#pragma unroll
        for (int i = 0; i < K; i++) r.d[i] = a.d[i] * b.v + a.v * b.d[i];
        return r;
    }
    __device__ __forceinline__ friend Dual operator/(const Dual& a, const Dual& b)
    {
        Dual r;
        r.v = a.v / b.v;
        const double ib = 1.0 / b.v;
#pragma unroll
        for (int i = 0; i < K; i++) r.d[i] = (a.d[i] - r.v * b.d[i]) * ib;
        return r;
    }
    // mixed forms (a double operand has zero derivative)
    __device__ __forceinline__ friend Dual operator+(const Dual& a, double b) { Dual r = a; r.v = a.v + b; return r; }
    __device__ __forceinline__ friend Dual operator+(double a, const Dual& b) { Dual r = b; r.v = a + b.v; return r; }
    __device__ __forceinline__ friend Dual operator-(const Dual& a, double b) { Dual r = a; r.v = a.v - b; return r; }
    __device__ __forceinline__ friend Dual operator-(double a, const Dual& b)
    {
        Dual r;
        r.v = a - b.v;
#pragma unroll
        for (int i = 0; i < K; i++) r.d[i] = -b.d[i];
        return r;
    }
    __device__ __forceinline__ friend Dual operator*(const Dual& a, double b)
    {
        Dual r;
        r.v = a.v * b;
#pragma unroll
        for (int i = 0; i < K; i++) r.d[i] = a.d[i] * b;
        return r;
    }
    __device__ __forceinline__ friend Dual operator*(double a, const Dual& b)
    {
        Dual r;
        r.v = a * b.v;
#pragma unroll
        for (int i = 0; i < K; i++) r.d[i] = a * b.d[i];
        return r;
    }
    __device__ __forceinline__ friend Dual operator/(const Dual& a, double b)
    {
        Dual r;
        r.v = a.v / b;
#pragma unroll
        for (int i = 0; i < K; i++) r.d[i] = a.d[i] / b;
        return r;
    }
    __device__ __forceinline__ friend Dual operator/(double a, const Dual& b)
    {
        Dual r;
        r.v = a / b.v;
        const double ib = 1.0 / b.v;
#pragma unroll
        for (int i = 0; i < K; i++) r.d[i] = -r.v * b.d[i] * ib;
        return r;
    }

#define RP_DUAL_CMP(op)                                                                        \
    __device__ __forceinline__ friend bool operator op(const Dual& a, const Dual& b) { return a.v op b.v; } \
    __device__ __forceinline__ friend bool operator op(const Dual& a, double b) { return a.v op b; }         \
    __device__ __forceinline__ friend bool operator op(double a, const Dual& b) { return a op b.v; }
    RP_DUAL_CMP(<) RP_DUAL_CMP(>) RP_DUAL_CMP(<=) RP_DUAL_CMP(>=) RP_DUAL_CMP(==) RP_DUAL_CMP(!=)
#undef RP_DUAL_CMP
};

template <int K>
__device__ __forceinline__ Dual<K> sqrt(const Dual<K>& a)
{
    Dual<K> r;
    r.v = ::sqrt(a.v);
    const double k = 0.5 / r.v;
#pragma unroll
    for (int i = 0; i < K; i++) r.d[i] = a.d[i] * k;
    return r;
}

template <int K>
__device__ __forceinline__ Dual<K> fabs(const Dual<K>& a)
{
    return a.v < 0.0 ? -a : a;
}

// value of a scalar of either kind
__device__ __forceinline__ double val(double x) { return x; }
template <int K>
__device__ __forceinline__ double val(const Dual<K>& x) { return x.v; }

// scalar conversion between kinds (derivative parts kept where both have them)
template <class To>
struct Conv;
template <>
struct Conv<double> {
    __device__ __forceinline__ static double from(double x) { return x; }
    template <int K>
    __device__ __forceinline__ static double from(const Dual<K>& x) { return x.v; }
};
template <int K>
struct Conv<Dual<K>> {
    __device__ __forceinline__ static Dual<K> from(double x) { return Dual<K>(x); }
    template <int J>
    __device__ __forceinline__ static Dual<K> from(const Dual<J>& x)
    {
        Dual<K> r(x.v);
#pragma unroll
        for (int i = 0; i < (J < K ? J : K); i++) r.d[i] = x.d[i];
        return r;
    }
};

}  // namespace rp
