// metric.h -- the metric arrays (dx1..dmx3) as the stencil formulas read
// them.  Included inside a form's namespace.  The exact form divides by the
// metric as the reference writes it; the fast forms (TCS_RECIP) multiply by
// its reciprocal from a table: x / dx1(i) becomes x * (1/dx1(i)), with
// 1/dx1(i) correctly rounded on the host, instead of a reciprocal and
// Newton steps at every division.
#pragma once

#include <type_traits>

//
// sc<S>(x): a constant of the formulas (a literal, a physical constant, dt,
// a mask) in the arithmetic of S: float for S = float, else the double
// itself (the double and dual-number instantiations are unchanged).
template <class S>
__device__ __forceinline__ constexpr auto sc(double x) {
    if constexpr (std::is_same<S, float>::value) return (float)x;
    else return x;
}
//
// metric<S>: the formulas' scalar type S selects the arithmetic; S = float
// (the mixed form's rows) reads the metric and its reciprocal rounded to
// float, every other S the doubles.
#ifdef TCS_RECIP
struct MVal {
    double v, r;
    __device__ __forceinline__ operator double() const { return v; }
};
struct RMetric {
    const double *v, *r;
    __device__ __forceinline__ MVal operator[](int i) const { return MVal{v[i], r[i]}; }
};
__device__ __forceinline__ double operator/(double x, MVal m) { return x * m.r; }
__device__ __forceinline__ double operator/(MVal a, MVal b) { return a.v * b.r; }
struct MValF {
    float v, r;
    __device__ __forceinline__ operator float() const { return v; }
};
struct RMetricF {
    const double *v, *r;
    __device__ __forceinline__ MValF operator[](int i) const { return MValF{(float)v[i], (float)r[i]}; }
};
__device__ __forceinline__ float operator/(float x, MValF m) { return x * m.r; }
__device__ __forceinline__ float operator/(MValF a, MValF b) { return a.v * b.r; }
template <class S>
__device__ __forceinline__ auto metric(const double *v, const double *r) {
    if constexpr (std::is_same<S, float>::value) return RMetricF{v, r};
    else return RMetric{v, r};
}
#else
template <class S>
__device__ __forceinline__ const double *metric(const double *v, const double *) { return v; }
#endif
