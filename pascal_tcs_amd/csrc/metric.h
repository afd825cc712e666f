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
// metric<S, M>(g): table M of the grid g as the formulas of scalar type S
// read it.  The tables in the order of the reciprocal and float copies:
enum MetricId : int { M_DX1, M_DMX1, M_DX2, M_DMX2, M_DX3, M_DMX3 };
template <int M>
__device__ __forceinline__ const double *mtab(const Grid &g) {
    if constexpr (M == M_DX1) return g.dx1;
    else if constexpr (M == M_DMX1) return g.dmx1;
    else if constexpr (M == M_DX2) return g.dx2;
    else if constexpr (M == M_DMX2) return g.dmx2;
    else if constexpr (M == M_DX3) return g.dx3;
    else return g.dmx3;
}
#ifdef TCS_RECIP
template <int M>
__device__ __forceinline__ const double *rtab(const Grid &g) {
    if constexpr (M == M_DX1) return g.rdx1;
    else if constexpr (M == M_DMX1) return g.rdmx1;
    else if constexpr (M == M_DX2) return g.rdx2;
    else if constexpr (M == M_DMX2) return g.rdmx2;
    else if constexpr (M == M_DX3) return g.rdx3;
    else return g.rdmx3;
}
// offset of table M in the back-to-back order (N1, N1, N2, N2, N3, N3 long)
template <int M>
__device__ __forceinline__ size_t moff(const Grid &g) {
    const size_t n1 = g.N1, n2 = g.N2, n3 = g.N3;
    if constexpr (M == M_DX1) return 0;
    else if constexpr (M == M_DMX1) return n1;
    else if constexpr (M == M_DX2) return 2 * n1;
    else if constexpr (M == M_DMX2) return 2 * n1 + n2;
    else if constexpr (M == M_DX3) return 2 * (n1 + n2);
    else return 2 * (n1 + n2) + n3;
}
//
// S = float (the mixed form's rows) reads the metric and its reciprocal
// rounded to float on the host (Grid::fmet, the values then the
// reciprocals), every other S the doubles.
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
    const float *v, *r;
    __device__ __forceinline__ MValF operator[](int i) const { return MValF{v[i], r[i]}; }
};
__device__ __forceinline__ float operator/(float x, MValF m) { return x * m.r; }
__device__ __forceinline__ float operator/(MValF a, MValF b) { return a.v * b.r; }
// the same values rounded to float from the double tables where read
struct RMetricFD {
    const double *v, *r;
    __device__ __forceinline__ MValF operator[](int i) const { return MValF{(float)v[i], (float)r[i]}; }
};
// D64: an S = float row reads the double tables and rounds where it reads
// (the same values; the rows whose compiled schedule is faster so)
template <class S, int M, bool D64 = false>
__device__ __forceinline__ auto metric(const Grid &g) {
    if constexpr (std::is_same<S, float>::value && D64) {
        return RMetricFD{mtab<M>(g), rtab<M>(g)};
    } else if constexpr (std::is_same<S, float>::value) {
        const float *v = g.fmet + moff<M>(g);
        return RMetricF{v, v + 2 * ((size_t)g.N1 + g.N2 + g.N3)};
    } else {
        return RMetric{mtab<M>(g), rtab<M>(g)};
    }
}
#else
template <class S, int M, bool D64 = false>
__device__ __forceinline__ const double *metric(const Grid &g) { return mtab<M>(g); }
#endif
