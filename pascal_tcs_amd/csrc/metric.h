// metric.h -- the metric arrays (dx1..dmx3) as the stencil formulas read
// them.  Included inside a form's namespace.  The exact form divides by the
// metric as the reference writes it; the fast forms (TCS_RECIP) multiply by
// its reciprocal from a table: x / dx1(i) becomes x * (1/dx1(i)), with
// 1/dx1(i) correctly rounded on the host, instead of a reciprocal and
// Newton steps at every division.
#pragma once

#ifdef TCS_RECIP
struct MVal {
    double v, r;
    __device__ __forceinline__ operator double() const { return v; }
};
struct RMetric {
    const double *v, *r;
    __device__ __forceinline__ MVal operator[](int i) const { return MVal{v[i], r[i]}; }
};
__device__ __forceinline__ RMetric metric(const double *v, const double *r) { return RMetric{v, r}; }
__device__ __forceinline__ double operator/(double x, MVal m) { return x * m.r; }
__device__ __forceinline__ double operator/(MVal a, MVal b) { return a.v * b.r; }
#else
__device__ __forceinline__ const double *metric(const double *v, const double *) { return v; }
#endif
