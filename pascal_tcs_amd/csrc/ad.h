// ad.h -- exact derivatives of the per-cell formulas of the step.
//
// The stage rows (thermal, momentum) and the cell updates (block
// corrections, pressure right-hand side, projection) are templates over the
// scalar type S and the field access P: a cell evaluates its formula with
// P[idx] loads from its input fields.  Instantiated with S = double and raw
// pointers they are the step itself.  The adjoint instantiates them twice
// more:
//   * RecW: doubles, every load recorded as (field, index) -- the cell's
//     exact input set, whatever its stencil and masks;
//   * SeedW: forward-mode dual numbers with the derivative seeded on one
//     recorded input, which gives that input's column of the cell's Jacobian
//     exactly (the same operations carried with their derivatives).
// k_cell_vjp combines them into the vector-Jacobian product of a whole
// kernel: for every cell, each input's column dotted with the cell's output
// adjoints is added (atomically: neighbouring cells share inputs) to that
// input's adjoint field.
#pragma once

#include "tcs_internal.h"

namespace tcs {

// ---- scalars ---------------------------------------------------------------
struct Dual {
    double v, d;
    __host__ __device__ Dual() : v(0.0), d(0.0) {}
    __host__ __device__ Dual(double x) : v(x), d(0.0) {}
    __host__ __device__ Dual(double x, double y) : v(x), d(y) {}
};
__device__ __forceinline__ Dual operator-(Dual a) { return Dual(-a.v, -a.d); }
__device__ __forceinline__ Dual operator+(Dual a, Dual b) { return Dual(a.v + b.v, a.d + b.d); }
__device__ __forceinline__ Dual operator-(Dual a, Dual b) { return Dual(a.v - b.v, a.d - b.d); }
__device__ __forceinline__ Dual operator*(Dual a, Dual b) { return Dual(a.v * b.v, a.d * b.v + a.v * b.d); }
__device__ __forceinline__ Dual operator/(Dual a, Dual b) {
    const double q = a.v / b.v;
    return Dual(q, (a.d - q * b.d) / b.v);
}
__device__ __forceinline__ Dual operator+(Dual a, double b) { return Dual(a.v + b, a.d); }
__device__ __forceinline__ Dual operator+(double a, Dual b) { return Dual(a + b.v, b.d); }
__device__ __forceinline__ Dual operator-(Dual a, double b) { return Dual(a.v - b, a.d); }
__device__ __forceinline__ Dual operator-(double a, Dual b) { return Dual(a - b.v, -b.d); }
__device__ __forceinline__ Dual operator*(Dual a, double b) { return Dual(a.v * b, a.d * b); }
__device__ __forceinline__ Dual operator*(double a, Dual b) { return Dual(a * b.v, a * b.d); }
__device__ __forceinline__ Dual operator/(Dual a, double b) { return Dual(a.v / b, a.d / b); }
__device__ __forceinline__ Dual operator/(double a, Dual b) {
    const double q = a / b.v;
    return Dual(q, -q * b.d / b.v);
}

// ---- field access ----------------------------------------------------------
// the step: raw pointers
struct RawW {
    __device__ __forceinline__ const double *operator()(int, const double *p) const { return p; }
};

// The mixed form's rows: FP32 arithmetic on the FP32 twins of the arrays
// they read.  In that form every array a stage row or block correction reads
// (state, increments, (j, k) tables, wall tables, stale scalars) lives in
// one FP64 arena and has an FP32 twin at the same element offset of an FP32
// arena, so the twin of p is f + (p - d).
struct F32Ptr {
    const float *p;
    __device__ __forceinline__ float operator[](size_t i) const { return p[i]; }
};
struct TwinW {
    const double *d;  // the FP64 arena
    const float *f;   // its FP32 twin
    __device__ __forceinline__ F32Ptr operator()(int, const double *p) const { return F32Ptr{f + (p - d)}; }
};
inline TwinW twin_w(const tcs_ctx *c) { return TwinW{c->a64, c->a32}; }
// the FP32 twin of an array of the mixed form's arena (null outside it)
inline float *twin(const tcs_ctx *c, const double *p) {
    if (!c->a64 || p < c->a64 || p >= c->a64 + c->an) return nullptr;
    return c->a32 + (p - c->a64);
}

// the cell's input set: (field, index) of every load, without repeats
constexpr int AD_MAXIN = 128;
struct Rec {
    int n;
    int fid[AD_MAXIN];
    long idx[AD_MAXIN];
    __device__ void add(int f, long i) {
        for (int r = 0; r < n; r++)
            if (fid[r] == f && idx[r] == i) return;
        if (n < AD_MAXIN) {
            fid[n] = f;
            idx[n] = i;
        }
        n++;  // n > AD_MAXIN reports an overflow
    }
};
struct RecPtr {
    const double *p;
    int f;
    Rec *rec;
    __device__ __forceinline__ double operator[](size_t i) const {
        rec->add(f, (long)i);
        return p[i];
    }
};
struct RecW {
    Rec *rec;
    __device__ __forceinline__ RecPtr operator()(int f, const double *p) const { return RecPtr{p, f, rec}; }
};

// one input seeded with derivative 1
struct SeedPtr {
    const double *p;
    long seed;
    __device__ __forceinline__ Dual operator[](size_t i) const { return Dual(p[i], (long)i == seed ? 1.0 : 0.0); }
};
struct SeedW {
    int f;
    long idx;
    __device__ __forceinline__ SeedPtr operator()(int ff, const double *p) const {
        return SeedPtr{p, ff == f ? idx : -1};
    }
};

// ---- vector-Jacobian product of a kernel -------------------------------------
// Output adjoints of a cell kernel: NOUT arrays, indexed in the interior
// layout (r3) or the full-field layout (ix); null = not an output.
struct OutBars {
    const double *p[4];
    int full;
};
// Input adjoints, one per field input of the cell (null = not wanted).
template <int NF>
struct InBars {
    double *p[NF];
};

// Cells (i, j, k), i in [1, n1m], j in [jlo, jhi], k in [1, n3m]; the cell's
// eval(w, i, j, k, a, b, c, d) gives up to four outputs.  flag[0] is raised
// if a cell reads more than AD_MAXIN distinct inputs.
template <class Cell>
__global__ void __launch_bounds__(64)
k_cell_vjp(Cell cell, int jlo, int jhi, OutBars ob, InBars<Cell::NF> ib, int *flag) {
    const Grid &g = cell.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y + jlo;
    const int k = blockIdx.z + 1;
    if (i > g.n1m || j > jhi) return;
    const size_t q = ob.full ? g.ix(i, j, k) : g.r3(i, j, k);
    double w[4];
    bool any = false;
#pragma unroll
    for (int o = 0; o < 4; o++) {
        w[o] = ob.p[o] ? ob.p[o][q] : 0.0;
        any = any || w[o] != 0.0;
    }
    if (!any) return;
    Rec rec;
    rec.n = 0;
    {
        double a, b, c, d = 0.0;
        cell.template eval<double>(RecW{&rec}, i, j, k, a, b, c, d);
    }
    if (rec.n > AD_MAXIN) {
        atomicOr(flag, 1);
        return;
    }
    for (int r = 0; r < rec.n; r++) {
        double *dst = ib.p[rec.fid[r]];
        if (!dst) continue;
        Dual a, b, c, d;
        cell.template eval<Dual>(SeedW{rec.fid[r], rec.idx[r]}, i, j, k, a, b, c, d);
        const double s = a.d * w[0] + b.d * w[1] + c.d * w[2] + d.d * w[3];
        if (s != 0.0) atomicAdd(dst + rec.idx[r], s);
    }
}

}  // namespace tcs
