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

#include <type_traits>

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

// Forward-mode numbers with K derivative components: one evaluation gives K
// Jacobian columns (K seeded inputs), the value arithmetic done once.
template <int K>
struct DualK {
    double v, d[K];
    __host__ __device__ DualK() : v(0.0) {
        for (int q = 0; q < K; q++) d[q] = 0.0;
    }
    __host__ __device__ DualK(double x) : v(x) {
        for (int q = 0; q < K; q++) d[q] = 0.0;
    }
};
template <int K>
__device__ __forceinline__ DualK<K> operator-(const DualK<K> &a) {
    DualK<K> r(-a.v);
#pragma unroll
    for (int q = 0; q < K; q++) r.d[q] = -a.d[q];
    return r;
}
template <int K>
__device__ __forceinline__ DualK<K> operator+(const DualK<K> &a, const DualK<K> &b) {
    DualK<K> r(a.v + b.v);
#pragma unroll
    for (int q = 0; q < K; q++) r.d[q] = a.d[q] + b.d[q];
    return r;
}
template <int K>
__device__ __forceinline__ DualK<K> operator-(const DualK<K> &a, const DualK<K> &b) {
    DualK<K> r(a.v - b.v);
#pragma unroll
    for (int q = 0; q < K; q++) r.d[q] = a.d[q] - b.d[q];
    return r;
}
template <int K>
__device__ __forceinline__ DualK<K> operator*(const DualK<K> &a, const DualK<K> &b) {
    DualK<K> r(a.v * b.v);
#pragma unroll
    for (int q = 0; q < K; q++) r.d[q] = a.d[q] * b.v + a.v * b.d[q];
    return r;
}
template <int K>
__device__ __forceinline__ DualK<K> operator/(const DualK<K> &a, const DualK<K> &b) {
    DualK<K> r(a.v / b.v);
#pragma unroll
    for (int q = 0; q < K; q++) r.d[q] = (a.d[q] - r.v * b.d[q]) / b.v;
    return r;
}
template <int K>
__device__ __forceinline__ DualK<K> operator+(const DualK<K> &a, double b) {
    DualK<K> r = a;
    r.v = a.v + b;
    return r;
}
template <int K>
__device__ __forceinline__ DualK<K> operator+(double a, const DualK<K> &b) {
    DualK<K> r = b;
    r.v = a + b.v;
    return r;
}
template <int K>
__device__ __forceinline__ DualK<K> operator-(const DualK<K> &a, double b) {
    DualK<K> r = a;
    r.v = a.v - b;
    return r;
}
template <int K>
__device__ __forceinline__ DualK<K> operator-(double a, const DualK<K> &b) {
    DualK<K> r(a - b.v);
#pragma unroll
    for (int q = 0; q < K; q++) r.d[q] = -b.d[q];
    return r;
}
template <int K>
__device__ __forceinline__ DualK<K> operator*(const DualK<K> &a, double b) {
    DualK<K> r(a.v * b);
#pragma unroll
    for (int q = 0; q < K; q++) r.d[q] = a.d[q] * b;
    return r;
}
template <int K>
__device__ __forceinline__ DualK<K> operator*(double a, const DualK<K> &b) {
    DualK<K> r(a * b.v);
#pragma unroll
    for (int q = 0; q < K; q++) r.d[q] = a * b.d[q];
    return r;
}
template <int K>
__device__ __forceinline__ DualK<K> operator/(const DualK<K> &a, double b) {
    DualK<K> r(a.v / b);
#pragma unroll
    for (int q = 0; q < K; q++) r.d[q] = a.d[q] / b;
    return r;
}
template <int K>
__device__ __forceinline__ DualK<K> operator/(double a, const DualK<K> &b) {
    DualK<K> r(a / b.v);
#pragma unroll
    for (int q = 0; q < K; q++) r.d[q] = -r.v * b.d[q] / b.v;
    return r;
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
// A value type of the adjoint's forward-mode passes (neither double nor
// float).  The stage rows keep their zero-weighted wall terms for these (a
// branch per row made the vector-Jacobian kernels spill: adjoint step 39.9
// -> 42.7 ms at 128^3); the inputs they record in double skip them.
template <class S>
constexpr bool is_dual() { return !std::is_same<S, double>::value && !std::is_same<S, float>::value; }

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
// inputs whose Jacobian columns one forward-mode evaluation carries
#ifndef AD_K_DEFAULT
#define AD_K_DEFAULT 4
#endif
constexpr int AD_K = AD_K_DEFAULT;
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
// The same set kept in LDS for the 64 lanes of a wavefront (k_cell_vjp):
// one 32-bit key (field << 28 | index) per entry, a lane's entry r at
// key[64 r]; the scan for repeats reads LDS instead of private memory.
// Fields < 16 and indexes < 2^28 (cell_vjp checks the arrays' sizes): 16 KB
// of LDS a wavefront, so the LDS holds as many wavefronts as the registers
// allow (64-bit keys, 32 KB, held five a CU).
constexpr int AD_MAXIN_L = 64;
constexpr int AD_KEY_FBITS = 4, AD_KEY_IBITS = 28;
struct RecL {
    unsigned *key;
    int n;
    __device__ void add(int f, long i) {
        const unsigned k = ((unsigned)f << AD_KEY_IBITS) | (unsigned)i;
        for (int r = 0; r < n && r < AD_MAXIN_L; r++)
            if (key[64 * r] == k) return;
        if (n < AD_MAXIN_L) key[64 * n] = k;
        n++;  // n > AD_MAXIN_L reports an overflow
    }
    __device__ __forceinline__ int fid(int r) const { return (int)(key[64 * r] >> AD_KEY_IBITS); }
    __device__ __forceinline__ long idx(int r) const { return (long)(key[64 * r] & ((1u << AD_KEY_IBITS) - 1)); }
};
template <class R>
struct RecPtrT {
    const double *p;
    int f;
    R *rec;
    __device__ __forceinline__ double operator[](size_t i) const {
        rec->add(f, (long)i);
        return p[i];
    }
};
template <class R>
struct RecWT {
    R *rec;
    __device__ __forceinline__ RecPtrT<R> operator()(int f, const double *p) const { return RecPtrT<R>{p, f, rec}; }
};
using RecPtr = RecPtrT<Rec>;
using RecW = RecWT<Rec>;

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

// K inputs seeded at once (field f[q], index idx[q] carries derivative
// component q; f[q] < 0: component q unused)
template <int K>
struct SeedKPtr {
    const double *p;
    long idx[K];
    __device__ __forceinline__ DualK<K> operator[](size_t i) const {
        DualK<K> r(p[i]);
#pragma unroll
        for (int q = 0; q < K; q++) r.d[q] = (long)i == idx[q] ? 1.0 : 0.0;
        return r;
    }
};
template <int K>
struct SeedKW {
    int f[K];
    long idx[K];
    __device__ __forceinline__ SeedKPtr<K> operator()(int ff, const double *p) const {
        SeedKPtr<K> s;
        s.p = p;
#pragma unroll
        for (int q = 0; q < K; q++) s.idx[q] = ff == f[q] ? idx[q] : -1;
        return s;
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

// *p += s from every lane, with the FP64 atomic add of the L2 (unsafeAtomicAdd:
// HIP's atomicAdd on a double is a compare-and-swap loop, which neighbouring
// cells adding to shared inputs kept retrying: 224 -> 11 ms per adjoint step
// at 128^3).  The lanes of a wavefront are 64 cells of one row (j, k); when
// all 64 are here and add to one address (an input every cell of the row
// reads: a (j, k) table entry, a stale scalar of the momentum stages) their
// sum is added once.  Otherwise each lane adds its own (skipping zeros).
// The adjoint arrays are device allocations (coarse-grained), which the
// L2's FP64 atomics serve.
__device__ __forceinline__ void add_to(double *p, double s) {
    const unsigned long long full = ~0ull;
    const unsigned long long first = (unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane((int)(size_t)p) |
                                     ((unsigned long long)(unsigned)__builtin_amdgcn_readfirstlane((int)((size_t)p >> 32)) << 32);
    const bool all_here = __builtin_amdgcn_read_exec() == full;
    if (all_here && __ballot((size_t)p == first) == full) {
#pragma unroll
        for (int o = 32; o >= 1; o >>= 1) s += __shfl_xor(s, o);
        if (threadIdx.x % 64 == 0 && s != 0.0) unsafeAtomicAdd(p, s);
    } else if (s != 0.0) {
        unsafeAtomicAdd(p, s);
    }
}

// Cells (i, j, k), i in [1, n1m], j in [jlo, jhi], k in [1, n3m]; the cell's
// eval(w, i, j, k, a, b, c, d) gives up to four outputs.  flag[0] is raised
// if a cell reads more than AD_MAXIN_L distinct inputs.
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
    static_assert(Cell::NF <= (1 << AD_KEY_FBITS), "input field ids must fit the LDS keys");
    __shared__ unsigned keys[64 * AD_MAXIN_L];
    RecL rec{keys + threadIdx.x, 0};
    {
        double a, b, c, d = 0.0;
        cell.template eval<double>(RecWT<RecL>{&rec}, i, j, k, a, b, c, d);
    }
    if (rec.n > AD_MAXIN_L) {
        atomicOr(flag, 1);
        return;
    }
    // AD_K inputs a pass: their Jacobian columns from one evaluation
    for (int r0 = 0; r0 < rec.n; r0 += AD_K) {
        SeedKW<AD_K> sw;
#pragma unroll
        for (int q = 0; q < AD_K; q++) {
            const int r = r0 + q;
            const bool in = r < rec.n && ib.p[rec.fid(r < rec.n ? r : 0)];
            sw.f[q] = in ? rec.fid(r) : -1;
            sw.idx[q] = in ? rec.idx(r) : -1;
        }
        DualK<AD_K> a, b, c, d;
        cell.template eval<DualK<AD_K>>(sw, i, j, k, a, b, c, d);
#pragma unroll
        for (int q = 0; q < AD_K; q++) {
            if (sw.f[q] < 0) continue;
            const double s = a.d[q] * w[0] + b.d[q] * w[1] + c.d[q] * w[2] + d.d[q] * w[3];
            add_to(ib.p[sw.f[q]] + sw.idx[q], s);
        }
    }
}

}  // namespace tcs
