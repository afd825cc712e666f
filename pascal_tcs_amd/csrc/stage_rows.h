// stage_rows.h -- how the ADI stages' tridiagonal rows reach the line solves.
//
// Every stage of the thermal and momentum solvers has a row functor:
// row(i, j, k, a, b, c, d) computes row (i, j, k) of the stage's systems
// (a, b, c: lower, diagonal and upper coefficient; d: the right-hand side,
// first stage only) from the module fields.  The rows go to the solve in one
// of two ways:
//
//   k_rows3d           one thread per interior cell writes the rows into
//                      the ADI scratch (am, ac, ap, rhs); a sweep then reads
//                      them back (single-rank line solves);
//   k_ptdma_part_src,  the split line solve computes each row where it needs
//   k_ptdma_part_rows_src  it (z and y lines: the thread of a part makes its
//                      own rows; x lines: the workgroup makes them block by
//                      block and passes them through LDS): no coefficient
//                      array is written to or read from memory.
//
// The rows are the same double values either way and the elimination is the
// split solve's, so both ways give bit-identical solutions.
#pragma once

#include <type_traits>

#include "ad.h"
#include "tdma_part.h"

namespace tcs {

// Row (i, j, k) of a stage in the arithmetic of the line solve's type S:
// FP64 solves take the rows as the step computes them; FP32 solves (the
// mixed form) make them in FP32 from the FP32 twins of the arrays they read
// (tw), with the metrics and constants rounded to float (the cell
// templates' sc<S> and metric<S>).  d is the row's right-hand side where
// the stage has one, else 0.
// WM >= 0 passes a wall-term mode to rows that have one (the z rows:
// 0 computes the wall terms, 2 leaves them out where wall_free(j) holds).
template <typename S, int WM = -1, class Row>
__device__ __forceinline__ void row_as(const Row &row, const TwinW &tw, int i, int j, int k, S &a, S &b, S &c,
                                       S &d) {
    if constexpr (std::is_same<S, float>::value) {
        d = 0.0f;
        if constexpr (WM >= 0) row.template eval<float, WM>(tw, i, j, k, a, b, c, d);
        else row.template eval<float>(tw, i, j, k, a, b, c, d);
    } else {
        double ta, tb, tc, td = 0.0;
        if constexpr (WM >= 0) row.template eval<double, WM>(RawW(), i, j, k, ta, tb, tc, td);
        else row(i, j, k, ta, tb, tc, td);
        a = (S)ta; b = (S)tb; c = (S)tc; d = (S)td;
    }
}

// One thread per interior (i, j) (i fastest, rows j >= jlo) and ROWS3D_KB
// consecutive k (dividing n3m, a power of two >= 4): row (i, j, k) of the
// stage's systems into am, ac, ap and, with RHS, its right-hand side.  The
// planes of a thread share their loads and the terms that do not depend on
// k (thermal x: the conductivity at the fixed k planes; 1.92 -> 1.47 ms at
// 512^3).
constexpr int ROWS3D_KB = 4;
constexpr int ROWS3D_IB = 1;  // consecutive i per thread (2: thermal x 1.53 -> 2.05 ms)
template <class Row, bool RHS, typename S>
__global__ void __launch_bounds__(256)
k_rows3d(Row row, TwinW tw, int jlo, S *__restrict__ am, S *__restrict__ ac, S *__restrict__ ap, S *__restrict__ rhs) {
    const int i0 = (blockIdx.x * blockDim.x + threadIdx.x) * ROWS3D_IB + 1;
    const int j = wave_row_j();
    if (i0 > row.g.n1m || j > row.g.n2m || j < jlo) return;
#pragma unroll
    for (int q = 0; q < ROWS3D_KB * ROWS3D_IB; q++) {
        const int i = i0 + q % ROWS3D_IB, k = blockIdx.z * ROWS3D_KB + q / ROWS3D_IB + 1;
        S a, b, c, d;
        row_as<S>(row, tw, i, j, k, a, b, c, d);
        const size_t o = row.g.r3(i, j, k);
        ac[o] = b;
        ap[o] = c;
        am[o] = a;
        if (RHS) rhs[o] = d;
    }
}

template <bool RHS, class Row, typename S>
int rows3d_launch(tcs_ctx *c, const char *label, const Row &row, int jlo, S *am, S *ac, S *ap, S *rhs) {
    const dim3 blk = interior_block(), grd = cell_grid(row.g.n1m, row.g.n2m, row.g.n3m, ROWS3D_KB, interior_block(), ROWS3D_IB);
    TCS_LAUNCH(c->stream, label, (k_rows3d<Row, RHS, S>), grd, blk, 0, row, twin_w(c), jlo, am, ac, ap, rhs);
    return TCS_OK;
}

// Occupancy and row loop of a stage's FP32 line solves (the mixed form):
// Row::W32 waves per SIMD at least (1: as many registers as the compiler
// takes, two waves per SIMD for these kernels) and Row::FULL32 for the
// unconditional row loop.  Stages that set neither keep 1 and the guarded
// loop.
template <class Row, class = void>
struct RowF32 {
    static constexpr int waves = 1;
    static constexpr bool full = false;
};
template <class Row>
struct RowF32<Row, std::void_t<decltype(Row::W32)>> {
    static constexpr int waves = Row::W32;
    static constexpr bool full = Row::FULL32;
};
template <class Row, typename S>
constexpr int line_waves() { return std::is_same<S, float>::value ? RowF32<Row>::waves : 1; }
// Lines per workgroup of a stage's FP32 z lines: Row::SX32 where set (32:
// 128-byte row segments at four waves per SIMD; 8: small workgroups, three
// of them a CU at three waves per SIMD), else PT_SX.
template <class Row, class = void>
struct RowSX32 {
    static constexpr int value = PT_SX;
};
template <class Row>
struct RowSX32<Row, std::void_t<decltype(Row::SX32)>> {
    static constexpr int value = Row::SX32;
};
// Lines per workgroup of a stage's FP64 z lines: Row::SX64 where set, else
// PT_SX.
template <class Row, class = void>
struct RowSX64 {
    static constexpr int value = PT_SX;
};
template <class Row>
struct RowSX64<Row, std::void_t<decltype(Row::SX64)>> {
    static constexpr int value = Row::SX64;
};

// Row::WALLS32: the row's FP32 z lines split their workgroups by the wall
// terms as the FP64 lines do (unconditional row loop rows only).
template <class Row, class = void>
struct RowWalls32 {
    static constexpr bool value = false;
};
template <class Row>
struct RowWalls32<Row, std::void_t<decltype(Row::WALLS32)>> {
    static constexpr bool value = Row::WALLS32;
};

// z lines as a split-solve source: line (s0, s1) = (i-1, j-jlo), row k-1;
// the right-hand side comes with the row.
template <class Row, typename S>
struct ZLineRows {
    // the split solve's unconditional row loop where the row asks for it
    // (measured per stage at 512^3: dU z 4.14 -> 3.14 ms; the thermal z rows
    // spill in it)
    static constexpr bool full_rows = Row::FULL;
    static constexpr bool full_rows32 = RowF32<Row>::full;
    static constexpr int waves = line_waves<Row, S>();
    // FP64 lines of the rows that take the unconditional row loop (and FP32
    // lines where the row sets WALLS32): the lines of one j (one workgroup)
    // away from the walls make their rows without the wall terms, the others
    // with them (dU z 3.17 -> 2.52, dV z 3.72 -> 2.71, dW z 3.78 -> 2.72 ms
    // at 512^3, dV and dW now in that loop).  The thermal z lines, which
    // spill with the second loop, and the other FP32 lines keep one loop and
    // their row's own wall-term mode.
    static constexpr bool walls =
        Row::FULL && (std::is_same<S, double>::value || (RowF32<Row>::full && RowWalls32<Row>::value));
    Row row;
    TwinW tw;
    int jlo;
    __device__ __forceinline__ void operator()(int s0, int s1, int r, S &a, S &b, S &c, S &d) const {
        row_as<S>(row, tw, s0 + 1, s1 + jlo, r + 1, a, b, c, d);
    }
    __device__ __forceinline__ bool wall_free(int s1) const { return row.wall_free(s1 + jlo); }
    template <int WM>
    __device__ __forceinline__ void rowm(int s0, int s1, int r, S &a, S &b, S &c, S &d) const {
        row_as<S, WM>(row, tw, s0 + 1, s1 + jlo, r + 1, a, b, c, d);
    }
};

// y lines as a split-solve source: line (s0, s1) = (i-1, k-1), row j-jlo;
// the right-hand side is the previous stage's solution in the scratch.
template <class Row, typename S>
struct YLineRows {
    static constexpr bool full_rows = true;  // the unconditional row loop pays for every y stage
    static constexpr bool full_rows32 = RowF32<Row>::full;
    static constexpr int waves = line_waves<Row, S>();
    Row row;
    TwinW tw;
    const S *rhs;
    int jlo;
    __device__ __forceinline__ void operator()(int s0, int s1, int r, S &a, S &b, S &c, S &d) const {
        const int i = s0 + 1, j = r + jlo, k = s1 + 1;
        row_as<S>(row, tw, i, j, k, a, b, c, d);
        d = rhs[row.g.r3(i, j, k)];
    }
};

// y lines whose rows j < jrow are identity rows (a = c = d = 0, b = 1), the
// rest as in YLineRows.  The split dV lines start at the wall row j = 1,
// where dV is zero, so that the 2^m - 1 rows j = 2..n2m become 2^m rows in
// whole parts.  Every row is made and the identity selected over it (the
// wall row's stencil reads ghost cells and may divide by the zero ghost
// width; the selection drops that), so the unconditional row loop stays
// branch-free.
template <class Row, typename S>
struct YLineRowsLead {
    static constexpr bool full_rows = true;
    static constexpr bool full_rows32 = RowF32<Row>::full;
    static constexpr int waves = line_waves<Row, S>();
    Row row;
    TwinW tw;
    const S *rhs;
    int jlo, jrow;
    __device__ __forceinline__ void operator()(int s0, int s1, int r, S &a, S &b, S &c, S &d) const {
        const int i = s0 + 1, j = r + jlo, k = s1 + 1;
        row_as<S>(row, tw, i, j, k, a, b, c, d);
        d = rhs[row.g.r3(i, j, k)];
        const bool wall = j < jrow;
        a = wall ? S(0) : a;
        b = wall ? S(1) : b;
        c = wall ? S(0) : c;
        d = wall ? S(0) : d;
    }
};

// x lines as a split-solve source: line (s0, s1) = (j-jlo, k-1), row i-1;
// the right-hand side is the previous stage's solution in the scratch.
template <class Row, typename S>
struct XLineRows {
    Row row;
    TwinW tw;
    const S *rhs;
    int jlo;
    __device__ __forceinline__ void operator()(int s0, int s1, int r, S &a, S &b, S &c, S &d) const {
        const int i = r + 1, j = s0 + jlo, k = s1 + 1;
        row_as<S>(row, tw, i, j, k, a, b, c, d);
        d = rhs[row.g.r3(i, j, k)];
    }
};

// Whether a stage's lines of n_row rows are solved with their rows computed
// in the solve: split solves are asked for (the mixed form always splits)
// and the lines are short enough to split.
template <typename S>
inline bool rows_in_solve(const tcs_ctx *c, int n_row) {
    int P, w1, w2;
    return (c->tdma_split || sizeof(S) == sizeof(float)) && ptdma_part_plan(n_row, P, w1, w2);
}

// z stage over the z lines of rows j >= jlo, solution into the scratch rhs:
// rows computed in the split solve (label_solve), or written by k_rows3d
// (label_rows) and swept.
template <class Row, typename S>
int z_stage(tcs_ctx *c, const char *label_rows, const char *label_solve, const Row &row, int jlo, S *am, S *ac,
            S *ap, S *rhs) {
    const long n1 = c->n1m, n12 = (long)c->n1m * c->n2m, off = (long)(jlo - 1) * n1;
    constexpr int SX = std::is_same<S, float>::value ? RowSX32<Row>::value : RowSX64<Row>::value;
    if (rows_in_solve<S>(c, c->n3m))
        return ptdma_part_src_launch<true, SX>(c->stream, label_solve, ZLineRows<Row, S>{row, twin_w(c), jlo},
                                               rhs + off, c->n1m, c->n2m - jlo + 1, n1, c->n3m, n12);
    TCS_TRY(rows3d_launch<true>(c, label_rows, row, jlo, am, ac, ap, rhs));
    return ptdma_sweep(c, am + off, ac + off, ap + off, rhs + off, c->n1m, 1, c->n2m - jlo + 1, n1, c->n3m, n12,
                       true);
}

// x stage over the x lines of rows j >= jlo (right-hand side and solution in
// the scratch rhs), in either of the two ways.
template <class Row, typename S>
int x_stage(tcs_ctx *c, const char *label_rows, const char *label_solve, const Row &row, int jlo, S *am, S *ac,
            S *ap, S *rhs) {
    const long n1 = c->n1m, n12 = (long)c->n1m * c->n2m, off = (long)(jlo - 1) * n1;
    const int nl = c->n2m - jlo + 1;
    if (rows_in_solve<S>(c, c->n1m))
        return ptdma_part_rows_src_launch<true>(c->stream, label_solve, XLineRows<Row, S>{row, twin_w(c), rhs, jlo},
                                                rhs + off, nl, n1, c->n3m, n12, c->n1m);
    TCS_TRY(rows3d_launch<false>(c, label_rows, row, jlo, am, ac, ap, rhs));
    return ptdma_sweep(c, am + off, ac + off, ap + off, rhs + off, nl, n1, c->n3m, n12, c->n1m, 1, true);
}

// y stage with its rows computed in the split solve (rows_in_solve holds):
// the y lines of rows j = jlo..jlo+nrows-1, right-hand side from the scratch
// rhs.  The solution is stored into (mode 1) or added to (mode 2) the
// interior of the field F, or, with F null, written back into rhs.  Rows
// j < jrow are identity rows (YLineRowsLead) when jrow > jlo.
template <int SX = PT_SX, class Row, typename S>
int y_stage_in_solve(tcs_ctx *c, const char *label, const Row &row, int jlo, int nrows, S *rhs, double *F,
                     int mode, int jrow = 0) {
    const long n1 = c->n1m, n12 = (long)c->n1m * c->n2m, off = (long)(jlo - 1) * n1;
    LineOut out;
    if (F) {
        out.F = F;
        out.mode = mode;
        out.base = 1 + (long)c->P1 * jlo + (long)c->P1 * c->N2;  // (i, j, k) = (1, jlo, 1)
        out.ss0 = 1;
        out.ss1 = (long)c->P1 * c->N2;
        out.rs = c->P1;
        if constexpr (sizeof(S) == sizeof(float)) {
            // the mixed form stores its increments into their FP32 twins
            out.G = twin(c, F);
            if (mode != 1 || !out.G) {
                set_error("y stage: the FP32 lines store increments into their FP32 twins only");
                return TCS_E_ARG;
            }
            out.F = nullptr;
            out.mode = 4;
        }
    }
    if (jrow > jlo)
        return ptdma_part_src_launch<false, SX>(c->stream, label, YLineRowsLead<Row, S>{row, twin_w(c), rhs, jlo, jrow},
                                                rhs + off, c->n1m, c->n3m, n12, nrows, n1, out);
    return ptdma_part_src_launch<false, SX>(c->stream, label, YLineRows<Row, S>{row, twin_w(c), rhs, jlo}, rhs + off,
                                            c->n1m, c->n3m, n12, nrows, n1, out);
}

}  // namespace tcs
