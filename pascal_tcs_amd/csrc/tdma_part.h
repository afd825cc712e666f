// tdma_part.h -- device building blocks of the split-line tridiagonal solve
// (PaScaL_TDMA's multi-rank algorithm with one GPU thread per rank), shared
// by the ADI sweeps (tdma_part.hip) and the Poisson y-solve (pressure.hip).
//
// A line of n_row <= 512 rows is split into P parts (P a power of two, rows
// per part by para_range, PaScaL_TDMA/src/para_range.f90:26-40); the thread
// of part p holds its m <= PT_MMAX rows in registers.  NR right-hand sides
// may share one matrix (the Poisson solve's real and imaginary parts): the
// matrix recurrences run once, every right-hand side gets exactly the
// operations a separate solve would give it.
#pragma once

#include <algorithm>

#include "tcs_internal.h"

namespace tcs {

constexpr int PT_MMAX = 16;               // rows per part
constexpr int PT_PMAX = 32;               // parts per line (n_row <= 512)
constexpr int PT_RS = 2 * PT_PMAX + 1;    // reduced-row stride in LDS (odd: no bank conflicts)
constexpr int PT_SX = 16;                 // lines per workgroup

namespace part {

__device__ __forceinline__ void range(int p, int w1, int w2, int &j0, int &m) {
    j0 = p * w1 + (p < w2 ? p : w2);
    m = w1 + (p < w2 ? 1 : 0);
}

// One step of the per-rank modified Thomas algorithm of
// pascal_tdma.f90:501-536 on a part held in registers: row r (coefficients
// ar, br, cr, dr) enters the forward elimination, rows < r already hold
// their eliminated form.  Rows 0 and 1 are only normalised.
template <typename T, int NR>
__device__ __forceinline__ void reduce_row(int r, T ar, T br, T cr, const T (&dr)[NR], T (&a)[PT_MMAX],
                                           T (&c)[PT_MMAX], T (&d)[NR][PT_MMAX]) {
    if (r < 2) {
        a[r] = ar / br;
#pragma unroll
        for (int k = 0; k < NR; k++) d[k][r] = dr[k] / br;
        c[r] = cr / br;
    } else {
        const T rr = T(1) / (br - ar * c[r - 1]);
#pragma unroll
        for (int k = 0; k < NR; k++) d[k][r] = rr * (dr[k] - ar * d[k][r - 1]);
        c[r] = rr * cr;
        a[r] = -rr * ar * a[r - 1];
    }
}

// The rest of the per-rank algorithm once rows 0..m-1 have been through
// reduce_row: backward elimination and the fix-up of row 0.  a, c, d come
// back as the reduced-form coefficients.
template <typename T, int NR>
__device__ __forceinline__ void reduce_back(T (&a)[PT_MMAX], T (&c)[PT_MMAX], T (&d)[NR][PT_MMAX], int m) {
#pragma unroll
    for (int r = PT_MMAX - 3; r >= 1; r--) {
        if (r <= m - 3) {
#pragma unroll
            for (int k = 0; k < NR; k++) d[k][r] = d[k][r] - c[r] * d[k][r + 1];
            a[r] = a[r] - c[r] * a[r + 1];
            c[r] = -c[r] * c[r + 1];
        }
    }
    const T rr = T(1) / (T(1) - a[1] * c[0]);
#pragma unroll
    for (int k = 0; k < NR; k++) d[k][0] = rr * (d[k][0] - c[0] * d[k][1]);
    a[0] = rr * a[0];
    c[0] = -rr * c[0] * c[1];
}

// One part (rows 0..m-1, m >= 2) in registers, all rows loaded: b is
// consumed; a, c, d come back as the reduced-form coefficients.
template <typename T, int NR>
__device__ __forceinline__ void reduce(T (&a)[PT_MMAX], T (&b)[PT_MMAX], T (&c)[PT_MMAX],
                                       T (&d)[NR][PT_MMAX], int m) {
#pragma unroll
    for (int r = 0; r < PT_MMAX; r++) {
        if (r < m) {
            T dr[NR];
#pragma unroll
            for (int k = 0; k < NR; k++) dr[k] = d[k][r];
            reduce_row<T, NR>(r, a[r], b[r], c[r], dr, a, c, d);
        }
    }
    reduce_back<T, NR>(a, c, d, m);
}

// a, c, d at row m-1 (a runtime index) in one unrolled pass: selecting all
// of them together keeps the arrays in registers (a separate select per
// array lets the compiler turn it into a dynamically indexed private array)
template <typename T, int NR>
__device__ __forceinline__ void lasts(const T (&a)[PT_MMAX], const T (&c)[PT_MMAX], const T (&d)[NR][PT_MMAX],
                                      int m, T &la, T &lc, T (&ld)[NR]) {
    la = a[0];
    lc = c[0];
#pragma unroll
    for (int k = 0; k < NR; k++) ld[k] = d[k][0];
#pragma unroll
    for (int r = 1; r < PT_MMAX; r++) {
        if (r == m - 1) {
            la = a[r];
            lc = c[r];
#pragma unroll
            for (int k = 0; k < NR; k++) ld[k] = d[k][r];
        }
    }
}

// x(r) of right-hand side k from the part's reduced solution xf (row 0)
// and xl (row m-1): pascal_tdma.f90:567-576
template <typename T, int NR>
__device__ __forceinline__ T x_of(const T (&a)[PT_MMAX], const T (&c)[PT_MMAX], const T (&d)[NR][PT_MMAX],
                                  int k, int r, int m, T xf, T xl) {
    if (r == 0) return xf;
    if (r == m - 1) return xl;
    return d[k][r] - a[r] * xf - c[r] * xl;
}

// Reduced arrays of a workgroup, [array][line][row]: a, c, b (or the e
// scratch of the serial cyclic solve), then the NR right-hand sides.
template <typename T, int SX, int NR>
struct RedLDS {
    T v[(3 + NR) * SX * PT_RS];
    __device__ T *q(int which, int line) { return v + ((size_t)which * SX + line) * PT_RS; }
};

// The 2-row reduced system of a single part, serially: tdma_many
// (tdmas.f90:113-146) or tdma_cycl_many (tdmas.f90:159-212, re its e
// scratch), the unit diagonal's divisions left out (exact).
template <typename T, bool CYC, int NR>
__device__ void reduced_serial(T *ra, T *rc, T *re, T *const (&rd)[NR], int n2) {
    if (!CYC) {
        for (int j = 1; j < n2; j++) {
            const T r = T(1) / (T(1) - ra[j] * rc[j - 1]);
#pragma unroll
            for (int k = 0; k < NR; k++) rd[k][j] = r * (rd[k][j] - ra[j] * rd[k][j - 1]);
            rc[j] = r * rc[j];
        }
        for (int j = n2 - 2; j >= 0; j--) {
#pragma unroll
            for (int k = 0; k < NR; k++) rd[k][j] = rd[k][j] - rc[j] * rd[k][j + 1];
        }
    } else {
        for (int j = 0; j < n2; j++) re[j] = T(0);
        re[1] = -ra[1];
        re[n2 - 1] = -rc[n2 - 1];
        for (int j = 2; j < n2; j++) {
            const T rr = T(1) / (T(1) - ra[j] * rc[j - 1]);
#pragma unroll
            for (int k = 0; k < NR; k++) rd[k][j] = rr * (rd[k][j] - ra[j] * rd[k][j - 1]);
            re[j] = rr * (re[j] - ra[j] * re[j - 1]);
            rc[j] = rr * rc[j];
        }
        for (int j = n2 - 2; j >= 1; j--) {
#pragma unroll
            for (int k = 0; k < NR; k++) rd[k][j] = rd[k][j] - rc[j] * rd[k][j + 1];
            re[j] = re[j] - rc[j] * re[j + 1];
        }
#pragma unroll
        for (int k = 0; k < NR; k++) {
            rd[k][0] = (rd[k][0] - ra[0] * rd[k][n2 - 1] - rc[0] * rd[k][1]) /
                       (T(1) + ra[0] * re[n2 - 1] + rc[0] * re[1]);
            for (int j = 1; j < n2; j++) rd[k][j] = rd[k][j] + rd[k][0] * re[j];
        }
    }
}

// After reduce(): publish the part's two reduced rows (2p, 2p+1 of the
// line's reduced system), solve the reduced system, return the part's
// solution at its first (xf) and last (xl) row for every right-hand side.
//
// One part (P = 1): the reference's serial tdma_many / tdma_cycl_many on the
// 2-row system.  P >= 2 (a power of two): parallel cyclic reduction by the
// line's P threads, two rows each.  Step with stride s: row i eliminates its
// couplings to rows i-s and i+s with their current values,
//   k1 = a_i/b_{i-s}, k2 = c_i/b_{i+s}, a_i' = -a_{i-s} k1, c_i' = -c_{i+s} k2,
//   b_i' = b_i - c_{i-s} k1 - a_{i+s} k2, d_i' = d_i - d_{i-s} k1 - d_{i+s} k2;
// outside a non-cyclic system the rows are identity rows.  Non-cyclic:
// strides 1..N/2 (N = 2P) and x_i = d_i/b_i.  Cyclic: strides 1..N/4, then
// row i couples only to j = i + N/2 and the pair is solved directly,
// x_i = (d_i b_j - e_i d_j)/(b_i b_j - e_i e_j), e = a + c.
// Every thread of the workgroup must call this (it synchronises).
template <typename T, bool CYC, int SX, int NR>
__device__ __forceinline__ void exchange(RedLDS<T, SX, NR> &red, int line, int p, int P, const T (&a)[PT_MMAX],
                                         const T (&c)[PT_MMAX], const T (&d)[NR][PT_MMAX], int m, T (&xf)[NR],
                                         T (&xl)[NR]) {
    T *ra = red.q(0, line), *rc = red.q(1, line), *rb = red.q(2, line);
    T *rd[NR];
#pragma unroll
    for (int k = 0; k < NR; k++) rd[k] = red.q(3 + k, line);
    T la, lc, ld[NR];
    lasts<T, NR>(a, c, d, m, la, lc, ld);
    ra[2 * p] = a[0]; rc[2 * p] = c[0];
    ra[2 * p + 1] = la; rc[2 * p + 1] = lc;
#pragma unroll
    for (int k = 0; k < NR; k++) { rd[k][2 * p] = d[k][0]; rd[k][2 * p + 1] = ld[k]; }
    if (P == 1) {
        __syncthreads();
        if (p == 0) reduced_serial<T, CYC, NR>(ra, rc, rb, rd, 2);
        __syncthreads();
#pragma unroll
        for (int k = 0; k < NR; k++) { xf[k] = rd[k][0]; xl[k] = rd[k][1]; }
        return;
    }
    rb[2 * p] = T(1); rb[2 * p + 1] = T(1);
    __syncthreads();
    const int N = 2 * P, lim = CYC ? N / 2 : N;
    T va[2] = {a[0], la}, vb[2] = {T(1), T(1)}, vc[2] = {c[0], lc}, vd[NR][2];
#pragma unroll
    for (int k = 0; k < NR; k++) { vd[k][0] = d[k][0]; vd[k][1] = ld[k]; }
    for (int s = 1; s < lim; s <<= 1) {
        T na[2], nb[2], nc[2], nd[NR][2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            int im = 2 * p + h - s, ip = 2 * p + h + s;
            T am = T(0), bm = T(1), cm = T(0), ap = T(0), bp = T(1), cp = T(0), dm[NR], dp[NR];
#pragma unroll
            for (int k = 0; k < NR; k++) { dm[k] = T(0); dp[k] = T(0); }
            if (CYC) { im &= N - 1; ip &= N - 1; }
            if (CYC || im >= 0) {
                am = ra[im]; bm = rb[im]; cm = rc[im];
#pragma unroll
                for (int k = 0; k < NR; k++) dm[k] = rd[k][im];
            }
            if (CYC || ip < N) {
                ap = ra[ip]; bp = rb[ip]; cp = rc[ip];
#pragma unroll
                for (int k = 0; k < NR; k++) dp[k] = rd[k][ip];
            }
            const T k1 = va[h] / bm, k2 = vc[h] / bp;
            na[h] = -am * k1;
            nc[h] = -cp * k2;
            nb[h] = vb[h] - cm * k1 - ap * k2;
#pragma unroll
            for (int k = 0; k < NR; k++) nd[k][h] = vd[k][h] - dm[k] * k1 - dp[k] * k2;
        }
        __syncthreads();
#pragma unroll
        for (int h = 0; h < 2; h++) {
            ra[2 * p + h] = na[h]; rb[2 * p + h] = nb[h]; rc[2 * p + h] = nc[h];
            va[h] = na[h]; vb[h] = nb[h]; vc[h] = nc[h];
#pragma unroll
            for (int k = 0; k < NR; k++) { rd[k][2 * p + h] = nd[k][h]; vd[k][h] = nd[k][h]; }
        }
        __syncthreads();
    }
#pragma unroll
    for (int h = 0; h < 2; h++) {
#pragma unroll
        for (int k = 0; k < NR; k++) {
            T x;
            if (CYC) {
                const int j = (2 * p + h + P) & (N - 1);
                const T ei = va[h] + vc[h], ej = ra[j] + rc[j];
                x = (vd[k][h] * rb[j] - ei * rd[k][j]) / (vb[h] * rb[j] - ei * ej);
            } else {
                x = vd[k][h] / vb[h];
            }
            if (h == 0) xf[k] = x;
            else xl[k] = x;
        }
    }
}

// The solution of a part of line (s0, s1): into D at dbase + row*rs
// (out.mode 0), or into the field of out (modes 1, 2, 3).
template <typename T>
__device__ __forceinline__ void store(const T (&a)[PT_MMAX], const T (&c)[PT_MMAX], const T (&d)[1][PT_MMAX],
                                      int j0, int m, T xf, T xl, T *D, long dbase, long rs, const LineOut &out,
                                      int s0, int s1) {
    if constexpr (sizeof(T) == sizeof(float)) {
        if (out.mode == 4) {  // FP32 lines only: the mixed form's increments
            const long fb = out.base + (long)s0 * out.ss0 + (long)s1 * out.ss1;
#pragma unroll
            for (int r = 0; r < PT_MMAX; r++)
                if (r < m) out.G[fb + (long)(j0 + r) * out.rs] = x_of<T, 1>(a, c, d, 0, r, m, xf, xl);
            return;
        }
    }
    if (out.mode == 0) {
#pragma unroll
        for (int r = 0; r < PT_MMAX; r++)
            if (r < m) D[dbase + (long)(j0 + r) * rs] = x_of<T, 1>(a, c, d, 0, r, m, xf, xl);
    } else {
        const long fb = out.base + (long)s0 * out.ss0 + (long)s1 * out.ss1;
#pragma unroll
        for (int r = 0; r < PT_MMAX; r++) {
            if (r < m) {
                const double x = (double)x_of<T, 1>(a, c, d, 0, r, m, xf, xl);
                const long o = fb + (long)(j0 + r) * out.rs;
                const double v = out.mode == 1 ? x : (out.mode == 2 ? out.F[o] : out.Fin[o]) + x;
                out.F[o] = v;
                if constexpr (sizeof(T) == sizeof(float)) {
                    if (out.G) out.G[o] = (float)v;  // the mixed form's FP32 copy of the field
                }
            }
        }
    }
}

}  // namespace part

// Workgroup order for the solves that compute their rows.  Workgroups go
// to the 8 XCDs round-robin in dispatch order, and the rows of a line group
// read the fields of its neighbour groups in both grid directions.  So each
// run of 32 workgroups an XCD receives in a row covers a compact tile of 4 x
// 8 line groups, and consecutive tiles go to consecutive XCDs: neighbouring
// groups mostly run at the same time on the same XCD and share its L2.
// Grids that do not divide into whole tiles keep the plain order.
constexpr int XT_X = 4, XT_Y = 8;
__device__ __forceinline__ void xcd_tiles(int &bx, int &by) {
    const int nbx = gridDim.x, nby = gridDim.y;
    bx = blockIdx.x;
    by = blockIdx.y;
    if (nbx % XT_X != 0 || nby % XT_Y != 0 || (nbx * nby) % (8 * XT_X * XT_Y) != 0) return;
    const int w = blockIdx.x + nbx * blockIdx.y;
    const int xcd = w & 7, q = w >> 3, u = q % (XT_X * XT_Y);
    const int tile = (q / (XT_X * XT_Y)) * 8 + xcd, ntx = nbx / XT_X;
    bx = (tile % ntx) * XT_X + u % XT_X;
    by = (tile / ntx) * XT_Y + u / XT_X;
}

// Row sources whose rows have a wall-term mode (Src::walls: the z lines).
template <class Src, class = void>
struct SrcWalls {
    static constexpr bool value = false;
};
template <class Src>
struct SrcWalls<Src, std::void_t<decltype(Src::walls)>> {
    static constexpr bool value = Src::walls;
};

// Row `row` of line (s0, s1): with WM >= 0 in that wall-term mode.
template <int WM, class Src, typename T>
__device__ __forceinline__ void src_row(const Src &src, int s0, int s1, int row, T &a, T &b, T &c, T &d) {
    if constexpr (WM >= 0) src.template rowm<WM>(s0, s1, row, a, b, c, d);
    else src(s0, s1, row, a, b, c, d);
}

// One workgroup's lines: rows in wall-term mode WM, each entering the
// forward elimination as it is made, then the reduced system and the
// solution.  FULL (every part holds PT_MMAX rows): one unconditional run of
// rows for the active lines, so consecutive rows make the loads and terms
// they share once; otherwise guarded by r < m.  (The row loop in a function
// of its own, taking the part arrays by reference, put them in scratch
// memory.)
template <int WM, typename T, bool CYC, class Src, int SX, bool FULL>
__device__ __forceinline__ void part_src_lines(const Src &src, part::RedLDS<T, SX, 1> &red, T *D, int n0, long ss1, long rs,
                                               int P, int w1, int w2, const LineOut &out, int line, int p, int s0,
                                               int s1) {
    using namespace part;
    const bool act = s0 < n0;
    int j0, m;
    range(p, w1, w2, j0, m);
    T a[PT_MMAX], c[PT_MMAX], d[1][PT_MMAX];
    if (FULL && act) {
#pragma unroll
        for (int r = 0; r < PT_MMAX; r++) {
            T ar, br, cr, dr[1];
            src_row<WM>(src, s0, s1, j0 + r, ar, br, cr, dr[0]);
            reduce_row<T, 1>(r, ar, br, cr, dr, a, c, d);
        }
    } else if (FULL && sizeof(T) == sizeof(float)) {
        // padding lines of FP32 solves: zero rows (the guarded loop below
        // puts the FP32 arrays on the stack in this instantiation; the FP64
        // ones keep it, dU y 1.02 against 1.14 ms with the zero rows)
#pragma unroll
        for (int r = 0; r < PT_MMAX; r++) { a[r] = T(0); c[r] = T(0); d[0][r] = T(0); }
    } else {
#pragma unroll
        for (int r = 0; r < PT_MMAX; r++) {
            if (r < m) {
                T ar = T(0), br = T(1), cr = T(0), dr[1] = {T(0)};
                if (act) src_row<WM>(src, s0, s1, j0 + r, ar, br, cr, dr[0]);
                reduce_row<T, 1>(r, ar, br, cr, dr, a, c, d);
            } else {
                a[r] = T(0); c[r] = T(0); d[0][r] = T(0);
            }
        }
    }
    reduce_back<T, 1>(a, c, d, m);
    T xf[1], xl[1];
    exchange<T, CYC, SX, 1>(red, line, p, P, a, c, d, m, xf, xl);
    if (!act) return;
    store<T>(a, c, d, j0, m, xf[0], xl[0], D, (long)s0 + (long)s1 * ss1, rs, out, s0, s1);
}

// Lines strided in memory whose rows are computed, not read: src(s0, s1,
// row, a, b, c, d) gives row `row` of line (s0, s1) (s0 < n0 along the
// adjacent lines, s1 < gridDim.y; workgroups of SX lines in xcd_tiles
// order).  Each row enters the forward elimination as soon as it is made,
// so no coefficient array goes through memory; the thread layout, the
// reduced solve and the output (D at s0 + s1*ss1 + row*rs, or a field) are
// those of the split solve of stored lines.  Sources with wall terms
// (SrcWalls) make the rows of a workgroup whose j is away from the walls
// without them.
template <typename T, bool CYC, class Src, int SX, bool FULL>
__global__ void __launch_bounds__(SX * PT_PMAX) __attribute__((amdgpu_waves_per_eu(Src::waves)))
k_ptdma_part_src(Src src, T *D, int n0, int n1, long ss1, long rs, int P, int w1, int w2, LineOut out) {
    __shared__ part::RedLDS<T, SX, 1> red;
    const int t = threadIdx.x;
    const int line = t % SX, p = t / SX;
    int bx, by;
    xcd_tiles(bx, by);
    if (by >= n1) return;  // a padding row of workgroups (whole workgroup)
    const int s0 = bx * SX + line, s1 = by;
    if constexpr (SrcWalls<Src>::value) {
        // s1 is the workgroup's: one branch for the whole workgroup
        if (src.wall_free(s1))
            part_src_lines<2, T, CYC, Src, SX, FULL>(src, red, D, n0, ss1, rs, P, w1, w2, out, line, p, s0, s1);
        else
            part_src_lines<0, T, CYC, Src, SX, FULL>(src, red, D, n0, ss1, rs, P, w1, w2, out, line, p, s0, s1);
    } else {
        part_src_lines<-1, T, CYC, Src, SX, FULL>(src, red, D, n0, ss1, rs, P, w1, w2, out, line, p, s0, s1);
    }
}

// padded LDS position of row j of a staged line (one pad value per 16 rows)
__device__ __forceinline__ int stage_pos(int j) { return j + (j >> 4); }
inline int stage_len(int rows) { return rows - 1 + ((rows - 1) >> 4) + 1; }

// Row blocks of the contiguous-line solve with computed rows: the parts of
// a line go to PT_RBLK blocks (fewer when P is smaller), block g holding
// parts [g*P/nb, (g+1)*P/nb) and so rows [row0(g), row0(g+1)).
constexpr int PT_RBLK = 4;
constexpr int PT_EU = 2;    // rows a thread makes at once
constexpr int PT_XRUN = 4;  // consecutive rows a thread makes when a block divides evenly
__host__ __device__ inline int rblocks(int P) { return P < PT_RBLK ? P : PT_RBLK; }
__host__ __device__ inline int rblock_row0(int g, int P, int w1, int w2) {
    const int p = g * (P / rblocks(P));
    return p * w1 + (p < w2 ? p : w2);
}

// Lines per workgroup of the contiguous-line solve with computed rows: the
// rows are made along the lines, so fewer lines cost no coalescing, and the
// small workgroups (PT_XSX * P threads, one wavefront for 512 rows) with
// their small stages share a CU, one's memory waits covered by the others'
// work (at 512^3: 2 lines 12-30% faster than 16).
constexpr int PT_XSX = 2;

// Contiguous lines (rows 1 apart, x sweeps) whose rows are computed:
// src(s0, s1, row, a, b, c, d) as in k_ptdma_part_src, line (s0, s1) at
// s0*ss0 + s1*ss1 of D.  The rows of the workgroup's PT_XSX lines are made
// one row block at a time by all its threads, consecutive rows on
// consecutive threads (coalesced stencil reads), and pass through LDS, the
// four arrays of the block together; the threads whose parts lie in the
// block take their rows from there.  Then the solve of k_ptdma_part_rows,
// whose LDS stage also takes the solution out.
template <typename T, bool CYC, class Src>
__global__ void __launch_bounds__(PT_XSX * PT_PMAX)
k_ptdma_part_rows_src(Src src, T *D, int n0, long ss0, long ss1, int n_row, int ls, int lsb, int P, int w1,
                      int w2) {
    using namespace part;
    extern __shared__ __align__(16) unsigned char stage_raw[];
    T *stage = reinterpret_cast<T *>(stage_raw);
    __shared__ RedLDS<T, PT_XSX, 1> red;
    const int t = threadIdx.x, nt = blockDim.x;
    const int line = t % PT_XSX, p = t / PT_XSX;
    int bx, s1;
    xcd_tiles(bx, s1);
    const int s0 = bx * PT_XSX + line;
    const bool act = s0 < n0;
    const int nl = min(PT_XSX, n0 - bx * PT_XSX);  // lines present in this workgroup
    int j0, m;
    range(p, w1, w2, j0, m);
    const int nb = rblocks(P), mygrp = p / (P / nb);

    T a[PT_MMAX], b[PT_MMAX], c[PT_MMAX], d[1][PT_MMAX];
    for (int g = 0; g < nb; g++) {
        const int r0 = rblock_row0(g, P, w1, w2), nr = rblock_row0(g + 1, P, w1, w2) - r0;
        if (g > 0) __syncthreads();  // the previous block has been read
        // stage[(q*PT_XSX + l)*lsb + stage_pos(r - r0)], q = a, b, c, d
        if (sizeof(T) == 8 && nr * PT_XSX == PT_XRUN * nt && nl == PT_XSX) {
            // a whole block, PT_XRUN consecutive rows of one line per thread,
            // made unconditionally: consecutive rows share their loads and
            // terms (FP64 dU/dV/dW x solves 1.58-1.72 -> 1.08-1.11 ms at
            // 512^3)
            const int per = nt / PT_XSX;  // threads per line
            const int l = t / per, rb = (t - l * per) * PT_XRUN;
            T ar[PT_XRUN], br[PT_XRUN], cr[PT_XRUN], dr[PT_XRUN];
#pragma unroll
            for (int u = 0; u < PT_XRUN; u++) src(bx * PT_XSX + l, s1, r0 + rb + u, ar[u], br[u], cr[u], dr[u]);
#pragma unroll
            for (int u = 0; u < PT_XRUN; u++) {
                const int o = l * lsb + stage_pos(rb + u);
                stage[o] = ar[u];
                stage[o + PT_XSX * lsb] = br[u];
                stage[o + 2 * PT_XSX * lsb] = cr[u];
                stage[o + 3 * PT_XSX * lsb] = dr[u];
            }
        } else {
        // otherwise a thread makes PT_EU rows (nt apart) before it stores any,
        // so their reads overlap
        for (int e0 = t; e0 < nr * PT_XSX; e0 += PT_EU * nt) {
            T ar[PT_EU], br[PT_EU], cr[PT_EU], dr[PT_EU];
            int o[PT_EU];
#pragma unroll
            for (int u = 0; u < PT_EU; u++) {
                const int e = e0 + u * nt;
                ar[u] = T(0); br[u] = T(1); cr[u] = T(0); dr[u] = T(0);
                o[u] = -1;
                if (e < nr * PT_XSX) {
                    const int l = e / nr, r = e - l * nr;
                    if (l < nl) src(bx * PT_XSX + l, s1, r0 + r, ar[u], br[u], cr[u], dr[u]);
                    o[u] = l * lsb + stage_pos(r);
                }
            }
#pragma unroll
            for (int u = 0; u < PT_EU; u++) {
                if (o[u] >= 0) {
                    stage[o[u]] = ar[u];
                    stage[o[u] + PT_XSX * lsb] = br[u];
                    stage[o[u] + 2 * PT_XSX * lsb] = cr[u];
                    stage[o[u] + 3 * PT_XSX * lsb] = dr[u];
                }
            }
        }
        }
        __syncthreads();
        if (mygrp == g) {
#pragma unroll
            for (int r = 0; r < PT_MMAX; r++) {
                const bool in = r < m;
                const T *q = stage + line * lsb + stage_pos(j0 - r0 + r);
                a[r] = in ? q[0] : T(0);
                b[r] = in ? q[PT_XSX * lsb] : T(1);
                c[r] = in ? q[2 * PT_XSX * lsb] : T(0);
                d[0][r] = in ? q[3 * PT_XSX * lsb] : T(0);
            }
        }
    }
    reduce<T, 1>(a, b, c, d, m);
    T xf[1], xl[1];
    exchange<T, CYC, PT_XSX, 1>(red, line, p, P, a, c, d, m, xf, xl);
    // every thread has read the stage before exchange's first barrier
    if (act) {
#pragma unroll
        for (int r = 0; r < PT_MMAX; r++)
            if (r < m) stage[line * ls + stage_pos(j0 + r)] = x_of<T, 1>(a, c, d, 0, r, m, xf[0], xl[0]);
    }
    __syncthreads();
    const long gbase = (long)bx * PT_XSX * ss0 + (long)s1 * ss1;
    for (int r = t; r < n_row; r += nt) {
#pragma unroll
        for (int l = 0; l < PT_XSX; l++)
            if (l < nl) D[gbase + l * ss0 + r] = stage[l * ls + stage_pos(r)];
    }
}

// Launch of k_ptdma_part_src over n0 x n1 lines of n_row rows (n_row in
// [2, 512]), SX lines per workgroup (16: four 128-byte row segments per
// wavefront load; 8: 64-byte segments, but two workgroups per CU); returns
// TCS_E_ARG outside the split range.
template <bool CYC, int SX = PT_SX, typename T, class Src>
int ptdma_part_src_launch(hipStream_t st, const char *label, const Src &src, T *D, int n0, int n1, long ss1,
                          int n_row, long rs, const LineOut &out = LineOut()) {
    int P, w1, w2;
    if (n0 <= 0 || n1 <= 0) return TCS_OK;
    if (!ptdma_part_plan(n_row, P, w1, w2) || n1 > 65535) {
        set_error("ptdma_part: n_row must be in [2, 512] and the outer line count <= 65535");
        return TCS_E_ARG;
    }
    // the outer line count padded to whole XCD tiles (the padding
    // workgroups return at once): a dV stage's 511 lines then take the tiled
    // order too
    const unsigned gx = (unsigned)((n0 + SX - 1) / SX);
    unsigned gy = (unsigned)n1;
    const unsigned gyp = (gy + XT_Y - 1) / XT_Y * XT_Y;
    if (gx % XT_X == 0 && (gx * gyp) % (8 * XT_X * XT_Y) == 0 && gyp <= 65535) gy = gyp;
    const dim3 blk(SX * P), grd(gx, gy);
    // FP64 lines whose parts all hold PT_MMAX rows take the unconditional
    // row loop where the source asks for it (Src::full_rows: every y stage,
    // thermal y 2.63 -> 2.00 ms, dU y 1.52 -> 1.02 ms, dW y 1.43 -> 1.29 ms
    // at 512^3; the dU z stage, 4.14 -> 3.14 ms); FP32 lines (the mixed
    // form) where the row asks for it as well (Src::full_rows32)
    if (Src::full_rows && (sizeof(T) == 8 || Src::full_rows32) && w1 == PT_MMAX && w2 == 0)
        TCS_LAUNCH(st, label, (k_ptdma_part_src<T, CYC, Src, SX, true>), grd, blk, 0, src, D, n0, n1, ss1, rs, P, w1, w2,
                   out);
    else
        TCS_LAUNCH(st, label, (k_ptdma_part_src<T, CYC, Src, SX, false>), grd, blk, 0, src, D, n0, n1, ss1, rs, P, w1, w2,
                   out);
    return TCS_OK;
}

// Launch of k_ptdma_part_rows_src over n0 x n1 contiguous lines of n_row
// rows (n_row in [2, 512]).
template <bool CYC, typename T, class Src>
int ptdma_part_rows_src_launch(hipStream_t st, const char *label, const Src &src, T *D, int n0, long ss0, int n1,
                               long ss1, int n_row) {
    int P, w1, w2;
    if (n0 <= 0 || n1 <= 0) return TCS_OK;
    if (!ptdma_part_plan(n_row, P, w1, w2) || n1 > 65535) {
        set_error("ptdma_part: n_row must be in [2, 512] and the outer line count <= 65535");
        return TCS_E_ARG;
    }
    int lsb = 0;
    for (int g = 0; g < rblocks(P); g++)
        lsb = std::max(lsb, stage_len(rblock_row0(g + 1, P, w1, w2) - rblock_row0(g, P, w1, w2)));
    const int ls = stage_len(n_row);
    const int lds = (int)sizeof(T) * std::max(PT_XSX * ls, 4 * PT_XSX * lsb);
    static bool lds_raised = false;  // more than 64 KiB of dynamic LDS (gfx950: 160 KiB per CU)
    if (!lds_raised) {
        const int most = std::max(PT_XSX * stage_len(PT_MMAX * PT_PMAX), 4 * PT_XSX * stage_len(PT_MMAX * PT_PMAX / PT_RBLK));
        TCS_HIP(hipFuncSetAttribute((const void *)k_ptdma_part_rows_src<T, CYC, Src>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)sizeof(T) * most));
        lds_raised = true;
    }
    const dim3 blk(PT_XSX * P), grd((unsigned)((n0 + PT_XSX - 1) / PT_XSX), (unsigned)n1);
    TCS_LAUNCH(st, label, (k_ptdma_part_rows_src<T, CYC, Src>), grd, blk, lds, src, D, n0, ss0, ss1, n_row, ls, lsb,
               P, w1, w2);
    return TCS_OK;
}

}  // namespace tcs
