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

// One part (rows 0..m-1, m >= 2) in registers: the per-rank modified Thomas
// algorithm of pascal_tdma.f90:501-536.  b is consumed; a, c, d come back as
// the reduced-form coefficients.
template <typename T, int NR>
__device__ __forceinline__ void reduce(T (&a)[PT_MMAX], T (&b)[PT_MMAX], T (&c)[PT_MMAX],
                                       T (&d)[NR][PT_MMAX], int m) {
    a[0] = a[0] / b[0];
#pragma unroll
    for (int k = 0; k < NR; k++) d[k][0] = d[k][0] / b[0];
    c[0] = c[0] / b[0];
    a[1] = a[1] / b[1];
#pragma unroll
    for (int k = 0; k < NR; k++) d[k][1] = d[k][1] / b[1];
    c[1] = c[1] / b[1];
#pragma unroll
    for (int r = 2; r < PT_MMAX; r++) {
        if (r < m) {
            const T rr = T(1) / (b[r] - a[r] * c[r - 1]);
#pragma unroll
            for (int k = 0; k < NR; k++) d[k][r] = rr * (d[k][r] - a[r] * d[k][r - 1]);
            c[r] = rr * c[r];
            a[r] = -rr * a[r] * a[r - 1];
        }
    }
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

}  // namespace part
}  // namespace tcs
