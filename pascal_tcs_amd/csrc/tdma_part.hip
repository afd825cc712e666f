// tdma_part.hip -- batched tridiagonal solves with PaScaL_TDMA's partitioned
// (multi-rank) algorithm, run inside one workgroup.
//
// PaScaL_TDMA splits the rows of every system over the ranks along the solve
// direction (para_range, PaScaL_TDMA/src/para_range.f90:26-40).  Each rank
// eliminates both off-diagonals of its own rows (modified Thomas,
// pascal_tdma.f90:501-536), the two boundary rows of every rank form a
// reduced system of 2*nprocs rows that is gathered in rank order
// (pascal_tdma.f90:243-266) and solved by tdma_many / tdma_cycl_many
// (tdmas.f90:113-212), and each rank updates its rows from the reduced
// solution (pascal_tdma.f90:567-576).
//
// Here a "rank" is one thread: a line of n_row <= 512 rows is split into
// P = ceil(n_row/16) parts, each thread keeps its part (at most 16 rows of
// A, C, D) in registers through the whole solve, the reduced systems of a
// workgroup sit in LDS and one thread per system solves its reduced system
// in the reference's order.  HBM traffic is the algorithmic minimum: A, B,
// C, D read once and the solution written once (40 bytes per unknown); A and
// C are not written back.
//
//   k_ptdma_part       lines strided in memory (y and z sweeps): 16
//                      neighbouring lines per workgroup, so every row step
//                      of a wavefront is four 128-byte segments;
//   k_ptdma_part_rows  contiguous lines (x sweeps): a workgroup stages 4
//                      lines of A, B, C, D through LDS with coalesced loads,
//                      each thread then reads its part from LDS.
//
// Every operation follows the oracle's ptdma_many_ranks in order, with FP
// contraction off, so results are bit-identical to it.
#include "tcs_internal.h"

namespace tcs {

constexpr int PT_MMAX = 16;               // rows per part
constexpr int PT_PMAX = 32;               // parts per line (n_row <= 512)
constexpr int PT_RS = 2 * PT_PMAX + 1;    // reduced-row stride in LDS (odd: no bank conflicts)
constexpr int PT_SX = 16;                 // lines per workgroup, strided kernel
constexpr int PT_SXR = 4;                 // lines per workgroup, rows kernel
constexpr int PT_UMAX = PT_MMAX / PT_SXR; // staged values per thread, line and array

bool ptdma_part_plan(int n_row, int &P, int &w1, int &w2) {
    if (n_row < 2 || n_row > PT_MMAX * PT_PMAX) return false;
    P = (n_row + PT_MMAX - 1) / PT_MMAX;
    w1 = n_row / P;
    w2 = n_row % P;
    return true;
}

namespace {

__device__ __forceinline__ void part_range(int p, int w1, int w2, int &j0, int &m) {
    j0 = p * w1 + (p < w2 ? p : w2);
    m = w1 + (p < w2 ? 1 : 0);
}

// One part (rows 0..m-1, m >= 2) in registers: pascal_tdma.f90:501-536.
// b is consumed; a, c, d come back as the reduced-form coefficients.
__device__ __forceinline__ void part_reduce(double (&a)[PT_MMAX], double (&b)[PT_MMAX], double (&c)[PT_MMAX],
                                            double (&d)[PT_MMAX], int m) {
    a[0] = a[0] / b[0];
    d[0] = d[0] / b[0];
    c[0] = c[0] / b[0];
    a[1] = a[1] / b[1];
    d[1] = d[1] / b[1];
    c[1] = c[1] / b[1];
#pragma unroll
    for (int r = 2; r < PT_MMAX; r++) {
        if (r < m) {
            const double rr = 1.0 / (b[r] - a[r] * c[r - 1]);
            d[r] = rr * (d[r] - a[r] * d[r - 1]);
            c[r] = rr * c[r];
            a[r] = -rr * a[r] * a[r - 1];
        }
    }
#pragma unroll
    for (int r = PT_MMAX - 3; r >= 1; r--) {
        if (r <= m - 3) {
            d[r] = d[r] - c[r] * d[r + 1];
            a[r] = a[r] - c[r] * a[r + 1];
            c[r] = -c[r] * c[r + 1];
        }
    }
    const double rr = 1.0 / (1.0 - a[1] * c[0]);
    d[0] = rr * (d[0] - c[0] * d[1]);
    a[0] = rr * a[0];
    c[0] = -rr * c[0] * c[1];
}

// values of row m-1 (a runtime index) without dynamic register indexing
__device__ __forceinline__ void part_last(const double (&a)[PT_MMAX], const double (&c)[PT_MMAX],
                                          const double (&d)[PT_MMAX], int m, double &la, double &lc, double &ld) {
    la = a[0]; lc = c[0]; ld = d[0];
#pragma unroll
    for (int r = 1; r < PT_MMAX; r++)
        if (r == m - 1) { la = a[r]; lc = c[r]; ld = d[r]; }
}

// x(r) of the part from its reduced solution xf (row 0) and xl (row m-1):
// pascal_tdma.f90:567-576
__device__ __forceinline__ double part_x(const double (&a)[PT_MMAX], const double (&c)[PT_MMAX],
                                         const double (&d)[PT_MMAX], int r, int m, double xf, double xl) {
    if (r == 0) return xf;
    if (r == m - 1) return xl;
    return d[r] - a[r] * xf - c[r] * xl;
}

// Reduced system of one line in LDS (unit diagonal, n2 = 2P rows), solved
// in place in ra/rc/rd by tdma_many (tdmas.f90:113-146) or tdma_cycl_many
// (tdmas.f90:159-212; re is its e scratch).  Division by the unit diagonal
// is exact and left out.
template <bool CYC>
__device__ void reduced_solve(double *ra, double *rc, double *rd, double *re, int n2) {
    if (!CYC) {
        for (int j = 1; j < n2; j++) {
            const double r = 1.0 / (1.0 - ra[j] * rc[j - 1]);
            rd[j] = r * (rd[j] - ra[j] * rd[j - 1]);
            rc[j] = r * rc[j];
        }
        for (int j = n2 - 2; j >= 0; j--) rd[j] = rd[j] - rc[j] * rd[j + 1];
    } else {
        for (int j = 0; j < n2; j++) re[j] = 0.0;
        re[1] = -ra[1];
        re[n2 - 1] = -rc[n2 - 1];
        for (int j = 2; j < n2; j++) {
            const double rr = 1.0 / (1.0 - ra[j] * rc[j - 1]);
            rd[j] = rr * (rd[j] - ra[j] * rd[j - 1]);
            re[j] = rr * (re[j] - ra[j] * re[j - 1]);
            rc[j] = rr * rc[j];
        }
        for (int j = n2 - 2; j >= 1; j--) {
            rd[j] = rd[j] - rc[j] * rd[j + 1];
            re[j] = re[j] - rc[j] * re[j + 1];
        }
        rd[0] = (rd[0] - ra[0] * rd[n2 - 1] - rc[0] * rd[1]) / (1.0 + ra[0] * re[n2 - 1] + rc[0] * re[1]);
        for (int j = 1; j < n2; j++) rd[j] = rd[j] + rd[0] * re[j];
    }
}

// Reduced arrays of a workgroup: [q][line][row], q = a, c, d, e.
template <int SX>
struct RedLDS {
    double v[4 * SX * PT_RS];
    __device__ double *q(int which, int line) { return v + ((size_t)which * SX + line) * PT_RS; }
};

// After part_reduce: publish the part's two reduced rows, solve, read back.
template <bool CYC, int SX>
__device__ __forceinline__ void part_exchange(RedLDS<SX> &red, int line, int p, int P, const double (&a)[PT_MMAX],
                                              const double (&c)[PT_MMAX], const double (&d)[PT_MMAX], int m,
                                              double &xf, double &xl) {
    double la, lc, ld;
    part_last(a, c, d, m, la, lc, ld);
    double *ra = red.q(0, line), *rc = red.q(1, line), *rd = red.q(2, line);
    ra[2 * p] = a[0]; rc[2 * p] = c[0]; rd[2 * p] = d[0];
    ra[2 * p + 1] = la; rc[2 * p + 1] = lc; rd[2 * p + 1] = ld;
    __syncthreads();
    if (p == 0) reduced_solve<CYC>(ra, rc, rd, red.q(3, line), 2 * P);
    __syncthreads();
    xf = rd[2 * p];
    xl = rd[2 * p + 1];
}

// Lines strided in memory: line (s0, s1) starts at s0 + s1*ss1 and its rows
// are rs apart.  blockDim.x = PT_SX * P, thread t = line + PT_SX * part.
template <bool CYC>
__global__ void __launch_bounds__(PT_SX * PT_PMAX)
k_ptdma_part(const double *__restrict__ A, const double *__restrict__ B, const double *__restrict__ C,
             double *__restrict__ D, int n0, long ss1, long rs, int P, int w1, int w2) {
    __shared__ RedLDS<PT_SX> red;
    const int t = threadIdx.x;
    const int line = t % PT_SX, p = t / PT_SX;
    const int s0 = blockIdx.x * PT_SX + line;
    const bool act = s0 < n0;
    const long base = (long)s0 + (long)blockIdx.y * ss1;
    int j0, m;
    part_range(p, w1, w2, j0, m);

    double a[PT_MMAX], c[PT_MMAX], d[PT_MMAX];
    {
        double b[PT_MMAX];
#pragma unroll
        for (int r = 0; r < PT_MMAX; r++) {
            if (act && r < m) {
                const long o = base + (long)(j0 + r) * rs;
                a[r] = A[o]; b[r] = B[o]; c[r] = C[o]; d[r] = D[o];
            } else {
                a[r] = 0.0; b[r] = 1.0; c[r] = 0.0; d[r] = 0.0;
            }
        }
        part_reduce(a, b, c, d, m);
    }
    double xf, xl;
    part_exchange<CYC, PT_SX>(red, line, p, P, a, c, d, m, xf, xl);
    if (!act) return;
#pragma unroll
    for (int r = 0; r < PT_MMAX; r++)
        if (r < m) D[base + (long)(j0 + r) * rs] = part_x(a, c, d, r, m, xf, xl);
}

// padded LDS position of row j of a staged line (one pad double per 16 rows)
__device__ __forceinline__ int stage_pos(int j) { return j + (j >> 4); }

// Contiguous lines (rows 1 apart): line (s0, s1) starts at s0*ss0 + s1*ss1.
// blockDim.x = PT_SXR * P, thread t = part + P * line.  Dynamic LDS: the 4
// staged arrays, 4 * PT_SXR * ls doubles with ls = stage_pos(n_row-1) + 1.
template <bool CYC>
__global__ void __launch_bounds__(PT_SXR * PT_PMAX)
k_ptdma_part_rows(const double *__restrict__ A, const double *__restrict__ B, const double *__restrict__ C,
                  double *__restrict__ D, int n0, long ss0, long ss1, int n_row, int ls, int P, int w1, int w2) {
    extern __shared__ double stage[];
    __shared__ RedLDS<PT_SXR> red;
    const int t = threadIdx.x, nt = blockDim.x;
    const int p = t % P, line = t / P;
    const int s0 = blockIdx.x * PT_SXR + line;
    const bool act = s0 < n0;
    int j0, m;
    part_range(p, w1, w2, j0, m);

    // Stage A, B, C, D of the workgroup's lines, coalesced along each line.
    // A line has n_row <= 16*P rows and the workgroup 4*P threads, so each
    // thread moves at most PT_UMAX values per line and array; all of its
    // loads are issued before the first LDS store.
    const double *src[4] = {A, B, C, D};
    {
        double v[4][PT_SXR][PT_UMAX];
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int l = 0; l < PT_SXR; l++)
#pragma unroll
                for (int u = 0; u < PT_UMAX; u++) {
                    const int s = blockIdx.x * PT_SXR + l, j = t + nt * u;
                    v[q][l][u] = (s < n0 && j < n_row) ? src[q][(long)s * ss0 + (long)blockIdx.y * ss1 + j] : 0.0;
                }
#pragma unroll
        for (int q = 0; q < 4; q++)
#pragma unroll
            for (int l = 0; l < PT_SXR; l++)
#pragma unroll
                for (int u = 0; u < PT_UMAX; u++) {
                    const int j = t + nt * u;
                    if (j < n_row) stage[((size_t)q * PT_SXR + l) * ls + stage_pos(j)] = v[q][l][u];
                }
    }
    __syncthreads();

    double a[PT_MMAX], c[PT_MMAX], d[PT_MMAX];
    {
        double b[PT_MMAX];
        const double *sa = stage + ((size_t)0 * PT_SXR + line) * ls;
        const double *sb = stage + ((size_t)1 * PT_SXR + line) * ls;
        const double *sc = stage + ((size_t)2 * PT_SXR + line) * ls;
        const double *sd = stage + ((size_t)3 * PT_SXR + line) * ls;
#pragma unroll
        for (int r = 0; r < PT_MMAX; r++) {
            if (act && r < m) {
                const int o = stage_pos(j0 + r);
                a[r] = sa[o]; b[r] = sb[o]; c[r] = sc[o]; d[r] = sd[o];
            } else {
                a[r] = 0.0; b[r] = 1.0; c[r] = 0.0; d[r] = 0.0;
            }
        }
        part_reduce(a, b, c, d, m);
    }
    double xf, xl;
    part_exchange<CYC, PT_SXR>(red, line, p, P, a, c, d, m, xf, xl);

    // solution back through LDS (the D slot) for coalesced stores
    double *sd = stage + ((size_t)3 * PT_SXR + line) * ls;
    if (act) {
#pragma unroll
        for (int r = 0; r < PT_MMAX; r++)
            if (r < m) sd[stage_pos(j0 + r)] = part_x(a, c, d, r, m, xf, xl);
    }
    __syncthreads();
#pragma unroll
    for (int l = 0; l < PT_SXR; l++)
#pragma unroll
        for (int u = 0; u < PT_UMAX; u++) {
            const int s = blockIdx.x * PT_SXR + l, j = t + nt * u;
            if (s < n0 && j < n_row)
                D[(long)s * ss0 + (long)blockIdx.y * ss1 + j] = stage[((size_t)3 * PT_SXR + l) * ls + stage_pos(j)];
        }
}

}  // namespace

int ptdma_part_lds_bytes(int n_row) {
    const int ls = n_row - 1 + ((n_row - 1) >> 4) + 1;
    return 4 * PT_SXR * ls * (int)sizeof(double);
}

// Partitioned solve of n0 x n1 lines (see the kernels for the layouts).
// Returns TCS_E_ARG when the line length is outside the partitioned range;
// callers then use the single-rank kernels.
int ptdma_part_launch(double *A, const double *B, double *C, double *D, int n0, long ss0, int n1, long ss1,
                      int n_row, long rs, bool cyclic, hipStream_t st) {
    int P, w1, w2;
    if (n0 <= 0 || n1 <= 0) return TCS_OK;
    if (!ptdma_part_plan(n_row, P, w1, w2) || n1 > 65535) {
        set_error("ptdma_part: n_row must be in [2, 512] and the outer line count <= 65535");
        return TCS_E_ARG;
    }
    if (rs == 1) {
        const int ls = n_row - 1 + ((n_row - 1) >> 4) + 1;
        const int lds = 4 * PT_SXR * ls * (int)sizeof(double);
        static bool lds_raised = false;  // allow more than 64 KiB of dynamic LDS (gfx950: 160 KiB per CU)
        if (!lds_raised) {
            const int most = ptdma_part_lds_bytes(PT_MMAX * PT_PMAX);
            TCS_HIP(hipFuncSetAttribute((const void *)k_ptdma_part_rows<false>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, most));
            TCS_HIP(hipFuncSetAttribute((const void *)k_ptdma_part_rows<true>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, most));
            lds_raised = true;
        }
        dim3 blk(PT_SXR * P), grd((unsigned)((n0 + PT_SXR - 1) / PT_SXR), (unsigned)n1);
        if (cyclic)
            TCS_LAUNCH(st, "ptdma_part_rows_cyclic", k_ptdma_part_rows<true>, grd, blk, lds, A, B, C, D, n0, ss0,
                       ss1, n_row, ls, P, w1, w2);
        else
            TCS_LAUNCH(st, "ptdma_part_rows", k_ptdma_part_rows<false>, grd, blk, lds, A, B, C, D, n0, ss0, ss1,
                       n_row, ls, P, w1, w2);
        return TCS_OK;
    }
    if (ss0 != 1) {
        set_error("ptdma_part: strided lines must be adjacent in memory (ss0 = 1)");
        return TCS_E_ARG;
    }
    dim3 blk(PT_SX * P), grd((unsigned)((n0 + PT_SX - 1) / PT_SX), (unsigned)n1);
    if (cyclic)
        TCS_LAUNCH(st, "ptdma_part_cyclic", k_ptdma_part<true>, grd, blk, 0, A, B, C, D, n0, ss1, rs, P, w1, w2);
    else
        TCS_LAUNCH(st, "ptdma_part", k_ptdma_part<false>, grd, blk, 0, A, B, C, D, n0, ss1, rs, P, w1, w2);
    return TCS_OK;
}

}  // namespace tcs

using namespace tcs;

extern "C" int ptdma_many_solve_part_dev(const double *A, const double *B, const double *C, double *D,
                                         int32_t n_sys, int32_t n_row, int32_t cyclic, void *stream) {
    if (!A || !B || !C || !D || n_sys < 0) {
        set_error("ptdma_many_solve_part_dev: null pointer or n_sys < 0");
        return TCS_E_ARG;
    }
    // (n_sys, n_row): the systems are the strided lines, rows n_sys apart
    return ptdma_part_launch((double *)A, B, (double *)C, D, n_sys, 1, 1, 0, n_row, n_sys, cyclic != 0,
                             (hipStream_t)stream);
}

extern "C" int ptdma_axis_solve_part_dev(const double *A, const double *B, const double *C, double *D, int32_t n1,
                                         int32_t n2, int32_t n3, int32_t axis, int32_t cyclic, void *stream) {
    if (!A || !B || !C || !D || n1 <= 0 || n2 <= 0 || n3 <= 0 || axis < 0 || axis > 2) {
        set_error("ptdma_axis_solve_part_dev: bad arguments");
        return TCS_E_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    double *a = (double *)A, *c = (double *)C;
    const long s1 = n1, s2 = (long)n1 * n2;
    switch (axis) {
    case 0:
        return ptdma_part_launch(a, B, c, D, n2, s1, n3, s2, n1, 1, cyclic != 0, st);
    case 1:
        return ptdma_part_launch(a, B, c, D, n1, 1, n3, s2, n2, s1, cyclic != 0, st);
    default:
        return ptdma_part_launch(a, B, c, D, n1, 1, n2, s1, n3, s2, cyclic != 0, st);
    }
}
