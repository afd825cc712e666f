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
// the power of two P >= n_row/16 parts, each thread keeps its part (at most
// 16 rows of A, C, D) in registers through the whole solve, and the reduced
// systems of a workgroup sit in LDS.  With one part the reduced 2-row system
// is solved in the reference's order; with P >= 2 the line's P threads solve
// it together by parallel cyclic reduction (log2(2P) steps instead of a
// serial sweep over 2P rows, whose latency otherwise dominates the kernel).
// HBM traffic is the algorithmic minimum: A, B, C, D read once and the
// solution written once (40 bytes per unknown in FP64); A and C are not
// written back.
//
//   k_ptdma_part       lines strided in memory (y and z sweeps): 16
//                      neighbouring lines per workgroup, so every row step
//                      of a wavefront is four 128-byte segments;
//   k_ptdma_part_rows  contiguous lines (x sweeps): a workgroup stages 16
//                      lines of A, B, C, D one array at a time through LDS
//                      with coalesced loads, each thread then reads its part.
//
// Every operation follows the oracle's orc_ptdma_many_solve_split in order,
// with FP contraction off, so results are bit-identical to it.
#include "tdma_part.h"

namespace tcs {

bool ptdma_part_plan(int n_row, int &P, int &w1, int &w2) {
    if (n_row < 2 || n_row > PT_MMAX * PT_PMAX) return false;
    P = 1;
    while (P * PT_MMAX < n_row) P <<= 1;  // power of two: the cyclic reduction pairs rows i, i + P
    w1 = n_row / P;
    w2 = n_row % P;
    return true;
}

namespace {

using namespace part;

// Workgroup -> group of SX lines.  With SX = 8 two workgroups share every
// 128-byte row segment of 16 lines; they are placed 8 apart in dispatch
// order so that both run on the same XCD (workgroups go to the 8 XCDs
// round-robin) and the second finds the segment in that XCD's L2.
template <int SX>
__device__ __forceinline__ int line_group(int bx, int nbx) {
    if (SX != 8 || (nbx & 15) != 0) return bx;
    const int q = bx >> 4, r = bx & 15;
    return ((q << 3) + (r & 7)) * 2 + (r >> 3);
}

// Lines strided in memory: line (s0, s1) starts at s0 + s1*ss1 and its rows
// are rs apart.  blockDim.x = SX * P, thread t = line + SX * part.
template <typename T, bool CYC, int SX>
__global__ void __launch_bounds__(SX * PT_PMAX)
k_ptdma_part(const T *__restrict__ A, const T *__restrict__ B, const T *__restrict__ C,
             T *__restrict__ D, int n0, long ss1, long rs, int P, int w1, int w2, LineOut out) {
    __shared__ RedLDS<T, SX, 1> red;
    const int t = threadIdx.x;
    const int line = t % SX, p = t / SX;
    const int s0 = line_group<SX>(blockIdx.x, gridDim.x) * SX + line;
    const bool act = s0 < n0;
    const long base = (long)s0 + (long)blockIdx.y * ss1;
    int j0, m;
    range(p, w1, w2, j0, m);

    T a[PT_MMAX], c[PT_MMAX], d[1][PT_MMAX];
    {
        T b[PT_MMAX];
#pragma unroll
        for (int r = 0; r < PT_MMAX; r++) {
            if (act && r < m) {
                const long o = base + (long)(j0 + r) * rs;
                a[r] = A[o]; b[r] = B[o]; c[r] = C[o]; d[0][r] = D[o];
            } else {
                a[r] = T(0); b[r] = T(1); c[r] = T(0); d[0][r] = T(0);
            }
        }
        reduce<T, 1>(a, b, c, d, m);
    }
    T xf[1], xl[1];
    exchange<T, CYC, SX, 1>(red, line, p, P, a, c, d, m, xf, xl);
    if (!act) return;
    store<T>(a, c, d, j0, m, xf[0], xl[0], D, base, rs, out, s0, blockIdx.y);
}

// Contiguous lines (rows 1 apart): line (s0, s1) starts at s0*ss0 + s1*ss1.
// The same thread layout as k_ptdma_part (thread t = line + PT_SX * part,
// blockDim.x = PT_SX * P >= n_row).  A, B, C, D pass one after the other
// through an LDS stage of PT_SX lines (dynamic LDS, PT_SX * ls values with
// ls = stage_pos(n_row-1) + 1): every thread loads row t of each of the 16
// lines (coalesced along the lines), the next array's loads are issued
// before the stage is read back part-wise.  The solution goes back out
// through the stage the same way.
template <typename T, bool CYC>
__global__ void __launch_bounds__(PT_SX * PT_PMAX)
k_ptdma_part_rows(const T *__restrict__ A, const T *__restrict__ B, const T *__restrict__ C, T *__restrict__ D,
                  int n0, long ss0, long ss1, int n_row, int ls, int P, int w1, int w2) {
    extern __shared__ __align__(16) unsigned char stage_raw[];
    T *stage = reinterpret_cast<T *>(stage_raw);
    __shared__ RedLDS<T, PT_SX, 1> red;
    const int t = threadIdx.x;
    const int line = t % PT_SX, p = t / PT_SX;
    const int s0 = blockIdx.x * PT_SX + line;
    const bool act = s0 < n0;
    int j0, m;
    range(p, w1, w2, j0, m);
    const long gbase = (long)blockIdx.x * PT_SX * ss0 + (long)blockIdx.y * ss1;
    const int nl = min(PT_SX, n0 - (int)blockIdx.x * PT_SX);  // lines present in this workgroup
    const bool row_ok = t < n_row;

    T a[PT_MMAX], b[PT_MMAX], c[PT_MMAX], d[1][PT_MMAX];
    T v[PT_SX];
#pragma unroll
    for (int l = 0; l < PT_SX; l++) v[l] = (l < nl && row_ok) ? A[gbase + l * ss0 + t] : T(0);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        if (q > 0) __syncthreads();  // the stage's previous array has been read
        if (row_ok) {
#pragma unroll
            for (int l = 0; l < PT_SX; l++) stage[l * ls + stage_pos(t)] = v[l];
        }
        if (q < 3) {
            const T *g = q == 0 ? B : (q == 1 ? C : D);
#pragma unroll
            for (int l = 0; l < PT_SX; l++) v[l] = (l < nl && row_ok) ? g[gbase + l * ss0 + t] : T(0);
        }
        __syncthreads();
        T *dst = q == 0 ? a : (q == 1 ? b : (q == 2 ? c : d[0]));
        const T pad = q == 1 ? T(1) : T(0);
#pragma unroll
        for (int r = 0; r < PT_MMAX; r++) dst[r] = (act && r < m) ? stage[line * ls + stage_pos(j0 + r)] : pad;
    }
    reduce<T, 1>(a, b, c, d, m);
    T xf[1], xl[1];
    exchange<T, CYC, PT_SX, 1>(red, line, p, P, a, c, d, m, xf, xl);
    // every thread has read the stage before exchange's first barrier
    if (act) {
#pragma unroll
        for (int r = 0; r < PT_MMAX; r++)
            if (r < m) stage[line * ls + stage_pos(j0 + r)] = x_of<T, 1>(a, c, d, 0, r, m, xf[0], xl[0]);
    }
    __syncthreads();
    if (row_ok) {
#pragma unroll
        for (int l = 0; l < PT_SX; l++)
            if (l < nl) D[gbase + l * ss0 + t] = stage[l * ls + stage_pos(t)];
    }
}

}  // namespace

template <typename T>
static int part_lds_bytes(int n_row) {
    return PT_SX * stage_len(n_row) * (int)sizeof(T);
}

template <typename T> struct PartLabels;
template <> struct PartLabels<double> {
    static constexpr const char *rows_cyc = "ptdma_part_rows_cyclic", *rows = "ptdma_part_rows",
                                *cyc = "ptdma_part_cyclic", *plain = "ptdma_part";
};
template <> struct PartLabels<float> {
    static constexpr const char *rows_cyc = "ptdma_part_rows_cyclic_f32", *rows = "ptdma_part_rows_f32",
                                *cyc = "ptdma_part_cyclic_f32", *plain = "ptdma_part_f32";
};

// Partitioned solve of n0 x n1 lines (see the kernels for the layouts).
// Returns TCS_E_ARG when the line length is outside the partitioned range.
template <typename T>
static int part_launch(T *A, const T *B, T *C, T *D, int n0, long ss0, int n1, long ss1, int n_row, long rs,
                       bool cyclic, hipStream_t st, const LineOut &out) {
    int P, w1, w2;
    if (n0 <= 0 || n1 <= 0) return TCS_OK;
    if (!ptdma_part_plan(n_row, P, w1, w2) || n1 > 65535) {
        set_error("ptdma_part: n_row must be in [2, 512] and the outer line count <= 65535");
        return TCS_E_ARG;
    }
    using L = PartLabels<T>;
    if (rs == 1 && out.mode != 0) {
        set_error("ptdma_part: field output is for strided lines only");
        return TCS_E_ARG;
    }
    if (rs == 1) {
        const int ls = stage_len(n_row);
        const int lds = part_lds_bytes<T>(n_row);
        static bool lds_raised = false;  // allow more than 64 KiB of dynamic LDS (gfx950: 160 KiB per CU)
        if (!lds_raised) {
            const int most = part_lds_bytes<T>(PT_MMAX * PT_PMAX);
            TCS_HIP(hipFuncSetAttribute((const void *)k_ptdma_part_rows<T, false>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, most));
            TCS_HIP(hipFuncSetAttribute((const void *)k_ptdma_part_rows<T, true>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, most));
            lds_raised = true;
        }
        dim3 blk(PT_SX * P), grd((unsigned)((n0 + PT_SX - 1) / PT_SX), (unsigned)n1);
        if (cyclic)
            TCS_LAUNCH(st, L::rows_cyc, (k_ptdma_part_rows<T, true>), grd, blk, lds, A, B, C, D, n0, ss0, ss1, n_row,
                       ls, P, w1, w2);
        else
            TCS_LAUNCH(st, L::rows, (k_ptdma_part_rows<T, false>), grd, blk, lds, A, B, C, D, n0, ss0, ss1, n_row, ls,
                       P, w1, w2);
        return TCS_OK;
    }
    if (ss0 != 1) {
        set_error("ptdma_part: strided lines must be adjacent in memory (ss0 = 1)");
        return TCS_E_ARG;
    }
    dim3 blk(PT_SX * P), grd((unsigned)((n0 + PT_SX - 1) / PT_SX), (unsigned)n1);
    if (cyclic)
        TCS_LAUNCH(st, L::cyc, (k_ptdma_part<T, true, PT_SX>), grd, blk, 0, A, B, C, D, n0, ss1, rs, P, w1, w2,
                   out);
    else
        TCS_LAUNCH(st, L::plain, (k_ptdma_part<T, false, PT_SX>), grd, blk, 0, A, B, C, D, n0, ss1, rs, P, w1, w2,
                   out);
    return TCS_OK;
}

int ptdma_part_launch(double *A, const double *B, double *C, double *D, int n0, long ss0, int n1, long ss1,
                      int n_row, long rs, bool cyclic, hipStream_t st, const LineOut &out) {
    return part_launch<double>(A, B, C, D, n0, ss0, n1, ss1, n_row, rs, cyclic, st, out);
}

int ptdma_part_launch(float *A, const float *B, float *C, float *D, int n0, long ss0, int n1, long ss1,
                      int n_row, long rs, bool cyclic, hipStream_t st, const LineOut &out) {
    return part_launch<float>(A, B, C, D, n0, ss0, n1, ss1, n_row, rs, cyclic, st, out);
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
