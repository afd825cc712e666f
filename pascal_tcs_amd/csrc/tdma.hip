// tdma.hip -- batched tridiagonal solves with PaScaL_TDMA's single-rank
// operation order (PaScaL_TDMA/src/pascal_tdma.f90:487-578 and :715-806,
// reduced 2-row solves of PaScaL_TDMA/src/tdmas.f90:113-212).
//
// One thread owns one system and walks its rows; consecutive threads own
// systems that are adjacent in memory wherever the caller's layout allows
// (the z and y sweeps of a field), so every row step is one coalesced
// 512-byte wavefront access per array.  The three passes of the modified
// Thomas algorithm (forward elimination with A fill-in, backward elimination
// with C fill-in, update from the reduced solution) keep their operands in
// place, exactly as the reference does with its A, C, D arrays.
#include "tcs_internal.h"

#include <mutex>

namespace tcs {

template <bool CYC>
__global__ void __launch_bounds__(256)
k_ptdma_exact(double *__restrict__ A, const double *__restrict__ B, double *__restrict__ C,
              double *__restrict__ D, int n0, long ss0, long ss1, int n_row, long rs) {
    const int s0 = blockIdx.x * blockDim.x + threadIdx.x;
    if (s0 >= n0) return;
    const long base = (long)s0 * ss0 + (long)blockIdx.y * ss1;
    double *a = A + base;
    const double *b = B + base;
    double *c = C + base;
    double *d = D + base;

    // rows 1 and 2 normalised by their own diagonal (:503-511)
    const double b0 = b[0], b1 = b[rs];
    double a0 = a[0] / b0;
    double d0 = d[0] / b0;
    double c0 = c[0] / b0;
    const double a1 = a[rs] / b1;
    const double d1 = d[rs] / b1;
    const double c1 = c[rs] / b1;
    a[rs] = a1; d[rs] = d1; c[rs] = c1;

    // forward elimination j = 3..n_row (:513-520)
    double ap = a1, cp = c1, dp = d1;
    for (int j = 2; j < n_row; j++) {
        const long o = (long)j * rs;
        const double aj = a[o], bj = b[o], cj = c[o], dj = d[o];
        const double r = 1.0 / (bj - aj * cp);
        const double dn = r * (dj - aj * dp);
        const double cn = r * cj;
        const double an = -r * aj * ap;
        a[o] = an; c[o] = cn; d[o] = dn;
        ap = an; cp = cn; dp = dn;
    }
    const double aN = ap, cN = cp, dN = dp;  // last row (forward values)

    // backward elimination j = n_row-2..2 (:523-529)
    double ax = a1, cx = c1, dx = d1;  // row 2 after the backward sweep
    if (n_row >= 3) {
        const long o2 = (long)(n_row - 2) * rs;
        ax = a[o2]; cx = c[o2]; dx = d[o2];
        for (int j = n_row - 3; j >= 1; j--) {
            const long o = (long)j * rs;
            const double aj = a[o], cj = c[o], dj = d[o];
            const double dn = dj - cj * dx;
            const double an = aj - cj * ax;
            const double cn = -cj * cx;
            a[o] = an; c[o] = cn; d[o] = dn;
            ax = an; cx = cn; dx = dn;
        }
    }

    // row 1 and the reduced 2-row system (:531-542)
    {
        const double r = 1.0 / (1.0 - ax * c0);
        d0 = r * (d0 - c0 * dx);
        a0 = r * a0;
        c0 = -r * c0 * cx;
    }
    a[0] = a0; c[0] = c0;

    double x1, xn;
    if (!CYC) {
        // tdma_many with n2=2, b=1 (tdmas.f90:113-146); x/1.0 is exact
        const double r = 1.0 / (1.0 - aN * c0);
        xn = r * (dN - aN * d0);
        x1 = d0 - c0 * xn;
    } else {
        // tdma_cycl_many with n2=2 (tdmas.f90:159-212): e(2) is set to -a(2)
        // and then overwritten by e(n2) = -c(n2)
        const double e2 = -cN;
        x1 = (d0 - a0 * dN - c0 * dN) / (1.0 + a0 * e2 + c0 * e2);
        xn = dN + x1 * e2;
    }
    d[0] = x1;
    d[(long)(n_row - 1) * rs] = xn;

    // update of the inner rows (:572-576)
    for (int j = 1; j < n_row - 1; j++) {
        const long o = (long)j * rs;
        d[o] = d[o] - a[o] * x1 - c[o] * xn;
    }
}

// Systems whose rows are contiguous in memory (the x sweep of an (i,j,k)
// array): one wavefront owns 64 systems and streams their rows through LDS
// in chunks of ROWS_R, so every global access is a coalesced run of
// ROWS_R doubles per line while each lane still walks its own system row by
// row in the reference's order.  LDS rows are padded to ROWS_R+1 doubles.
constexpr int ROWS_R = 16;
constexpr int ROWS_P = ROWS_R + 1;

// tile slot a holds array a of {A, B, C, D}; MASK selects the arrays moved
template <int MASK>
__device__ __forceinline__ void rows_load(double (*tile)[64 * ROWS_P], double *A, double *B, double *C, double *D,
                                          long lbase0, long ss0, int nlines, int r0, int nr) {
    const int t = threadIdx.x;
    double *const src[4] = {A, B, C, D};
    for (int q = 0; q < ROWS_R; q++) {
        const int idx = q * 64 + t;
        const int l = idx / ROWS_R, rr = idx - l * ROWS_R;
        if (l < nlines && rr < nr) {
            const long g = lbase0 + (long)l * ss0 + r0 + rr;
#pragma unroll
            for (int a = 0; a < 4; a++)
                if (MASK & (1 << a)) tile[a][l * ROWS_P + rr] = src[a][g];
        }
    }
}

// stores tile rows [lo, hi) of the chunk starting at global row r0
template <int MASK>
__device__ __forceinline__ void rows_store(double (*tile)[64 * ROWS_P], double *A, double *B, double *C, double *D,
                                           long lbase0, long ss0, int nlines, int r0, int lo, int hi) {
    const int t = threadIdx.x;
    double *const dst[4] = {A, B, C, D};
    for (int q = 0; q < ROWS_R; q++) {
        const int idx = q * 64 + t;
        const int l = idx / ROWS_R, rr = idx - l * ROWS_R;
        if (l < nlines && rr >= lo && rr < hi) {
            const long g = lbase0 + (long)l * ss0 + r0 + rr;
#pragma unroll
            for (int a = 0; a < 4; a++)
                if (MASK & (1 << a)) dst[a][g] = tile[a][l * ROWS_P + rr];
        }
    }
}

constexpr int M_ABCD = 15, M_ACD = 13, M_D = 8;

template <bool CYC>
__global__ void __launch_bounds__(64)
k_ptdma_rows(double *__restrict__ A, double *__restrict__ B, double *__restrict__ C, double *__restrict__ D,
             int n0, long ss0, long ss1, int n_row) {
    __shared__ double tile[4][64 * ROWS_P];
    const int t = threadIdx.x;
    const int s0a = blockIdx.x * 64;
    const int nlines = min(64, n0 - s0a);
    const long lbase0 = (long)s0a * ss0 + (long)blockIdx.y * ss1;
    const bool active = t < nlines;
    double *TA = tile[0] + t * ROWS_P;
    double *TB = tile[1] + t * ROWS_P;
    double *TC = tile[2] + t * ROWS_P;
    double *TD = tile[3] + t * ROWS_P;
    const int nch = (n_row + ROWS_R - 1) / ROWS_R;

    // forward elimination (rows 1,2 normalised; j = 3..n_row)
    double a0 = 0, c0 = 0, d0 = 0, a1 = 0, c1 = 0, d1 = 0;
    double ap = 0, cp = 0, dp = 0;
    for (int ch = 0; ch < nch; ch++) {
        const int r0 = ch * ROWS_R, nr = min(ROWS_R, n_row - r0);
        rows_load<M_ABCD>(tile, A, B, C, D, lbase0, ss0, nlines, r0, nr);
        __syncthreads();
        if (active) {
            for (int rr = 0; rr < nr; rr++) {
                const int j = r0 + rr;
                const double aj = TA[rr], bj = TB[rr], cj = TC[rr], dj = TD[rr];
                if (j == 0) {
                    a0 = aj / bj; d0 = dj / bj; c0 = cj / bj;
                } else if (j == 1) {
                    a1 = aj / bj; d1 = dj / bj; c1 = cj / bj;
                    TA[rr] = a1; TC[rr] = c1; TD[rr] = d1;
                    ap = a1; cp = c1; dp = d1;
                } else {
                    const double r = 1.0 / (bj - aj * cp);
                    const double dn = r * (dj - aj * dp);
                    const double cn = r * cj;
                    const double an = -r * aj * ap;
                    TA[rr] = an; TC[rr] = cn; TD[rr] = dn;
                    ap = an; cp = cn; dp = dn;
                }
            }
        }
        __syncthreads();
        // row 0 is written back after the row-1 fix-up; keep its input bits here
        rows_store<M_ACD>(tile, A, B, C, D, lbase0, ss0, nlines, r0, ch == 0 ? 1 : 0, nr);
        __syncthreads();
    }
    (void)a0;
    const double aN = ap, cN = cp, dN = dp;

    // backward elimination j = n_row-2..2 (1-based)
    double ax = a1, cx = c1, dx = d1;
    if (n_row >= 3) {
        const int jtop = n_row - 2;  // 0-based row whose forward values seed the sweep
        for (int ch = jtop / ROWS_R; ch >= 0; ch--) {
            const int r0 = ch * ROWS_R, nr = min(ROWS_R, n_row - r0);
            rows_load<M_ACD>(tile, A, B, C, D, lbase0, ss0, nlines, r0, nr);
            __syncthreads();
            if (active) {
                const int hi = min(r0 + nr - 1, jtop);
                for (int j = hi; j >= max(r0, 1); j--) {
                    const int rr = j - r0;
                    if (j == jtop) {
                        ax = TA[rr]; cx = TC[rr]; dx = TD[rr];
                        continue;
                    }
                    const double aj = TA[rr], cj = TC[rr], dj = TD[rr];
                    const double dn = dj - cj * dx;
                    const double an = aj - cj * ax;
                    const double cn = -cj * cx;
                    TA[rr] = an; TC[rr] = cn; TD[rr] = dn;
                    ax = an; cx = cn; dx = dn;
                }
            }
            __syncthreads();
            rows_store<M_ACD>(tile, A, B, C, D, lbase0, ss0, nlines, r0, ch == 0 ? 1 : 0, nr);
            __syncthreads();
        }
    }

    // row 1 fix-up and the reduced 2-row system
    double x1 = 0, xn = 0;
    {
        const double r = 1.0 / (1.0 - ax * c0);
        d0 = r * (d0 - c0 * dx);
        a0 = r * a0;
        c0 = -r * c0 * cx;
    }
    if (!CYC) {
        const double r = 1.0 / (1.0 - aN * c0);
        xn = r * (dN - aN * d0);
        x1 = d0 - c0 * xn;
    } else {
        const double e2 = -cN;
        x1 = (d0 - a0 * dN - c0 * dN) / (1.0 + a0 * e2 + c0 * e2);
        xn = dN + x1 * e2;
    }

    // update pass: rows 2..n_row-1 (1-based), rows 1 and n_row from the reduced solve
    for (int ch = 0; ch < nch; ch++) {
        const int r0 = ch * ROWS_R, nr = min(ROWS_R, n_row - r0);
        rows_load<M_ACD>(tile, A, B, C, D, lbase0, ss0, nlines, r0, nr);
        __syncthreads();
        if (active) {
            for (int rr = 0; rr < nr; rr++) {
                const int j = r0 + rr;
                if (j == 0) {
                    TA[rr] = a0; TC[rr] = c0; TD[rr] = x1;
                } else if (j == n_row - 1) {
                    TD[rr] = xn;
                } else {
                    TD[rr] = TD[rr] - TA[rr] * x1 - TC[rr] * xn;
                }
            }
        }
        __syncthreads();
        if (ch == 0) rows_store<M_ACD>(tile, A, B, C, D, lbase0, ss0, nlines, r0, 0, 1);  // row 1: A, C, D
        rows_store<M_D>(tile, A, B, C, D, lbase0, ss0, nlines, r0, ch == 0 ? 1 : 0, nr);
        __syncthreads();
    }
}

int ptdma_exact_launch(double *A, const double *B, double *C, double *D,
                       int n0, long ss0, int n1, long ss1, int n_row, long rs,
                       bool cyclic, hipStream_t st) {
    if (n0 <= 0 || n1 <= 0) return TCS_OK;
    if (n_row < 2 || n1 > 65535) {
        set_error("ptdma: n_row must be >= 2 and the outer system count <= 65535");
        return TCS_E_ARG;
    }
    if (rs == 1 && n_row >= 2) {
        dim3 rblk(64), rgrd((unsigned)((n0 + 63) / 64), (unsigned)n1);
        if (cyclic)
            TCS_LAUNCH(st, "ptdma_rows_cyclic", k_ptdma_rows<true>, rgrd, rblk, 0, A, (double *)B, C, D, n0, ss0, ss1, n_row);
        else
            TCS_LAUNCH(st, "ptdma_rows", k_ptdma_rows<false>, rgrd, rblk, 0, A, (double *)B, C, D, n0, ss0, ss1, n_row);
        return TCS_OK;
    }
    dim3 blk(256), grd((unsigned)((n0 + 255) / 256), (unsigned)n1);
    if (cyclic)
        TCS_LAUNCH(st, "ptdma_cyclic", k_ptdma_exact<true>, grd, blk, 0, A, B, C, D, n0, ss0, ss1, n_row, rs);
    else
        TCS_LAUNCH(st, "ptdma", k_ptdma_exact<false>, grd, blk, 0, A, B, C, D, n0, ss0, ss1, n_row, rs);
    return TCS_OK;
}

}  // namespace tcs

using namespace tcs;

static int many_solve(double *A, const double *B, double *C, double *D, int32_t n_sys,
                      int32_t n_row, bool cyc, void *stream) {
    if (!A || !B || !C || !D || n_sys < 0 || n_row < 2) {
        set_error("ptdma_many_solve: null pointer or n_row < 2");
        return TCS_E_ARG;
    }
    // (n_sys, n_row): systems contiguous; split n_sys into (n0 = chunk) x n1
    // so the outer count stays within the grid's y limit.
    const int chunk = 65536;
    int n0 = n_sys < chunk ? n_sys : chunk;
    int n1 = (n_sys + chunk - 1) / chunk;
    if (n_sys % n0 != 0) {
        // ragged tail: two launches
        int full = n_sys / n0;
        TCS_TRY(ptdma_exact_launch(A, B, C, D, n0, 1, full, n0, n_row, n_sys, cyc, (hipStream_t)stream));
        long off = (long)full * n0;
        return ptdma_exact_launch(A + off, B + off, C + off, D + off, n_sys - (int)off, 1, 1, 0, n_row,
                                  n_sys, cyc, (hipStream_t)stream);
    }
    return ptdma_exact_launch(A, B, C, D, n0, 1, n1, n0, n_row, n_sys, cyc, (hipStream_t)stream);
}

extern "C" int ptdma_many_solve_dev(double *A, const double *B, double *C, double *D, int32_t n_sys,
                                    int32_t n_row, void *stream) {
    return many_solve(A, B, C, D, n_sys, n_row, false, stream);
}

extern "C" int ptdma_many_solve_cycle_dev(double *A, const double *B, double *C, double *D,
                                          int32_t n_sys, int32_t n_row, void *stream) {
    return many_solve(A, B, C, D, n_sys, n_row, true, stream);
}

extern "C" int ptdma_many_solve_host(double *A, double *B, double *C, double *D, int32_t n_sys,
                                     int32_t n_row, int32_t cyclic) {
    if (!A || !B || !C || !D || n_sys <= 0 || n_row < 2) {
        set_error("ptdma_many_solve_host: bad arguments");
        return TCS_E_ARG;
    }
    size_t bytes = (size_t)n_sys * (size_t)n_row * sizeof(double);
    double *d[4] = {nullptr, nullptr, nullptr, nullptr};
    double *h[4] = {A, B, C, D};
    int rc = TCS_OK;
    for (int q = 0; q < 4 && rc == TCS_OK; q++) {
        if (hipMalloc(&d[q], bytes) != hipSuccess) { set_error("hipMalloc failed"); rc = TCS_E_NOMEM; }
        else if (hipMemcpy(d[q], h[q], bytes, hipMemcpyHostToDevice) != hipSuccess) { set_error("H2D copy failed"); rc = TCS_E_HIP; }
    }
    if (rc == TCS_OK) rc = many_solve(d[0], d[1], d[2], d[3], n_sys, n_row, cyclic != 0, nullptr);
    if (rc == TCS_OK && hipDeviceSynchronize() != hipSuccess) { set_error("ptdma kernel failed"); rc = TCS_E_HIP; }
    if (rc == TCS_OK) {
        for (int q : {0, 2, 3})
            if (hipMemcpy(h[q], d[q], bytes, hipMemcpyDeviceToHost) != hipSuccess) { set_error("D2H copy failed"); rc = TCS_E_HIP; }
    }
    for (int q = 0; q < 4; q++)
        if (d[q]) (void)hipFree(d[q]);
    return rc;
}

// ---- host solve without device allocation ---------------------------------
// ptdma_many_solve_host stages the four host arrays in device buffers it
// allocates and frees on every call.  The variant below stages them in one
// file-scope device array instead (each GPU the library runs on holds its
// own copy until the process ends), so a caller that must not allocate --
// or that calls the solve plane by plane in a loop -- pays only the copies.
// Four arrays of at most NOALLOC_MAX doubles each (4 x 16 MiB).  Every call
// on a GPU shares that GPU's array: calls that overlap on one GPU (from
// several host threads) race.  The array needs no clearing: each call
// overwrites the part it uses before any kernel reads it.
constexpr size_t NOALLOC_MAX = (size_t)1 << 21;
static __device__ __align__(256) double scratch[4 * NOALLOC_MAX];

constexpr int NOALLOC_DEVICES = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[NOALLOC_DEVICES];  // device address of `scratch` per device

// the current device's address of `scratch`, looked up once per device
static int scratch_address(double **out) {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= NOALLOC_DEVICES) {
        set_error("ptdma_many_solve_host_noalloc: no usable current device");
        return TCS_E_HIP;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error(std::string("ptdma_many_solve_host_noalloc: hipGetSymbolAddress failed: ") +
                      hipGetErrorString(e));
            return TCS_E_HIP;
        }
        g_scratch_ptr[dev] = (double *)p;
    }
    *out = g_scratch_ptr[dev];
    return TCS_OK;
}

extern "C" int ptdma_many_solve_host_noalloc(double *A, double *B, double *C, double *D, int32_t n_sys,
                                             int32_t n_row, int32_t cyclic) {
    if (!A || !B || !C || !D || n_sys <= 0 || n_row < 2) {
        set_error("ptdma_many_solve_host_noalloc: bad arguments");
        return TCS_E_ARG;
    }
    const size_t n = (size_t)n_sys * (size_t)n_row;
    if (n > NOALLOC_MAX) {
        set_error("ptdma_many_solve_host_noalloc: n_sys*n_row = " + std::to_string(n) + " exceeds the maximum " +
                  std::to_string(NOALLOC_MAX) + " of the static staging array (use ptdma_many_solve_host)");
        return TCS_E_ARG;
    }
    double *base = nullptr;
    TCS_TRY(scratch_address(&base));
    const size_t bytes = n * sizeof(double);
    double *d[4] = {base, base + NOALLOC_MAX, base + 2 * NOALLOC_MAX, base + 3 * NOALLOC_MAX};
    double *h[4] = {A, B, C, D};
    for (int q = 0; q < 4; q++)
        if (hipMemcpy(d[q], h[q], bytes, hipMemcpyHostToDevice) != hipSuccess) {
            set_error("ptdma_many_solve_host_noalloc: H2D copy failed");
            return TCS_E_HIP;
        }
    TCS_TRY(many_solve(d[0], d[1], d[2], d[3], n_sys, n_row, cyclic != 0, nullptr));
    if (hipStreamSynchronize(nullptr) != hipSuccess) {
        set_error("ptdma_many_solve_host_noalloc: kernel failed");
        return TCS_E_HIP;
    }
    for (int q : {0, 2, 3})
        if (hipMemcpy(h[q], d[q], bytes, hipMemcpyDeviceToHost) != hipSuccess) {
            set_error("ptdma_many_solve_host_noalloc: D2H copy failed");
            return TCS_E_HIP;
        }
    return TCS_OK;
}

extern "C" int ptdma_single_solve_dev(double *A, const double *B, double *C, double *D, int32_t n_row,
                                      int32_t cyclic, void *stream) {
    return many_solve(A, B, C, D, 1, n_row, cyclic != 0, stream);
}

extern "C" int ptdma_single_solve_host(double *A, double *B, double *C, double *D, int32_t n_row, int32_t cyclic) {
    return ptdma_many_solve_host(A, B, C, D, 1, n_row, cyclic);
}

extern "C" int ptdma_axis_solve_dev(double *A, const double *B, double *C, double *D, int32_t n1,
                                    int32_t n2, int32_t n3, int32_t axis, int32_t cyclic, void *stream) {
    if (!A || !B || !C || !D || n1 <= 0 || n2 <= 0 || n3 <= 0 || axis < 0 || axis > 2) {
        set_error("ptdma_axis_solve_dev: bad arguments");
        return TCS_E_ARG;
    }
    hipStream_t st = (hipStream_t)stream;
    long s1 = n1, s2 = (long)n1 * n2;
    switch (axis) {
    case 0:  // rows along i: systems (j,k)
        return ptdma_exact_launch(A, B, C, D, n2, s1, n3, s2, n1, 1, cyclic != 0, st);
    case 1:  // rows along j: systems (i,k)
        return ptdma_exact_launch(A, B, C, D, n1, 1, n3, s2, n2, s1, cyclic != 0, st);
    default:  // rows along k: systems (i,j)
        return ptdma_exact_launch(A, B, C, D, n1, 1, n2, s1, n3, s2, cyclic != 0, st);
    }
}
