// tcs_internal.h -- context layout and device helpers shared by the HIP
// translation units of libpascal_tcs.  Not part of the C-ABI.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <initializer_list>
#include <string>
#include <utility>

#include "../../include/pascal_tcs.h"

namespace tcs {

void set_error(const std::string &msg);

// Per-label HIP-event timing of kernel launches and stage entry points
// (the device counterpart of the reference's labelled timers,
// PaScaL_TDMA/src/timer.f90).  Active only while a context has it enabled;
// every launch in this library goes through TCS_LAUNCH.
struct Profiler;
struct AdjWork;  // adjoint fields and the step's kept values (adjoint.hip)
void prof_begin(const char *label, hipStream_t s);
void prof_end(hipStream_t s);

#define TCS_HIP(expr)                                                          \
    do {                                                                       \
        hipError_t _e = (expr);                                                \
        if (_e != hipSuccess) {                                                \
            ::tcs::set_error(std::string(#expr) + ": " + hipGetErrorString(_e)); \
            return TCS_E_HIP;                                                  \
        }                                                                      \
    } while (0)

#define TCS_LAUNCH_CHECK()                                                     \
    do {                                                                       \
        hipError_t _e = hipGetLastError();                                     \
        if (_e != hipSuccess) {                                                \
            ::tcs::set_error(std::string("kernel launch: ") + hipGetErrorString(_e)); \
            return TCS_E_HIP;                                                  \
        }                                                                      \
    } while (0)

#define TCS_LAUNCH(st, label, kern, grd, blk, lds, ...)                        \
    do {                                                                       \
        ::tcs::prof_begin(label, st);                                          \
        hipLaunchKernelGGL(kern, grd, blk, lds, st, __VA_ARGS__);              \
        ::tcs::prof_end(st);                                                   \
        TCS_LAUNCH_CHECK();                                                    \
    } while (0)

#define TCS_TRY(expr)                                                          \
    do {                                                                       \
        int _r = (expr);                                                       \
        if (_r != TCS_OK) return _r;                                           \
    } while (0)

// Grid description passed by value to every stencil kernel.
// Full fields: (0:n1m+1, 0:n2m+1, 0:n3m+1), i fastest, rows P1 >= N1
// elements apart (a multiple of 16) and each field allocated FIELD_FRONT
// elements after a 128-byte boundary, so the interior run i = 1.. of every
// row starts on a 128-byte line: a 16-cell row segment is one cache line.
// Interior work arrays: (1:n1m, 1:n2m, 1:n3m), i fastest.
constexpr int FIELD_FRONT = 15;
inline int field_pitch(int N1) { return (N1 + 15) / 16 * 16; }
struct Grid {
    int n1m, n2m, n3m;
    int N1, N2, N3;  // full extents n?m+2
    int P1;          // row pitch of the full fields
    const double *dx1, *dx2, *dx3, *dmx1, *dmx2, *dmx3;
    const double *rdx1, *rdx2, *rdx3, *rdmx1, *rdmx2, *rdmx3;  // their reciprocals
    // the six tables (dx1, dmx1, dx2, dmx2, dx3, dmx3) back to back rounded to
    // float, then their reciprocals in the same order (the mixed form's rows)
    const float *fmet;
    const int *jC, *jS;
    __device__ __forceinline__ size_t ix(int i, int j, int k) const {
        return (size_t)i + (size_t)P1 * ((size_t)j + (size_t)N2 * (size_t)k);
    }
    __device__ __forceinline__ size_t b2(int i, int k) const { return (size_t)i + (size_t)N1 * (size_t)k; }
    __device__ __forceinline__ size_t t2(int j, int k) const { return (size_t)j + (size_t)N2 * (size_t)k; }
    __device__ __forceinline__ size_t r3(int i, int j, int k) const {
        return (size_t)(i - 1) + (size_t)n1m * ((size_t)(j - 1) + (size_t)n2m * (size_t)(k - 1));
    }
};

// Polynomial property coefficients and nondimensional constants.
struct Phys {
    double Ct, Cmu, Cmt, Cmp, DeltaT;
    double a10, a11, a12, a13, a14, a15;
    double b10, b11, b12, b13, b14, b15;
    double c10, c11, c12, c13, c14, c15;
    double d10, d11, d12, d13, d14, d15;
    double KPP0, Rho0, Cp0, Mu0, a12pera11;
    // the fast forms' property polynomials, from the coefficients above:
    // c10/KPP0, Rho0*Cp0, and the numerator of d(1/(rho cp)) as a quintic
    double kppS, rcS, dircN[6];
};

}  // namespace tcs

struct tcs_ctx {
    tcs_params prm;
    tcs::Phys phys;
    int n1m, n2m, n3m;
    int N1, N2, N3;
    int P1;  // row pitch of the full fields (field_pitch(N1))
    size_t nfull, nint, nwall, ntab;
    int h1p, h3p;
    int device;
    int tdma_split;          // tcs_params.tdma_split
    hipStream_t stream;

    // host copies of the metrics
    double *h_dx1, *h_dx2, *h_dx3, *h_dmx1, *h_dmx2, *h_dmx3;
    // device metrics / masks
    double *dx1, *dx2, *dx3, *dmx1, *dmx2, *dmx3;
    double *rmet;            // 1/dx1, 1/dmx1, 1/dx2, 1/dmx2, 1/dx3, 1/dmx3 back to back,
                             // then the float tables of Grid::fmet
    int *jC, *jS;

    double *f[TCS_NFIELD];  // INVRHO / MU hold (0:n2sub,0:n3sub) tables
    // second buffers of T, U, V, W, dPhat, dP0: tcs_step writes their
    // updates there (the reads of the old values stay undisturbed) and swaps;
    // outside tcs_step their contents are dead (field uploads and downloads
    // stage through spare[TCS_T])
    double *spare[TCS_NFIELD];
    double *wall[TCS_NWALL];

    // ADI scratch (interior size each)
    double *am, *ac, *ap, *rhs;
    // stale scalars of the momentum stages (mu1,mu2 of dV; mu1..mu4 of dW)
    double *stale;

    // Poisson
    double *prhs;            // interior
    double *spec_x;          // complex (h1p, n2m, n3m), kx fastest
    double *spec_y;          // complex (h1p, n3m, n2m), kx fastest then kz
    double *pA, *pC;         // TDMA scratch (h1p*n3m, n2m)
    double *dxk2, *dzk;      // wave numbers
    double *ptab;            // y coefficients am(j), ap(j), ac(j) of the Poisson systems
    double *tw1, *tw3;       // twiddles (n/2 complex each)
    double *linesum, *planesum, *scal;  // mean reduction
    double *red;             // reduction scratch
    double *h_pinned;        // pinned host scalars
    tcs::Profiler *prof;     // null unless profiling is enabled
    tcs::AdjWork *adj;       // null until an adjoint entry is called
    int cfl_valid;           // cfl_max holds mpi_Post_CFL's maximum for the current fields
    double cfl_max;
    // routine calls held back so that the next call can fuse with them
    // (PEND_* bits; tcs::flush_pending runs them as called)
    int pend;
    double pend_dt_ldv, pend_dt_ldu, pend_dt_rhs;
    // The mixed form (precision 2): every array its stage rows and block
    // corrections read (T, U, V, W, P and the second buffers of T, U, V, W;
    // dU, dV, dW; the Mu and invRho tables; the wall tables; the stale
    // scalars) is carved from one FP64 arena a64 of an elements, and the
    // FP32 arena a32 holds each one's twin at the same element offset.  The
    // twins of the increments dU, dV, dW are where the mixed form keeps them
    // (the FP64 arrays are filled from them on download); the others are
    // FP32 copies of the FP64 arrays, current while their TWIN_* bit is set
    // in twin_ok.  Null / 0 in the FP64 forms.
    double *a64;
    float *a32;
    size_t an, aused;
    unsigned twin_ok;

    tcs::Grid grid() const {
        tcs::Grid g;
        g.n1m = n1m; g.n2m = n2m; g.n3m = n3m;
        g.N1 = N1; g.N2 = N2; g.N3 = N3;
        g.P1 = P1;
        g.dx1 = dx1; g.dx2 = dx2; g.dx3 = dx3;
        g.rdx1 = rmet; g.rdmx1 = rmet + N1; g.rdx2 = rmet + 2 * N1; g.rdmx2 = g.rdx2 + N2;
        g.rdx3 = g.rdx2 + 2 * N2; g.rdmx3 = g.rdx3 + N3;
        g.fmet = reinterpret_cast<const float *>(rmet + 2 * (N1 + N2 + N3));
        g.dmx1 = dmx1; g.dmx2 = dmx2; g.dmx3 = dmx3;
        g.jC = jC; g.jS = jS;
        return g;
    }
};

namespace tcs {

// launch helpers -------------------------------------------------------------
inline dim3 interior_block() { return dim3(64, 4, 1); }
// Row j = blockIdx.y * 4 + threadIdx.y + 1 of a kernel launched with
// interior_block(): the 64 threads of a wavefront share threadIdx.y, so j is
// the same on every lane, and readfirstlane lets the compiler see it (the
// (j, k) tables and j's metrics are then loaded once per wavefront).
__device__ __forceinline__ int wave_row_j() {
    return __builtin_amdgcn_readfirstlane((int)(blockIdx.y * blockDim.y + threadIdx.y + 1));
}
inline dim3 interior_grid(int nx, int ny, int nz) {
    return dim3((unsigned)((nx + 63) / 64), (unsigned)((ny + 3) / 4), (unsigned)nz);
}
// Cell-update stencil kernels where a thread takes KB consecutive k planes
// of one (i, j) (KB <= 4 divides n3m, a power of two >= 4): the values a
// plane shares with the next are loaded once.
inline dim3 cell_grid(int nx, int ny, int nz, int kb, dim3 blk = interior_block(), int ib = 1) {
    return dim3((unsigned)((nx / ib + blk.x - 1) / blk.x), (unsigned)((ny + blk.y - 1) / blk.y), (unsigned)(nz / kb));
}

// TDMA launcher (exact PaScaL order) on strided systems:
//   system (s0, s1) starts at base + s0*ss0 + s1*ss1, rows step rs.
int ptdma_exact_launch(double *A, const double *B, double *C, double *D,
                       int n0, long ss0, int n1, long ss1, int n_row, long rs,
                       bool cyclic, hipStream_t st);
// Where a split solve puts its solution: back into D (mode 0), or into an
// FP64 field F at F[base + s0*ss0 + s1*ss1 + row*rs], stored (mode 1),
// added to what F holds (mode 2), or added to what Fin holds at the same
// place (mode 3: F a second buffer of the field), or stored into the FP32
// field G (mode 4: the mixed form's increments) -- the last sweep of an ADI
// stage writes the increment straight into its module array.
struct LineOut {
    double *F = nullptr;
    const double *Fin = nullptr;
    float *G = nullptr;  // mode 4: the FP32 field G at the same place, stored; modes 1-3 of FP32
                         // lines: G (when set) also receives the FP32 copy of the FP64 value
    long base = 0, ss0 = 0, ss1 = 0, rs = 0;
    int mode = 0;
};
// The same systems split into parts (PaScaL_TDMA's multi-rank algorithm,
// one thread per rank); n_row in [2, 512]; strided lines need ss0 = 1.
bool ptdma_part_plan(int n_row, int &P, int &w1, int &w2);
int ptdma_part_launch(double *A, const double *B, double *C, double *D, int n0, long ss0, int n1, long ss1,
                      int n_row, long rs, bool cyclic, hipStream_t st, const LineOut &out = LineOut());
// FP32 lines (the mixed form's ADI solves), same algorithm and layouts.
int ptdma_part_launch(float *A, const float *B, float *C, float *D, int n0, long ss0, int n1, long ss1,
                      int n_row, long rs, bool cyclic, hipStream_t st, const LineOut &out = LineOut());
// ADI sweep of a context: split solve when the context asks for it and the
// line length allows, else the single-rank order.
inline int ptdma_sweep(const tcs_ctx *c, double *A, const double *B, double *C, double *D, int n0, long ss0,
                       int n1, long ss1, int n_row, long rs, bool cyclic) {
    int P, w1, w2;
    if (c->tdma_split && ptdma_part_plan(n_row, P, w1, w2) && (rs == 1 || ss0 == 1))
        return ptdma_part_launch(A, B, C, D, n0, ss0, n1, ss1, n_row, rs, cyclic, c->stream);
    return ptdma_exact_launch(A, B, C, D, n0, ss0, n1, ss1, n_row, rs, cyclic, c->stream);
}
// FP32 sweep of the mixed form: always the split solve (lines <= 512 rows).
inline int ptdma_sweep(const tcs_ctx *c, float *A, const float *B, float *C, float *D, int n0, long ss0, int n1,
                       long ss1, int n_row, long rs, bool cyclic) {
    return ptdma_part_launch(A, B, C, D, n0, ss0, n1, ss1, n_row, rs, cyclic, c->stream);
}

// Stage entry points of the kernel translation units (thermal, momentum,
// pressure).  Each of them is compiled three times: into namespace `exact`
// with FP contraction off (the reference's operation order, precision 0),
// into namespace `fast` with contraction and reciprocal arithmetic allowed
// (precision 1), and into namespace `mixed`, the fast arithmetic with FP32
// ADI scratch and FP32 line solves (precision 2).  TCS_FORM names the
// namespace a translation unit builds, TCS_SCRATCH its scratch type.
// The *_into / *_pseudo routines are tcs_step's fused forms: thermal's last
// sweep writes T + dT into a second buffer (when the lines are split); the
// block corrections also write the pseudo-update U + dU, V + dV, W + dW and
// the projection the new dPhat, dP0 into second buffers.
#define TCS_STAGE_DECLS                                                                 \
    int thermal_solver(tcs_ctx *c, double dt);                                          \
    int thermal_coeffi(tcs_ctx *c);                                                     \
    int momentum_coeffi(tcs_ctx *c);                                                    \
    int momentum_solvedU(tcs_ctx *c, double dt);                                        \
    int momentum_solvedV(tcs_ctx *c, double dt);                                        \
    int momentum_solvedW(tcs_ctx *c, double dt);                                        \
    int momentum_blockLdV(tcs_ctx *c, double dt);                                       \
    int momentum_blockLdU(tcs_ctx *c, double dt);                                       \
    int momentum_pseudoupdate(tcs_ctx *c);                                              \
    int pressure_rhs(tcs_ctx *c, double dt);                                            \
    int pressure_poisson(tcs_ctx *c);                                                   \
    int pressure_poisson_rhs(tcs_ctx *c, double dt);                                    \
    int pressure_projection(tcs_ctx *c, double dt);                                     \
    int thermal_solver_into(tcs_ctx *c, double dt, double *Tn, bool *written);          \
    int momentum_blockLdV_pseudo(tcs_ctx *c, double dt, double *Vn);                    \
    int momentum_blockLdU_pseudo(tcs_ctx *c, double dt, double *Un, double *Wn);        \
    int pressure_projection_into(tcs_ctx *c, double dt, double *dPhatn, double *dP0n);
namespace exact { TCS_STAGE_DECLS }
namespace fast { TCS_STAGE_DECLS }
namespace mixed { TCS_STAGE_DECLS }
#undef TCS_STAGE_DECLS
// up to FieldSet::MAX full fields for one launch (blockIdx.y picks one)
struct FieldSet {
    static constexpr int MAX = 6;
    double *f[MAX];
};
struct FieldSet32 {
    float *f[FieldSet::MAX];
};
int halo(tcs_ctx *c, double *f);
int halo(tcs_ctx *c, std::initializer_list<double *> fields);
// the same ghost-cell update of FP32 fields (the mixed form's increments)
int halo(tcs_ctx *c, std::initializer_list<float *> fields);
// The mixed form's FP32 twins of FP64 arrays (tcs_ctx::a32): which are
// current (tcs_ctx::twin_ok), making them current, and the increment of a
// field as the form keeps it (the FP64 array, or its FP32 twin).
enum {
    TWIN_T = 1, TWIN_U = 2, TWIN_V = 4, TWIN_W = 8, TWIN_P = 16,
    TWIN_TAB = 32,    // Mu, invRho
    TWIN_WALL = 64,   // the eight wall tables
    TWIN_STALE = 128  // the stale scalars of the momentum stages
};
int ensure_twins(tcs_ctx *c, unsigned bits);
// FP32 increment twin -> FP64 array (downloads of dU, dV, dW in the mixed
// form) and back (uploads)
int increment_to_fp64(tcs_ctx *c, int id);
int increment_from_fp64(tcs_ctx *c, int id);
// copy the cells outside the interior (x/z ghosts, wall rows) of a full field
int copy_shell(tcs_ctx *c, const double *src, double *dst);
int copy_shell(tcs_ctx *c, std::initializer_list<std::pair<const double *, double *>> pairs);
// Routine entries that hold their launches back: tcs_momentum_blockLdV, the
// ghost-cell update of dV after it and tcs_momentum_blockLdU wait for
// tcs_momentum_pseudoupdateUVW (which then runs the block corrections with
// the pseudo-update fused in), tcs_pressure_rhs waits for
// tcs_pressure_poisson_fft2 (the right-hand side made inside the x
// transform).  Every other entry point runs what is held first, unfused.
enum { PEND_LDV = 1, PEND_HALO_DV = 2, PEND_LDU = 4, PEND_RHS = 8 };
int flush_pending(tcs_ctx *c);
int post_div(tcs_ctx *c, double *out);
int post_cfl(tcs_ctx *c, double *maxuovx);
int nusselt(tcs_ctx *c, double Tup, double Tbt, double *nb, double *nt);
void adjoint_free(tcs_ctx *c);

}  // namespace tcs
