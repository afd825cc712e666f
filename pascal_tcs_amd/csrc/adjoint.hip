// adjoint.hip -- the discrete adjoint of one MPM-STD step (src/main.f90:131-203
// in the plain FP64 form, precision 0).
//
// The stage units are compiled here a fourth time, into namespace tcs::adj,
// so the adjoint uses the step's own row and cell formulas.  tcs_adjoint_step
// first runs the step forward from the context's state, keeping the inputs
// every nonlinear part needs and the three ADI solutions of every stage;
// then it walks the step backwards:
//   * halos, wall extrapolation, the dP mean: their transposes (the ghost
//     adjoints go to the periodic sources, the mean projection is symmetric);
//   * cell updates (projection, dPhat history, block corrections, pressure
//     RHS, pseudo-update): vector-Jacobian products of the cell formulas
//     (ad.h: each cell's recorded inputs, one dual-number pass per input);
//   * each ADI stage x = A^-1 d: lambda = A^-T xbar by the transposed line
//     solve (the stage's rows, transposed in place along the line, solved by
//     the single-rank PaScaL_TDMA kernel), row adjoints abar_i = -lambda_i
//     x_{i-1}, bbar_i = -lambda_i x_i, cbar_i = -lambda_i x_{i+1},
//     dbar_i = lambda_i, then the rows' vector-Jacobian product;
//   * the Poisson solve: the same FFT pipeline with every wavenumber's
//     y-system transposed (the pinned mean mode's row projection after the
//     solve, the zeroed imaginary parts before it), which is the exact
//     transpose because the modified wavenumbers are even in kx and kz.
// The adjoint fields of the state (T, U, V, W, P, dPhat, dP0) and of the
// temperature wall arrays are kept in the context.
#define TCS_FORM adj
#include "momentum.hip"
#include "pressure.hip"
#include "thermal.hip"

#include <algorithm>
#include <utility>

namespace tcs {
namespace adj {

constexpr int NSTATE = 7;
constexpr int STATE_IDS[NSTATE] = {TCS_T, TCS_U, TCS_V, TCS_W, TCS_P, TCS_DPHAT, TCS_DP0};
enum { ST_T, ST_U, ST_V, ST_W, ST_P, ST_DPHAT, ST_DP0 };
enum { STG_T, STG_U, STG_V, STG_W };

}  // namespace adj

struct AdjWork {
    double *b[adj::NSTATE];         // adjoints of the state fields
    double *bdU, *bdV, *bdW, *bdP;  // of the increments (within a step)
    double *bMu, *biR, *bstale;     // of the (j,k) tables and the stale scalars
    double *bwall[TCS_NWALL];       // of the wall arrays
    double *tmp[4];                 // full-size work
    double *s0[adj::NSTATE];        // state at the start of the step
    double *sT1, *sdU1, *sdV1, *sdW1, *sdV2, *sdU2, *sU1, *sV1, *sW1, *sdP, *sMu, *siR, *sstale;
    double *X[4][3];                // ADI stage solutions (interior layout)
    double *lam, *abar, *bbar, *cbar, *aT, *cT, *rint;  // interior layout
    double *scal;
    int *flag;
    std::vector<void *> all;
};

namespace adj {

// ---- transposes of the linear pieces -----------------------------------------

// k_halo transposed: every ghost's adjoint goes to its periodic source (the
// same enumeration as k_halo; sources are never x or z ghosts)
__global__ void k_halo_adj(Grid g, double *__restrict__ b) {
    const int N1 = g.N1, N2 = g.N2, N3 = g.N3;
    const int n1m = g.n1m, n3m = g.n3m;
    const size_t nx = (size_t)2 * N2 * N3, nz = (size_t)2 * N1 * N2;
    const size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t dst, src;
    if (q < nx) {
        const int side = (int)(q % 2);
        const size_t r = q / 2;
        const int j = (int)(r % N2), k = (int)(r / N2);
        if (k == 0 || k == N3 - 1) return;
        dst = g.ix(side ? n1m + 1 : 0, j, k);
        src = g.ix(side ? 1 : n1m, j, k);
    } else if (q < nx + nz) {
        const size_t r = q - nx;
        const int side = (int)(r % 2);
        const size_t t = r / 2;
        const int i = (int)(t % N1), j = (int)(t / N1);
        const int is = (i == 0) ? n1m : (i == n1m + 1 ? 1 : i);
        dst = g.ix(i, j, side ? n3m + 1 : 0);
        src = g.ix(is, j, side ? 1 : n3m);
    } else {
        return;
    }
    const double v = b[dst];
    b[dst] = 0.0;
    if (v != 0.0) unsafeAtomicAdd(b + src, v);  // the L2's FP64 atomic (not a compare-and-swap loop)
}

int halo_adj(tcs_ctx *c, double *b) {
    const size_t n = (size_t)2 * c->N2 * c->N3 + (size_t)2 * c->N1 * c->N2;
    TCS_LAUNCH(c->stream, "adj_halo", k_halo_adj, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->grid(), b);
    return TCS_OK;
}

// k_wall_extrap transposed
__global__ void k_wall_extrap_adj(Grid g, double *__restrict__ b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int k = blockIdx.y + 1;
    if (i > g.n1m) return;
    const double *dmx2 = g.dmx2;
    const int n2sub = g.n2m + 1, n2msub = g.n2m;
    const double s = dmx2[1] + dmx2[2];
    const double a0 = (s * s) / ((s * s) - (dmx2[1] * dmx2[1]));
    const double b0 = (dmx2[1] * dmx2[1]) / ((s * s) - (dmx2[1] * dmx2[1]));
    const double t = dmx2[n2sub] + dmx2[n2msub];
    const double a1 = (t * t) / ((t * t) - (dmx2[n2sub] * dmx2[n2sub]));
    const double b1 = (dmx2[n2sub] * dmx2[n2sub]) / ((t * t) - (dmx2[n2sub] * dmx2[n2sub]));
    const double gb = b[g.ix(i, 0, k)], gt = b[g.ix(i, n2sub, k)];
    b[g.ix(i, 0, k)] = 0.0;
    b[g.ix(i, n2sub, k)] = 0.0;
    b[g.ix(i, 1, k)] += a0 * gb;
    b[g.ix(i, 2, k)] -= b0 * gb;
    b[g.ix(i, n2msub, k)] += a1 * gt;
    b[g.ix(i, n2sub - 2, k)] -= b1 * gt;
}

int wall_extrap_adj(tcs_ctx *c, double *b) {
    TCS_LAUNCH(c->stream, "adj_wall_extrap", k_wall_extrap_adj, dim3((c->n1m + 63) / 64, c->n3m), dim3(64), 0,
               c->grid(), b);
    return TCS_OK;
}

// interior sum (one atomic per workgroup) and the mean projection
__global__ void k_interior_sum(Grid g, const double *__restrict__ b, double *__restrict__ sum) {
    __shared__ double s[256];
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    const int t = threadIdx.x + blockDim.x * threadIdx.y;
    s[t] = (i <= g.n1m && j <= g.n2m) ? b[g.ix(i, j, k)] : 0.0;
    __syncthreads();
    for (int h = 128; h >= 1; h >>= 1) {
        if (t < h) s[t] += s[t + h];
        __syncthreads();
    }
    if (t == 0) unsafeAtomicAdd(sum, s[0]);
}
__global__ void k_interior_sub(Grid g, const double *__restrict__ sum, double *__restrict__ b) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.n1m || j > g.n2m) return;
    b[g.ix(i, j, k)] -= sum[0] / ((double)g.n1m * (double)g.n2m * (double)g.n3m);
}

// full-field layout <-> interior layout
__global__ void k_full_to_int(Grid g, int jlo, const double *__restrict__ f, double *__restrict__ r) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.n1m || j > g.n2m) return;
    r[g.r3(i, j, k)] = j >= jlo ? f[g.ix(i, j, k)] : 0.0;
}

// zero the interior cells j >= jlo of a full field
__global__ void k_zero_interior(Grid g, int jlo, double *__restrict__ f) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.n1m || j > g.n2m || j < jlo) return;
    f[g.ix(i, j, k)] = 0.0;
}

// transposed Poisson y-systems, one (kx, kz) system per thread (the rows of
// k_poisson_tdma: A = amj, B = acj - dxk2 - dzk, C = apj, the mean mode's
// first row pinned); imaginary parts zeroed before, the pinned row after
__global__ void k_poisson_tdma_T(int n3m, int h1p, int n2m, int h3p, const double *__restrict__ amj,
                                 const double *__restrict__ apj, const double *__restrict__ acj,
                                 const double *__restrict__ dxk2, const double *__restrict__ dzk,
                                 double2 *__restrict__ S, double *__restrict__ cw) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int nsys = n3m * h1p;
    if (s >= nsys) return;
    const int kx = s / n3m, kz = s - kx * n3m;
    const size_t rs = (size_t)nsys;
    double2 *d = S + s;
    double *cp = cw + s;
    const double shift = dxk2[kx], zk = dzk[kz];
    const bool pin = (kx == 0 && kz == 0);
    const bool zim = (kx == 0 && (kz == 0 || kz == h3p - 1));
    auto A = [&](int jj) { return (pin && jj == 0) ? 0.0 : amj[jj]; };
    auto B = [&](int jj) { return (pin && jj == 0) ? 1.0 : acj[jj] - shift - zk; };
    auto C = [&](int jj) { return (pin && jj == 0) ? 0.0 : apj[jj]; };
    // row jj of the transpose: (C(jj-1), B(jj), A(jj+1)); ends dropped (non-cyclic)
    double cprev = 0.0, dr = 0.0, di = 0.0;
    for (int jj = 0; jj < n2m; jj++) {
        const double at = jj > 0 ? C(jj - 1) : 0.0;
        const double ct = jj < n2m - 1 ? A(jj + 1) : 0.0;
        const double m = 1.0 / (B(jj) - at * cprev);
        const double2 v = d[(size_t)jj * rs];
        const double vi = zim ? 0.0 : v.y;
        dr = (v.x - at * dr) * m;
        di = (vi - at * di) * m;
        cprev = ct * m;
        cp[(size_t)jj * rs] = cprev;
        d[(size_t)jj * rs] = make_double2(dr, di);
    }
    double xr = dr, xi = di;
    for (int jj = n2m - 2; jj >= 0; jj--) {
        const double2 v = d[(size_t)jj * rs];
        const double c_ = cp[(size_t)jj * rs];
        xr = v.x - c_ * xr;
        xi = v.y - c_ * xi;
        d[(size_t)jj * rs] = make_double2(xr, xi);
    }
    if (pin) d[0] = make_double2(0.0, d[0].y);
}

// ---- ADI stage transposes ------------------------------------------------------

// the line through interior cell q along `axis` (0 x, 1 y, 2 z), rows j >= jlo:
// position, length and stride of the line
struct LinePos {
    int pos, n;
    long stride;
};
__device__ __forceinline__ LinePos line_pos(const Grid &g, int axis, int jlo, int i, int j, int k) {
    const long n1 = g.n1m, n12 = (long)g.n1m * g.n2m;
    if (axis == 0) return {i - 1, g.n1m, 1};
    if (axis == 1) return {j - jlo, g.n2m - jlo + 1, n1};
    return {k - 1, g.n3m, n12};
}

// transposed rows in place of the stage's rows: aT_i = c_{i-1}, cT_i = a_{i+1}
__global__ void k_coef_T(Grid g, int axis, int jlo, int cyc, const double *__restrict__ am, const double *__restrict__ ap,
                         double *__restrict__ aT, double *__restrict__ cT) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.n1m || j > g.n2m || j < jlo) return;
    const size_t q = g.r3(i, j, k);
    const LinePos L = line_pos(g, axis, jlo, i, j, k);
    const bool hp = L.pos > 0 || cyc, hn = L.pos < L.n - 1 || cyc;
    const long qp = L.pos > 0 ? (long)q - L.stride : (long)q + (long)(L.n - 1) * L.stride;
    const long qn = L.pos < L.n - 1 ? (long)q + L.stride : (long)q - (long)(L.n - 1) * L.stride;
    aT[q] = hp ? ap[qp] : 0.0;
    cT[q] = hn ? am[qn] : 0.0;
}

// row adjoints of a solved stage from lambda = A^-T xbar and the solution x
__global__ void k_row_bars(Grid g, int axis, int jlo, int cyc, const double *__restrict__ lam,
                           const double *__restrict__ x, double *__restrict__ ab, double *__restrict__ bb,
                           double *__restrict__ cb) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.n1m || j > g.n2m || j < jlo) return;
    const size_t q = g.r3(i, j, k);
    const LinePos L = line_pos(g, axis, jlo, i, j, k);
    const bool hp = L.pos > 0 || cyc, hn = L.pos < L.n - 1 || cyc;
    const long qp = L.pos > 0 ? (long)q - L.stride : (long)q + (long)(L.n - 1) * L.stride;
    const long qn = L.pos < L.n - 1 ? (long)q + L.stride : (long)q - (long)(L.n - 1) * L.stride;
    const double l = lam[q];
    ab[q] = hp ? -l * x[qp] : 0.0;
    bb[q] = -l * x[q];
    cb[q] = hn ? -l * x[qn] : 0.0;
}

template <class Cell>
int cell_vjp(tcs_ctx *c, AdjWork *w, const Cell &cell, int jlo, OutBars ob, InBars<Cell::NF> ib) {
    // the LDS input sets keep an index in 28 bits: the largest array (a full
    // field) must be indexed within it
    if (c->nfull + FIELD_FRONT + 1 > ((size_t)1 << AD_KEY_IBITS)) return TCS_E_ARG;
    const dim3 blk(64), grd((unsigned)((c->n1m + 63) / 64), (unsigned)(c->n2m - jlo + 1), (unsigned)c->n3m);
    TCS_LAUNCH(c->stream, "adj_cell_vjp", (k_cell_vjp<Cell>), grd, blk, 0, cell, jlo, c->n2m, ob, ib, w->flag);
    return TCS_OK;
}

// Line geometry of a stage's solves over the rows j >= jlo (as the forward
// stages launch them): (n0, ss0, n1, ss1, n_row, rs) and the offset of row jlo.
struct LineGeom {
    long off;
    int n0;
    long ss0;
    int n1;
    long ss1;
    int n_row;
    long rs;
};
static LineGeom geom(const tcs_ctx *c, int axis, int jlo) {
    const long n1 = c->n1m, n12 = (long)c->n1m * c->n2m, off = (long)(jlo - 1) * n1;
    const int nl = c->n2m - jlo + 1;
    if (axis == 2) return {off, c->n1m, 1, nl, n1, c->n3m, n12};
    if (axis == 0) return {off, nl, n1, c->n3m, n12, c->n1m, 1};
    return {off, c->n1m, 1, c->n3m, n12, nl, n1};
}

static int sweep(tcs_ctx *c, double *A, const double *B, double *C, double *D, int axis, int jlo, bool cyc) {
    const LineGeom L = geom(c, axis, jlo);
    return ptdma_sweep(c, A + L.off, B + L.off, C + L.off, D + L.off, L.n0, L.ss0, L.n1, L.ss1, L.n_row, L.rs, cyc);
}

// One stage backwards.  On entry w->lam holds xbar (interior layout, rows
// j >= jlo); on exit it holds lambda = A^-T xbar, which is the adjoint of the
// stage's right-hand side (the previous stage's solution for x and y
// stages; the rows' own d for the z stage, HAS_D).
template <bool HAS_D, class Row>
int stage_adj(tcs_ctx *c, AdjWork *w, const Row &row, int axis, int jlo, bool cyc, const double *X,
              InBars<Row::NF> ib) {
    const Grid g = c->grid();
    const dim3 blk = interior_block(), grd = interior_grid(c->n1m, c->n2m, c->n3m);
    double *am = c->am, *ac = c->ac, *ap = c->ap;
    TCS_TRY(rows3d_launch<HAS_D>(c, "adj_rows", row, jlo, am, ac, ap, w->rint));
    TCS_LAUNCH(c->stream, "adj_coef_T", k_coef_T, grd, blk, 0, g, axis, jlo, (int)cyc, am, ap, w->aT, w->cT);
    {
        const LineGeom L = geom(c, axis, jlo);
        TCS_TRY(ptdma_exact_launch(w->aT + L.off, ac + L.off, w->cT + L.off, w->lam + L.off, L.n0, L.ss0, L.n1, L.ss1,
                                   L.n_row, L.rs, cyc, c->stream));
    }
    TCS_LAUNCH(c->stream, "adj_row_bars", k_row_bars, grd, blk, 0, g, axis, jlo, (int)cyc, w->lam, X, w->abar, w->bbar,
               w->cbar);
    OutBars ob{{w->abar, w->bbar, w->cbar, HAS_D ? w->lam : nullptr}, 0};
    return cell_vjp(c, w, row, jlo, ob, ib);
}

// ---- small kernels -----------------------------------------------------------

// U += dU, V += dV (j >= 2; times 0 at j = 1), W += dW transposed
__global__ void k_pseudo_adj(Grid g, const double *__restrict__ bU, const double *__restrict__ bV,
                             const double *__restrict__ bW, double *__restrict__ bdU, double *__restrict__ bdV,
                             double *__restrict__ bdW) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.n1m || j > g.n2m) return;
    const size_t q = g.ix(i, j, k);
    bdU[q] += bU[q];
    bdV[q] += j >= 2 ? bV[q] : bV[q] * 0.0;
    bdW[q] += bW[q];
}

// Mu, invRho tables from T(n1sub, j, k) transposed
__global__ void k_coeffi_adj(Grid g, Phys p, const double *__restrict__ T, const double *__restrict__ bMu,
                             const double *__restrict__ biR, double *__restrict__ bT) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (j >= g.N2) return;
    const size_t q = g.ix(g.n1m + 1, j, k);
    Dual mu, ir;
    mu_ir_of_T(p, Dual(T[q], 1.0), mu, ir);
    bT[q] += mu.d * bMu[g.t2(j, k)] + ir.d * biR[g.t2(j, k)];
}

// the stale viscosities transposed (one thread; six outputs, few inputs)
__global__ void k_stale_adj(Grid g, const double *__restrict__ Mu, const double *__restrict__ bstale,
                            double *__restrict__ bMu) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    Rec rec;
    rec.n = 0;
    double v[6];
    stale_values<double>(g, RecPtr{Mu, 0, &rec}, v);
    for (int r = 0; r < rec.n && r < AD_MAXIN; r++) {
        Dual o[6];
        stale_values<Dual>(g, SeedPtr{Mu, rec.idx[r]}, o);
        double s = 0.0;
        for (int q = 0; q < 6; q++) s += o[q].d * bstale[q];
        bMu[rec.idx[r]] += s;
    }
}

// the wall Nusselt number's gradient, times weight: Nu_bottom =
// sum over the interior (TBCbt - T(i,1,k)) / dmx2(1) * L2 / (n1m n3m dT)
__global__ void k_nusselt_adj(Grid g, double wgt, double *__restrict__ bT, double *__restrict__ bTbt) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int k = blockIdx.y + 1;
    if (i > g.n1m) return;
    bT[g.ix(i, 1, k)] -= wgt;
    bTbt[g.b2(i, k)] += wgt;
}

// ---- the step forward, keeping what the adjoint needs ----------------------------

static int copy(tcs_ctx *c, double *dst, const double *src, size_t n) {
    TCS_HIP(hipMemcpyAsync(dst, src, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    return TCS_OK;
}

// ADI of one increment (or of T), forward: the three solutions into X[0..2]
template <class RZ, class RX, class RY>
static int adi_forward(tcs_ctx *c, const RZ &rz, const RX &rx, const RY &ry, int jlo, double *const (&X)[3]) {
    double *am = c->am, *ac = c->ac, *ap = c->ap, *rhs = c->rhs;
    TCS_TRY(rows3d_launch<true>(c, "adj_fwd_rows", rz, jlo, am, ac, ap, rhs));
    TCS_TRY(sweep(c, am, ac, ap, rhs, 2, jlo, true));
    TCS_TRY(copy(c, X[0], rhs, c->nint));
    TCS_TRY(rows3d_launch<false>(c, "adj_fwd_rows", rx, jlo, am, ac, ap, rhs));
    TCS_TRY(sweep(c, am, ac, ap, rhs, 0, jlo, true));
    TCS_TRY(copy(c, X[1], rhs, c->nint));
    TCS_TRY(rows3d_launch<false>(c, "adj_fwd_rows", ry, jlo, am, ac, ap, rhs));
    TCS_TRY(sweep(c, am, ac, ap, rhs, 1, jlo, false));
    TCS_TRY(copy(c, X[2], rhs, c->nint));
    return TCS_OK;
}

static int forward_keep(tcs_ctx *c, AdjWork *w, double dt) {
    const Grid g = c->grid();
    double **f = c->f;
    const dim3 blk = interior_block(), grd = interior_grid(c->n1m, c->n2m, c->n3m);
    for (int s = 0; s < NSTATE; s++) TCS_TRY(copy(c, w->s0[s], f[STATE_IDS[s]], c->nfull));
    // thermal
    TCS_TRY(adi_forward(c, ThermalZ{g, c->phys, dt, f[TCS_T], f[TCS_U], f[TCS_V], f[TCS_W], c->wall[TCS_TBCUP],
                                    c->wall[TCS_TBCBT]},
                        ThermalX{g, c->phys, dt, f[TCS_T], f[TCS_U]}, ThermalY{g, c->phys, dt, f[TCS_T], f[TCS_V]}, 1,
                        w->X[STG_T]));
    TCS_LAUNCH(c->stream, "adj_fwd_add", k_interior_out<2>, grd, blk, 0, g, 1, f[TCS_T], w->X[STG_T][2]);
    TCS_TRY(halo(c, f[TCS_T]));
    TCS_TRY(copy(c, w->sT1, f[TCS_T], c->nfull));
    TCS_TRY(momentum_coeffi(c));
    TCS_TRY(copy(c, w->sMu, f[TCS_MU], c->ntab));
    TCS_TRY(copy(c, w->siR, f[TCS_INVRHO], c->ntab));
    // solvedU
    TCS_TRY(adi_forward(c, DUZ{g, c->phys, dt, f[TCS_U], f[TCS_V], f[TCS_W], f[TCS_P], f[TCS_T], f[TCS_MU],
                               f[TCS_INVRHO], c->wall[TCS_UBCUP], c->wall[TCS_UBCBT], c->wall[TCS_VBCUP],
                               c->wall[TCS_VBCBT]},
                        DUX{g, c->phys, dt, f[TCS_U], f[TCS_MU], f[TCS_INVRHO]},
                        DUY{g, c->phys, dt, f[TCS_V], f[TCS_MU], f[TCS_INVRHO]}, 1, w->X[STG_U]));
    TCS_LAUNCH(c->stream, "adj_fwd_store", k_interior_out<1>, grd, blk, 0, g, 1, f[TCS_DU], w->X[STG_U][2]);
    TCS_TRY(halo(c, f[TCS_DU]));
    TCS_TRY(copy(c, w->sdU1, f[TCS_DU], c->nfull));
    // solvedV (rows j >= 2)
    TCS_TRY(launch_stale(c));
    TCS_TRY(copy(c, w->sstale, c->stale, 6));
    TCS_TRY(adi_forward(c, DVZ{g, c->phys, dt, f[TCS_U], f[TCS_V], f[TCS_W], f[TCS_P], f[TCS_DU], f[TCS_MU],
                               f[TCS_INVRHO], c->wall[TCS_VBCUP], c->wall[TCS_VBCBT]},
                        DVX{g, c->phys, dt, f[TCS_U], f[TCS_INVRHO], c->stale},
                        DVY{g, c->phys, dt, f[TCS_V], f[TCS_MU], f[TCS_INVRHO]}, 2, w->X[STG_V]));
    TCS_LAUNCH(c->stream, "adj_fwd_zero", k_dV_zero_rows, dim3((c->N1 + 63) / 64, c->N3), dim3(64), 0, g, f[TCS_DV]);
    TCS_LAUNCH(c->stream, "adj_fwd_store", k_interior_out<1>, grd, blk, 0, g, 2, f[TCS_DV], w->X[STG_V][2]);
    TCS_TRY(halo(c, f[TCS_DV]));
    TCS_TRY(copy(c, w->sdV1, f[TCS_DV], c->nfull));
    // solvedW
    TCS_TRY(launch_stale(c));
    TCS_TRY(adi_forward(c, DWZ{g, c->phys, dt, f[TCS_U], f[TCS_V], f[TCS_W], f[TCS_P], f[TCS_DU], f[TCS_DV],
                               f[TCS_MU], f[TCS_INVRHO], c->wall[TCS_VBCUP], c->wall[TCS_VBCBT],
                               c->wall[TCS_WBCUP], c->wall[TCS_WBCBT]},
                        DWX{g, c->phys, dt, f[TCS_U], f[TCS_INVRHO], c->stale},
                        DWY{g, c->phys, dt, f[TCS_V], f[TCS_INVRHO], c->stale}, 1, w->X[STG_W]));
    TCS_LAUNCH(c->stream, "adj_fwd_store", k_interior_out<1>, grd, blk, 0, g, 1, f[TCS_DW], w->X[STG_W][2]);
    TCS_TRY(halo(c, f[TCS_DW]));
    TCS_TRY(copy(c, w->sdW1, f[TCS_DW], c->nfull));
    // block corrections, pseudo-update
    TCS_TRY(momentum_blockLdV(c, dt));
    TCS_TRY(halo(c, f[TCS_DV]));
    TCS_TRY(copy(c, w->sdV2, f[TCS_DV], c->nfull));
    TCS_TRY(momentum_blockLdU(c, dt));
    TCS_TRY(copy(c, w->sdU2, f[TCS_DU], c->nfull));
    TCS_TRY(momentum_pseudoupdate(c));
    TCS_TRY(halo(c, f[TCS_U]));
    TCS_TRY(halo(c, f[TCS_V]));
    TCS_TRY(halo(c, f[TCS_W]));
    TCS_TRY(copy(c, w->sU1, f[TCS_U], c->nfull));
    TCS_TRY(copy(c, w->sV1, f[TCS_V], c->nfull));
    TCS_TRY(copy(c, w->sW1, f[TCS_W], c->nfull));
    // pressure
    TCS_TRY(pressure_rhs(c, dt));
    TCS_TRY(pressure_poisson(c));
    TCS_TRY(halo(c, f[TCS_DP]));
    TCS_TRY(copy(c, w->sdP, f[TCS_DP], c->nfull));
    TCS_TRY(pressure_projection(c, dt));
    TCS_TRY(halo(c, f[TCS_U]));
    TCS_TRY(halo(c, f[TCS_V]));
    TCS_TRY(halo(c, f[TCS_W]));
    TCS_TRY(halo(c, f[TCS_P]));
    TCS_TRY(halo(c, f[TCS_DPHAT]));
    return TCS_OK;
}

// ---- the step backwards ----------------------------------------------------------

// Before a cell update that overwrites field F's interior rows j >= jlo: its
// new adjoint starts as the old one outside those cells (untouched values
// pass through) and zero inside (filled by the update's product).
static int overwrite_begin(tcs_ctx *c, double *tmp, const double *bar, int jlo) {
    TCS_TRY(copy(c, tmp, bar, c->nfull));
    TCS_LAUNCH(c->stream, "adj_zero", k_zero_interior, interior_grid(c->n1m, c->n2m, c->n3m), interior_block(), 0,
               c->grid(), jlo, tmp);
    return TCS_OK;
}

static int backward(tcs_ctx *c, AdjWork *w, double dt) {
    const Grid g = c->grid();
    const dim3 blk = interior_block(), grd = interior_grid(c->n1m, c->n2m, c->n3m);
    double **b = w->b;
    const double *TBCup = c->wall[TCS_TBCUP], *TBCbt = c->wall[TCS_TBCBT];
    // halos after the projection
    TCS_TRY(halo_adj(c, b[ST_DPHAT]));
    TCS_TRY(halo_adj(c, b[ST_P]));
    TCS_TRY(halo_adj(c, b[ST_W]));
    TCS_TRY(halo_adj(c, b[ST_V]));
    TCS_TRY(halo_adj(c, b[ST_U]));
    // P wall extrapolation; dPhat / dP0 history; projection
    TCS_TRY(wall_extrap_adj(c, b[ST_P]));
    {
        const DPhatUpdate cell{g, w->sdP, w->s0[ST_DPHAT], w->s0[ST_DP0]};
        TCS_TRY(overwrite_begin(c, w->tmp[0], b[ST_DPHAT], 1));
        TCS_TRY(overwrite_begin(c, w->tmp[1], b[ST_DP0], 1));
        TCS_TRY(cell_vjp(c, w, cell, 1, OutBars{{b[ST_DPHAT], b[ST_DP0], nullptr, nullptr}, 1},
                         InBars<3>{{w->bdP, w->tmp[0], w->tmp[1]}}));
        std::swap(b[ST_DPHAT], w->tmp[0]);
        std::swap(b[ST_DP0], w->tmp[1]);
    }
    {
        const Projection cell{g, c->phys, dt, w->siR, w->sdP, w->s0[ST_DPHAT], w->s0[ST_DP0],
                              w->sU1, w->sV1, w->sW1, w->s0[ST_P]};
        for (int q = 0; q < 4; q++) TCS_TRY(overwrite_begin(c, w->tmp[q], b[ST_U + q], 1));
        TCS_TRY(cell_vjp(c, w, cell, 1, OutBars{{b[ST_U], b[ST_V], b[ST_W], b[ST_P]}, 1},
                         InBars<8>{{w->biR, w->bdP, b[ST_DPHAT], b[ST_DP0], w->tmp[0], w->tmp[1], w->tmp[2],
                                    w->tmp[3]}}));
        for (int q = 0; q < 4; q++) std::swap(b[ST_U + q], w->tmp[q]);
    }
    // Poisson: halo(dP), wall extrapolation, mean, then the transposed solve
    TCS_TRY(halo_adj(c, w->bdP));
    TCS_TRY(wall_extrap_adj(c, w->bdP));
    TCS_HIP(hipMemsetAsync(w->scal, 0, sizeof(double), c->stream));
    TCS_LAUNCH(c->stream, "adj_sum", k_interior_sum, grd, blk, 0, g, w->bdP, w->scal);
    TCS_LAUNCH(c->stream, "adj_sub", k_interior_sub, grd, blk, 0, g, w->scal, w->bdP);
    {
        const int n2m = c->n2m, n3m = c->n3m, h1p = c->h1p;
        double2 *spec = (double2 *)c->spec_x;
        TCS_LAUNCH(c->stream, "adj_pack", k_full_to_int, grd, blk, 0, g, 1, w->bdP, w->rint);
        TCS_TRY(fft_x_fwd_launch(c, "adj_fft_x_fwd", w->rint, spec));
        TCS_TRY(fft_z_launch<-1>(c, "adj_fft_z", spec));
        const int nsys = n3m * h1p;
        TCS_LAUNCH(c->stream, "adj_poisson_T", k_poisson_tdma_T, dim3((nsys + 255) / 256), dim3(256), 0, n3m, h1p, n2m,
                   c->h3p, c->ptab, c->ptab + n2m, c->ptab + 2 * n2m, c->dxk2, c->dzk, spec, c->pC);
        TCS_TRY(fft_z_launch<+1>(c, "adj_fft_z", spec));
        TCS_TRY(fft_x_inv_launch(c, "adj_fft_x_inv", spec, w->tmp[0]));
        TCS_LAUNCH(c->stream, "adj_pack", k_full_to_int, grd, blk, 0, g, 1, w->tmp[0], w->rint);
        TCS_HIP(hipMemsetAsync(w->bdP, 0, c->nfull * sizeof(double), c->stream));
    }
    // pressure right-hand side
    {
        const PressureRHS cell{g, c->phys, dt, w->sU1, w->sV1, w->sW1, w->s0[ST_DPHAT], w->siR,
                               c->wall[TCS_VBCUP], c->wall[TCS_VBCBT]};
        TCS_TRY(cell_vjp(c, w, cell, 1, OutBars{{w->rint, nullptr, nullptr, nullptr}, 0},
                         InBars<7>{{b[ST_U], b[ST_V], b[ST_W], b[ST_DPHAT], w->biR, nullptr, nullptr}}));
    }
    // halos of U, V, W; pseudo-update
    TCS_TRY(halo_adj(c, b[ST_W]));
    TCS_TRY(halo_adj(c, b[ST_V]));
    TCS_TRY(halo_adj(c, b[ST_U]));
    TCS_LAUNCH(c->stream, "adj_pseudo", k_pseudo_adj, grd, blk, 0, g, b[ST_U], b[ST_V], b[ST_W], w->bdU, w->bdV,
               w->bdW);
    // blockLdU, halo(dV), blockLdV, halo(dW)
    {
        const BlockLdU cell{g, c->phys, dt, w->s0[ST_U], w->sdV2, w->sdW1, w->sMu, w->siR, w->sdU1};
        TCS_TRY(overwrite_begin(c, w->tmp[0], w->bdU, 1));
        TCS_TRY(cell_vjp(c, w, cell, 1, OutBars{{w->bdU, nullptr, nullptr, nullptr}, 1},
                         InBars<6>{{b[ST_U], w->bdV, w->bdW, w->bMu, w->biR, w->tmp[0]}}));
        std::swap(w->bdU, w->tmp[0]);
    }
    TCS_TRY(halo_adj(c, w->bdV));
    {
        const BlockLdV cell{g, c->phys, dt, w->s0[ST_V], w->sdW1, w->sMu, w->siR, w->sdV1};
        TCS_TRY(overwrite_begin(c, w->tmp[0], w->bdV, 2));
        TCS_TRY(cell_vjp(c, w, cell, 2, OutBars{{w->bdV, nullptr, nullptr, nullptr}, 1},
                         InBars<5>{{b[ST_V], w->bdW, w->bMu, w->biR, w->tmp[0]}}));
        std::swap(w->bdV, w->tmp[0]);
    }
    TCS_TRY(halo_adj(c, w->bdW));
    // solvedW: dW's interior is the y stage's solution
    TCS_LAUNCH(c->stream, "adj_pack", k_full_to_int, grd, blk, 0, g, 1, w->bdW, w->lam);
    TCS_HIP(hipMemsetAsync(w->bdW, 0, c->nfull * sizeof(double), c->stream));
    {
        const DWZ rz{g, c->phys, dt, w->s0[ST_U], w->s0[ST_V], w->s0[ST_W], w->s0[ST_P], w->sdU1, w->sdV1, w->sMu,
                     w->siR, c->wall[TCS_VBCUP], c->wall[TCS_VBCBT], c->wall[TCS_WBCUP], c->wall[TCS_WBCBT]};
        const DWX rx{g, c->phys, dt, w->s0[ST_U], w->siR, w->sstale};
        const DWY ry{g, c->phys, dt, w->s0[ST_V], w->siR, w->sstale};
        TCS_TRY(stage_adj<false>(c, w, ry, 1, 1, false, w->X[STG_W][2], InBars<3>{{b[ST_V], w->biR, w->bstale}}));
        TCS_TRY(stage_adj<false>(c, w, rx, 0, 1, true, w->X[STG_W][1], InBars<3>{{b[ST_U], w->biR, w->bstale}}));
        TCS_TRY(stage_adj<true>(c, w, rz, 2, 1, true, w->X[STG_W][0],
                                InBars<12>{{b[ST_U], b[ST_V], b[ST_W], b[ST_P], w->bdU, w->bdV, w->bMu, w->biR,
                                            nullptr, nullptr, nullptr, nullptr}}));
    }
    // solvedV: rows j >= 2; dV's other rows are zeroed in the forward step
    TCS_TRY(halo_adj(c, w->bdV));
    TCS_LAUNCH(c->stream, "adj_pack", k_full_to_int, grd, blk, 0, g, 2, w->bdV, w->lam);
    TCS_HIP(hipMemsetAsync(w->bdV, 0, c->nfull * sizeof(double), c->stream));
    {
        const DVZ rz{g, c->phys, dt, w->s0[ST_U], w->s0[ST_V], w->s0[ST_W], w->s0[ST_P], w->sdU1, w->sMu, w->siR,
                     c->wall[TCS_VBCUP], c->wall[TCS_VBCBT]};
        const DVX rx{g, c->phys, dt, w->s0[ST_U], w->siR, w->sstale};
        const DVY ry{g, c->phys, dt, w->s0[ST_V], w->sMu, w->siR};
        TCS_TRY(stage_adj<false>(c, w, ry, 1, 2, false, w->X[STG_V][2], InBars<3>{{b[ST_V], w->bMu, w->biR}}));
        TCS_TRY(stage_adj<false>(c, w, rx, 0, 2, true, w->X[STG_V][1], InBars<3>{{b[ST_U], w->biR, w->bstale}}));
        TCS_TRY(stage_adj<true>(c, w, rz, 2, 2, true, w->X[STG_V][0],
                                InBars<9>{{b[ST_U], b[ST_V], b[ST_W], b[ST_P], w->bdU, w->bMu, w->biR, nullptr,
                                           nullptr}}));
    }
    // solvedU
    TCS_TRY(halo_adj(c, w->bdU));
    TCS_LAUNCH(c->stream, "adj_pack", k_full_to_int, grd, blk, 0, g, 1, w->bdU, w->lam);
    TCS_HIP(hipMemsetAsync(w->bdU, 0, c->nfull * sizeof(double), c->stream));
    {
        const DUZ rz{g, c->phys, dt, w->s0[ST_U], w->s0[ST_V], w->s0[ST_W], w->s0[ST_P], w->sT1, w->sMu, w->siR,
                     c->wall[TCS_UBCUP], c->wall[TCS_UBCBT], c->wall[TCS_VBCUP], c->wall[TCS_VBCBT]};
        const DUX rx{g, c->phys, dt, w->s0[ST_U], w->sMu, w->siR};
        const DUY ry{g, c->phys, dt, w->s0[ST_V], w->sMu, w->siR};
        TCS_TRY(stage_adj<false>(c, w, ry, 1, 1, false, w->X[STG_U][2], InBars<3>{{b[ST_V], w->bMu, w->biR}}));
        TCS_TRY(stage_adj<false>(c, w, rx, 0, 1, true, w->X[STG_U][1], InBars<3>{{b[ST_U], w->bMu, w->biR}}));
        TCS_TRY(stage_adj<true>(c, w, rz, 2, 1, true, w->X[STG_U][0],
                                InBars<11>{{b[ST_U], b[ST_V], b[ST_W], b[ST_P], b[ST_T], w->bMu, w->biR, nullptr,
                                            nullptr, nullptr, nullptr}}));
    }
    // stale viscosities and the Mu / invRho tables (from T after the thermal stage)
    TCS_LAUNCH(c->stream, "adj_stale", k_stale_adj, dim3(1), dim3(64), 0, g, w->sMu, w->bstale, w->bMu);
    TCS_LAUNCH(c->stream, "adj_coeffi", k_coeffi_adj, dim3((c->N2 + 63) / 64, c->N3), dim3(64), 0, g, c->phys, w->sT1,
               w->bMu, w->biR, b[ST_T]);
    TCS_TRY(halo_adj(c, b[ST_T]));
    // thermal: T1 = T0 + x3 on the interior
    TCS_LAUNCH(c->stream, "adj_pack", k_full_to_int, grd, blk, 0, g, 1, b[ST_T], w->lam);
    {
        const ThermalZ rz{g, c->phys, dt, w->s0[ST_T], w->s0[ST_U], w->s0[ST_V], w->s0[ST_W], TBCup, TBCbt};
        const ThermalX rx{g, c->phys, dt, w->s0[ST_T], w->s0[ST_U]};
        const ThermalY ry{g, c->phys, dt, w->s0[ST_T], w->s0[ST_V]};
        TCS_TRY(stage_adj<false>(c, w, ry, 1, 1, false, w->X[STG_T][2], InBars<2>{{b[ST_T], b[ST_V]}}));
        TCS_TRY(stage_adj<false>(c, w, rx, 0, 1, true, w->X[STG_T][1], InBars<2>{{b[ST_T], b[ST_U]}}));
        TCS_TRY(stage_adj<true>(c, w, rz, 2, 1, true, w->X[STG_T][0],
                                InBars<6>{{b[ST_T], b[ST_U], b[ST_V], b[ST_W], w->bwall[TCS_TBCUP],
                                           w->bwall[TCS_TBCBT]}}));
    }
    // the step's own tables and scalars start from zero next time
    TCS_HIP(hipMemsetAsync(w->bMu, 0, c->ntab * sizeof(double), c->stream));
    TCS_HIP(hipMemsetAsync(w->biR, 0, c->ntab * sizeof(double), c->stream));
    TCS_HIP(hipMemsetAsync(w->bstale, 0, 6 * sizeof(double), c->stream));
    return TCS_OK;
}

}  // namespace adj

void adjoint_free(tcs_ctx *c) {
    if (!c->adj) return;
    for (void *p : c->adj->all) (void)hipFree(p);
    delete c->adj;
    c->adj = nullptr;
}

static int adjoint_alloc(tcs_ctx *c) {
    if (c->adj) return TCS_OK;
    AdjWork *w = new AdjWork();
    c->adj = w;
    // full fields get the step's layout (FIELD_FRONT elements in front)
    auto get = [&](double **p, size_t n) -> int {
        const size_t front = n == c->nfull ? FIELD_FRONT + 1 : 0;
        void *q = nullptr;
        if (hipMalloc(&q, (std::max<size_t>(n, 1) + front) * sizeof(double)) != hipSuccess) return TCS_E_HIP;
        w->all.push_back(q);
        if (hipMemset(q, 0, (std::max<size_t>(n, 1) + front) * sizeof(double)) != hipSuccess) return TCS_E_HIP;
        *p = (double *)q + (front ? FIELD_FRONT : 0);
        return TCS_OK;
    };
    int rc = TCS_OK;
    for (int s = 0; s < adj::NSTATE && rc == TCS_OK; s++) rc |= get(&w->b[s], c->nfull) | get(&w->s0[s], c->nfull);
    for (double **p : {&w->bdU, &w->bdV, &w->bdW, &w->bdP, &w->sT1, &w->sdU1, &w->sdV1, &w->sdW1, &w->sdV2, &w->sdU2,
                       &w->sU1, &w->sV1, &w->sW1, &w->sdP, &w->tmp[0], &w->tmp[1], &w->tmp[2], &w->tmp[3]})
        rc |= get(p, c->nfull);
    for (double **p : {&w->bMu, &w->biR, &w->sMu, &w->siR}) rc |= get(p, c->ntab);
    rc |= get(&w->bstale, 6) | get(&w->sstale, 6) | get(&w->scal, 1);
    for (int q = 0; q < TCS_NWALL; q++) rc |= get(&w->bwall[q], c->nwall);
    for (int s = 0; s < 4; s++)
        for (int q = 0; q < 3; q++) rc |= get(&w->X[s][q], c->nint);
    for (double **p : {&w->lam, &w->abar, &w->bbar, &w->cbar, &w->aT, &w->cT, &w->rint}) rc |= get(p, c->nint);
    {
        void *q = nullptr;
        if (hipMalloc(&q, sizeof(int)) != hipSuccess || hipMemset(q, 0, sizeof(int)) != hipSuccess) rc = TCS_E_HIP;
        else {
            w->all.push_back(q);
            w->flag = (int *)q;
        }
    }
    if (rc != TCS_OK) {
        adjoint_free(c);
        set_error("tcs_adjoint: device allocation failed");
        return TCS_E_HIP;
    }
    return TCS_OK;
}

}  // namespace tcs

using namespace tcs;

extern "C" int tcs_adjoint_zero(tcs_ctx *c) {
    if (!c) return TCS_E_ARG;
    TCS_TRY(flush_pending(c));
    TCS_TRY(adjoint_alloc(c));
    AdjWork *w = c->adj;
    for (int s = 0; s < adj::NSTATE; s++) TCS_HIP(hipMemsetAsync(w->b[s], 0, c->nfull * sizeof(double), c->stream));
    for (int q = 0; q < TCS_NWALL; q++)
        TCS_HIP(hipMemsetAsync(w->bwall[q], 0, c->nwall * sizeof(double), c->stream));
    for (double *p : {w->bdU, w->bdV, w->bdW, w->bdP})
        TCS_HIP(hipMemsetAsync(p, 0, c->nfull * sizeof(double), c->stream));
    return TCS_OK;
}

static double *adjoint_array(tcs_ctx *c, int32_t kind, int32_t id, size_t &n) {
    AdjWork *w = c->adj;
    if (kind == 0) {
        for (int s = 0; s < adj::NSTATE; s++)
            if (adj::STATE_IDS[s] == id) {
                n = c->nfull;
                return w->b[s];
            }
        return nullptr;
    }
    if (kind == 1 && id >= 0 && id < TCS_NWALL) {
        n = c->nwall;
        return w->bwall[id];
    }
    return nullptr;
}

extern "C" int tcs_adjoint_copy(tcs_ctx *c, int32_t kind, int32_t id, int32_t to_adjoint, double *dev) {
    if (!c || !dev) return TCS_E_ARG;
    TCS_TRY(flush_pending(c));
    TCS_TRY(adjoint_alloc(c));
    size_t n = 0;
    double *a = adjoint_array(c, kind, id, n);
    if (!a) {
        set_error("tcs_adjoint_copy: kind 0 takes a state field id (T, U, V, W, P, dPhat, dP0), kind 1 a wall id");
        return TCS_E_ARG;
    }
    if (kind == 0) {
        // dev holds the reference layout (rows N1 apart), the adjoint field rows P1 apart
        const size_t pd = (size_t)c->N1 * sizeof(double), pa = (size_t)c->P1 * sizeof(double);
        const size_t rows = (size_t)c->N2 * c->N3;
        if (to_adjoint) TCS_HIP(hipMemcpy2DAsync(a, pa, dev, pd, pd, rows, hipMemcpyDeviceToDevice, c->stream));
        else TCS_HIP(hipMemcpy2DAsync(dev, pd, a, pa, pd, rows, hipMemcpyDeviceToDevice, c->stream));
        return TCS_OK;
    }
    if (to_adjoint) TCS_HIP(hipMemcpyAsync(a, dev, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    else TCS_HIP(hipMemcpyAsync(dev, a, n * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    return TCS_OK;
}

extern "C" int tcs_adjoint_add_nusselt(tcs_ctx *c, double weight, double Tup, double Tbt) {
    if (!c) return TCS_E_ARG;
    TCS_TRY(flush_pending(c));
    TCS_TRY(adjoint_alloc(c));
    const Grid g = c->grid();
    const double wgt = weight * c->prm.L2 / (Tbt - Tup) / c->h_dmx2[1] / ((double)c->n1m * (double)c->n3m);
    TCS_LAUNCH(c->stream, "adj_nusselt", adj::k_nusselt_adj, dim3((c->n1m + 63) / 64, c->n3m), dim3(64), 0, g, wgt,
               c->adj->b[adj::ST_T], c->adj->bwall[TCS_TBCBT]);
    return TCS_OK;
}

extern "C" int tcs_adjoint_step(tcs_ctx *c, double dt) {
    if (!c) return TCS_E_ARG;
    TCS_TRY(flush_pending(c));
    if (c->prm.precision != 0) {
        set_error("tcs_adjoint_step: the adjoint differentiates the plain FP64 step (precision 0)");
        return TCS_E_ARG;
    }
    TCS_TRY(adjoint_alloc(c));
    c->cfl_valid = 0;
    TCS_TRY(adj::forward_keep(c, c->adj, dt));
    TCS_TRY(adj::backward(c, c->adj, dt));
    int flag = 0;
    TCS_HIP(hipMemcpyAsync(&flag, c->adj->flag, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    TCS_HIP(hipStreamSynchronize(c->stream));
    if (flag) {
        set_error("tcs_adjoint_step: a cell read more distinct inputs than the recorder holds");
        return TCS_E_STATE;
    }
    return TCS_OK;
}
