// momentum.hip -- mpi_momentum_* (src/module_solve_momentum.f90) as HIP kernels.
//
// Mu and invRho are (0:n2sub,0:n3sub) tables: the reference assigns
// Mu(:,j,k) inside its i loop (:156-171), so every x-row holds the value of
// its last point i=n1sub.  The kernels index the tables by (j,k) and keep the
// reference's interpolation arithmetic, which therefore sees equal
// neighbours but rounds exactly as the reference does.
//
// Stage-2/3 quirks reproduced: dV x-stage uses mu1,mu2 of stage 1's last
// point (:821-826); dW x-stage mu1,mu2 (:1177-1182); dW y-stage mu3,mu4 of
// stage 1's last point and km = n3m-1 from the x-stage's last k (:1216-1246).
#include "stage_rows.h"

#ifndef TCS_FORM
#define TCS_FORM exact
#endif

namespace tcs {
namespace TCS_FORM {

#include "metric.h"

#ifndef TCS_SCRATCH
#define TCS_SCRATCH double
#endif
// ADI scratch (am, ac, ap and the right-hand side / increment): FP64, or
// FP32 in the mixed form, whose line solves then run in FP32 too.
using scr_t = TCS_SCRATCH;
#define SCR(p) (reinterpret_cast<scr_t *>(p))

int sweep_y_field(tcs_ctx *c, int j0, int nrows, double *F, int mode);

#define MU(j, k) Mu[g.t2((j), (k))]
#define IR(j, k) iR[g.t2((j), (k))]

// mpi_momentum_coeffi (:145-173): Mu and invRho of one temperature
template <class S>
__device__ __forceinline__ void mu_ir_of_T(const Phys &p, S T, S &mu, S &ir) {
    const S tmp1 = sc<S>(p.DeltaT) * T;
    const S tmp2 = tmp1 * tmp1;
    const S tmp3 = tmp2 * tmp1;
    const S tmp4 = tmp3 * tmp1;
    const S tmp5 = tmp4 * tmp1;
    mu = (sc<S>(p.a10) + sc<S>(p.a10) * (sc<S>(p.a11) * tmp1 + sc<S>(p.a12) * tmp2 + sc<S>(p.a13) * tmp3 + sc<S>(p.a14) * tmp4 + sc<S>(p.a15) * tmp5))
       * (sc<S>(p.d10) + sc<S>(p.d10) * (sc<S>(p.d11) * tmp1 + sc<S>(p.d12) * tmp2 + sc<S>(p.d13) * tmp3 + sc<S>(p.d14) * tmp4 + sc<S>(p.d15) * tmp5)) / sc<S>(p.Mu0);
    ir = sc<S>(p.Rho0) / (sc<S>(p.a10) + sc<S>(p.a10) * (sc<S>(p.a11) * tmp1 + sc<S>(p.a12) * tmp2 + sc<S>(p.a13) * tmp3 + sc<S>(p.a14) * tmp4 + sc<S>(p.a15) * tmp5));
}

// on the (j,k) table from T(n1sub,j,k)
__global__ void k_momentum_coeffi(Grid g, Phys p, const double *__restrict__ T, double *__restrict__ Mu,
                                  double *__restrict__ iR) {
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (j >= g.N2) return;
    double mu, ir;
    mu_ir_of_T(p, T[g.ix(g.n1m + 1, j, k)], mu, ir);
    Mu[g.t2(j, k)] = mu;
    iR[g.t2(j, k)] = ir;
}

// the stage-1 last-point viscosities that later stages reuse
template <class S, class FP>
__device__ __forceinline__ void stale_values(const Grid &g, FP Mu, S (&out)[6]) {
    const double *dx1 = g.dx1, *dx2 = g.dx2, *dx3 = g.dx3, *dmx1 = g.dmx1, *dmx2 = g.dmx2, *dmx3 = g.dmx3;
    const int i = g.n1m, j = g.n2m, k = g.n3m;
    const int im = i - 1, ip = i + 1, jm = j - 1, jp = j + 1, km = k - 1;
    // solvedV stage 1 (:661-665)
    {
        const S mua = sc<S>(0.5) * (dx2[jm] * MU(j, k) + dx2[j] * MU(jm, k)) / dmx2[j];
        const S muc = sc<S>(0.5) * (dx2[jm] * MU(j, k) + dx2[j] * MU(jm, k)) / dmx2[j];
        const S mub = sc<S>(0.5) * (dx2[jm] * MU(j, k) + dx2[j] * MU(jm, k)) / dmx2[j];
        out[0] = sc<S>(0.5) * (dx1[im] * muc + dx1[i] * mua) / dmx1[i];
        out[1] = sc<S>(0.5) * (dx1[ip] * muc + dx1[i] * mub) / dmx1[ip];
    }
    // solvedW stage 1 (:1022-1032)
    {
        S mua = sc<S>(0.5) * (dx3[km] * MU(j, k) + dx3[k] * MU(j, km)) / dmx3[k];
        S muc = sc<S>(0.5) * (dx3[km] * MU(j, k) + dx3[k] * MU(j, km)) / dmx3[k];
        S mub = sc<S>(0.5) * (dx3[km] * MU(j, k) + dx3[k] * MU(j, km)) / dmx3[k];
        out[2] = sc<S>(0.5) * (dx1[im] * muc + dx1[i] * mua) / dmx1[i];
        out[3] = sc<S>(0.5) * (dx1[ip] * muc + dx1[i] * mub) / dmx1[ip];
        mua = sc<S>(0.5) * (dx3[km] * MU(jm, k) + dx3[k] * MU(jm, km)) / dmx3[k];
        muc = sc<S>(0.5) * (dx3[km] * MU(j, k) + dx3[k] * MU(j, km)) / dmx3[k];
        mub = sc<S>(0.5) * (dx3[km] * MU(jp, k) + dx3[k] * MU(jp, km)) / dmx3[k];
        out[4] = sc<S>(0.5) * (dx2[jm] * muc + dx2[j] * mua) / dmx2[j];
        out[5] = sc<S>(0.5) * (dx2[jp] * muc + dx2[j] * mub) / dmx2[jp];
    }
}

__global__ void k_momentum_stale(Grid g, const double *__restrict__ Mu, double *__restrict__ out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double v[6];
    stale_values<double>(g, Mu, v);
    for (int q = 0; q < 6; q++) out[q] = v[q];
}

#define KERNEL_IJK                                                            \
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;                  \
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;                  \
    const int k = blockIdx.z + 1;                                             \
    if (i > g.n1m || j > g.n2m) return;                                       \
    const double *dx1 = g.dx1, *dx2 = g.dx2, *dx3 = g.dx3;                    \
    const double *dmx1 = g.dmx1, *dmx2 = g.dmx2, *dmx3 = g.dmx3;              \
    (void)dx1; (void)dx2; (void)dx3; (void)dmx1; (void)dmx2; (void)dmx3;

// the metric arrays under their reference names, in a row routine
#define ROW_METRICS                                                                                   \
    const auto dx1 = metric<S, M_DX1>(g), dx2 = metric<S, M_DX2>(g), dx3 = metric<S, M_DX3>(g);     \
    const auto dmx1 = metric<S, M_DMX1>(g), dmx2 = metric<S, M_DMX2>(g), dmx3 = metric<S, M_DMX3>(g); \
    (void)dx1; (void)dx2; (void)dx3; (void)dmx1; (void)dmx2; (void)dmx3;
// the same, an FP32 row rounding the double tables where it reads them
#define ROW_METRICS_D64                                                                                          \
    const auto dx1 = metric<S, M_DX1, true>(g), dx2 = metric<S, M_DX2, true>(g), dx3 = metric<S, M_DX3, true>(g); \
    const auto dmx1 = metric<S, M_DMX1, true>(g), dmx2 = metric<S, M_DMX2, true>(g),                             \
               dmx3 = metric<S, M_DMX3, true>(g);                                                                \
    (void)dx1; (void)dx2; (void)dx3; (void)dmx1; (void)dmx2; (void)dmx3;

// Each stage's row routine gives row (i, j, k) of the stage's tridiagonal
// systems (am, ac, ap: lower, diagonal, upper) and, in the first stage, the
// right-hand side; the functor after it carries its arguments.

// ---------------------------------------------------------------- dU -------
template <class S, int WM, class FP>
__device__ __forceinline__ void dU_z_row(const Grid &g, const Phys &p, double dt, FP U,
    FP V, FP W, FP P,
    FP T, FP Mu, FP iR,
    FP UBCup, FP UBCbt, FP VBCup,
    FP VBCbt,
    int i, int j, int k, S &am, S &ac, S &ap, S &rhsout) {
    ROW_METRICS
    const int im = i - 1, ip = i + 1, jm = j - 1, jp = j + 1, km = k - 1, kp = k + 1;
    const int jum = g.jC[jm], jup = g.jC[jp];
    const S Cmu = sc<S>(p.Cmu);
#define Uf(a, b, c) U[g.ix(a, b, c)]
#define Vf(a, b, c) V[g.ix(a, b, c)]
#define Wf(a, b, c) W[g.ix(a, b, c)]
    const S u1 = sc<S>(0.5) * (Uf(im, j, k) + Uf(i, j, k));
    const S u2 = sc<S>(0.5) * (Uf(ip, j, k) + Uf(i, j, k));
    const S v3 = sc<S>(0.5) * (dx1[im] * Vf(i, j, k) + dx1[i] * Vf(im, j, k)) / dmx1[i];
    const S v4 = sc<S>(0.5) * (dx1[im] * Vf(i, jp, k) + dx1[i] * Vf(im, jp, k)) / dmx1[i];
    const S w5 = sc<S>(0.5) * (dx1[im] * Wf(i, j, k) + dx1[i] * Wf(im, j, k)) / dmx1[i];
    const S w6 = sc<S>(0.5) * (dx1[im] * Wf(i, j, kp) + dx1[i] * Wf(im, j, kp)) / dmx1[i];

    const S dudx1 = (Uf(i, j, k) - Uf(im, j, k)) / dx1[im];
    const S dudx2 = (Uf(ip, j, k) - Uf(i, j, k)) / dx1[i];
    const S dudy3 = (Uf(i, j, k) - Uf(i, jm, k)) / dmx2[j];
    const S dudy4 = (Uf(i, jp, k) - Uf(i, j, k)) / dmx2[jp];
    const S dudz5 = (Uf(i, j, k) - Uf(i, j, km)) / dmx3[k];
    const S dudz6 = (Uf(i, j, kp) - Uf(i, j, k)) / dmx3[kp];
    const S dvdx3 = (Vf(i, j, k) - Vf(im, j, k)) / dmx1[i];
    const S dvdx4 = (Vf(i, jp, k) - Vf(im, jp, k)) / dmx1[i];
    const S dwdx5 = (Wf(i, j, k) - Wf(im, j, k)) / dmx1[i];
    const S dwdx6 = (Wf(i, j, kp) - Wf(im, j, kp)) / dmx1[i];

    S mua = sc<S>(0.5) * (dx1[im] * MU(jm, k) + dx1[i] * MU(jm, k)) / dmx1[i];
    S muc = sc<S>(0.5) * (dx1[im] * MU(j, k) + dx1[i] * MU(j, k)) / dmx1[i];
    S mub = sc<S>(0.5) * (dx1[im] * MU(jp, k) + dx1[i] * MU(jp, k)) / dmx1[i];
    const S mu3 = sc<S>(0.5) * (dx2[jm] * muc + dx2[j] * mua) / dmx2[j];
    const S mu4 = sc<S>(0.5) * (dx2[jp] * muc + dx2[j] * mub) / dmx2[jp];
    mua = sc<S>(0.5) * (dx1[im] * MU(j, km) + dx1[i] * MU(j, km)) / dmx1[i];
    muc = sc<S>(0.5) * (dx1[im] * MU(j, k) + dx1[i] * MU(j, k)) / dmx1[i];
    mub = sc<S>(0.5) * (dx1[im] * MU(j, kp) + dx1[i] * MU(j, kp)) / dmx1[i];
    const S mu5 = sc<S>(0.5) * (dx3[km] * muc + dx3[k] * mua) / dmx3[k];
    const S mu6 = sc<S>(0.5) * (dx3[kp] * muc + dx3[k] * mub) / dmx3[kp];

    const S invRhoc = sc<S>(0.5) * (dx1[im] * IR(j, k) + dx1[i] * IR(j, k)) / dmx1[i];
    const S invRhocCmu_half = sc<S>(0.5) * Cmu * invRhoc;

    const S viscous_u1 = sc<S>(1.0) * (MU(j, k) * dudx2 - MU(j, k) * dudx1) / dmx1[i];
    const S viscous_u2 = sc<S>(1.0) * (mu4 * dudy4 - mu3 * dudy3) / dx2[j];
    const S viscous_u3 = sc<S>(1.0) * (mu6 * dudz6 - mu5 * dudz5) / dx3[k];
    const S viscous_u12 = sc<S>(1.0) * (mu4 * dvdx4 - mu3 * dvdx3) / dx2[j];
    const S viscous_u13 = sc<S>(1.0) * (mu6 * dwdx6 - mu5 * dwdx5) / dx3[k];

    S rhs = invRhocCmu_half * (sc<S>(2.0) * viscous_u1 + viscous_u2 + viscous_u3 + viscous_u12 + viscous_u13);
    rhs = rhs - sc<S>(p.Cmp) * invRhoc * (P[g.ix(i, j, k)] - P[g.ix(im, j, k)]) / dmx1[i];
    const S Tc = sc<S>(0.5) * (dx1[im] * T[g.ix(i, j, k)] + dx1[i] * T[g.ix(im, j, k)]) / dmx1[i];
    rhs = rhs + sc<S>(p.Cmt) * (Tc + sc<S>(p.a12pera11) * (Tc * Tc) * sc<S>(p.DeltaT)) * invRhoc;

    // The wall terms are multiplied by 1 - jum and 1 - jup, zero away from the
    // walls, where rhs + 0 * term = rhs for finite terms.  WM says how a row
    // treats them: 0 computes them, 1 skips them where j is away from the
    // walls (a branch on j, uniform in a wavefront), 2 leaves them out (the
    // caller has checked that j is away from the walls, wall_free); skipped,
    // the loads of the wall arrays are not issued either.
    if (WM == 0 || (WM == 1 && (jum != 1 || jup != 1))) {
    S ubc_down = sc<S>(0.25) * v3 / dmx2[j] * UBCbt[g.b2(i, k)]
                    - sc<S>(0.25) * dudy3 * (dx1[i] * VBCbt[g.b2(im, k)] + dx1[im] * VBCbt[g.b2(i, k)]) / dmx1[i] * sc<S>(0.5);
    S ubc_up = -sc<S>(0.25) * v4 / dmx2[jp] * UBCup[g.b2(i, k)]
                  - sc<S>(0.25) * dudy4 * (dx1[i] * VBCup[g.b2(im, k)] + dx1[im] * VBCup[g.b2(i, k)]) / dmx1[i] * sc<S>(0.5);
    ubc_down = ubc_down + invRhocCmu_half / dx2[j] * mu3 / dmx2[j] * UBCbt[g.b2(i, k)];
    ubc_down = ubc_down - invRhocCmu_half / dx2[j] * mu3 / dmx1[i] * (VBCbt[g.b2(i, k)] - VBCbt[g.b2(im, k)]);
    ubc_up = ubc_up + invRhocCmu_half / dx2[j] * mu4 / dmx2[jp] * UBCup[g.b2(i, k)];
    ubc_up = ubc_up + invRhocCmu_half / dx2[j] * mu4 / dmx1[i] * (VBCup[g.b2(i, k)] - VBCup[g.b2(im, k)]);
    rhs = rhs + (sc<S>(1.0) - sc<S>(jum)) * ubc_down + (sc<S>(1.0) - sc<S>(jup)) * ubc_up;
    }

    const S mACI = invRhocCmu_half / dmx1[i] * (MU(j, k) / dx1[i] + MU(j, k) / dx1[im]) * sc<S>(2.0)
                      + sc<S>(0.25) * (dudx2 * sc<S>(0.5) + dudx1 * sc<S>(0.5) - u2 / dx1[i] + u1 / dx1[im]);
    const S mAPI = -invRhocCmu_half / dmx1[i] * MU(j, k) / dx1[i] * sc<S>(2.0) + sc<S>(0.25) * (u2 / dx1[i] + dudx2 * sc<S>(0.5));
    const S mAMI = -invRhocCmu_half / dmx1[i] * MU(j, k) / dx1[im] * sc<S>(2.0) + sc<S>(0.25) * (-u1 / dx1[im] + dudx1 * sc<S>(0.5));
    const S mACJ = invRhocCmu_half / dx2[j] * (mu4 / dmx2[jp] + mu3 / dmx2[j]) + sc<S>(0.25) * (-v4 / dmx2[jp] + v3 / dmx2[j]);
    S mAPJ = -invRhocCmu_half / dx2[j] * mu4 / dmx2[jp] + sc<S>(0.25) * (v4 / dmx2[jp]);
    S mAMJ = -invRhocCmu_half / dx2[j] * mu3 / dmx2[j] + sc<S>(0.25) * (-v3 / dmx2[j]);
    mAPJ = mAPJ * sc<S>(jup);
    mAMJ = mAMJ * sc<S>(jum);
    const S mACK = invRhocCmu_half / dx3[k] * (mu6 / dmx3[kp] + mu5 / dmx3[k]) + sc<S>(0.25) * (-w6 / dmx3[kp] + w5 / dmx3[k]);
    const S mAPK = -invRhocCmu_half / dx3[k] * mu6 / dmx3[kp] + sc<S>(0.25) * (w6 / dmx3[kp]);
    const S mAMK = -invRhocCmu_half / dx3[k] * mu5 / dmx3[k] + sc<S>(0.25) * (-w5 / dmx3[k]);

    rhs = rhs - (mAPI * Uf(ip, j, k) + mACI * Uf(i, j, k) + mAMI * Uf(im, j, k)
                 + mAPJ * Uf(i, jp, k) + mACJ * Uf(i, j, k) + mAMJ * Uf(i, jm, k)
                 + mAPK * Uf(i, j, kp) + mACK * Uf(i, j, k) + mAMK * Uf(i, j, km));
    rhs = rhs - (sc<S>(0.25) * (v4 * sc<S>(jup) * dudy4 + v3 * sc<S>(jum) * dudy3)
                 - invRhocCmu_half / dx2[j] * (mu4 * dvdx4 * sc<S>(jup) - mu3 * dvdx3 * sc<S>(jum)));
    rhs = rhs - (sc<S>(0.25) * (w6 * dudz6 + w5 * dudz5) - invRhocCmu_half / dx3[k] * (mu6 * dwdx6 - mu5 * dwdx5));

    ac = mACK * sc<S>(dt) + sc<S>(1.0);
    ap = mAPK * sc<S>(dt);
    am = mAMK * sc<S>(dt);
    rhsout = rhs * sc<S>(dt);
}
struct DUZ {
    // FP32 lines of the mixed form: the unconditional row loop (2.14 -> 2.03 ms
    // at 512^3)
    static constexpr int W32 = 1;
    static constexpr bool FULL32 = true;
    Grid g;
    Phys p;
    double dt;
    const double *U, *V, *W, *P, *T, *Mu, *iR, *UBCup, *UBCbt, *VBCup, *VBCbt;
    static constexpr int NF = 11;  // field inputs, in the order below
    static constexpr bool FULL = true;  // unconditional row loop in the FP64 z solve
    template <class S, int WM = -1, class Wr>
    __device__ __forceinline__ void eval(const Wr &w, int i, int j, int k, S &a, S &b, S &c, S &d) const {
        // WM < 0 (the row's own wall-term mode): terms kept; the FP64 z lines
        // pass 2 or 0 per workgroup (wall_free), the FP32 ones keep them
        dU_z_row<S, (WM >= 0 ? WM : 0)>(g, p, dt, w(0, U), w(1, V), w(2, W), w(3, P), w(4, T), w(5, Mu), w(6, iR), w(7, UBCup), w(8, UBCbt), w(9, VBCup), w(10, VBCbt), i, j, k, a, b, c, d);
    }
    // no wall term of row j is weighted (the wall terms' flags are all 1)
    __device__ __forceinline__ bool wall_free(int j) const { return g.jC[j - 1] == 1 && g.jC[j + 1] == 1; }
    __device__ __forceinline__ void operator()(int i, int j, int k, double &a, double &b, double &c,
                                               double &d) const {
        eval<double>(RawW(), i, j, k, a, b, c, d);
    }
};

template <class S, class FP>
__device__ __forceinline__ void dU_x_row(const Grid &g, const Phys &p, double dt, FP U,
    FP Mu, FP iR,
    int i, int j, int k, S &am, S &ac, S &ap) {
    ROW_METRICS
    const int im = i - 1, ip = i + 1;
    const S u1 = sc<S>(0.5) * (Uf(im, j, k) + Uf(i, j, k));
    const S u2 = sc<S>(0.5) * (Uf(ip, j, k) + Uf(i, j, k));
    const S dudx1 = (Uf(i, j, k) - Uf(im, j, k)) / dx1[im];
    const S dudx2 = (Uf(ip, j, k) - Uf(i, j, k)) / dx1[i];
    const S invRhoc = sc<S>(0.5) * (dx1[im] * IR(j, k) + dx1[i] * IR(j, k)) / dmx1[i];
    const S invRhocCmu_half = sc<S>(0.5) * sc<S>(p.Cmu) * invRhoc;
    const S mACI = invRhocCmu_half / dmx1[i] * (MU(j, k) / dx1[i] + MU(j, k) / dx1[im]) * sc<S>(2.0)
                      + sc<S>(0.25) * (dudx2 * sc<S>(0.5) + dudx1 * sc<S>(0.5) - u2 / dx1[i] + u1 / dx1[im]);
    const S mAPI = -invRhocCmu_half / dmx1[i] * MU(j, k) / dx1[i] * sc<S>(2.0) + sc<S>(0.25) * (u2 / dx1[i] + dudx2 * sc<S>(0.5));
    const S mAMI = -invRhocCmu_half / dmx1[i] * MU(j, k) / dx1[im] * sc<S>(2.0) + sc<S>(0.25) * (-u1 / dx1[im] + dudx1 * sc<S>(0.5));
    ac = mACI * sc<S>(dt) + sc<S>(1.0);
    ap = mAPI * sc<S>(dt);
    am = mAMI * sc<S>(dt);
}
struct DUX {
    Grid g;
    Phys p;
    double dt;
    const double *U, *Mu, *iR;
    static constexpr int NF = 3;  // field inputs, in the order below
    template <class S, class Wr>
    __device__ __forceinline__ void eval(const Wr &w, int i, int j, int k, S &a, S &b, S &c, S &d) const {
        dU_x_row<S>(g, p, dt, w(0, U), w(1, Mu), w(2, iR), i, j, k, a, b, c);
    }
    __device__ __forceinline__ void operator()(int i, int j, int k, double &a, double &b, double &c,
                                               double &d) const {
        eval<double>(RawW(), i, j, k, a, b, c, d);
    }
};

template <class S, class FP>
__device__ __forceinline__ void dU_y_row(const Grid &g, const Phys &p, double dt, FP V,
    FP Mu, FP iR,
    int i, int j, int k, S &am, S &ac, S &ap) {
    ROW_METRICS
    const int im = i - 1, jm = j - 1, jp = j + 1;
    const int jum = g.jC[jm], jup = g.jC[jp];
    const S v3 = sc<S>(0.5) * (dx1[im] * Vf(i, j, k) + dx1[i] * Vf(im, j, k)) / dmx1[i];
    const S v4 = sc<S>(0.5) * (dx1[im] * Vf(i, jp, k) + dx1[i] * Vf(im, jp, k)) / dmx1[i];
    const S mua = sc<S>(0.5) * (dx1[im] * MU(jm, k) + dx1[i] * MU(jm, k)) / dmx1[i];
    const S muc = sc<S>(0.5) * (dx1[im] * MU(j, k) + dx1[i] * MU(j, k)) / dmx1[i];
    const S mub = sc<S>(0.5) * (dx1[im] * MU(jp, k) + dx1[i] * MU(jp, k)) / dmx1[i];
    const S mu3 = sc<S>(0.5) * (dx2[jm] * muc + dx2[j] * mua) / dmx2[j];
    const S mu4 = sc<S>(0.5) * (dx2[jp] * muc + dx2[j] * mub) / dmx2[jp];
    const S invRhoc = sc<S>(0.5) * (dx1[im] * IR(j, k) + dx1[i] * IR(j, k)) / dmx1[i];
    const S invRhocCmu_half = sc<S>(0.5) * sc<S>(p.Cmu) * invRhoc;
    const S mACJ = invRhocCmu_half / dx2[j] * (mu4 / dmx2[jp] + mu3 / dmx2[j]) + sc<S>(0.25) * (-v4 / dmx2[jp] + v3 / dmx2[j]);
    S mAPJ = -invRhocCmu_half / dx2[j] * mu4 / dmx2[jp] + sc<S>(0.25) * (v4 / dmx2[jp]);
    S mAMJ = -invRhocCmu_half / dx2[j] * mu3 / dmx2[j] + sc<S>(0.25) * (-v3 / dmx2[j]);
    mAPJ = mAPJ * sc<S>(jup);
    mAMJ = mAMJ * sc<S>(jum);
    ac = mACJ * sc<S>(dt) + sc<S>(1.0);
    ap = mAPJ * sc<S>(dt);
    am = mAMJ * sc<S>(dt);
}
struct DUY {
    // FP32 lines of the mixed form: the unconditional row loop (0.96 -> 0.56 ms
    // at 512^3)
    static constexpr int W32 = 1;
    static constexpr bool FULL32 = true;
    Grid g;
    Phys p;
    double dt;
    const double *V, *Mu, *iR;
    static constexpr int NF = 3;  // field inputs, in the order below
    template <class S, class Wr>
    __device__ __forceinline__ void eval(const Wr &w, int i, int j, int k, S &a, S &b, S &c, S &d) const {
        dU_y_row<S>(g, p, dt, w(0, V), w(1, Mu), w(2, iR), i, j, k, a, b, c);
    }
    __device__ __forceinline__ void operator()(int i, int j, int k, double &a, double &b, double &c,
                                               double &d) const {
        eval<double>(RawW(), i, j, k, a, b, c, d);
    }
};

// interior copy: F(i,j,k) = RHS for j in [j0, n2m]
// dV(0:n1sub,0:1,0:n3sub) = 0 and dV(:,n2sub,:) = 0 (:907-909)
template <typename E>
__global__ void k_dV_zero_rows(Grid g, E *__restrict__ dV) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = blockIdx.y;
    if (i >= g.N1) return;
    dV[g.ix(i, 0, k)] = E(0);
    dV[g.ix(i, 1, k)] = E(0);
    dV[g.ix(i, g.n2m + 1, k)] = E(0);
}

// ---------------------------------------------------------------- dV -------
template <class S, int WM, class FP>
__device__ __forceinline__ void dV_z_row(const Grid &g, const Phys &p, double dt, FP U,
    FP V, FP W, FP P,
    FP dU, FP Mu, FP iR,
    FP VBCup, FP VBCbt,
    int i, int j, int k, S &am, S &ac, S &ap, S &rhsout) {
    ROW_METRICS_D64
    const int im = i - 1, ip = i + 1, jm = j - 1, jp = j + 1, km = k - 1, kp = k + 1;
    const int jvp = g.jS[jm], jvm = g.jS[jp];
    const S Cmu = sc<S>(p.Cmu);
#define dUf(a, b, c) dU[g.ix(a, b, c)]
    const S v3 = sc<S>(0.5) * (Vf(i, jm, k) + Vf(i, j, k));
    const S v4 = sc<S>(0.5) * (Vf(i, j, k) + Vf(i, jp, k));
    const S u1 = (dx2[jm] * Uf(i, j, k) + dx2[j] * Uf(i, jm, k)) / dmx2[j] * sc<S>(0.5);
    const S u2 = (dx2[jm] * Uf(ip, j, k) + dx2[j] * Uf(ip, jm, k)) / dmx2[j] * sc<S>(0.5);
    const S w5 = (dx2[jm] * Wf(i, j, k) + dx2[j] * Wf(i, jm, k)) / dmx2[j] * sc<S>(0.5);
    const S w6 = (dx2[jm] * Wf(i, j, kp) + dx2[j] * Wf(i, jm, kp)) / dmx2[j] * sc<S>(0.5);
    const S ddu1 = (dx2[jm] * dUf(i, j, k) + dx2[j] * dUf(i, jm, k)) / dmx2[j] * sc<S>(0.5);
    const S ddu2 = (dx2[jm] * dUf(ip, j, k) + dx2[j] * dUf(ip, jm, k)) / dmx2[j] * sc<S>(0.5);

    const S dvdx1 = (Vf(i, j, k) - Vf(im, j, k)) / dmx1[i];
    const S dvdx2 = (Vf(ip, j, k) - Vf(i, j, k)) / dmx1[ip];
    const S dvdy3 = (Vf(i, j, k) - Vf(i, jm, k)) / dx2[jm];
    const S dvdy4 = (Vf(i, jp, k) - Vf(i, j, k)) / dx2[j];
    const S dvdz5 = (Vf(i, j, k) - Vf(i, j, km)) / dmx3[k];
    const S dvdz6 = (Vf(i, j, kp) - Vf(i, j, k)) / dmx3[kp];
    const S dudy1 = (Uf(i, j, k) - Uf(i, jm, k)) / dmx2[j];
    const S dudy2 = (Uf(ip, j, k) - Uf(ip, jm, k)) / dmx2[j];
    const S dwdy5 = (Wf(i, j, k) - Wf(i, jm, k)) / dmx2[j];
    const S dwdy6 = (Wf(i, j, kp) - Wf(i, jm, kp)) / dmx2[j];
    const S ddudy1 = (dUf(i, j, k) - dUf(i, jm, k)) / dmx2[j];
    const S ddudy2 = (dUf(ip, j, k) - dUf(ip, jm, k)) / dmx2[j];

    S mua = sc<S>(0.5) * (dx2[jm] * MU(j, k) + dx2[j] * MU(jm, k)) / dmx2[j];
    S muc = sc<S>(0.5) * (dx2[jm] * MU(j, k) + dx2[j] * MU(jm, k)) / dmx2[j];
    S mub = sc<S>(0.5) * (dx2[jm] * MU(j, k) + dx2[j] * MU(jm, k)) / dmx2[j];
    const S mu1 = sc<S>(0.5) * (dx1[im] * muc + dx1[i] * mua) / dmx1[i];
    const S mu2 = sc<S>(0.5) * (dx1[ip] * muc + dx1[i] * mub) / dmx1[ip];
    mua = sc<S>(0.5) * (dx2[jm] * MU(j, km) + dx2[j] * MU(jm, km)) / dmx2[j];
    muc = sc<S>(0.5) * (dx2[jm] * MU(j, k) + dx2[j] * MU(jm, k)) / dmx2[j];
    mub = sc<S>(0.5) * (dx2[jm] * MU(j, kp) + dx2[j] * MU(jm, kp)) / dmx2[j];
    const S mu5 = sc<S>(0.5) * (dx3[km] * muc + dx3[k] * mua) / dmx3[k];
    const S mu6 = sc<S>(0.5) * (dx3[kp] * muc + dx3[k] * mub) / dmx3[kp];

    const S invRhoc = sc<S>(0.5) * (dx2[jm] * IR(j, k) + dx2[j] * IR(jm, k)) / dmx2[j];
    const S invRhocCmu_half = sc<S>(0.5) * Cmu * invRhoc;

    const S viscous_v1 = sc<S>(1.0) * (mu2 * dvdx2 - mu1 * dvdx1) / dx1[i];
    const S viscous_v2 = sc<S>(1.0) * (MU(j, k) * dvdy4 - MU(jm, k) * dvdy3) / dmx2[j];
    const S viscous_v3 = sc<S>(1.0) * (mu6 * dvdz6 - mu5 * dvdz5) / dx3[k];
    const S viscous_v21 = sc<S>(1.0) * (mu2 * dudy2 - mu1 * dudy1) / dx1[i];
    const S viscous_v23 = sc<S>(1.0) * (mu6 * dwdy6 - mu5 * dwdy5) / dx3[k];

    S rhs = invRhocCmu_half * (viscous_v1 + sc<S>(2.0) * viscous_v2 + viscous_v3 + viscous_v21 + viscous_v23);
    rhs = rhs - sc<S>(p.Cmp) * invRhoc * (P[g.ix(i, j, k)] - P[g.ix(i, jm, k)]) / dmx2[j];

    if (WM == 0 || (WM == 1 && (jvm != 1 || jvp != 1))) {  // wall terms, as in the dU rows
    const S vbt = VBCbt[g.b2(i, k)], vup = VBCup[g.b2(i, k)];
    S vbc_down = (sc<S>(0.25) / dx2[jm] * v3 * vbt - sc<S>(0.25) * sc<S>(0.5) * dvdy3 * vbt);
    S vbc_up = (-sc<S>(0.25) / dx2[j] * v4 * vup - sc<S>(0.25) * sc<S>(0.5) * dvdy4 * vup);
    vbc_down = vbc_down + invRhocCmu_half / dmx2[j] * MU(jm, k) / dx2[jm] * vbt * sc<S>(2.0);
    vbc_up = vbc_up + invRhocCmu_half / dmx2[j] * MU(j, k) / dx2[j] * vup * sc<S>(2.0);
    rhs = rhs + (sc<S>(1.0) - sc<S>(jvm)) * vbc_down + (sc<S>(1.0) - sc<S>(jvp)) * vbc_up;
    }

    const S mACI = invRhocCmu_half / dx1[i] * (mu2 / dmx1[ip] + mu1 / dmx1[i]) + sc<S>(0.25) * (-u2 / dmx1[ip] + u1 / dmx1[i]);
    const S mAPI = -invRhocCmu_half / dx1[i] * mu2 / dmx1[ip] + sc<S>(0.25) * (u2 / dmx1[ip]);
    const S mAMI = -invRhocCmu_half / dx1[i] * mu1 / dmx1[i] + sc<S>(0.25) * (-u1 / dmx1[i]);
    const S mACJ = invRhocCmu_half / dmx2[j] * (MU(j, k) / dx2[j] + MU(jm, k) / dx2[jm]) * sc<S>(2.0)
                      + sc<S>(0.25) * (dvdy4 * sc<S>(0.5) + dvdy3 * sc<S>(0.5) - v4 / dx2[j] + v3 / dx2[jm]);
    S mAPJ = -invRhocCmu_half / dmx2[j] * MU(j, k) / dx2[j] * sc<S>(2.0) + sc<S>(0.25) * (v4 / dx2[j] + dvdy4 * sc<S>(0.5));
    mAPJ = mAPJ * sc<S>(jvp);
    S mAMJ = -invRhocCmu_half / dmx2[j] * MU(jm, k) / dx2[jm] * sc<S>(2.0) + sc<S>(0.25) * (-v3 / dx2[jm] + dvdy3 * sc<S>(0.5));
    mAMJ = mAMJ * sc<S>(jvm);
    const S mACK = invRhocCmu_half / dx3[k] * (mu6 / dmx3[kp] + mu5 / dmx3[k]) + sc<S>(0.25) * (-w6 / dmx3[kp] + w5 / dmx3[k]);
    const S mAPK = -invRhocCmu_half / dx3[k] * mu6 / dmx3[kp] + sc<S>(0.25) * (w6 / dmx3[kp]);
    const S mAMK = -invRhocCmu_half / dx3[k] * mu5 / dmx3[k] + sc<S>(0.25) * (-w5 / dmx3[k]);

    rhs = rhs - (mAPI * Vf(ip, j, k) + mACI * Vf(i, j, k) + mAMI * Vf(im, j, k)
                 + mAPJ * Vf(i, jp, k) + mACJ * Vf(i, j, k) + mAMJ * Vf(i, jm, k)
                 + mAPK * Vf(i, j, kp) + mACK * Vf(i, j, k) + mAMK * Vf(i, j, km));
    rhs = rhs - (sc<S>(0.25) * (u2 * dvdx2 + u1 * dvdx1) - invRhocCmu_half / dx1[i] * (mu2 * dudy2 - mu1 * dudy1));
    rhs = rhs - (sc<S>(0.25) * (w6 * dvdz6 + w5 * dvdz5) - invRhocCmu_half / dx3[k] * (mu6 * dwdy6 - mu5 * dwdy5));
    rhs = rhs - (sc<S>(0.25) * (ddu2 * dvdx2 + ddu1 * dvdx1) - invRhocCmu_half / dx1[i] * (mu2 * ddudy2 - mu1 * ddudy1));

    ac = mACK * sc<S>(dt) + sc<S>(1.0);
    ap = mAPK * sc<S>(dt);
    am = mAMK * sc<S>(dt);
    rhsout = rhs * sc<S>(dt);
}
struct DVZ {
    // FP32 lines of the mixed form: the unconditional row loop with the
    // workgroups split by the wall terms (2.05 -> 1.83 ms at 512^3)
    static constexpr int W32 = 1;
    static constexpr bool FULL32 = true;
    static constexpr bool WALLS32 = true;
    Grid g;
    Phys p;
    double dt;
    const double *U, *V, *W, *P, *dU, *Mu, *iR, *VBCup, *VBCbt;
    static constexpr int NF = 9;  // field inputs, in the order below
    static constexpr bool FULL = true;  // unconditional row loop in the FP64 z solve
    template <class S, int WM = -1, class Wr>
    __device__ __forceinline__ void eval(const Wr &w, int i, int j, int k, S &a, S &b, S &c, S &d) const {
        // WM < 0 (the row's own wall-term mode): a branch on j in FP32 rows,
        // terms kept otherwise; the z lines pass 2 or 0 per workgroup
        // (wall_free)
        dV_z_row<S, (WM >= 0 ? WM : (std::is_same<S, float>::value ? 1 : 0))>(g, p, dt, w(0, U), w(1, V), w(2, W), w(3, P), w(4, dU), w(5, Mu), w(6, iR), w(7, VBCup), w(8, VBCbt), i, j, k, a, b, c, d);
    }
    // no wall term of row j is weighted (the wall terms' flags are all 1)
    __device__ __forceinline__ bool wall_free(int j) const { return g.jS[j - 1] == 1 && g.jS[j + 1] == 1; }
    __device__ __forceinline__ void operator()(int i, int j, int k, double &a, double &b, double &c,
                                               double &d) const {
        eval<double>(RawW(), i, j, k, a, b, c, d);
    }
};

template <class S, class FP>
__device__ __forceinline__ void dV_x_row(const Grid &g, const Phys &p, double dt, FP U,
    FP iR, FP stale,
    int i, int j, int k, S &am, S &ac, S &ap) {
    ROW_METRICS
    const int ip = i + 1, jm = j - 1;
    const S mu1 = stale[0], mu2 = stale[1];
    const S u1 = (dx2[jm] * Uf(i, j, k) + dx2[j] * Uf(i, jm, k)) / dmx2[j] * sc<S>(0.5);
    const S u2 = (dx2[jm] * Uf(ip, j, k) + dx2[j] * Uf(ip, jm, k)) / dmx2[j] * sc<S>(0.5);
    const S invRhoc = sc<S>(0.5) * (dx2[jm] * IR(j, k) + dx2[j] * IR(jm, k)) / dmx2[j];
    const S invRhocCmu_half = sc<S>(0.5) * sc<S>(p.Cmu) * invRhoc;
    const S mACI = invRhocCmu_half / dx1[i] * (mu2 / dmx1[ip] + mu1 / dmx1[i]) + sc<S>(0.25) * (-u2 / dmx1[ip] + u1 / dmx1[i]);
    const S mAPI = -invRhocCmu_half / dx1[i] * mu2 / dmx1[ip] + sc<S>(0.25) * (u2 / dmx1[ip]);
    const S mAMI = -invRhocCmu_half / dx1[i] * mu1 / dmx1[i] + sc<S>(0.25) * (-u1 / dmx1[i]);
    ac = mACI * sc<S>(dt) + sc<S>(1.0);
    ap = mAPI * sc<S>(dt);
    am = mAMI * sc<S>(dt);
}
struct DVX {
    Grid g;
    Phys p;
    double dt;
    const double *U, *iR, *stale;
    static constexpr int NF = 3;  // field inputs, in the order below
    template <class S, class Wr>
    __device__ __forceinline__ void eval(const Wr &w, int i, int j, int k, S &a, S &b, S &c, S &d) const {
        dV_x_row<S>(g, p, dt, w(0, U), w(1, iR), w(2, stale), i, j, k, a, b, c);
    }
    __device__ __forceinline__ void operator()(int i, int j, int k, double &a, double &b, double &c,
                                               double &d) const {
        eval<double>(RawW(), i, j, k, a, b, c, d);
    }
};

template <class S, class FP>
__device__ __forceinline__ void dV_y_row(const Grid &g, const Phys &p, double dt, FP V,
    FP Mu, FP iR,
    int i, int j, int k, S &am, S &ac, S &ap) {
    ROW_METRICS
    const int jm = j - 1, jp = j + 1;
    const int jvp = g.jS[jm], jvm = g.jS[jp];
    const S v3 = sc<S>(0.5) * (Vf(i, jm, k) + Vf(i, j, k));
    const S v4 = sc<S>(0.5) * (Vf(i, j, k) + Vf(i, jp, k));
    const S dvdy3 = (Vf(i, j, k) - Vf(i, jm, k)) / dx2[jm];
    const S dvdy4 = (Vf(i, jp, k) - Vf(i, j, k)) / dx2[j];
    const S invRhoc = sc<S>(0.5) * (dx2[jm] * IR(j, k) + dx2[j] * IR(jm, k)) / dmx2[j];
    const S invRhocCmu_half = sc<S>(0.5) * sc<S>(p.Cmu) * invRhoc;
    const S mACJ = invRhocCmu_half / dmx2[j] * (MU(j, k) / dx2[j] + MU(jm, k) / dx2[jm]) * sc<S>(2.0)
                      + sc<S>(0.25) * (dvdy4 * sc<S>(0.5) + dvdy3 * sc<S>(0.5) - v4 / dx2[j] + v3 / dx2[jm]);
    S mAPJ = -invRhocCmu_half / dmx2[j] * MU(j, k) / dx2[j] * sc<S>(2.0) + sc<S>(0.25) * (v4 / dx2[j] + dvdy4 * sc<S>(0.5));
    mAPJ = mAPJ * sc<S>(jvp);
    S mAMJ = -invRhocCmu_half / dmx2[j] * MU(jm, k) / dx2[jm] * sc<S>(2.0) + sc<S>(0.25) * (-v3 / dx2[jm] + dvdy3 * sc<S>(0.5));
    mAMJ = mAMJ * sc<S>(jvm);
    ac = mACJ * sc<S>(dt) + sc<S>(1.0);
    ap = mAPJ * sc<S>(dt);
    am = mAMJ * sc<S>(dt);
}
struct DVY {
    // FP32 lines of the mixed form: the unconditional row loop (0.75 -> 0.53 ms
    // at 512^3)
    static constexpr int W32 = 1;
    static constexpr bool FULL32 = true;
    Grid g;
    Phys p;
    double dt;
    const double *V, *Mu, *iR;
    static constexpr int NF = 3;  // field inputs, in the order below
    template <class S, class Wr>
    __device__ __forceinline__ void eval(const Wr &w, int i, int j, int k, S &a, S &b, S &c, S &d) const {
        dV_y_row<S>(g, p, dt, w(0, V), w(1, Mu), w(2, iR), i, j, k, a, b, c);
    }
    __device__ __forceinline__ void operator()(int i, int j, int k, double &a, double &b, double &c,
                                               double &d) const {
        eval<double>(RawW(), i, j, k, a, b, c, d);
    }
};

// ---------------------------------------------------------------- dW -------
template <class S, int WM, class FP>
__device__ __forceinline__ void dW_z_row(const Grid &g, const Phys &p, double dt, FP U,
    FP V, FP W, FP P,
    FP dU, FP dV, FP Mu,
    FP iR, FP VBCup, FP VBCbt,
    FP WBCup, FP WBCbt,
    int i, int j, int k, S &am, S &ac, S &ap, S &rhsout) {
    ROW_METRICS
    const int im = i - 1, ip = i + 1, jm = j - 1, jp = j + 1, km = k - 1, kp = k + 1;
    const int jwp = g.jC[jp], jwm = g.jC[jm];
    const S Cmu = sc<S>(p.Cmu);
#define dVf(a, b, c) dV[g.ix(a, b, c)]
    const S w5 = sc<S>(0.5) * (Wf(i, j, k) + Wf(i, j, km));
    const S w6 = sc<S>(0.5) * (Wf(i, j, k) + Wf(i, j, kp));
    const S u1 = (dx3[km] * Uf(i, j, k) + dx3[k] * Uf(i, j, km)) / dmx3[k] * sc<S>(0.5);
    const S u2 = (dx3[km] * Uf(ip, j, k) + dx3[k] * Uf(ip, j, km)) / dmx3[k] * sc<S>(0.5);
    const S v3 = (dx3[km] * Vf(i, j, k) + dx3[k] * Vf(i, j, km)) / dmx3[k] * sc<S>(0.5);
    const S v4 = (dx3[km] * Vf(i, jp, k) + dx3[k] * Vf(i, jp, km)) / dmx3[k] * sc<S>(0.5);

    const S dwdx1 = (Wf(i, j, k) - Wf(im, j, k)) / dmx1[i];
    const S dwdx2 = (Wf(ip, j, k) - Wf(i, j, k)) / dmx1[ip];
    const S dwdy3 = (Wf(i, j, k) - Wf(i, jm, k)) / dmx2[j];
    const S dwdy4 = (Wf(i, jp, k) - Wf(i, j, k)) / dmx2[jp];
    const S dwdz5 = (Wf(i, j, k) - Wf(i, j, km)) / dx3[km];
    const S dwdz6 = (Wf(i, j, kp) - Wf(i, j, k)) / dx3[k];
    const S dudz1 = (Uf(i, j, k) - Uf(i, j, km)) / dmx3[k];
    const S dudz2 = (Uf(ip, j, k) - Uf(ip, j, km)) / dmx3[k];
    const S dvdz3 = (Vf(i, j, k) - Vf(i, j, km)) / dmx3[k];
    const S dvdz4 = (Vf(i, jp, k) - Vf(i, jp, km)) / dmx3[k];
    const S ddudz1 = (dUf(i, j, k) - dUf(i, j, km)) / dmx3[k];
    const S ddudz2 = (dUf(ip, j, k) - dUf(ip, j, km)) / dmx3[k];
    const S ddvdz3 = (dVf(i, j, k) - dVf(i, j, km)) / dmx3[k];
    const S ddvdz4 = (dVf(i, jp, k) - dVf(i, jp, km)) / dmx3[k];
    const S ddu1 = (dx3[km] * dUf(i, j, k) + dx3[k] * dUf(i, j, km)) / dmx3[k] * sc<S>(0.5);
    const S ddu2 = (dx3[km] * dUf(ip, j, k) + dx3[k] * dUf(ip, j, km)) / dmx3[k] * sc<S>(0.5);
    const S ddv3 = (dx3[km] * dVf(i, j, k) + dx3[k] * dVf(i, j, km)) / dmx3[k] * sc<S>(0.5) * sc<S>(jwm);
    const S ddv4 = (dx3[km] * dVf(i, jp, k) + dx3[k] * dVf(i, jp, km)) / dmx3[k] * sc<S>(0.5) * sc<S>(jwp);

    S mua = sc<S>(0.5) * (dx3[km] * MU(j, k) + dx3[k] * MU(j, km)) / dmx3[k];
    S muc = sc<S>(0.5) * (dx3[km] * MU(j, k) + dx3[k] * MU(j, km)) / dmx3[k];
    S mub = sc<S>(0.5) * (dx3[km] * MU(j, k) + dx3[k] * MU(j, km)) / dmx3[k];
    const S mu1 = sc<S>(0.5) * (dx1[im] * muc + dx1[i] * mua) / dmx1[i];
    const S mu2 = sc<S>(0.5) * (dx1[ip] * muc + dx1[i] * mub) / dmx1[ip];
    mua = sc<S>(0.5) * (dx3[km] * MU(jm, k) + dx3[k] * MU(jm, km)) / dmx3[k];
    muc = sc<S>(0.5) * (dx3[km] * MU(j, k) + dx3[k] * MU(j, km)) / dmx3[k];
    mub = sc<S>(0.5) * (dx3[km] * MU(jp, k) + dx3[k] * MU(jp, km)) / dmx3[k];
    const S mu3 = sc<S>(0.5) * (dx2[jm] * muc + dx2[j] * mua) / dmx2[j];
    const S mu4 = sc<S>(0.5) * (dx2[jp] * muc + dx2[j] * mub) / dmx2[jp];

    const S invRhoc = sc<S>(0.5) * (dx3[km] * IR(j, k) + dx3[k] * IR(j, km)) / dmx3[k];
    const S invRhocCmu_half = sc<S>(0.5) * Cmu * invRhoc;

    const S viscous_w1 = sc<S>(1.0) * (mu2 * dwdx2 - mu1 * dwdx1) / dx1[i];
    const S viscous_w2 = sc<S>(1.0) * (mu4 * dwdy4 - mu3 * dwdy3) / dx2[j];
    const S viscous_w3 = sc<S>(1.0) * (MU(j, k) * dwdz6 - MU(j, km) * dwdz5) / dmx3[k];
    const S viscous_w31 = sc<S>(1.0) * (mu2 * dudz2 - mu1 * dudz1) / dx1[i];
    const S viscous_w32 = sc<S>(1.0) * (mu4 * dvdz4 - mu3 * dvdz3) / dx2[j];

    S rhs = invRhocCmu_half * (viscous_w1 + viscous_w2 + sc<S>(2.0) * viscous_w3 + viscous_w31 + viscous_w32);
    rhs = rhs - sc<S>(p.Cmp) * invRhoc * (P[g.ix(i, j, k)] - P[g.ix(i, j, km)]) / dmx3[k];

    if (WM == 0 || (WM == 1 && (jwm != 1 || jwp != 1))) {  // wall terms, as in the dU rows
    S zbc_down = sc<S>(0.25) * v3 / dmx2[j] * WBCbt[g.b2(i, k)]
                    - sc<S>(0.25) * dwdy3 * (dx3[km] * VBCbt[g.b2(i, k)] + dx3[k] * VBCbt[g.b2(i, km)]) / dmx3[k] * sc<S>(0.5);
    S zbc_up = -sc<S>(0.25) * v4 / dmx2[jp] * WBCup[g.b2(i, k)]
                  - sc<S>(0.25) * dwdy4 * (dx3[km] * VBCup[g.b2(i, k)] + dx3[k] * VBCup[g.b2(i, km)]) / dmx3[k] * sc<S>(0.5);
    zbc_down = zbc_down + invRhocCmu_half / dx2[j] * mu3 / dmx2[jp] * WBCbt[g.b2(i, k)];
    zbc_up = zbc_up + invRhocCmu_half / dx2[j] * mu4 / dmx2[j] * WBCup[g.b2(i, k)];
    zbc_down = zbc_down - invRhocCmu_half / dx2[j] * mu3 / dmx3[k] * (VBCbt[g.b2(i, k)] - VBCbt[g.b2(i, km)]);
    zbc_up = zbc_up + invRhocCmu_half / dx2[j] * mu4 / dmx3[k] * (VBCup[g.b2(i, k)] - VBCup[g.b2(i, km)]);
    rhs = rhs + (sc<S>(1.0) - sc<S>(jwm)) * zbc_down + (sc<S>(1.0) - sc<S>(jwp)) * zbc_up;
    }

    const S mACI = invRhocCmu_half / dx1[i] * (mu2 / dmx1[ip] + mu1 / dmx1[i]) + sc<S>(0.25) * (-u2 / dmx1[ip] + u1 / dmx1[i]);
    const S mAPI = -invRhocCmu_half / dx1[i] * mu2 / dmx1[ip] + sc<S>(0.25) * (u2 / dmx1[ip]);
    const S mAMI = -invRhocCmu_half / dx1[i] * mu1 / dmx1[i] + sc<S>(0.25) * (-u1 / dmx1[i]);
    const S mACJ = invRhocCmu_half / dx2[j] * (mu4 / dmx2[jp] + mu3 / dmx2[j]) + sc<S>(0.25) * (-v4 / dmx2[jp] + v3 / dmx2[j]);
    S mAPJ = -invRhocCmu_half / dx2[j] * mu4 / dmx2[jp] + sc<S>(0.25) * (v4 / dmx2[jp]);
    mAPJ = mAPJ * sc<S>(jwp);
    S mAMJ = -invRhocCmu_half / dx2[j] * mu3 / dmx2[j] + sc<S>(0.25) * (-v3 / dmx2[j]);
    mAMJ = mAMJ * sc<S>(jwm);
    const S mACK = invRhocCmu_half / dmx3[k] * (MU(j, k) / dx3[k] + MU(j, km) / dx3[km]) * sc<S>(2.0)
                      + sc<S>(0.25) * (-w6 / dx3[k] + w5 / dx3[km] + sc<S>(0.5) * dwdz6 + sc<S>(0.5) * dwdz5);
    const S mAPK = -invRhocCmu_half / dmx3[k] * MU(j, k) / dx3[k] * sc<S>(2.0) + sc<S>(0.25) * (w6 / dx3[k] + sc<S>(0.5) * dwdz6);
    const S mAMK = -invRhocCmu_half / dmx3[k] * MU(j, km) / dx3[km] * sc<S>(2.0) + sc<S>(0.25) * (-w5 / dx3[km] + sc<S>(0.5) * dwdz5);

    rhs = rhs - (mAPI * Wf(ip, j, k) + mACI * Wf(i, j, k) + mAMI * Wf(im, j, k)
                 + mAPJ * Wf(i, jp, k) + mACJ * Wf(i, j, k) + mAMJ * Wf(i, jm, k)
                 + mAPK * Wf(i, j, kp) + mACK * Wf(i, j, k) + mAMK * Wf(i, j, km));
    rhs = rhs - (sc<S>(0.25) * (u2 * dwdx2 + u1 * dwdx1) - invRhocCmu_half / dx1[i] * (mu2 * dudz2 - mu1 * dudz1));
    rhs = rhs - (sc<S>(0.25) * (v3 * sc<S>(jwm) * dwdy3 + v4 * sc<S>(jwp) * dwdy4)
                 - invRhocCmu_half / dx2[j] * (mu4 * dvdz4 * sc<S>(jwp) - mu3 * dvdz3 * sc<S>(jwm)));
    rhs = rhs - (sc<S>(0.25) * (ddu2 * dwdx2 + ddu1 * dwdx1) - invRhocCmu_half / dx1[i] * (mu2 * ddudz2 - mu1 * ddudz1));
    rhs = rhs - (sc<S>(0.25) * (ddv3 * sc<S>(jwm) * dwdy3 + ddv4 * sc<S>(jwp) * dwdy4)
                 - invRhocCmu_half / dx2[j] * (mu4 * ddvdz4 * sc<S>(jwp) - mu3 * ddvdz3 * sc<S>(jwm)));

    ac = mACK * sc<S>(dt) + sc<S>(1.0);
    ap = mAPK * sc<S>(dt);
    am = mAMK * sc<S>(dt);
    rhsout = rhs * sc<S>(dt);
}
struct DWZ {
    // FP32 lines of the mixed form: the unconditional row loop with the
    // workgroups split by the wall terms (2.20 -> 1.88 ms at 512^3)
    static constexpr int W32 = 1;
    static constexpr bool FULL32 = true;
    static constexpr bool WALLS32 = true;
    Grid g;
    Phys p;
    double dt;
    const double *U, *V, *W, *P, *dU, *dV, *Mu, *iR, *VBCup, *VBCbt, *WBCup, *WBCbt;
    static constexpr int NF = 12;  // field inputs, in the order below
    static constexpr bool FULL = true;  // unconditional row loop in the FP64 z solve
    template <class S, int WM = -1, class Wr>
    __device__ __forceinline__ void eval(const Wr &w, int i, int j, int k, S &a, S &b, S &c, S &d) const {
        // WM < 0 (the row's own wall-term mode): a branch on j, terms kept in
        // the adjoint's dual passes; the z lines pass 2 or 0 per workgroup
        // (wall_free)
        dW_z_row<S, (WM >= 0 ? WM : (is_dual<S>() ? 0 : 1))>(g, p, dt, w(0, U), w(1, V), w(2, W), w(3, P), w(4, dU), w(5, dV), w(6, Mu), w(7, iR), w(8, VBCup), w(9, VBCbt), w(10, WBCup), w(11, WBCbt), i, j, k, a, b, c, d);
    }
    // no wall term of row j is weighted (the wall terms' flags are all 1)
    __device__ __forceinline__ bool wall_free(int j) const { return g.jC[j - 1] == 1 && g.jC[j + 1] == 1; }
    __device__ __forceinline__ void operator()(int i, int j, int k, double &a, double &b, double &c,
                                               double &d) const {
        eval<double>(RawW(), i, j, k, a, b, c, d);
    }
};

template <class S, class FP>
__device__ __forceinline__ void dW_x_row(const Grid &g, const Phys &p, double dt, FP U,
    FP iR, FP stale,
    int i, int j, int k, S &am, S &ac, S &ap) {
    ROW_METRICS
    const int ip = i + 1, km = k - 1;
    const S mu1 = stale[2], mu2 = stale[3];
    const S u1 = (dx3[km] * Uf(i, j, k) + dx3[k] * Uf(i, j, km)) / dmx3[k] * sc<S>(0.5);
    const S u2 = (dx3[km] * Uf(ip, j, k) + dx3[k] * Uf(ip, j, km)) / dmx3[k] * sc<S>(0.5);
    const S invRhoc = sc<S>(0.5) * (dx3[km] * IR(j, k) + dx3[k] * IR(j, km)) / dmx3[k];
    const S invRhocCmu_half = sc<S>(0.5) * sc<S>(p.Cmu) * invRhoc;
    const S mACI = invRhocCmu_half / dx1[i] * (mu2 / dmx1[ip] + mu1 / dmx1[i]) + sc<S>(0.25) * (-u2 / dmx1[ip] + u1 / dmx1[i]);
    const S mAPI = -invRhocCmu_half / dx1[i] * mu2 / dmx1[ip] + sc<S>(0.25) * (u2 / dmx1[ip]);
    const S mAMI = -invRhocCmu_half / dx1[i] * mu1 / dmx1[i] + sc<S>(0.25) * (-u1 / dmx1[i]);
    ac = mACI * sc<S>(dt) + sc<S>(1.0);
    ap = mAPI * sc<S>(dt);
    am = mAMI * sc<S>(dt);
}
struct DWX {
    Grid g;
    Phys p;
    double dt;
    const double *U, *iR, *stale;
    static constexpr int NF = 3;  // field inputs, in the order below
    template <class S, class Wr>
    __device__ __forceinline__ void eval(const Wr &w, int i, int j, int k, S &a, S &b, S &c, S &d) const {
        dW_x_row<S>(g, p, dt, w(0, U), w(1, iR), w(2, stale), i, j, k, a, b, c);
    }
    __device__ __forceinline__ void operator()(int i, int j, int k, double &a, double &b, double &c,
                                               double &d) const {
        eval<double>(RawW(), i, j, k, a, b, c, d);
    }
};

template <class S, class FP>
__device__ __forceinline__ void dW_y_row(const Grid &g, const Phys &p, double dt, FP V,
    FP iR, FP stale,
    int i, int j, int k, S &am, S &ac, S &ap) {
    ROW_METRICS_D64
    const int jp = j + 1, jm = j - 1;
    const int km = g.n3m - 1;  // stale from the x-stage's last k
    const int jwp = g.jC[jp], jwm = g.jC[jm];
    const S mu3 = stale[4], mu4 = stale[5];
    const S v3 = (dx3[km] * Vf(i, j, k) + dx3[k] * Vf(i, j, km)) / dmx3[k] * sc<S>(0.5);
    const S v4 = (dx3[km] * Vf(i, jp, k) + dx3[k] * Vf(i, jp, km)) / dmx3[k] * sc<S>(0.5);
    const S invRhoc = sc<S>(0.5) * (dx3[km] * IR(j, k) + dx3[k] * IR(j, km)) / dmx3[k];
    const S invRhocCmu_half = sc<S>(0.5) * sc<S>(p.Cmu) * invRhoc;
    const S mACJ = invRhocCmu_half / dx2[j] * (mu4 / dmx2[jp] + mu3 / dmx2[j]) + sc<S>(0.25) * (-v4 / dmx2[jp] + v3 / dmx2[j]);
    S mAPJ = -invRhocCmu_half / dx2[j] * mu4 / dmx2[jp] + sc<S>(0.25) * (v4 / dmx2[jp]);
    mAPJ = mAPJ * sc<S>(jwp);
    S mAMJ = -invRhocCmu_half / dx2[j] * mu3 / dmx2[j] + sc<S>(0.25) * (-v3 / dmx2[j]);
    mAMJ = mAMJ * sc<S>(jwm);
    ac = mACJ * sc<S>(dt) + sc<S>(1.0);
    ap = mAPJ * sc<S>(dt);
    am = mAMJ * sc<S>(dt);
}
struct DWY {
    // FP32 lines of the mixed form: the unconditional row loop (0.74 -> 0.56 ms
    // at 512^3)
    static constexpr int W32 = 1;
    static constexpr bool FULL32 = true;
    Grid g;
    Phys p;
    double dt;
    const double *V, *iR, *stale;
    static constexpr int NF = 3;  // field inputs, in the order below
    template <class S, class Wr>
    __device__ __forceinline__ void eval(const Wr &w, int i, int j, int k, S &a, S &b, S &c, S &d) const {
        dW_y_row<S>(g, p, dt, w(0, V), w(1, iR), w(2, stale), i, j, k, a, b, c);
    }
    __device__ __forceinline__ void operator()(int i, int j, int k, double &a, double &b, double &c,
                                               double &d) const {
        eval<double>(RawW(), i, j, k, a, b, c, d);
    }
};

// ------------------------------------------------------- block corrections --
// mpi_momentum_blockLdV (:1278-1356): the new dV of cell (i,j,k), j >= 2
template <class S, class FP>
__device__ __forceinline__ void blockLdV_cell(const Grid &g, const Phys &p, double dt, FP V, FP dW, FP Mu,
                                             FP iR, FP dV, int i, int j, int k, S &out) {
    ROW_METRICS
    const int jm = j - 1, km = k - 1, kp = k + 1;
#define dWf(a, b, c) dW[g.ix(a, b, c)]
    const S dwm5 = (dx2[jm] * dWf(i, j, k) + dx2[j] * dWf(i, jm, k)) / dmx2[j] * sc<S>(0.5);
    const S dwm6 = (dx2[jm] * dWf(i, j, kp) + dx2[j] * dWf(i, jm, kp)) / dmx2[j] * sc<S>(0.5);
    const S ddwdy5 = (dWf(i, j, k) - dWf(i, jm, k)) / dmx2[j];
    const S ddwdy6 = (dWf(i, j, kp) - dWf(i, jm, kp)) / dmx2[j];
    const S dvdz5 = (Vf(i, j, k) - Vf(i, j, km)) / dmx3[k];
    const S dvdz6 = (Vf(i, j, kp) - Vf(i, j, k)) / dmx3[kp];
    const S invRhoc = sc<S>(0.5) * (dx2[jm] * IR(j, k) + dx2[j] * IR(jm, k)) / dmx2[j];
    const S mua = sc<S>(0.5) * (dx2[jm] * MU(j, km) + dx2[j] * MU(jm, km)) / dmx2[j];
    const S muc = sc<S>(0.5) * (dx2[jm] * MU(j, k) + dx2[j] * MU(jm, k)) / dmx2[j];
    const S mub = sc<S>(0.5) * (dx2[jm] * MU(j, kp) + dx2[j] * MU(jm, kp)) / dmx2[j];
    const S mu5 = sc<S>(0.5) * (dx3[km] * muc + dx3[k] * mua) / dmx3[k];
    const S mu6 = sc<S>(0.5) * (dx3[kp] * muc + dx3[k] * mub) / dmx3[kp];
    out = dV[g.ix(i, j, k)] - sc<S>(dt) * (sc<S>(0.25) * (dwm5 * dvdz5 + dwm6 * dvdz6)
                          - sc<S>(0.5) * sc<S>(p.Cmu) * invRhoc / dx3[k] * (mu6 * ddwdy6 - mu5 * ddwdy5));
}
struct BlockLdV {
    Grid g;
    Phys p;
    double dt;
    const double *V, *dW, *Mu, *iR, *dV;
    static constexpr int NF = 5;
    template <class S, class Wr>
    __device__ __forceinline__ void eval(const Wr &w, int i, int j, int k, S &a, S &b, S &c, S &d) const {
        blockLdV_cell<S>(g, p, dt, w(0, V), w(1, dW), w(2, Mu), w(3, iR), w(4, dV), i, j, k, a);
    }
};

// E: the increments' type and the arithmetic (double with the fields
// themselves; float with their FP32 twins in the mixed form)
template <typename E, class Acc>
__global__ void __launch_bounds__(256) k_blockLdV(BlockLdV cell, Acc acc, E *dV) {
    const Grid &g = cell.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.n1m || j > g.n2m || j < 2) return;
    E o, x;
    cell.eval<E>(acc, i, j, k, o, x, x, x);
    dV[g.ix(i, j, k)] = o;
}

// mpi_momentum_blockLdU (:1362-1460): the new dU of cell (i,j,k)
template <class S, class FP>
__device__ __forceinline__ void blockLdU_cell(const Grid &g, const Phys &p, double dt, FP U, FP dV, FP dW,
                                             FP Mu, FP iR, FP dU, int i, int j, int k, S &out) {
    ROW_METRICS
    const int im = i - 1, jm = j - 1, jp = j + 1, km = k - 1, kp = k + 1;
    const int jum = g.jC[jm], jup = g.jC[jp];
    const S invRhoc = sc<S>(0.5) * (dx1[im] * IR(j, k) + dx1[i] * IR(j, k)) / dmx1[i];
    const S dvm3 = (dx1[i] * dVf(im, j, k) + dx1[im] * dVf(i, j, k)) / dmx1[i] * sc<S>(0.5) * sc<S>(jum);
    const S dvm4 = (dx1[i] * dVf(im, jp, k) + dx1[im] * dVf(i, jp, k)) / dmx1[i] * sc<S>(0.5) * sc<S>(jup);
    const S dudy3 = (Uf(i, j, k) - Uf(i, jm, k)) / dmx2[j];
    const S dudy4 = (Uf(i, jp, k) - Uf(i, j, k)) / dmx2[jp];
    const S dvmdx3 = (dVf(i, j, k) - dVf(im, j, k)) / dmx1[i] * sc<S>(jum);
    const S dvmdx4 = (dVf(i, jp, k) - dVf(im, jp, k)) / dmx1[i] * sc<S>(jup);
    S mua = sc<S>(0.5) * (dx1[im] * MU(jm, k) + dx1[i] * MU(jm, k)) / dmx1[i];
    S muc = sc<S>(0.5) * (dx1[im] * MU(j, k) + dx1[i] * MU(j, k)) / dmx1[i];
    S mub = sc<S>(0.5) * (dx1[im] * MU(jp, k) + dx1[i] * MU(jp, k)) / dmx1[i];
    const S mu3 = sc<S>(0.5) * (dx2[jm] * muc + dx2[j] * mua) / dmx2[j];
    const S mu4 = sc<S>(0.5) * (dx2[jp] * muc + dx2[j] * mub) / dmx2[jp];
    const S dwm5 = (dx1[i] * dWf(im, j, k) + dx1[im] * dWf(i, j, k)) / dmx1[i] * sc<S>(0.5);
    const S dwm6 = (dx1[i] * dWf(im, j, kp) + dx1[im] * dWf(i, j, kp)) / dmx1[i] * sc<S>(0.5);
    const S dudz5 = (Uf(i, j, k) - Uf(i, j, km)) / dmx3[k];
    const S dudz6 = (Uf(i, j, kp) - Uf(i, j, k)) / dmx3[kp];
    const S dwmdx5 = (dWf(i, j, k) - dWf(im, j, k)) / dmx1[i];
    const S dwmdx6 = (dWf(i, j, kp) - dWf(im, j, kp)) / dmx1[i];
    mua = sc<S>(0.5) * (dx1[im] * MU(j, km) + dx1[i] * MU(j, km)) / dmx1[i];
    muc = sc<S>(0.5) * (dx1[im] * MU(j, k) + dx1[i] * MU(j, k)) / dmx1[i];
    mub = sc<S>(0.5) * (dx1[im] * MU(j, kp) + dx1[i] * MU(j, kp)) / dmx1[i];
    const S mu5 = sc<S>(0.5) * (dx3[km] * muc + dx3[k] * mua) / dmx3[k];
    const S mu6 = sc<S>(0.5) * (dx3[kp] * muc + dx3[k] * mub) / dmx3[kp];
    out = dU[g.ix(i, j, k)]
          - sc<S>(dt) * (sc<S>(0.25) * (dvm4 * dudy4 + dvm3 * dudy3) - sc<S>(0.5) * sc<S>(p.Cmu) * invRhoc / dx2[j] * (mu4 * dvmdx4 - mu3 * dvmdx3))
          - sc<S>(dt) * (sc<S>(0.25) * (dwm6 * dudz6 + dwm5 * dudz5) - sc<S>(0.5) * sc<S>(p.Cmu) * invRhoc / dx3[k] * (mu6 * dwmdx6 - mu5 * dwmdx5));
}
struct BlockLdU {
    Grid g;
    Phys p;
    double dt;
    const double *U, *dV, *dW, *Mu, *iR, *dU;
    static constexpr int NF = 6;
    template <class S, class Wr>
    __device__ __forceinline__ void eval(const Wr &w, int i, int j, int k, S &a, S &b, S &c, S &d) const {
        blockLdU_cell<S>(g, p, dt, w(0, U), w(1, dV), w(2, dW), w(3, Mu), w(4, iR), w(5, dU), i, j, k, a);
    }
};

// planes and consecutive cells in i per thread of the block corrections
// (measured at 512^3: blockLdU 2.27 -> 1.76 ms with two cells in i, blockLdV
// the same either way)
constexpr int LDV_KB = 1, LDU_KB = 4, LDV_IB = 1, LDU_IB = 2;

// blockLdV with the pseudo-update of V: dV(j >= 2) = the new dV and Vn =
// V + dV; at j = 1 Vn = V + dV * 0 (:178-208)
template <typename E, class Acc>
__global__ void __launch_bounds__(256) k_blockLdV_pseudo(BlockLdV cell, Acc acc, E *dV, double *__restrict__ Vn) {
    const Grid &g = cell.g;
    const int i0 = (blockIdx.x * blockDim.x + threadIdx.x) * LDV_IB + 1;  // n1m is a multiple of LDV_IB
    const int j = wave_row_j();
    if (i0 > g.n1m || j > g.n2m) return;
    // loads of every cell before the stores (dV is updated in place, each
    // cell reading only its own point of it)
    constexpr int NC = LDV_IB * LDV_KB;
    E o[NC];
    double vn[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const int i = i0 + c % LDV_IB, k = blockIdx.z * LDV_KB + c / LDV_IB + 1;
        const size_t q = g.ix(i, j, k);
        if (j < 2) {
            o[c] = E(0);  // not stored
            vn[c] = cell.V[q] + (double)dV[q] * 0.0;
        } else {
            E x;
            cell.eval<E>(acc, i, j, k, o[c], x, x, x);
            vn[c] = cell.V[q] + (double)o[c];
        }
    }
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const size_t q = g.ix(i0 + c % LDV_IB, j, blockIdx.z * LDV_KB + c / LDV_IB + 1);
        if (j >= 2) dV[q] = o[c];
        Vn[q] = vn[c];
    }
}

template <typename E, class Acc>
__global__ void __launch_bounds__(256) k_blockLdU(BlockLdU cell, Acc acc, E *dU) {
    const Grid &g = cell.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.n1m || j > g.n2m) return;
    E o, x;
    cell.eval<E>(acc, i, j, k, o, x, x, x);
    dU[g.ix(i, j, k)] = o;
}

template <typename E>
__global__ void __launch_bounds__(256)
k_pseudoupdate(Grid g, double *__restrict__ U, double *__restrict__ V, double *__restrict__ W,
               const E *__restrict__ dU, const E *__restrict__ dV, const E *__restrict__ dW) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.n1m || j > g.n2m) return;
    const size_t q = g.ix(i, j, k);
    U[q] = U[q] + (double)dU[q];
    if (j >= 2) V[q] = V[q] + (double)dV[q];
    else V[q] = V[q] + (double)dV[q] * 0.0;  // (2 - j_indexS) with j_indexS = 2
    W[q] = W[q] + (double)dW[q];
}

// ------------------------------------------------------------- host side ---
// The increments dU, dV, dW as this form keeps them, and the access its
// rows and block corrections read through: the FP64 module arrays
// themselves, or (the mixed form) their FP32 twins.
using inc_t = scr_t;
template <typename E = inc_t>
static inline E *inc(tcs_ctx *c, int q) {
    if constexpr (sizeof(E) == sizeof(float)) return twin(c, c->f[q]);
    else return c->f[q];
}
template <typename E = inc_t>
static inline auto cell_access(const tcs_ctx *c) {
    if constexpr (sizeof(E) == sizeof(float)) return twin_w(c);
    else return RawW();
}
// what the momentum stages read besides the increments
constexpr unsigned MOMENTUM_TWINS = TWIN_U | TWIN_V | TWIN_W | TWIN_P | TWIN_T | TWIN_TAB | TWIN_WALL;

int momentum_coeffi(tcs_ctx *c) {
    c->twin_ok &= ~(unsigned)TWIN_TAB;
    TCS_LAUNCH(c->stream, "momentum_coeffi", k_momentum_coeffi, dim3((c->N2 + 63) / 64, c->N3), dim3(64), 0, c->grid(),
                       c->phys, c->f[TCS_T], c->f[TCS_MU], c->f[TCS_INVRHO]);
    return TCS_OK;
}

static int launch_stale(tcs_ctx *c) {
    c->twin_ok &= ~(unsigned)TWIN_STALE;
    TCS_LAUNCH(c->stream, "momentum_stale", k_momentum_stale, dim3(1), dim3(64), 0, c->grid(), c->f[TCS_MU], c->stale);
    return TCS_OK;
}

// y stage of a momentum increment: rows j = jlo..n2m stored into the field
// F; SX lines per workgroup of the split solve (measured at 512^3: the
// light dU and dV y rows run faster in two workgroups of 8 lines per CU,
// the dW rows in one of 16; FP32 lines keep 16, 64-byte segments).  With
// jlo = 2 (dV) the split lines start at the wall row as an identity row
// when all n2m rows fit a split line (dV y at 512^3: 1.56 -> 1.0 ms, whole
// 16-row parts and the unconditional row loop); dV's wall row is zero
// either way.
template <int SX, class Row>
static int y_stage(tcs_ctx *c, const char *label_rows, const char *label_solve, const Row &row, int jlo, double *F) {
    scr_t *am = SCR(c->am), *ac = SCR(c->ac), *ap = SCR(c->ap), *rhs = SCR(c->rhs);
    const int nrows = c->n2m - jlo + 1;
    if (jlo == 2 && rows_in_solve<scr_t>(c, c->n2m))
        return y_stage_in_solve<SX>(c, label_solve, row, 1, c->n2m, rhs, F, 1, jlo);
    if (rows_in_solve<scr_t>(c, nrows)) return y_stage_in_solve<SX>(c, label_solve, row, jlo, nrows, rhs, F, 1);
    TCS_TRY(rows3d_launch<false>(c, label_rows, row, jlo, am, ac, ap, rhs));
    return sweep_y_field(c, jlo, nrows, F, 1);
}

int momentum_solvedU(tcs_ctx *c, double dt) {
    TCS_TRY(ensure_twins(c, MOMENTUM_TWINS));
    const Grid g = c->grid();
    double **f = c->f;
    scr_t *am = SCR(c->am), *ac = SCR(c->ac), *ap = SCR(c->ap), *rhs = SCR(c->rhs);
    const DUZ z{g, c->phys, dt, f[TCS_U], f[TCS_V], f[TCS_W], f[TCS_P], f[TCS_T], f[TCS_MU], f[TCS_INVRHO],
                c->wall[TCS_UBCUP], c->wall[TCS_UBCBT], c->wall[TCS_VBCUP], c->wall[TCS_VBCBT]};
    TCS_TRY(z_stage(c, "dU_z", "dU_z_solve", z, 1, am, ac, ap, rhs));
    TCS_TRY(x_stage(c, "dU_x", "dU_x_solve", DUX{g, c->phys, dt, f[TCS_U], f[TCS_MU], f[TCS_INVRHO]}, 1, am, ac, ap,
                    rhs));
    return y_stage<sizeof(scr_t) == 8 ? 8 : 16>(c, "dU_y", "dU_y_solve", DUY{g, c->phys, dt, f[TCS_V], f[TCS_MU], f[TCS_INVRHO]}, 1, f[TCS_DU]);
}

int momentum_solvedV(tcs_ctx *c, double dt) {
    const Grid g = c->grid();
    double **f = c->f;
    scr_t *am = SCR(c->am), *ac = SCR(c->ac), *ap = SCR(c->ap), *rhs = SCR(c->rhs);
    TCS_TRY(launch_stale(c));
    TCS_TRY(ensure_twins(c, MOMENTUM_TWINS | TWIN_STALE));
    // rows j = 1 are not part of the dV systems: every stage covers j = 2..n2m
    const DVZ z{g, c->phys, dt, f[TCS_U], f[TCS_V], f[TCS_W], f[TCS_P], f[TCS_DU], f[TCS_MU], f[TCS_INVRHO],
                c->wall[TCS_VBCUP], c->wall[TCS_VBCBT]};
    TCS_TRY(z_stage(c, "dV_z", "dV_z_solve", z, 2, am, ac, ap, rhs));
    TCS_TRY(x_stage(c, "dV_x", "dV_x_solve", DVX{g, c->phys, dt, f[TCS_U], f[TCS_INVRHO], c->stale}, 2, am, ac, ap,
                    rhs));
    TCS_LAUNCH(c->stream, "dV_zero_rows", k_dV_zero_rows<inc_t>, dim3((c->N1 + 63) / 64, c->N3), dim3(64), 0, g,
               inc(c, TCS_DV));
    return y_stage<sizeof(scr_t) == 8 ? 8 : 16>(c, "dV_y", "dV_y_solve", DVY{g, c->phys, dt, f[TCS_V], f[TCS_MU], f[TCS_INVRHO]}, 2, f[TCS_DV]);
}

int momentum_solvedW(tcs_ctx *c, double dt) {
    const Grid g = c->grid();
    double **f = c->f;
    scr_t *am = SCR(c->am), *ac = SCR(c->ac), *ap = SCR(c->ap), *rhs = SCR(c->rhs);
    TCS_TRY(launch_stale(c));
    TCS_TRY(ensure_twins(c, MOMENTUM_TWINS | TWIN_STALE));
    const DWZ z{g, c->phys, dt, f[TCS_U], f[TCS_V], f[TCS_W], f[TCS_P], f[TCS_DU], f[TCS_DV], f[TCS_MU],
                f[TCS_INVRHO], c->wall[TCS_VBCUP], c->wall[TCS_VBCBT], c->wall[TCS_WBCUP], c->wall[TCS_WBCBT]};
    TCS_TRY(z_stage(c, "dW_z", "dW_z_solve", z, 1, am, ac, ap, rhs));
    TCS_TRY(x_stage(c, "dW_x", "dW_x_solve", DWX{g, c->phys, dt, f[TCS_U], f[TCS_INVRHO], c->stale}, 1, am, ac, ap,
                    rhs));
    return y_stage<16>(c, "dW_y", "dW_y_solve", DWY{g, c->phys, dt, f[TCS_V], f[TCS_INVRHO], c->stale}, 1, f[TCS_DW]);
}

int momentum_blockLdV(tcs_ctx *c, double dt) {
    TCS_TRY(ensure_twins(c, TWIN_V | TWIN_TAB));
    const dim3 blk = interior_block(), grd = interior_grid(c->n1m, c->n2m, c->n3m);
    const BlockLdV cell{c->grid(), c->phys, dt, c->f[TCS_V], c->f[TCS_DW], c->f[TCS_MU], c->f[TCS_INVRHO], c->f[TCS_DV]};
    auto acc = cell_access(c);
    TCS_LAUNCH(c->stream, "blockLdV", (k_blockLdV<inc_t, decltype(acc)>), grd, blk, 0, cell, acc, inc(c, TCS_DV));
    return TCS_OK;
}

int momentum_blockLdU(tcs_ctx *c, double dt) {
    TCS_TRY(ensure_twins(c, TWIN_U | TWIN_TAB));
    const dim3 blk = interior_block(), grd = interior_grid(c->n1m, c->n2m, c->n3m);
    const BlockLdU cell{c->grid(), c->phys, dt,          c->f[TCS_U],      c->f[TCS_DV], c->f[TCS_DW],
                        c->f[TCS_MU], c->f[TCS_INVRHO], c->f[TCS_DU]};
    auto acc = cell_access(c);
    TCS_LAUNCH(c->stream, "blockLdU", (k_blockLdU<inc_t, decltype(acc)>), grd, blk, 0, cell, acc, inc(c, TCS_DU));
    return TCS_OK;
}

// blockLdU with the pseudo-update of U and W: dU = the new dU, Un = U + dU,
// Wn = W + dW (:178-208)
template <typename E, class Acc>
__global__ void __launch_bounds__(256) k_blockLdU_pseudo(BlockLdU cell, Acc acc, const double *__restrict__ W,
                                                        const E *__restrict__ dW, E *dU,
                                                        double *__restrict__ Un, double *__restrict__ Wn) {
    const Grid &g = cell.g;
    const int i0 = (blockIdx.x * blockDim.x + threadIdx.x) * LDU_IB + 1;  // n1m is a multiple of LDU_IB
    const int j = wave_row_j();
    if (i0 > g.n1m || j > g.n2m) return;
    // loads of every cell before the stores (dU is updated in place, each
    // cell reading only its own point of it)
    constexpr int NC = LDU_IB * LDU_KB;
    E o[NC];
    double un[NC], wn[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const int i = i0 + c % LDU_IB, k = blockIdx.z * LDU_KB + c / LDU_IB + 1;
        E x;
        cell.eval<E>(acc, i, j, k, o[c], x, x, x);
        const size_t q = g.ix(i, j, k);
        un[c] = cell.U[q] + (double)o[c];
        wn[c] = W[q] + (double)dW[q];
    }
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const size_t q = g.ix(i0 + c % LDU_IB, j, blockIdx.z * LDU_KB + c / LDU_IB + 1);
        dU[q] = o[c];
        Un[q] = un[c];
        Wn[q] = wn[c];
    }
}

int momentum_blockLdV_pseudo(tcs_ctx *c, double dt, double *Vn) {
    const dim3 blk = interior_block(), grd = cell_grid(c->n1m, c->n2m, c->n3m, LDV_KB, interior_block(), LDV_IB);
    const BlockLdV cell{c->grid(), c->phys, dt, c->f[TCS_V], c->f[TCS_DW], c->f[TCS_MU], c->f[TCS_INVRHO], c->f[TCS_DV]};
    TCS_TRY(ensure_twins(c, TWIN_V | TWIN_TAB));
    c->twin_ok &= ~(unsigned)TWIN_V;  // Vn becomes V
    TCS_TRY(copy_shell(c, c->f[TCS_V], Vn));
    auto acc = cell_access(c);
    TCS_LAUNCH(c->stream, "blockLdV", (k_blockLdV_pseudo<inc_t, decltype(acc)>), grd, blk, 0, cell, acc,
               inc(c, TCS_DV), Vn);
    return TCS_OK;
}

int momentum_blockLdU_pseudo(tcs_ctx *c, double dt, double *Un, double *Wn) {
    const dim3 blk = interior_block(), grd = cell_grid(c->n1m, c->n2m, c->n3m, LDU_KB, interior_block(), LDU_IB);
    const BlockLdU cell{c->grid(), c->phys, dt, c->f[TCS_U], c->f[TCS_DV], c->f[TCS_DW],
                        c->f[TCS_MU], c->f[TCS_INVRHO], c->f[TCS_DU]};
    TCS_TRY(ensure_twins(c, TWIN_U | TWIN_TAB));
    c->twin_ok &= ~(unsigned)(TWIN_U | TWIN_W);  // Un, Wn become U, W
    TCS_TRY(copy_shell(c, {{c->f[TCS_U], Un}, {c->f[TCS_W], Wn}}));
    auto acc = cell_access(c);
    TCS_LAUNCH(c->stream, "blockLdU", (k_blockLdU_pseudo<inc_t, decltype(acc)>), grd, blk, 0, cell, acc,
               c->f[TCS_W], inc(c, TCS_DW), inc(c, TCS_DU), Un, Wn);
    return TCS_OK;
}

int momentum_pseudoupdate(tcs_ctx *c) {
    c->twin_ok &= ~(unsigned)(TWIN_U | TWIN_V | TWIN_W);
    const dim3 blk = interior_block(), grd = interior_grid(c->n1m, c->n2m, c->n3m);
    TCS_LAUNCH(c->stream, "pseudoupdate", k_pseudoupdate<inc_t>, grd, blk, 0, c->grid(), c->f[TCS_U], c->f[TCS_V],
               c->f[TCS_W], inc(c, TCS_DU), inc(c, TCS_DV), inc(c, TCS_DW));
    return TCS_OK;
}

}  // namespace TCS_FORM
}  // namespace tcs
