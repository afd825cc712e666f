// thermal.hip -- mpi_thermal_coeffi / mpi_thermal_solver
// (src/module_solve_thermal.f90:116-156, :175-532) as HIP kernels.
//
// The property polynomials KPP, dKPP, invRhoCp, dinvRhoCp are pure
// functions of T, so the assembly kernels evaluate them from T in
// registers instead of reading four stored 3-D arrays; the arithmetic is
// the reference's, so the values are identical to the stored ones.
// Each ADI stage is an assembly kernel (one thread per cell, i along the
// wavefront) that writes am/ac/ap (and the RHS in stage 1) in the interior
// (i,j,k) layout, followed by the exact-order PaScaL sweep along the
// stage's axis with strided system access (no transposes).
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

// Conductivity polynomial and its derivative (:116-156).  The exact form
// evaluates the power sums as the reference writes them; the fast forms in
// Horner's form (fewer products and live temporaries: the thermal z rows
// evaluate seven of these a row).
template <class S>
__device__ __forceinline__ void th_kpp(const Phys &p, S T, S &kpp, S &dkpp) {
#ifdef TCS_RECIP
    if constexpr (!is_dual<S>()) {
        const S t = sc<S>(p.DeltaT) * T;
        const S s = sc<S>(p.kppS);
        kpp = s * (sc<S>(1.0) + t * (sc<S>(p.c11) + t * (sc<S>(p.c12) + t * (sc<S>(p.c13) + t * (sc<S>(p.c14) + t * sc<S>(p.c15))))));
        dkpp = s * (sc<S>(p.c11) + t * (sc<S>(2.0) * sc<S>(p.c12) + t * (sc<S>(3.0) * sc<S>(p.c13) + t * (sc<S>(4.0) * sc<S>(p.c14) + t * (sc<S>(5.0) * sc<S>(p.c15))))));
        return;
    }
#endif
    const S tmp1 = sc<S>(p.DeltaT) * T;
    const S tmp2 = tmp1 * tmp1;
    const S tmp3 = tmp2 * tmp1;
    const S tmp4 = tmp3 * tmp1;
    const S tmp5 = tmp4 * tmp1;
    kpp = sc<S>(p.c10) * (sc<S>(1.0) + sc<S>(p.c11) * tmp1 + sc<S>(p.c12) * tmp2 + sc<S>(p.c13) * tmp3 + sc<S>(p.c14) * tmp4 + sc<S>(p.c15) * tmp5) / sc<S>(p.KPP0);
    dkpp = sc<S>(p.c10) * (sc<S>(p.c11) + sc<S>(2.0) * sc<S>(p.c12) * tmp1 + sc<S>(3.0) * sc<S>(p.c13) * tmp2 + sc<S>(4.0) * sc<S>(p.c14) * tmp3 + sc<S>(5.0) * sc<S>(p.c15) * tmp4) / sc<S>(p.KPP0);
}

template <class S>
__device__ __forceinline__ S th_kpp_only(const Phys &p, S T) {
    S k, dk;
    th_kpp(p, T, k, dk);
    return k;
}

template <class S>
__device__ __forceinline__ S th_dkpp_only(const Phys &p, S T) {
    S k, dk;
    th_kpp(p, T, k, dk);
    return dk;
}

// 1/(rho cp) and its derivative term (:116-156); the fast forms take the
// polynomials in Horner's form, the derivative term's numerator collected
// by powers of DeltaT*T on the host (Phys::dircN).
template <class S>
__device__ __forceinline__ void th_rhocp(const Phys &p, S T, S &irc, S &dirc) {
#ifdef TCS_RECIP
    if constexpr (!is_dual<S>()) {
        const S t = sc<S>(p.DeltaT) * T;
        irc = sc<S>(p.rcS)
            / (sc<S>(p.a10) * (sc<S>(1.0) + t * (sc<S>(p.a11) + t * (sc<S>(p.a12) + t * (sc<S>(p.a13) + t * (sc<S>(p.a14) + t * sc<S>(p.a15)))))))
            / (sc<S>(p.b10) * (sc<S>(1.0) + t * (sc<S>(p.b11) + t * (sc<S>(p.b12) + t * (sc<S>(p.b13) + t * (sc<S>(p.b14) + t * sc<S>(p.b15)))))));
        const S da = sc<S>(1.0) + t * (sc<S>(p.a11) + t * (sc<S>(p.a12) + t * sc<S>(p.a13)));
        const S db = sc<S>(1.0) + t * (sc<S>(p.b11) + t * (sc<S>(p.b12) + t * sc<S>(p.b13)));
        const S num = sc<S>(p.dircN[0]) + t * (sc<S>(p.dircN[1]) + t * (sc<S>(p.dircN[2]) + t * (sc<S>(p.dircN[3])
                      + t * (sc<S>(p.dircN[4]) + t * sc<S>(p.dircN[5])))));
        dirc = -num / (da * da) / (db * db);
        return;
    }
#endif
    const S tmp1 = sc<S>(p.DeltaT) * T;
    const S tmp2 = tmp1 * tmp1;
    const S tmp3 = tmp2 * tmp1;
    {
        const S tmp4 = tmp3 * tmp1;
        const S tmp5 = tmp4 * tmp1;
        irc = sc<S>(p.Rho0) * sc<S>(p.Cp0)
            / (sc<S>(p.a10) + sc<S>(p.a10) * (sc<S>(p.a11) * tmp1 + sc<S>(p.a12) * tmp2 + sc<S>(p.a13) * tmp3 + sc<S>(p.a14) * tmp4 + sc<S>(p.a15) * tmp5))
            / (sc<S>(p.b10) + sc<S>(p.b10) * (sc<S>(p.b11) * tmp1 + sc<S>(p.b12) * tmp2 + sc<S>(p.b13) * tmp3 + sc<S>(p.b14) * tmp4 + sc<S>(p.b15) * tmp5));
    }
    const S da = sc<S>(1.0) + sc<S>(p.a11) * tmp1 + sc<S>(p.a12) * tmp2 + sc<S>(p.a13) * tmp3;
    const S db = sc<S>(1.0) + sc<S>(p.b11) * tmp1 + sc<S>(p.b12) * tmp2 + sc<S>(p.b13) * tmp3;
    dirc = -(sc<S>(p.b11) * (sc<S>(1.0) + sc<S>(3.0) * sc<S>(p.a12) * tmp2 + sc<S>(4.0) * sc<S>(p.a13) * tmp3)
             + sc<S>(p.a11) * (sc<S>(1.0) + sc<S>(2.0) * sc<S>(p.b11) * tmp1 + sc<S>(3.0) * sc<S>(p.b12) * tmp2 + sc<S>(4.0) * sc<S>(p.b13) * tmp3)
             + tmp1 * (sc<S>(p.b12) * (sc<S>(2.0) + sc<S>(5.0) * sc<S>(p.a13) * tmp3)
                       + sc<S>(p.a12) * (sc<S>(2.0) + sc<S>(4.0) * sc<S>(p.b12) * tmp2 + sc<S>(5.0) * sc<S>(p.b13) * tmp3)
                       + sc<S>(3.0) * tmp1 * (sc<S>(p.a13) + sc<S>(p.b13) + sc<S>(2.0) * sc<S>(p.a13) * sc<S>(p.b13) * tmp3)))
         / (da * da) / (db * db);
}

// mpi_thermal_coeffi: materialise the four property arrays over the full field
__global__ void k_thermal_coeffi(Grid g, Phys p, const double *__restrict__ T, double *__restrict__ KPP,
                                 double *__restrict__ dKPP, double *__restrict__ iRC, double *__restrict__ diRC) {
    const size_t n = (size_t)g.P1 * g.N2 * g.N3;  // the row padding too
    for (size_t q = (size_t)blockIdx.x * blockDim.x + threadIdx.x; q < n; q += (size_t)gridDim.x * blockDim.x) {
        double k, dk, irc, dirc;
        th_kpp(p, T[q], k, dk);
        th_rhocp(p, T[q], irc, dirc);
        KPP[q] = k; dKPP[q] = dk; iRC[q] = irc; diRC[q] = dirc;
    }
}

// stage 1: z (:236-361)
template <class S, int WM, class FP>
__device__ __forceinline__ void thermal_z_row(const Grid &g, const Phys &p, double dt,
    FP T, FP U, FP V,
    FP W, FP TBCup, FP TBCbt,
    int i, int j, int k, S &am, S &ac, S &ap, S &rhsout) {
    const int im = i - 1, ip = i + 1, jm = j - 1, jp = j + 1, km = k - 1, kp = k + 1;
    const int jem = g.jC[jm], jep = g.jC[jp];
    const auto dx1 = metric<S, M_DX1>(g), dx2 = metric<S, M_DX2>(g), dx3 = metric<S, M_DX3>(g);
    const auto dmx1 = metric<S, M_DMX1>(g), dmx2 = metric<S, M_DMX2>(g), dmx3 = metric<S, M_DMX3>(g);
    const S Ct = sc<S>(p.Ct);

    const S Tc = T[g.ix(i, j, k)];
    const S Tim = T[g.ix(im, j, k)], Tip = T[g.ix(ip, j, k)];
    const S Tjm = T[g.ix(i, jm, k)], Tjp = T[g.ix(i, jp, k)];
    const S Tkm = T[g.ix(i, j, km)], Tkp = T[g.ix(i, j, kp)];

    S irc, dirc;
    th_rhocp(p, Tc, irc, dirc);
    S Kc, dKc, Kim, dKim, Kip, dKip, Kjm, dKjm, Kjp, dKjp, Kkm, dKkm, Kkp, dKkp;
    th_kpp(p, Tc, Kc, dKc);
    th_kpp(p, Tim, Kim, dKim);
    th_kpp(p, Tip, Kip, dKip);
    th_kpp(p, Tjm, Kjm, dKjm);
    th_kpp(p, Tjp, Kjp, dKjp);
    th_kpp(p, Tkm, Kkm, dKkm);
    th_kpp(p, Tkp, Kkp, dKkp);

    const S invRhoCpCt_half = sc<S>(0.5) * Ct * irc;
    const S dinvRhoCpCt_half = sc<S>(0.5) * Ct * dirc;

    const S e1 = (dx1[i] * Tim + dx1[im] * Tc) * (sc<S>(0.5) / dmx1[i]);
    const S e2 = (dx1[ip] * Tc + dx1[i] * Tip) * (sc<S>(0.5) / dmx1[ip]);
    const S e3 = (dx2[j] * Tjm + dx2[jm] * Tc) * (sc<S>(0.5) / dmx2[j]);
    const S e4 = (dx2[jp] * Tc + dx2[j] * Tjp) * (sc<S>(0.5) / dmx2[jp]);
    const S e5 = (dx3[k] * Tkm + dx3[km] * Tc) * (sc<S>(0.5) / dmx3[k]);
    const S e6 = (dx3[kp] * Tc + dx3[k] * Tkp) * (sc<S>(0.5) / dmx3[kp]);
    const S dedx1 = (Tc - Tim) / dmx1[i];
    const S dedx2 = (Tip - Tc) / dmx1[ip];
    const S dedy3 = (Tc - Tjm) / dmx2[j];
    const S dedy4 = (Tjp - Tc) / dmx2[jp];
    const S dedz5 = (Tc - Tkm) / dmx3[k];
    const S dedz6 = (Tkp - Tc) / dmx3[kp];

    const S Uc = U[g.ix(i, j, k)], Uip = U[g.ix(ip, j, k)];
    const S Vc = V[g.ix(i, j, k)], Vjp = V[g.ix(i, jp, k)];
    const S Wc = W[g.ix(i, j, k)], Wkp = W[g.ix(i, j, kp)];
    const S convect_e1 = (Uc * dedx1 + Uip * dedx2) * sc<S>(0.5);
    const S convect_e2 = (Vc * dedy3 + Vjp * dedy4) * sc<S>(0.5);
    const S convect_e3 = (Wc * dedz5 + Wkp * dedz6) * sc<S>(0.5);
    S rhs = -sc<S>(0.5) * (convect_e1 + convect_e2 + convect_e3);

    const S kPP1 = sc<S>(0.5) / dmx1[i] * (dx1[i] * Kim + dx1[im] * Kc);
    const S kPP2 = sc<S>(0.5) / dmx1[ip] * (dx1[ip] * Kc + dx1[i] * Kip);
    const S kPP3 = sc<S>(0.5) / dmx2[j] * (dx2[j] * Kjm + dx2[jm] * Kc);
    const S kPP4 = sc<S>(0.5) / dmx2[jp] * (dx2[jp] * Kc + dx2[j] * Kjp);
    const S kPP5 = sc<S>(0.5) / dmx3[k] * (dx3[k] * Kkm + dx3[km] * Kc);
    const S kPP6 = sc<S>(0.5) / dmx3[kp] * (dx3[kp] * Kc + dx3[k] * Kkp);
    const S dkPP1 = sc<S>(0.5) / dmx1[i] * (dx1[i] * dKim + dx1[im] * dKc);
    const S dkPP2 = sc<S>(0.5) / dmx1[ip] * (dx1[ip] * dKc + dx1[i] * dKip);
    const S dkPP3 = sc<S>(0.5) / dmx2[j] * (dx2[j] * dKjm + dx2[jm] * dKc);
    const S dkPP4 = sc<S>(0.5) / dmx2[jp] * (dx2[jp] * dKc + dx2[j] * dKjp);
    const S dkPP5 = sc<S>(0.5) / dmx3[k] * (dx3[k] * dKkm + dx3[km] * dKc);
    const S dkPP6 = sc<S>(0.5) / dmx3[kp] * (dx3[kp] * dKc + dx3[k] * dKkp);

    const S viscous_e1 = (kPP2 * dedx2 - kPP1 * dedx1 - e2 * dkPP2 * dedx2 + e1 * dkPP1 * dedx1) / dx1[i];
    const S viscous_e2 = (kPP4 * dedy4 - kPP3 * dedy3 - e4 * dkPP4 * dedy4 + e3 * dkPP3 * dedy3) / dx2[j];
    const S viscous_e3 = (kPP6 * dedz6 - kPP5 * dedz5 - e6 * dkPP6 * dedz6 + e5 * dkPP5 * dedz5) / dx3[k];

    rhs = rhs
        + invRhoCpCt_half * (viscous_e1 + viscous_e2 + viscous_e3)
        - dinvRhoCpCt_half * Tc * ((kPP2 * dedx2 - kPP1 * dedx1) / dx1[i]
                                   + (kPP4 * dedy4 - kPP3 * dedy3) / dx2[j]
                                   + (kPP6 * dedz6 - kPP5 * dedz5) / dx3[k]);

    // the wall terms are multiplied by 1 - jem and 1 - jep, zero away from the
    // walls (WM as in the momentum z rows: 0 computes them, 1 skips them
    // there, 2 leaves them out)
    if (WM == 0 || (WM == 1 && (jem != 1 || jep != 1))) {
    const S tup = TBCup[g.b2(i, k)], tbt = TBCbt[g.b2(i, k)];
    S ebc_up = -sc<S>(0.25) * Vjp / dmx2[jp] * tup;
    S ebc_down = sc<S>(0.25) * Vc / dmx2[j] * tbt;
    ebc_up = ebc_up
           + invRhoCpCt_half / dx2[j] * kPP4 / dmx2[jp] * tup
           + invRhoCpCt_half / dx2[j] * dkPP4 * dedy4 * tup;
    ebc_down = ebc_down
             + invRhoCpCt_half / dx2[j] * kPP3 / dmx2[j] * tbt
             - invRhoCpCt_half / dx2[j] * dkPP3 * dedy3 * tbt;
    rhs = rhs + sc<S>(1 - jem) * ebc_down + sc<S>(1 - jep) * ebc_up;
    }

    const S eACK = invRhoCpCt_half / dx3[k] * (kPP6 / dmx3[kp] + kPP5 / dmx3[k]
                                                    - dx3[kp] * (sc<S>(0.5) / dmx3[kp]) * dkPP6 * dedz6
                                                    + dx3[km] * (sc<S>(0.5) / dmx3[k]) * dkPP5 * dedz5)
                      + (-sc<S>(0.25) * Wkp / dmx3[kp] + sc<S>(0.25) * Wc / dmx3[k]);
    const S eAPK = -invRhoCpCt_half / dx3[k] * (kPP6 / dmx3[kp] + dx3[k] * (sc<S>(0.5) / dmx3[kp]) * dkPP6 * dedz6) + sc<S>(0.25) * Wkp / dmx3[kp];
    const S eAMK = -invRhoCpCt_half / dx3[k] * (kPP5 / dmx3[k] - dx3[k] * (sc<S>(0.5) / dmx3[k]) * dkPP5 * dedz5) - sc<S>(0.25) * Wc / dmx3[k];

    const S eACI = invRhoCpCt_half / dx1[i] * (kPP2 / dmx1[ip] + kPP1 / dmx1[i]
                                                    - dx1[ip] * (sc<S>(0.5) / dmx1[ip]) * dkPP2 * dedx2
                                                    + dx1[im] * (sc<S>(0.5) / dmx1[i]) * dkPP1 * dedx1)
                      + (sc<S>(0.25) * Uc / dmx1[i] - sc<S>(0.25) * Uip / dmx1[ip]);
    const S eAPI = -invRhoCpCt_half / dx1[i] * (kPP2 / dmx1[ip] + dx1[i] * (sc<S>(0.5) / dmx1[ip]) * dkPP2 * dedx2) + sc<S>(0.25) * Uip / dmx1[ip];
    const S eAMI = -invRhoCpCt_half / dx1[i] * (kPP1 / dmx1[i] - dx1[i] * (sc<S>(0.5) / dmx1[i]) * dkPP1 * dedx1) - sc<S>(0.25) * Uc / dmx1[i];

    const S eACJ = invRhoCpCt_half / dx2[j] * (kPP4 / dmx2[jp] + kPP3 / dmx2[j]
                                                    - dx2[jp] * (sc<S>(0.5) / dmx2[jp]) * dkPP4 * dedy4 * sc<S>(jep)
                                                    + dx2[jm] * (sc<S>(0.5) / dmx2[j]) * dkPP3 * dedy3 * sc<S>(jem))
                      + (sc<S>(0.25) * Vc / dmx2[j] - sc<S>(0.25) * Vjp / dmx2[jp]);
    S eAPJ = -invRhoCpCt_half / dx2[j] * (kPP4 / dmx2[jp] + dx2[j] * (sc<S>(0.5) / dmx2[jp]) * dkPP4 * dedy4) + sc<S>(0.25) * Vjp / dmx2[jp];
    S eAMJ = -invRhoCpCt_half / dx2[j] * (kPP3 / dmx2[j] - dx2[j] * (sc<S>(0.5) / dmx2[j]) * dkPP3 * dedy3) - sc<S>(0.25) * Vc / dmx2[j];
    eAPJ = eAPJ * sc<S>(jep);
    eAMJ = eAMJ * sc<S>(jem);

    rhs = rhs
        - (eAPK * Tkp + eACK * Tc + eAMK * Tkm
           + eAPJ * Tjp + eACJ * Tc + eAMJ * Tjm
           + eAPI * Tip + eACI * Tc + eAMI * Tim
           - dinvRhoCpCt_half * Tc * ((kPP2 * dedx2 - kPP1 * dedx1) / dx1[i]
                                      + (kPP4 * dedy4 - kPP3 * dedy3) / dx2[j]
                                      + (kPP6 * dedz6 - kPP5 * dedz5) / dx3[k]));

    ac = eACK * sc<S>(dt)
          - sc<S>(0.5) * dinvRhoCpCt_half * ((kPP2 * dedx2 - kPP1 * dedx1) / dx1[i]
                                      + (kPP4 * dedy4 - kPP3 * dedy3) / dx2[j]
                                      + (kPP6 * dedz6 - kPP5 * dedz5) / dx3[k]) * sc<S>(dt)
          + sc<S>(1.0);
    ap = eAPK * sc<S>(dt);
    am = eAMK * sc<S>(dt);
    rhsout = rhs * sc<S>(dt);
}
struct ThermalZ {
    // FP32 lines of the mixed form: 8 lines a workgroup at three waves per
    // SIMD (2.41 -> 2.33 ms at 512^3)
    static constexpr int W32 = 3;
    static constexpr bool FULL32 = false;
    static constexpr int SX32 = 8;
    // FP64 lines 8 a workgroup as well (3.08 -> 3.00 ms at 512^3; the
    // momentum z lines lose with 8: dU 2.49 -> 3.67, dW 2.71 -> 4.13 ms)
    static constexpr int SX64 = 8;
    Grid g;
    Phys p;
    double dt;
    const double *T, *U, *V, *W, *TBCup, *TBCbt;
    static constexpr int NF = 6;  // field inputs, in the order below
    static constexpr bool FULL = false;  // unconditional row loop in the FP64 z solve
    template <class S, int WM = -1, class Wr>
    __device__ __forceinline__ void eval(const Wr &w, int i, int j, int k, S &a, S &b, S &c, S &d) const {
        // WM < 0 (the row's own wall-term mode): a branch on j, skipping the
        // terms away from the walls (FP64 3.44 -> 3.23 ms, FP32 2.33 -> 2.06
        // ms at 512^3), terms kept in the adjoint's dual passes
        thermal_z_row<S, (WM >= 0 ? WM : (is_dual<S>() ? 0 : 1))>(g, p, dt, w(0, T), w(1, U), w(2, V), w(3, W), w(4, TBCup), w(5, TBCbt), i, j, k, a, b, c, d);
    }
    // no wall term of row j is weighted (the wall terms' flags are all 1)
    __device__ __forceinline__ bool wall_free(int j) const { return g.jC[j - 1] == 1 && g.jC[j + 1] == 1; }
    __device__ __forceinline__ void operator()(int i, int j, int k, double &a, double &b, double &c,
                                               double &d) const {
        eval<double>(RawW(), i, j, k, a, b, c, d);
    }
};

// stage 2: x (:388-429).  jm,jp,km,kp are the values the reference's loop
// variables keep after stage 1: n2m-1, n2m+1, n3m-1, n3m+1.
template <class S, class FP>
__device__ __forceinline__ void thermal_x_row(const Grid &g, const Phys &p, double dt,
    FP T, FP U,
    int i, int j, int k, S &am, S &ac, S &ap) {
    const int im = i - 1, ip = i + 1;
    const int jm = g.n2m - 1, jp = g.n2m + 1, km = g.n3m - 1, kp = g.n3m + 1;
    // the double tables rounded where read (the FP32 rows' faster schedule)
    const auto dx1 = metric<S, M_DX1, true>(g), dx2 = metric<S, M_DX2, true>(g), dx3 = metric<S, M_DX3, true>(g);
    const auto dmx1 = metric<S, M_DMX1, true>(g), dmx2 = metric<S, M_DMX2, true>(g), dmx3 = metric<S, M_DMX3, true>(g);

    const S Tc = T[g.ix(i, j, k)];
    const S Tim = T[g.ix(im, j, k)], Tip = T[g.ix(ip, j, k)];
    const S Tjm = T[g.ix(i, jm, k)], Tjp = T[g.ix(i, jp, k)];
    const S Tkm = T[g.ix(i, j, km)], Tkp = T[g.ix(i, j, kp)];
    S irc, dirc;
    th_rhocp(p, Tc, irc, dirc);
    S Kc, dKc, Kim, dKim, Kip, dKip;
    th_kpp(p, Tc, Kc, dKc);
    th_kpp(p, Tim, Kim, dKim);
    th_kpp(p, Tip, Kip, dKip);
    const S Kjm = th_kpp_only(p, Tjm), Kjp = th_kpp_only(p, Tjp);
    const S Kkm = th_kpp_only(p, Tkm), Kkp = th_kpp_only(p, Tkp);

    const S invRhoCpCt_half = sc<S>(0.5) * sc<S>(p.Ct) * irc;
    const S kPP1 = sc<S>(0.5) / dmx1[i] * (dx1[i] * Kim + dx1[im] * Kc);
    const S kPP2 = sc<S>(0.5) / dmx1[ip] * (dx1[ip] * Kc + dx1[i] * Kip);
    const S kPP3 = sc<S>(0.5) / dmx2[j] * (dx2[j] * Kjm + dx2[jm] * Kc);
    const S kPP4 = sc<S>(0.5) / dmx2[jp] * (dx2[jp] * Kc + dx2[j] * Kjp);
    const S kPP5 = sc<S>(0.5) / dmx3[k] * (dx3[k] * Kkm + dx3[km] * Kc);
    const S kPP6 = sc<S>(0.5) / dmx3[kp] * (dx3[kp] * Kc + dx3[k] * Kkp);
    const S dkPP1 = sc<S>(0.5) / dmx1[i] * (dx1[i] * dKim + dx1[im] * dKc);
    const S dkPP2 = sc<S>(0.5) / dmx1[ip] * (dx1[ip] * dKc + dx1[i] * dKip);
    const S dedx1 = (Tc - Tim) / dmx1[i];
    const S dedx2 = (Tip - Tc) / dmx1[ip];
    const S dedy3 = (Tc - Tjm) / dmx2[j];
    const S dedy4 = (Tjp - Tc) / dmx2[jp];
    const S dedz5 = (Tc - Tkm) / dmx3[k];
    const S dedz6 = (Tkp - Tc) / dmx3[kp];
    const S Uc = U[g.ix(i, j, k)], Uip = U[g.ix(ip, j, k)];

    const S eACI = invRhoCpCt_half / dx1[i] * (kPP2 / dmx1[ip] + kPP1 / dmx1[i]
                                                    - dx1[ip] * (sc<S>(0.5) / dmx1[ip]) * dkPP2 * dedx2
                                                    + dx1[im] * (sc<S>(0.5) / dmx1[i]) * dkPP1 * dedx1)
                      + (sc<S>(0.25) * Uc / dmx1[i] - sc<S>(0.25) * Uip / dmx1[ip])
                      - sc<S>(0.5) * invRhoCpCt_half * ((kPP2 * dedx2 - kPP1 * dedx1) / dx1[i]
                                                 + (kPP4 * dedy4 - kPP3 * dedy3) / dx2[j]
                                                 + (kPP6 * dedz6 - kPP5 * dedz5) / dx3[k]);
    const S eAPI = -invRhoCpCt_half / dx1[i] * (kPP2 / dmx1[ip] + dx1[i] * (sc<S>(0.5) / dmx1[ip]) * dkPP2 * dedx2) + sc<S>(0.25) * Uip / dmx1[ip];
    const S eAMI = -invRhoCpCt_half / dx1[i] * (kPP1 / dmx1[i] - dx1[i] * (sc<S>(0.5) / dmx1[i]) * dkPP1 * dedx1) - sc<S>(0.25) * Uc / dmx1[i];
    ac = eACI * sc<S>(dt) + sc<S>(1.0);
    ap = eAPI * sc<S>(dt);
    am = eAMI * sc<S>(dt);
}
struct ThermalX {
    Grid g;
    Phys p;
    double dt;
    const double *T, *U;
    static constexpr int NF = 2;  // field inputs, in the order below
    template <class S, class Wr>
    __device__ __forceinline__ void eval(const Wr &w, int i, int j, int k, S &a, S &b, S &c, S &d) const {
        thermal_x_row<S>(g, p, dt, w(0, T), w(1, U), i, j, k, a, b, c);
    }
    __device__ __forceinline__ void operator()(int i, int j, int k, double &a, double &b, double &c,
                                               double &d) const {
        eval<double>(RawW(), i, j, k, a, b, c, d);
    }
};

// stage 3: y (:456-507); km,kp keep their stage-1 values n3m-1, n3m+1.
template <class S, class FP>
__device__ __forceinline__ void thermal_y_row(const Grid &g, const Phys &p, double dt,
    FP T, FP V,
    int i, int j, int k, S &am, S &ac, S &ap) {
    const int im = i - 1, ip = i + 1, jm = j - 1, jp = j + 1;
    const int km = g.n3m - 1, kp = g.n3m + 1;
    const int jem = g.jC[jm], jep = g.jC[jp];
    const auto dx1 = metric<S, M_DX1>(g), dx2 = metric<S, M_DX2>(g), dx3 = metric<S, M_DX3>(g);
    const auto dmx1 = metric<S, M_DMX1>(g), dmx2 = metric<S, M_DMX2>(g), dmx3 = metric<S, M_DMX3>(g);

    const S Tc = T[g.ix(i, j, k)];
    const S Tim = T[g.ix(im, j, k)], Tip = T[g.ix(ip, j, k)];
    const S Tjm = T[g.ix(i, jm, k)], Tjp = T[g.ix(i, jp, k)];
    const S Tkm = T[g.ix(i, j, km)], Tkp = T[g.ix(i, j, kp)];
    S irc, dirc;
    th_rhocp(p, Tc, irc, dirc);
    S Kc, dKc, Kjm, dKjm, Kjp, dKjp;
    th_kpp(p, Tc, Kc, dKc);
    th_kpp(p, Tjm, Kjm, dKjm);
    th_kpp(p, Tjp, Kjp, dKjp);
    const S Kim = th_kpp_only(p, Tim), Kip = th_kpp_only(p, Tip);
    const S Kkm = th_kpp_only(p, Tkm), Kkp = th_kpp_only(p, Tkp);

    const S invRhoCpCt_half = sc<S>(0.5) * sc<S>(p.Ct) * irc;
    const S kPP1 = sc<S>(0.5) / dmx1[i] * (dx1[i] * Kim + dx1[im] * Kc);
    const S kPP2 = sc<S>(0.5) / dmx1[ip] * (dx1[ip] * Kc + dx1[i] * Kip);
    const S kPP3 = sc<S>(0.5) / dmx2[j] * (dx2[j] * Kjm + dx2[jm] * Kc);
    const S kPP4 = sc<S>(0.5) / dmx2[jp] * (dx2[jp] * Kc + dx2[j] * Kjp);
    const S kPP5 = sc<S>(0.5) / dmx3[k] * (dx3[k] * Kkm + dx3[km] * Kc);
    const S kPP6 = sc<S>(0.5) / dmx3[kp] * (dx3[kp] * Kc + dx3[k] * Kkp);
    const S dkPP3 = sc<S>(0.5) / dmx2[j] * (dx2[j] * dKjm + dx2[jm] * dKc);
    const S dkPP4 = sc<S>(0.5) / dmx2[jp] * (dx2[jp] * dKc + dx2[j] * dKjp);
    const S dedx1 = (Tc - Tim) / dmx1[i];
    const S dedx2 = (Tip - Tc) / dmx1[ip];
    const S dedy3 = (Tc - Tjm) / dmx2[j];
    const S dedy4 = (Tjp - Tc) / dmx2[jp];
    const S dedz5 = (Tc - Tkm) / dmx3[k];
    const S dedz6 = (Tkp - Tc) / dmx3[kp];
    const S Vc = V[g.ix(i, j, k)], Vjp = V[g.ix(i, jp, k)];

    const S eACJ = invRhoCpCt_half / dx2[j] * (kPP4 / dmx2[jp] + kPP3 / dmx2[j]
                                                    - dx2[jp] * (sc<S>(0.5) / dmx2[jp]) * dkPP4 * dedy4 * sc<S>(jep)
                                                    + dx2[jm] * (sc<S>(0.5) / dmx2[j]) * dkPP3 * dedy3 * sc<S>(jem))
                      + (sc<S>(0.25) * Vc / dmx2[j] - sc<S>(0.25) * Vjp / dmx2[jp])
                      - sc<S>(0.5) * invRhoCpCt_half * ((kPP2 * dedx2 - kPP1 * dedx1) / dx1[i]
                                                 + (kPP4 * dedy4 - kPP3 * dedy3) / dx2[j]
                                                 + (kPP6 * dedz6 - kPP5 * dedz5) / dx3[k]);
    S eAPJ = -invRhoCpCt_half / dx2[j] * (kPP4 / dmx2[jp] + dx2[j] * (sc<S>(0.5) / dmx2[jp]) * dkPP4 * dedy4) + sc<S>(0.25) * Vjp / dmx2[jp];
    S eAMJ = -invRhoCpCt_half / dx2[j] * (kPP3 / dmx2[j] - dx2[j] * (sc<S>(0.5) / dmx2[j]) * dkPP3 * dedy3) - sc<S>(0.25) * Vc / dmx2[j];
    eAPJ = eAPJ * sc<S>(jep);
    eAMJ = eAMJ * sc<S>(jem);
    ac = eACJ * sc<S>(dt) + sc<S>(1.0);
    ap = eAPJ * sc<S>(dt);
    am = eAMJ * sc<S>(dt);
}
struct ThermalY {
    // FP32 lines of the mixed form at four waves per SIMD (2.54 -> 1.84 ms at
    // 512^3; the other stages lose or gain nothing with it)
    static constexpr int W32 = 4;
    static constexpr bool FULL32 = false;
    Grid g;
    Phys p;
    double dt;
    const double *T, *V;
    static constexpr int NF = 2;  // field inputs, in the order below
    template <class S, class Wr>
    __device__ __forceinline__ void eval(const Wr &w, int i, int j, int k, S &a, S &b, S &c, S &d) const {
        thermal_y_row<S>(g, p, dt, w(0, T), w(1, V), i, j, k, a, b, c);
    }
    __device__ __forceinline__ void operator()(int i, int j, int k, double &a, double &b, double &c,
                                               double &d) const {
        eval<double>(RawW(), i, j, k, a, b, c, d);
    }
};

// T(i,j,k) += RHSJ (:518-524)
// F(i,j,k) = inc (MODE 1) or F(i,j,k) += inc (MODE 2) over interior rows j >= j0
template <int MODE>
__global__ void k_interior_out(Grid g, int j0, double *__restrict__ F, const scr_t *__restrict__ inc) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.n1m || j > g.n2m || j < j0) return;
    const size_t q = g.ix(i, j, k);
    if (MODE == 1) F[q] = inc[g.r3(i, j, k)];
    else F[q] = F[q] + inc[g.r3(i, j, k)];
}

// y sweep of rows j = j0..j0+nrows-1 (1-based) whose solution ends the
// stage: stored into (mode 1) or added to (mode 2) the field F's interior.
// With split lines the solve writes F directly; otherwise a second kernel
// moves the solution from the scratch.
int sweep_y_field(tcs_ctx *c, int j0, int nrows, double *F, int mode) {
    const long n1 = c->n1m, n12 = (long)c->n1m * c->n2m;
    const long off = (long)(j0 - 1) * n1;
    int P, w1, w2;
    if (c->tdma_split && ptdma_part_plan(nrows, P, w1, w2)) {
        LineOut out;
        out.F = F;
        out.mode = mode;
        out.base = 1 + (long)c->P1 * j0 + (long)c->P1 * c->N2;  // (i,j,k) = (1, j0, 1)
        out.ss0 = 1;
        out.ss1 = (long)c->P1 * c->N2;
        out.rs = c->P1;
        return ptdma_part_launch(SCR(c->am) + off, SCR(c->ac) + off, SCR(c->ap) + off, SCR(c->rhs) + off, c->n1m, 1,
                                 c->n3m, n12, nrows, n1, false, c->stream, out);
    }
    TCS_TRY(ptdma_sweep(c, SCR(c->am) + off, SCR(c->ac) + off, SCR(c->ap) + off, SCR(c->rhs) + off, c->n1m, 1,
                        c->n3m, n12, nrows, n1, false));
    const dim3 blk = interior_block(), grd = interior_grid(c->n1m, c->n2m, c->n3m);
    if (mode == 1)
        TCS_LAUNCH(c->stream, "store_interior", k_interior_out<1>, grd, blk, 0, c->grid(), j0, F, SCR(c->rhs));
    else
        TCS_LAUNCH(c->stream, "add_interior", k_interior_out<2>, grd, blk, 0, c->grid(), j0, F, SCR(c->rhs));
    return TCS_OK;
}

int thermal_coeffi(tcs_ctx *c) {
    TCS_LAUNCH(c->stream, "thermal_coeffi", k_thermal_coeffi, dim3(2048), dim3(256), 0, c->grid(), c->phys,
                       c->f[TCS_T], c->f[TCS_KPP], c->f[TCS_DKPP], c->f[TCS_INVRHOCP], c->f[TCS_DINVRHOCP]);
    return TCS_OK;
}

int thermal_solver(tcs_ctx *c, double dt) {
    TCS_TRY(ensure_twins(c, TWIN_T | TWIN_U | TWIN_V | TWIN_W | TWIN_WALL));  // the mixed form's FP32 rows
    c->twin_ok &= ~(unsigned)TWIN_T;
    const Grid g = c->grid();
    double *T = c->f[TCS_T];
    scr_t *am = SCR(c->am), *ac = SCR(c->ac), *ap = SCR(c->ap), *rhs = SCR(c->rhs);
    const ThermalZ z{g, c->phys, dt, T, c->f[TCS_U], c->f[TCS_V], c->f[TCS_W], c->wall[TCS_TBCUP], c->wall[TCS_TBCBT]};
    TCS_TRY(z_stage(c, "thermal_z", "thermal_z_solve", z, 1, am, ac, ap, rhs));
    // The x rows evaluate seven property polynomials each: made by the
    // workgroup inside the contiguous-line solve they leave it waiting on
    // them, so here they go through the scratch as a kernel of their own.
    TCS_TRY(rows3d_launch<false>(c, "thermal_x", ThermalX{g, c->phys, dt, T, c->f[TCS_U]}, 1, am, ac, ap, rhs));
    TCS_TRY(ptdma_sweep(c, am, ac, ap, rhs, c->n2m, c->n1m, c->n3m, (long)c->n1m * c->n2m, c->n1m, 1, true));
    const ThermalY y{g, c->phys, dt, T, c->f[TCS_V]};
    if (rows_in_solve<scr_t>(c, c->n2m)) {
        // the y rows read T around every cell, so the increment goes through
        // the scratch and is added once every line is solved
        TCS_TRY(y_stage_in_solve(c, "thermal_y_solve", y, 1, c->n2m, rhs, nullptr, 0));
        TCS_LAUNCH(c->stream, "add_interior", k_interior_out<2>, interior_grid(c->n1m, c->n2m, c->n3m),
                   interior_block(), 0, g, 1, T, rhs);
        return TCS_OK;
    }
    TCS_TRY(rows3d_launch<false>(c, "thermal_y", y, 1, am, ac, ap, rhs));
    return sweep_y_field(c, 1, c->n2m, T, 2);
}

// thermal_solver whose last sweep writes T + dT into the second buffer Tn
// (with the cells outside the interior copied) when the y lines are split;
// otherwise the plain routine (*written = false, T updated in place).
int thermal_solver_into(tcs_ctx *c, double dt, double *Tn, bool *written) {
    *written = false;
    if (!rows_in_solve<scr_t>(c, c->n2m)) return thermal_solver(c, dt);
    TCS_TRY(ensure_twins(c, TWIN_T | TWIN_U | TWIN_V | TWIN_W | TWIN_WALL));  // the mixed form's FP32 rows
    c->twin_ok &= ~(unsigned)TWIN_T;
    const Grid g = c->grid();
    double *T = c->f[TCS_T];
    scr_t *am = SCR(c->am), *ac = SCR(c->ac), *ap = SCR(c->ap), *rhs = SCR(c->rhs);
    const ThermalZ z{g, c->phys, dt, T, c->f[TCS_U], c->f[TCS_V], c->f[TCS_W], c->wall[TCS_TBCUP], c->wall[TCS_TBCBT]};
    TCS_TRY(z_stage(c, "thermal_z", "thermal_z_solve", z, 1, am, ac, ap, rhs));
    TCS_TRY(rows3d_launch<false>(c, "thermal_x", ThermalX{g, c->phys, dt, T, c->f[TCS_U]}, 1, am, ac, ap, rhs));
    TCS_TRY(ptdma_sweep(c, am, ac, ap, rhs, c->n2m, c->n1m, c->n3m, (long)c->n1m * c->n2m, c->n1m, 1, true));
    TCS_TRY(copy_shell(c, T, Tn));
    const long n1 = c->n1m, n12 = (long)c->n1m * c->n2m;
    LineOut out;
    out.F = Tn;
    out.Fin = T;
    out.G = twin(c, Tn);  // the mixed form: T's FP32 copy of the updated cells
    out.mode = 3;
    out.base = 1 + (long)c->P1 + (long)c->P1 * c->N2;  // (i, j, k) = (1, 1, 1)
    out.ss0 = 1;
    out.ss1 = (long)c->P1 * c->N2;
    out.rs = c->P1;
    // FP64 lines 8 a workgroup (1.96 -> 1.91 ms at 512^3), FP32 lines 16
    // (8: 1.52 -> 1.96 ms)
    constexpr int SX = sizeof(scr_t) == 8 ? 8 : PT_SX;
    const int rc = ptdma_part_src_launch<false, SX>(c->stream, "thermal_y_solve",
                                                    YLineRows<ThermalY, scr_t>{ThermalY{g, c->phys, dt, T, c->f[TCS_V]},
                                                                               twin_w(c), rhs, 1},
                                                    rhs, c->n1m, c->n3m, n12, c->n2m, n1, out);
    if (rc != TCS_OK) return rc;
    *written = true;
    return TCS_OK;
}

}  // namespace TCS_FORM
}  // namespace tcs
