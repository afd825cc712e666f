// pressure.hip -- mpi_pressure_RHS / _Poisson_FFT2 / _Projection
// (src/module_solve_pressure.f90:140-231, :489-744, :756-874) on one GPU.
//
// The reference moves the RHS through six MPI_Alltoallw pencil transposes
// around FFTW calls.  Here the spectrum lives in one device buffer laid out
// [j][kx][k] (k fastest), the layout of the reference's RHSIKhat_Kline after
// its z transform, so:
//   * the x r2c kernel transforms TX consecutive k-lines of one j per
//     workgroup in LDS and writes each wavenumber row as TX contiguous
//     complex values;
//   * the z FFTs run along the contiguous axis in place;
//   * the y tridiagonal solve takes one (kx,kz) system per thread with
//     consecutive threads on consecutive kz, its coefficients generated from
//     1-D tables (no coefficient arrays are stored), real and imaginary
//     parts sharing one elimination (same matrix, so identical to solving
//     them in two calls);
//   * the inverse x c2r kernel mirrors the forward one and also produces
//     the per-line partial sums for the mean.
// FFTs are radix-2 Stockham with a host twiddle table, computed in place in
// LDS with the butterfly outputs staged in registers across barriers.
#include "ad.h"
#include "tdma_part.h"

#ifndef TCS_FORM
#define TCS_FORM exact
#endif

namespace tcs {
namespace TCS_FORM {

#include "metric.h"

// butterflies per thread per stage: 8 for the z transforms and the inverse
// x transform, 4 for the forward x transform (one radix-8 group a thread per
// pass at 512 points: fewer registers while its right-hand side's stencil
// loads are in flight, 2.42 -> 1.85 ms at 512^3; the other transforms lost
// 0.05-0.08 ms each with 4)
constexpr int FFT_MAXB = 8, FFT_MAXB_FWD = 4;

// In-place radix-2 Stockham on nl lines (re/im at [l*n + t]) in LDS by NT
// threads, n a power of two, nl*n <= 2*MB*NT (MB butterflies per thread).  sign<0 forward
// (exp(-i...)), sign>0 backward.  Caller synchronises before.
//
// Stage s (Ns = 2^s) combines x[jb] and x[jb + n/2] with the twiddle
// tw[(jb mod Ns) * n/(2Ns)] into positions 2Ns*floor(jb/Ns) + jb mod Ns
// (+Ns).  Three consecutive stages s, s+1, s+2 only mix the 8 values
// x[t*n/8 + g], t = 0..7, of a group g, so each thread runs them in
// registers and touches LDS once per three stages: the same butterflies with
// the same operands as the stage-by-stage loop, so the same results bit for
// bit.  The group's outputs land at 8(g - kk) + kk + r*Ns, kk = g mod Ns,
// with r the three result bits (first stage's bit lowest).
// LDS position of transform value x: one pad value per 32, so the strided
// stores of the first passes (8 and 64 values apart) spread over the banks
__device__ __forceinline__ int fpad(int x) { return x + (x >> 5); }
// LDS values that hold `count` padded transform values
__host__ __device__ constexpr int fft_lds_len(int count) { return count + (count >> 5); }

template <int SIGN, int R, int NT, int MB>
__device__ __forceinline__ void fft_pass(double *re, double *im, int n, int nl, int s,
                                         const double2 *__restrict__ tw) {
    constexpr int M = 1 << R;
    constexpr int GMAX = 2 * MB / M;  // groups per thread
    const int G = n >> R;
    const int total = G * nl;
    const int Ns = 1 << s;
    const int tid = threadIdx.x;
    double xr[GMAX][M], xi[GMAX][M];
    int obase[GMAX];
#pragma unroll
    for (int u = 0; u < GMAX; u++) {
        const int e = tid + u * NT;
        obase[u] = -1;
        if (e < total) {
            const int l = e / G;
            const int g = e - l * G;
            const int kk = g & (Ns - 1);
#pragma unroll
            for (int t = 0; t < M; t++) {
                xr[u][t] = re[fpad(l * n + t * G + g)];
                xi[u][t] = im[fpad(l * n + t * G + g)];
            }
#pragma unroll
            for (int i = 0; i < R; i++) {
                double yr[M], yi[M];
#pragma unroll
                for (int a = 0; a < M / 2; a++) {
                    // result bits r0..r_{i-1} of the pair: the low i bits of a, reversed
                    int rs = 0;
#pragma unroll
                    for (int m = 0; m < i; m++) rs |= ((a >> (i - 1 - m)) & 1) << m;
                    const double2 w = tw[(kk + rs * Ns) * (n >> (s + i + 1))];
                    const double ar = xr[u][a], ai = xi[u][a];
                    const double br = xr[u][a + M / 2], bi = xi[u][a + M / 2];
                    double tr, ti;
                    if (SIGN < 0) {
                        tr = br * w.x - bi * w.y;
                        ti = br * w.y + bi * w.x;
                    } else {
                        tr = br * w.x + bi * w.y;
                        ti = bi * w.x - br * w.y;
                    }
                    yr[2 * a] = ar + tr; yi[2 * a] = ai + ti;
                    yr[2 * a + 1] = ar - tr; yi[2 * a + 1] = ai - ti;
                }
#pragma unroll
                for (int t = 0; t < M; t++) { xr[u][t] = yr[t]; xi[u][t] = yi[t]; }
            }
            obase[u] = l * n + (g - kk) * M + kk;
        }
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < GMAX; u++) {
        if (obase[u] >= 0) {
#pragma unroll
            for (int t = 0; t < M; t++) {
                int r = 0;  // register index t holds result bits (r0..r_{R-1}) with r0 highest
#pragma unroll
                for (int m = 0; m < R; m++) r |= ((t >> (R - 1 - m)) & 1) << m;
                re[fpad(obase[u] + r * Ns)] = xr[u][t];
                im[fpad(obase[u] + r * Ns)] = xi[u][t];
            }
        }
    }
    __syncthreads();
}

template <int SIGN, int NT, int MB = FFT_MAXB>
__device__ __forceinline__ void lds_fft(double *re, double *im, int n, int nl, const double2 *__restrict__ tw) {
    int L = 0;
    while ((1 << L) < n) L++;
    int s = 0;
    for (; s + 3 <= L; s += 3) fft_pass<SIGN, 3, NT, MB>(re, im, n, nl, s, tw);
    if (L - s == 2) fft_pass<SIGN, 2, NT, MB>(re, im, n, nl, s, tw);
    else if (L - s == 1) fft_pass<SIGN, 1, NT, MB>(re, im, n, nl, s, tw);
}

// PRHS (:182-229) into the interior layout.  The wall term cbc is
// multiplied by 1 - jvm and 1 - jvp, zero away from the walls: with SKIPW
// the cells there skip it and its two wall-array loads (DivUm - 0 = DivUm;
// 1.85 -> 1.63 ms for the right-hand side and forward x FFT at 512^3).
template <class S, bool SKIPW = false, class FP>
__device__ __forceinline__ void pressure_rhs_cell(const Grid &g, const Phys &p, double dt, FP U, FP V, FP W,
                                                 FP dPhat, FP iR, FP VBCup, FP VBCbt, int i, int j, int k,
                                                 S &out) {
    const auto dx1 = metric<S, M_DX1>(g), dx2 = metric<S, M_DX2>(g), dx3 = metric<S, M_DX3>(g);
    const auto dmx1 = metric<S, M_DMX1>(g), dmx2 = metric<S, M_DMX2>(g), dmx3 = metric<S, M_DMX3>(g);
    const int ip = i + 1, im = i - 1, jp = j + 1, jm = j - 1, kp = k + 1, km = k - 1;
    const int jvm = g.jC[jm], jvp = g.jC[jp];
#define IRt(jj, kk) iR[g.t2((jj), (kk))]
#define dPh(a, b, c) dPhat[g.ix(a, b, c)]
    const S DivUm = (U[g.ix(ip, j, k)] - U[g.ix(i, j, k)]) / dx1[i]
                       + ((double)jvp * V[g.ix(i, jp, k)] - (double)jvm * V[g.ix(i, j, k)]) / dx2[j]
                       + (W[g.ix(i, j, kp)] - W[g.ix(i, j, k)]) / dx3[k];
    S div = DivUm;
    if (!SKIPW || jvm != 1 || jvp != 1) {
        const S cbc = (1.0 - (double)jvm) * VBCbt[g.b2(i, k)] / dx2[j]
                         - (1.0 - (double)jvp) * VBCup[g.b2(i, k)] / dx2[j];
        div = DivUm - cbc;
    }
    const S ddpdx1 = (dPh(i, j, k) - dPh(im, j, k)) / dmx1[i];
    const S ddpdx2 = (dPh(ip, j, k) - dPh(i, j, k)) / dmx1[ip];
    const S ddpdy3 = (dPh(i, j, k) - dPh(i, jm, k)) / dmx2[j];
    const S ddpdy4 = (dPh(i, jp, k) - dPh(i, j, k)) / dmx2[jp];
    const S ddpdz5 = (dPh(i, j, k) - dPh(i, j, km)) / dmx3[k];
    const S ddpdz6 = (dPh(i, j, kp) - dPh(i, j, k)) / dmx3[kp];
    const S invRho1 = 0.5 / dmx1[i] * (dx1[im] * IRt(j, k) + dx1[i] * IRt(j, k));
    const S invRho2 = 0.5 / dmx1[ip] * (dx1[ip] * IRt(j, k) + dx1[i] * IRt(j, k));
    const S invRho3 = 0.5 / dmx2[j] * (dx2[jm] * IRt(j, k) + dx2[j] * IRt(jm, k));
    const S invRho4 = 0.5 / dmx2[jp] * (dx2[jp] * IRt(j, k) + dx2[j] * IRt(jp, k));
    const S invRho5 = 0.5 / dmx3[k] * (dx3[km] * IRt(j, k) + dx3[k] * IRt(j, km));
    const S invRho6 = 0.5 / dmx3[kp] * (dx3[kp] * IRt(j, k) + dx3[k] * IRt(j, kp));
    const S ExtraTerm = ((1.0 - invRho2) * ddpdx2 - (1.0 - invRho1) * ddpdx1) / dx1[i]
                           + ((1.0 - invRho4) * ddpdy4 - (1.0 - invRho3) * ddpdy3) / dx2[j]
                           + ((1.0 - invRho6) * ddpdz6 - (1.0 - invRho5) * ddpdz5) / dx3[k];
    out = div / dt / p.Cmp + ExtraTerm;
}
struct PressureRHS {
    Grid g;
    Phys p;
    double dt;
    const double *U, *V, *W, *dPhat, *iR, *VBCup, *VBCbt;
    static constexpr int NF = 7;
    template <class S, class Wr>
    __device__ __forceinline__ void eval(const Wr &w, int i, int j, int k, S &a, S &b, S &c, S &d) const {
        pressure_rhs_cell<S, !is_dual<S>()>(g, p, dt, w(0, U), w(1, V), w(2, W), w(3, dPhat), w(4, iR), w(5, VBCup), w(6, VBCbt), i,
                             j, k, a);
    }
};

// planes per thread (measured at 512^3: 4 for the right-hand side, 2 for the
// projection) and consecutive cells in i per thread of the projection (1:
// with 2 it took 3.80 ms against 3.34)
constexpr int RHS_KB = 4, PROJ_KB = 2, PROJ_IB = 1;

__global__ void __launch_bounds__(256) k_pressure_rhs(PressureRHS cell, double *__restrict__ out) {
    const Grid &g = cell.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    if (i > g.n1m || j > g.n2m) return;
#pragma unroll
    for (int q = 0; q < RHS_KB; q++) {
        const int k = blockIdx.z * RHS_KB + q + 1;
        double o, x;
        cell.eval<double>(RawW(), i, j, k, o, x, x, x);
        out[g.r3(i, j, k)] = o;
    }
}

// Sources of the x r2c: the right-hand side read from the interior-layout
// array, or made where it is needed by the right-hand side's cell formula
// (no round trip through memory)
struct RhsArray {
    Grid g;
    const double *prhs;
    __device__ __forceinline__ double operator()(int i, int j, int k) const { return prhs[g.r3(i, j, k)]; }
};
struct RhsCell {
    PressureRHS cell;
    __device__ __forceinline__ double operator()(int i, int j, int k) const {
        double o, x;
        cell.eval<double>(RawW(), i, j, k, o, x, x, x);
        return o;
    }
};

// x r2c: TX k-lines (k0..k0+TX-1) of row j per workgroup.
// spec[(k + n3m*(kx + h1p*(j-1)))] complex, k 0-based.
template <int NT, class Src>
__global__ void __launch_bounds__(NT)
k_fft_x_fwd(Grid g, int h1p, int TX, Src src, const double2 *__restrict__ tw, double2 *__restrict__ spec) {
    extern __shared__ double lds[];
    const int n = g.n1m;
    double *re = lds, *im = lds + fft_lds_len(TX * n);
    int kb, jb;
    xcd_tiles(kb, jb);  // neighbouring rows j on one XCD: the right-hand side's stencil
    const int j = jb + 1;
    const int k0 = kb * TX;  // 0-based
    const int tid = threadIdx.x;
    if (TX == 4 && n == 2 * NT) {
        // the 512-point shape: a thread's two x positions one after the
        // other, each over the four k-lines, so the k-lines share their
        // loads (both positions at once hold more registers: 145 against
        // 121, 2.42 against 1.85 ms with the transform's 4 butterflies)
#pragma unroll 1
        for (int h = 0; h < 2; h++) {
#pragma unroll
            for (int l = 0; l < 4; l++) {
                const int t = tid + h * NT;
                re[fpad(l * n + t)] = src(t + 1, j, k0 + l + 1);
                im[fpad(l * n + t)] = 0.0;
            }
        }
    } else {
        for (int e = tid; e < TX * n; e += NT) {
            const int l = e / n, t = e - l * n;
            re[fpad(e)] = src(t + 1, j, k0 + l + 1);
            im[fpad(e)] = 0.0;
        }
    }
    __syncthreads();
    lds_fft<-1, NT, FFT_MAXB_FWD>(re, im, n, TX, tw);
    const size_t rowbase = (size_t)g.n3m * h1p * (j - 1);
    for (int e = tid; e < TX * h1p; e += NT) {
        const int kx = e / TX, l = e - kx * TX;
        spec[rowbase + (size_t)g.n3m * kx + k0 + l] = make_double2(re[fpad(l * n + kx)], im[fpad(l * n + kx)]);
    }
}

// z c2c in place along the contiguous axis: one (kx, j) line per workgroup
template <int SIGN, int NT>
__global__ void __launch_bounds__(NT)
k_fft_z(int n3m, int h1p, const double2 *__restrict__ tw, double2 *__restrict__ spec) {
    extern __shared__ double lds[];
    double *re = lds, *im = lds + fft_lds_len(n3m);
    const size_t base = (size_t)n3m * ((size_t)blockIdx.x + (size_t)h1p * blockIdx.y);
    const int tid = threadIdx.x;
    for (int t = tid; t < n3m; t += NT) {
        const double2 v = spec[base + t];
        re[fpad(t)] = v.x; im[fpad(t)] = v.y;
    }
    __syncthreads();
    lds_fft<SIGN, NT>(re, im, n3m, 1, tw);
    for (int t = tid; t < n3m; t += NT) spec[base + t] = make_double2(re[fpad(t)], im[fpad(t)]);
}

// y tridiagonal solves per (kx,kz), PaScaL single-rank order (:599-652).
// system s = kz + n3m*kx, row j stride n3m*h1p.  amj/apj/acj: 1-D tables.
__global__ void __launch_bounds__(256)
k_poisson_tdma(int n3m, int h1p, int n2m, int h3p, const double *__restrict__ amj, const double *__restrict__ apj,
               const double *__restrict__ acj, const double *__restrict__ dxk2, const double *__restrict__ dzk,
               double2 *__restrict__ S, double *__restrict__ pA, double *__restrict__ pC) {
    const int s = blockIdx.x * blockDim.x + threadIdx.x;
    const int nsys = n3m * h1p;
    if (s >= nsys) return;
    const int kx = s / n3m, kz = s - kx * n3m;
    const size_t rs = (size_t)nsys;
    double2 *d = S + s;
    double *a = pA + s, *c = pC + s;
    const double shift = dxk2[kx];
    const double zk = dzk[kz];
    const bool pin = (kx == 0 && kz == 0);
    // row coefficients: A = amj, B = (acj - dxk2) - dzk, C = apj; (1,1,1) pinned
    auto coef = [&](int jj, double &A, double &B, double &C) {
        A = amj[jj]; C = apj[jj]; B = acj[jj] - shift - zk;
        if (pin && jj == 0) { A = 0.0; B = 1.0; C = 0.0; }
    };
    double A0, B0, C0, A1, B1, C1;
    coef(0, A0, B0, C0);
    coef(1, A1, B1, C1);
    double2 D0 = d[0], D1 = d[rs];
    if (pin) D0.x = 0.0;
    double a0 = A0 / B0;
    double dr0 = D0.x / B0, di0 = D0.y / B0;
    double c0 = C0 / B0;
    const double a1 = A1 / B1;
    const double dr1 = D1.x / B1, di1 = D1.y / B1;
    const double c1 = C1 / B1;
    a[rs] = a1; c[rs] = c1; d[rs] = make_double2(dr1, di1);
    double ap = a1, cp = c1, drp = dr1, dip = di1;
    for (int jj = 2; jj < n2m; jj++) {
        double A, B, C;
        coef(jj, A, B, C);
        const double2 Dj = d[(size_t)jj * rs];
        const double r = 1.0 / (B - A * cp);
        const double dr = r * (Dj.x - A * drp);
        const double di = r * (Dj.y - A * dip);
        const double cn = r * C;
        const double an = -r * A * ap;
        a[(size_t)jj * rs] = an; c[(size_t)jj * rs] = cn; d[(size_t)jj * rs] = make_double2(dr, di);
        ap = an; cp = cn; drp = dr; dip = di;
    }
    const double aN = ap, drN = drp, diN = dip;
    double ax = a1, cx = c1, drx = dr1, dix = di1;
    if (n2m >= 3) {
        const size_t o2 = (size_t)(n2m - 2) * rs;
        ax = a[o2]; cx = c[o2];
        const double2 v = d[o2];
        drx = v.x; dix = v.y;
        for (int jj = n2m - 3; jj >= 1; jj--) {
            const size_t o = (size_t)jj * rs;
            const double aj = a[o], cj = c[o];
            const double2 dj = d[o];
            const double dr = dj.x - cj * drx;
            const double di = dj.y - cj * dix;
            const double an = aj - cj * ax;
            const double cn = -cj * cx;
            a[o] = an; c[o] = cn; d[o] = make_double2(dr, di);
            ax = an; cx = cn; drx = dr; dix = di;
        }
    }
    {
        const double r = 1.0 / (1.0 - ax * c0);
        dr0 = r * (dr0 - c0 * drx);
        di0 = r * (di0 - c0 * dix);
        a0 = r * a0;
        c0 = -r * c0 * cx;
    }
    const double rr = 1.0 / (1.0 - aN * c0);
    const double xnr = rr * (drN - aN * dr0);
    const double xni = rr * (diN - aN * di0);
    const double x1r = dr0 - c0 * xnr;
    const double x1i = di0 - c0 * xni;
    // imaginary parts zeroed at kx=0 for kz in {0, h3p-1} (:643-650)
    const bool zim = (kx == 0 && (kz == 0 || kz == h3p - 1));
    d[0] = make_double2(x1r, zim ? 0.0 : x1i);
    d[(size_t)(n2m - 1) * rs] = make_double2(xnr, zim ? 0.0 : xni);
    for (int jj = 1; jj < n2m - 1; jj++) {
        const size_t o = (size_t)jj * rs;
        const double aj = a[o], cj = c[o];
        const double2 dj = d[o];
        const double vr = dj.x - aj * x1r - cj * xnr;
        const double vi = dj.y - aj * x1i - cj * xni;
        d[o] = make_double2(vr, zim ? 0.0 : vi);
    }
}

// The same y solves with split lines (tdma_split): PaScaL_TDMA's multi-rank
// algorithm, one thread per part, reduced systems by parallel cyclic
// reduction (tdma_part.h).  Line s = kz + n3m*kx, row j stride n3m*h1p; the
// real and imaginary parts are two right-hand sides of one matrix (the
// reference solves them in two calls with equal coefficients).  The (0,0)
// system is pinned at j=1 as in the reference's real-part call; its
// imaginary part is zeroed afterwards like the reference's.
__global__ void __launch_bounds__(PT_SX * PT_PMAX)
k_poisson_part(int n3m, int h1p, int h3p, const double *__restrict__ amj, const double *__restrict__ apj,
               const double *__restrict__ acj, const double *__restrict__ dxk2, const double *__restrict__ dzk,
               double2 *__restrict__ S, int P, int w1, int w2) {
    __shared__ part::RedLDS<double, PT_SX, 2> red;
    const int t = threadIdx.x;
    const int line = t % PT_SX, p = t / PT_SX;
    const int nsys = n3m * h1p;
    const int s = blockIdx.x * PT_SX + line;
    const bool act = s < nsys;
    const int kx = act ? s / n3m : 0, kz = act ? s - kx * n3m : 0;
    const size_t rs = (size_t)nsys;
    int j0, m;
    part::range(p, w1, w2, j0, m);
    const double shift = dxk2[kx], zk = dzk[kz];
    const bool pin = act && kx == 0 && kz == 0;
    double a[PT_MMAX], c[PT_MMAX], d[2][PT_MMAX];
    {
        double b[PT_MMAX];
#pragma unroll
        for (int r = 0; r < PT_MMAX; r++) {
            a[r] = 0.0; b[r] = 1.0; c[r] = 0.0; d[0][r] = 0.0; d[1][r] = 0.0;
            if (act && r < m) {
                const int jj = j0 + r;
                const double2 v = S[s + (size_t)jj * rs];
                if (pin && jj == 0) {
                    d[1][r] = v.y;  // real part pinned to 0 with a unit row
                } else {
                    a[r] = amj[jj]; c[r] = apj[jj]; b[r] = acj[jj] - shift - zk;
                    d[0][r] = v.x; d[1][r] = v.y;
                }
            }
        }
        part::reduce<double, 2>(a, b, c, d, m);
    }
    double xf[2], xl[2];
    part::exchange<double, false, PT_SX, 2>(red, line, p, P, a, c, d, m, xf, xl);
    if (!act) return;
    const bool zim = (kx == 0 && (kz == 0 || kz == h3p - 1));
#pragma unroll
    for (int r = 0; r < PT_MMAX; r++) {
        if (r < m) {
            const double vr = part::x_of<double, 2>(a, c, d, 0, r, m, xf[0], xl[0]);
            const double vi = part::x_of<double, 2>(a, c, d, 1, r, m, xf[1], xl[1]);
            S[s + (size_t)(j0 + r) * rs] = make_double2(vr, zim ? 0.0 : vi);
        }
    }
}

// How the mean of dP is found.  The exact form sums the transformed field
// as the reference does (per-line pairwise sums, then planes, then the
// total) and subtracts the mean in a pass of its own.  The fast forms take
// it from the spectrum before the x c2r: the x sum of a line of the c2r
// output is n1m times its kx = 0 coefficient, so the line sums are the
// kx = 0 coefficients over n3m, and the c2r subtracts the mean as it writes
// dP (no per-line reduction, no subtraction pass: 0.4 ms at 512^3).
#ifdef TCS_RECIP
constexpr bool SPECTRAL_MEAN = true;
#else
constexpr bool SPECTRAL_MEAN = false;
#endif

// x c2r of TX k-lines of row j, scaling by 1/n1m/n3m into dP's interior,
// and the per-line pairwise sums for the mean (SUB: the mean scal[0] is
// known and subtracted instead).
template <int NT, bool SUB>
__global__ void __launch_bounds__(NT)
k_fft_x_inv(Grid g, int h1p, int TX, const double2 *__restrict__ spec, const double2 *__restrict__ tw,
            double *__restrict__ dP, double *__restrict__ linesum, const double *__restrict__ scal) {
    extern __shared__ double lds[];
    const int n = g.n1m;
    double *re = lds, *im = lds + fft_lds_len(TX * n);
    const int j = blockIdx.y + 1;
    const int k0 = blockIdx.x * TX;
    const int tid = threadIdx.x;
    const int nh = n / 2;
    const size_t rowbase = (size_t)g.n3m * h1p * (j - 1);
    for (int e = tid; e < TX * h1p; e += NT) {
        const int kx = e / TX, l = e - kx * TX;
        const double2 v = spec[rowbase + (size_t)g.n3m * kx + k0 + l];
        re[fpad(l * n + kx)] = v.x;
        im[fpad(l * n + kx)] = (kx == 0 || kx == nh) ? 0.0 : v.y;
        if (kx > 0 && kx < nh) {
            re[fpad(l * n + n - kx)] = v.x;
            im[fpad(l * n + n - kx)] = -v.y;
        }
    }
    __syncthreads();
    lds_fft<+1, NT>(re, im, n, TX, tw);
    const double fn1 = (double)g.n1m, fn3 = (double)g.n3m;
    if constexpr (SUB) {
        const double mean = scal[0];
        for (int e = tid; e < TX * n; e += NT) {
            const int l = e / n, t = e - l * n;
            dP[g.ix(t + 1, j, k0 + l + 1)] = re[fpad(e)] / fn1 / fn3 - mean;
        }
        return;
    }
    for (int e = tid; e < TX * n; e += NT) {
        const int l = e / n, t = e - l * n;
        const double v = re[fpad(e)] / fn1 / fn3;
        dP[g.ix(t + 1, j, k0 + l + 1)] = v;
        im[e] = v;  // reuse as the reduction buffer (unpadded)
    }
    __syncthreads();
    for (int h = n / 2; h >= 1; h >>= 1) {
        for (int e = tid; e < TX * h; e += NT) {
            const int l = e / h, t = e - l * h;
            im[l * n + t] = im[l * n + t] + im[l * n + t + h];
        }
        __syncthreads();
    }
    for (int l = tid; l < TX; l += NT) linesum[(size_t)(j - 1) + (size_t)g.n2m * (k0 + l)] = im[l * n];
}

// pairwise sum of `cnt` values at src + q*stride, zero padded to a power of two
__device__ double block_tree_sum(const double *__restrict__ src, int cnt, size_t stride, double *buf) {
    int P = 1;
    while (P < cnt) P <<= 1;
    for (int t = threadIdx.x; t < P; t += blockDim.x) buf[t] = (t < cnt) ? src[(size_t)t * stride] : 0.0;
    __syncthreads();
    for (int h = P / 2; h >= 1; h >>= 1) {
        for (int t = threadIdx.x; t < h; t += blockDim.x) buf[t] = buf[t] + buf[t + h];
        __syncthreads();
    }
    return buf[0];
}

// The fast forms' line sums from the spectrum after the z c2c: line (j, k)
// sums to the kx = 0 coefficient over n3m (see SPECTRAL_MEAN).
__global__ void k_spec_linesum(int n2m, int n3m, int h1p, const double2 *__restrict__ spec,
                               double *__restrict__ linesum) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y;  // 0-based
    if (k >= n3m) return;
    linesum[(size_t)j + (size_t)n2m * k] = spec[(size_t)n3m * h1p * j + k].x / (double)n3m;
}

__global__ void k_plane_sum(int n2m, const double *__restrict__ linesum, double *__restrict__ planesum) {
    extern __shared__ double buf[];
    const int k = blockIdx.x;
    const double s = block_tree_sum(linesum + (size_t)n2m * k, n2m, 1, buf);
    if (threadIdx.x == 0) planesum[k] = s;
}

__global__ void k_total_mean(int n1m, int n2m, int n3m, const double *__restrict__ planesum,
                             double *__restrict__ scal) {
    extern __shared__ double buf[];
    const double s = block_tree_sum(planesum, n3m, 1, buf);
    if (threadIdx.x == 0) scal[0] = s / (double)n1m / (double)n2m / (double)n3m;
}

__global__ void k_dp_subtract(Grid g, const double *__restrict__ scal, double *__restrict__ dP) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.n1m || j > g.n2m) return;
    const size_t q = g.ix(i, j, k);
    dP[q] = dP[q] - scal[0];
}

// quadratic wall extrapolation (:719-736) of dP (mode 0) or P (mode 1)
__global__ void k_wall_extrap(Grid g, double *__restrict__ F) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int k = blockIdx.y + 1;
    if (i > g.n1m) return;
    const double *dmx2 = g.dmx2;
    const int n2sub = g.n2m + 1, n2msub = g.n2m;
    const double s = dmx2[1] + dmx2[2];
    const double a0 = (s * s) / ((s * s) - (dmx2[1] * dmx2[1]));
    const double b0 = (dmx2[1] * dmx2[1]) / ((s * s) - (dmx2[1] * dmx2[1]));
    F[g.ix(i, 0, k)] = a0 * F[g.ix(i, 1, k)] - b0 * F[g.ix(i, 2, k)];
    const double t = dmx2[n2sub] + dmx2[n2msub];
    const double a1 = (t * t) / ((t * t) - (dmx2[n2sub] * dmx2[n2sub]));
    const double b1 = (dmx2[n2sub] * dmx2[n2sub]) / ((t * t) - (dmx2[n2sub] * dmx2[n2sub]));
    F[g.ix(i, n2sub, k)] = a1 * F[g.ix(i, n2msub, k)] - b1 * F[g.ix(i, n2sub - 2, k)];
}

// projection (:793-851).  The reference overwrites dPhat(i,j,k) inside the
// same loop nest (k outer, j = 2..n2m then j = 1, i inner), so the
// neighbours it reads at (i-1,j,k) [i>=2], (i,j-1,k) [j>=3] and (i,j,k-1)
// [k>=2] already hold the new value 2 dP - (dPhat + dP0)/2.  This kernel
// reads only old values and recomputes those neighbours' new values; a
// second kernel then stores the new dPhat and dP0.
template <class S>
__device__ __forceinline__ S dphat_new(S dP, S dPh, S dP0) {
    const S d0 = (dPh + dP0) / 2.0;
    return 2.0 * dP - d0;
}

template <class S, class FP>
__device__ __forceinline__ void projection_cell(const Grid &g, const Phys &p, double dt, FP iR, FP dP,
                                               FP dPhat, FP dP0, FP U, FP V, FP W, FP P, int i, int j, int k,
                                               S &uo, S &vo, S &wo, S &po) {
    const auto dx1 = metric<S, M_DX1>(g), dx2 = metric<S, M_DX2>(g), dx3 = metric<S, M_DX3>(g);
    const auto dmx1 = metric<S, M_DMX1>(g), dmx2 = metric<S, M_DMX2>(g), dmx3 = metric<S, M_DMX3>(g);
    const int im = i - 1, jm = j - 1, km = k - 1;
    const S Cmp = p.Cmp;
    const size_t q = g.ix(i, j, k);
    const S dPc = dP[q], dPhc = dPhat[q];
    const size_t qi = g.ix(im, j, k), qj = g.ix(i, jm, k), qk = g.ix(i, j, km);
    const S hi = (i >= 2) ? dphat_new(dP[qi], dPhat[qi], dP0[qi]) : dPhat[qi];
    const S hj = (j >= 3) ? dphat_new(dP[qj], dPhat[qj], dP0[qj]) : dPhat[qj];
    const S hk = (k >= 2) ? dphat_new(dP[qk], dPhat[qk], dP0[qk]) : dPhat[qk];
    S invRhoc = 0.5 / dmx1[i] * (dx1[i] * IRt(j, k) + dx1[im] * IRt(j, k));
    uo = U[q] - dt * Cmp * ((dPc - dP[qi]) / dmx1[i] + (invRhoc - 1.0) * (dPhc - hi) / dmx1[i]);
    invRhoc = 0.5 / dmx2[j] * (dx2[j] * IRt(jm, k) + dx2[jm] * IRt(j, k));
    if (j >= 2)
        vo = V[q] - dt * Cmp * ((dPc - dP[qj]) / dmx2[j] + (invRhoc - 1.0) * (dPhc - hj) / dmx2[j]);
    else
        vo = V[q] - dt * Cmp * ((dPc - dP[qj]) / dmx2[j] + (invRhoc - 1.0) * (dPhc - hj) / dmx2[j]) * 0.0;
    invRhoc = 0.5 / dmx3[k] * (dx3[k] * IRt(j, km) + dx3[km] * IRt(j, k));
    wo = W[q] - dt * Cmp * ((dPc - dP[qk]) / dmx3[k] + (invRhoc - 1.0) * (dPhc - hk) / dmx3[k]);
    po = P[q] + dPc;
}
struct Projection {
    Grid g;
    Phys p;
    double dt;
    const double *iR, *dP, *dPhat, *dP0, *U, *V, *W, *P;
    static constexpr int NF = 8;
    template <class S, class Wr>
    __device__ __forceinline__ void eval(const Wr &w, int i, int j, int k, S &a, S &b, S &c, S &d) const {
        projection_cell<S>(g, p, dt, w(0, iR), w(1, dP), w(2, dPhat), w(3, dP0), w(4, U), w(5, V), w(6, W), w(7, P), i,
                           j, k, a, b, c, d);
    }
};

__global__ void __launch_bounds__(256) k_projection(Projection cell, double *U, double *V, double *W, double *P) {
    const Grid &g = cell.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.n1m || j > g.n2m) return;
    double u, v, w, q;
    cell.eval<double>(RawW(), i, j, k, u, v, w, q);
    const size_t o = g.ix(i, j, k);
    U[o] = u;
    V[o] = v;
    W[o] = w;
    P[o] = q;
}

// the dPhat and dP0 history after the projection (:846-850)
template <class S, class FP>
__device__ __forceinline__ void dphat_cell(const Grid &g, FP dP, FP dPhat, FP dP0, int i, int j, int k, S &hn,
                                          S &d0n) {
    const size_t q = g.ix(i, j, k);
    const S d0 = (dPhat[q] + dP0[q]) / 2.0;
    d0n = d0;
    hn = 2.0 * dP[q] - d0;
}
struct DPhatUpdate {
    Grid g;
    const double *dP, *dPhat, *dP0;
    static constexpr int NF = 3;
    template <class S, class Wr>
    __device__ __forceinline__ void eval(const Wr &w, int i, int j, int k, S &a, S &b, S &c, S &d) const {
        dphat_cell<S>(g, w(0, dP), w(1, dPhat), w(2, dP0), i, j, k, a, b);
    }
};

// the projection with the new dPhat and dP0 written to second buffers
// TW (the mixed form): U, V, W, P's FP32 copies (tw[0..3]) of the updated
// cells written as well
template <bool TW>
__global__ void __launch_bounds__(256) k_projection_into(Projection cell, DPhatUpdate hist, double *U, double *V,
                                                        double *W, double *P, double *__restrict__ dPhatn,
                                                        double *__restrict__ dP0n, FieldSet32 tw) {
    const Grid &g = cell.g;
    const int i0 = (blockIdx.x * blockDim.x + threadIdx.x) * PROJ_IB + 1;  // n1m is a multiple of PROJ_IB
    const int j = wave_row_j();
    if (i0 > g.n1m || j > g.n2m) return;
    // every cell's loads before any store: U, V, W, P are updated in place
    // (each cell reads only its own point of them), and the loads the cells
    // share are made once
    constexpr int NC = PROJ_IB * PROJ_KB;
    double u[NC], v[NC], w[NC], q[NC], h[NC], d0[NC];
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const int i = i0 + c % PROJ_IB, k = blockIdx.z * PROJ_KB + c / PROJ_IB + 1;
        double x;
        cell.eval<double>(RawW(), i, j, k, u[c], v[c], w[c], q[c]);
        hist.eval<double>(RawW(), i, j, k, h[c], d0[c], x, x);
    }
#pragma unroll
    for (int c = 0; c < NC; c++) {
        const size_t o = g.ix(i0 + c % PROJ_IB, j, blockIdx.z * PROJ_KB + c / PROJ_IB + 1);
        U[o] = u[c];
        V[o] = v[c];
        W[o] = w[c];
        P[o] = q[c];
        dP0n[o] = d0[c];
        dPhatn[o] = h[c];
        if (TW) {
            tw.f[0][o] = (float)u[c];
            tw.f[1][o] = (float)v[c];
            tw.f[2][o] = (float)w[c];
            tw.f[3][o] = (float)q[c];
        }
    }
}

__global__ void __launch_bounds__(256) k_dphat_update(DPhatUpdate cell, double *dPhat, double *dP0) {
    const Grid &g = cell.g;
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > g.n1m || j > g.n2m) return;
    double h, d0, x;
    cell.eval<double>(RawW(), i, j, k, h, d0, x, x);
    const size_t q = g.ix(i, j, k);
    dP0[q] = d0;
    dPhat[q] = h;
}

// ---------------------------------------------------------------- host -----
// x FFTs: 256 threads (the strided gather of the k-lines wants them all)
// over TX k-lines of n1m points per workgroup, at most 4 (more leaves too
// few workgroups per CU), within one pass's butterflies and 64 KiB of LDS
constexpr int FFT_X_THREADS = 256;
static int fft_tile(int n1m, int n3m, int mb) {
    int tx = 1;
    while (tx < 4 && tx * 2 <= n3m && (tx * 2) * (n1m / 2) <= mb * FFT_X_THREADS &&
           (size_t)fft_lds_len(tx * 2 * n1m) * 16 <= 65536)
        tx *= 2;
    return tx;
}

// r2c along x of the interior-layout values src(i, j, k) into the spectrum
template <class Src>
static int fft_x_fwd_src_launch(tcs_ctx *c, const char *label, const Src &src, double2 *spec) {
    const int TX = fft_tile(c->n1m, c->n3m, FFT_MAXB_FWD);
    const size_t lds = (size_t)fft_lds_len(TX * c->n1m) * 2 * sizeof(double);
    TCS_LAUNCH(c->stream, label, (k_fft_x_fwd<FFT_X_THREADS, Src>), dim3(c->n3m / TX, c->n2m), dim3(FFT_X_THREADS),
               lds, c->grid(), c->h1p, TX, src, (const double2 *)c->tw1, spec);
    return TCS_OK;
}
static int fft_x_fwd_launch(tcs_ctx *c, const char *label, const double *in, double2 *spec) {
    return fft_x_fwd_src_launch(c, label, RhsArray{c->grid(), in}, spec);
}

// c2r along x into the full-layout field `out`, with the per-line sums
// (sub: subtracting the mean c->scal[0] instead)
static int fft_x_inv_launch(tcs_ctx *c, const char *label, const double2 *spec, double *out, bool sub = false) {
    const int TX = fft_tile(c->n1m, c->n3m, FFT_MAXB);
    const size_t lds = (size_t)fft_lds_len(TX * c->n1m) * 2 * sizeof(double);
    if (sub)
        TCS_LAUNCH(c->stream, label, (k_fft_x_inv<FFT_X_THREADS, true>), dim3(c->n3m / TX, c->n2m),
                   dim3(FFT_X_THREADS), lds, c->grid(), c->h1p, TX, spec, (const double2 *)c->tw1, out, c->linesum,
                   c->scal);
    else
        TCS_LAUNCH(c->stream, label, (k_fft_x_inv<FFT_X_THREADS, false>), dim3(c->n3m / TX, c->n2m),
                   dim3(FFT_X_THREADS), lds, c->grid(), c->h1p, TX, spec, (const double2 *)c->tw1, out, c->linesum,
                   c->scal);
    return TCS_OK;
}

// z FFTs: the fewest threads (64, 128 or 256) that hold a pass's
// butterflies at FFT_MAXB per thread (one wave for 512 points: small
// workgroups, more of them per CU)
static int fft_threads(int butterflies) {
    int nt = 64;
    while (nt < 256 && FFT_MAXB * nt < butterflies) nt <<= 1;
    return nt;
}

// c2c along z in place, one (kx, j) line per workgroup
template <int SIGN>
static int fft_z_launch(tcs_ctx *c, const char *label, double2 *spec) {
    const int NT = fft_threads(c->n3m / 2);
    const size_t lds = (size_t)fft_lds_len(c->n3m) * 2 * sizeof(double);
    const dim3 grd(c->h1p, c->n2m);
    const double2 *tw = (const double2 *)c->tw3;
    if (NT == 64) TCS_LAUNCH(c->stream, label, (k_fft_z<SIGN, 64>), grd, dim3(64), lds, c->n3m, c->h1p, tw, spec);
    else if (NT == 128) TCS_LAUNCH(c->stream, label, (k_fft_z<SIGN, 128>), grd, dim3(128), lds, c->n3m, c->h1p, tw, spec);
    else TCS_LAUNCH(c->stream, label, (k_fft_z<SIGN, 256>), grd, dim3(256), lds, c->n3m, c->h1p, tw, spec);
    return TCS_OK;
}

static PressureRHS rhs_cell(tcs_ctx *c, double dt) {
    return PressureRHS{c->grid(),        c->phys,           dt, c->f[TCS_U], c->f[TCS_V], c->f[TCS_W], c->f[TCS_DPHAT],
                       c->f[TCS_INVRHO], c->wall[TCS_VBCUP], c->wall[TCS_VBCBT]};
}

int pressure_rhs(tcs_ctx *c, double dt) {
    const dim3 blk = interior_block(), grd = cell_grid(c->n1m, c->n2m, c->n3m, RHS_KB);
    TCS_LAUNCH(c->stream, "pressure_rhs", k_pressure_rhs, grd, blk, 0, rhs_cell(c, dt), c->prhs);
    return TCS_OK;
}

// the solve after the x r2c (in spec_x): z transform, y solves, inverse
// transforms into dP, mean removal, wall extrapolation
static int poisson_after_x(tcs_ctx *c) {
    const Grid g = c->grid();
    const int n1m = c->n1m, n2m = c->n2m, n3m = c->n3m, h1p = c->h1p;
    double2 *spec = (double2 *)c->spec_x;
    TCS_TRY(fft_z_launch<-1>(c, "fft_z<-1>", spec));
    const int nsys = n3m * h1p;
    int P, w1, w2;
    if (c->tdma_split && ptdma_part_plan(n2m, P, w1, w2)) {
        TCS_LAUNCH(c->stream, "poisson_part", k_poisson_part, dim3((nsys + PT_SX - 1) / PT_SX), dim3(PT_SX * P), 0,
                   n3m, h1p, c->h3p, c->ptab, c->ptab + n2m, c->ptab + 2 * n2m, c->dxk2, c->dzk, spec, P, w1, w2);
    } else {
        TCS_LAUNCH(c->stream, "poisson_tdma", k_poisson_tdma, dim3((nsys + 255) / 256), dim3(256), 0, n3m, h1p, n2m,
                   c->h3p, c->ptab, c->ptab + n2m, c->ptab + 2 * n2m, c->dxk2, c->dzk, spec, c->pA, c->pC);
    }
    TCS_TRY(fft_z_launch<+1>(c, "fft_z<+1>", spec));
    if (SPECTRAL_MEAN)
        TCS_LAUNCH(c->stream, "spec_linesum", k_spec_linesum, dim3((n3m + 255) / 256, n2m), dim3(256), 0, n2m, n3m,
                   h1p, spec, c->linesum);
    else
        TCS_TRY(fft_x_inv_launch(c, "fft_x_inv", spec, c->f[TCS_DP]));
    int P2 = 1;
    while (P2 < n2m) P2 <<= 1;
    TCS_LAUNCH(c->stream, "plane_sum", k_plane_sum, dim3(n3m), dim3(256), P2 * sizeof(double), n2m, c->linesum,
                       c->planesum);
    int P3 = 1;
    while (P3 < n3m) P3 <<= 1;
    TCS_LAUNCH(c->stream, "total_mean", k_total_mean, dim3(1), dim3(256), P3 * sizeof(double), n1m, n2m, n3m,
                       c->planesum, c->scal);
    if (SPECTRAL_MEAN) {
        TCS_TRY(fft_x_inv_launch(c, "fft_x_inv", spec, c->f[TCS_DP], true));
    } else {
        const dim3 blk = interior_block(), grd = interior_grid(n1m, n2m, n3m);
        TCS_LAUNCH(c->stream, "dp_subtract", k_dp_subtract, grd, blk, 0, g, c->scal, c->f[TCS_DP]);
    }
    TCS_LAUNCH(c->stream, "wall_extrap", k_wall_extrap, dim3((n1m + 63) / 64, n3m), dim3(64), 0, g, c->f[TCS_DP]);
    return TCS_OK;
}

int pressure_poisson(tcs_ctx *c) {
    TCS_TRY(fft_x_fwd_launch(c, "fft_x_fwd", c->prhs, (double2 *)c->spec_x));
    return poisson_after_x(c);
}

// pressure_rhs and pressure_poisson in one: the x r2c makes the right-hand
// side as it loads it (the interior array is not written)
int pressure_poisson_rhs(tcs_ctx *c, double dt) {
    TCS_TRY(fft_x_fwd_src_launch(c, "rhs_fft_x_fwd", RhsCell{rhs_cell(c, dt)}, (double2 *)c->spec_x));
    return poisson_after_x(c);
}

int pressure_projection(tcs_ctx *c, double dt) {
    c->twin_ok &= ~(unsigned)(TWIN_U | TWIN_V | TWIN_W | TWIN_P);  // the mixed form's FP32 copies
    const Grid g = c->grid();
    const dim3 blk = interior_block(), grd = interior_grid(c->n1m, c->n2m, c->n3m);
    const Projection cell{g, c->phys, dt, c->f[TCS_INVRHO], c->f[TCS_DP], c->f[TCS_DPHAT], c->f[TCS_DP0],
                          c->f[TCS_U], c->f[TCS_V], c->f[TCS_W], c->f[TCS_P]};
    TCS_LAUNCH(c->stream, "projection", k_projection, grd, blk, 0, cell, c->f[TCS_U], c->f[TCS_V], c->f[TCS_W],
               c->f[TCS_P]);
    TCS_LAUNCH(c->stream, "dphat_update", k_dphat_update, grd, blk, 0, DPhatUpdate{g, c->f[TCS_DP], c->f[TCS_DPHAT],
               c->f[TCS_DP0]}, c->f[TCS_DPHAT], c->f[TCS_DP0]);
    TCS_LAUNCH(c->stream, "wall_extrap", k_wall_extrap, dim3((c->n1m + 63) / 64, c->n3m), dim3(64), 0, g, c->f[TCS_P]);
    return TCS_OK;
}

int pressure_projection_into(tcs_ctx *c, double dt, double *dPhatn, double *dP0n) {
    c->twin_ok &= ~(unsigned)(TWIN_U | TWIN_V | TWIN_W | TWIN_P);  // the mixed form's FP32 copies
    const Grid g = c->grid();
    const dim3 blk = interior_block(), grd = cell_grid(c->n1m, c->n2m, c->n3m, PROJ_KB, interior_block(), PROJ_IB);
    const Projection cell{g, c->phys, dt, c->f[TCS_INVRHO], c->f[TCS_DP], c->f[TCS_DPHAT], c->f[TCS_DP0],
                          c->f[TCS_U], c->f[TCS_V], c->f[TCS_W], c->f[TCS_P]};
    TCS_TRY(copy_shell(c, {{c->f[TCS_DPHAT], dPhatn}, {c->f[TCS_DP0], dP0n}}));
    const DPhatUpdate hist{g, c->f[TCS_DP], c->f[TCS_DPHAT], c->f[TCS_DP0]};
    if (c->a32) {
        const FieldSet32 tw{{twin(c, c->f[TCS_U]), twin(c, c->f[TCS_V]), twin(c, c->f[TCS_W]), twin(c, c->f[TCS_P])}};
        TCS_LAUNCH(c->stream, "projection", k_projection_into<true>, grd, blk, 0, cell, hist, c->f[TCS_U], c->f[TCS_V],
                   c->f[TCS_W], c->f[TCS_P], dPhatn, dP0n, tw);
    } else {
        TCS_LAUNCH(c->stream, "projection", k_projection_into<false>, grd, blk, 0, cell, hist, c->f[TCS_U], c->f[TCS_V],
                   c->f[TCS_W], c->f[TCS_P], dPhatn, dP0n, FieldSet32{});
    }
    TCS_LAUNCH(c->stream, "wall_extrap", k_wall_extrap, dim3((c->n1m + 63) / 64, c->n3m), dim3(64), 0, g, c->f[TCS_P]);
    return TCS_OK;
}

}  // namespace TCS_FORM
}  // namespace tcs
