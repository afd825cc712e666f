/*
 * tcs_oracle.c -- CPU restatement of the PaScaL_TCS MPM-STD step on one rank.
 *
 * TEST INFRASTRUCTURE ONLY (see tcs_oracle.h).  Build: -O2 -ffp-contract=off.
 *
 * Each routine follows the cited reference loop nest statement by statement
 * with the same local-variable scoping, so values that the reference reads
 * from loop variables left over from an earlier stage (stale jm/jp/km/kp,
 * mu1..mu4) are reproduced by construction.
 *
 * Deliberate deviations from the reference:
 *  D1. FFTs: MKL/FFTW (closed) is replaced by a radix-2 Stockham FFT with a
 *      host twiddle table (orc_fft); x r2c is a length-n1m complex FFT of
 *      (x,0) keeping bins 0..n1m/2; c2r builds the Hermitian extension with
 *      zero imaginary DC/Nyquist parts.  Power-of-two n1m, n3m only.
 *  D2. The dP mean (src/module_solve_pressure.f90:703-710, a compiler-chosen
 *      Fortran SUM) is summed as a pairwise tree per x-line, then per k-plane
 *      over lines, then over planes (orc_tree_sum).
 *  D3. The dP wall extrapolation loops k=1..n3msub; the reference's k=n3sub
 *      plane reads unfilled memory and is overwritten by the next halo
 *      update (src/main.f90:189), so it is not observable.
 *  D4. Arrays the reference allocates uninitialised are zero-filled.
 */
#include "tcs_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

struct orc {
    int n1m, n2m, n3m;
    int n1sub, n2sub, n3sub, n1msub, n2msub, n3msub;
    int N1, N2, N3;          /* full extents n?sub+1 */
    int i_indexS, j_indexS, k_indexS;
    int h1p, h3p;
    double *dx1, *dx2, *dx3, *dmx1, *dmx2, *dmx3;
    int *jC_BC, *jS_BC;
    double prm[OP_N];
    double *f[OF_NFIELD];
    double *bc[OB_N];
    double *dzk, *dxk2;
    int tdma_split;          /* 0: single-rank line solves; 1: lines split into parts */
    /* test hook: records of the line solves a routine makes */
    int cap_on, ncap;
    struct orc_cap { int n_sys, n_row, cyclic; double *a, *b, *c, *d, *x; } cap[ORC_MAXCAP];
};

/* Planted mistakes (test hook): the oracle built with -DORC_PLANTS takes the
   expression `wrong` at site `id` while orc_set_plant(id) is in force, so a
   test can show that its closed form rejects that mistake.  The ordinary
   build compiles `right` only. */
#ifdef ORC_PLANTS
static int orc_plant_id = 0;
#define PLANT(id, wrong, right) (orc_plant_id == (id) ? (wrong) : (right))
void orc_set_plant(int id) { orc_plant_id = id; }
int orc_plants_built(void) { return 1; }
#else
#define PLANT(id, wrong, right) (right)
void orc_set_plant(int id) { (void)id; }
int orc_plants_built(void) { return 0; }
#endif

static void cap_clear(orc *o);

/* The oracle built with -fopenmp (liboracle_omp.so, the CPU baseline's
   build) splits the step's loop nests and line solves over threads; every
   cell and every line is computed exactly as in the serial build, and the
   loop variables a later stage reads from an earlier one (the stale
   indices) leave the parallel loops as the sequentially last iteration
   leaves them (lastprivate), so both builds give the same bits. */
#ifdef _OPENMP
#include <omp.h>
#define OMP(x) _Pragma(#x)
#else
#define OMP(x)
#endif

#define IX(i, j, k) ((size_t)(i) + (size_t)o->N1 * ((size_t)(j) + (size_t)o->N2 * (size_t)(k)))
#define B2(i, k) ((size_t)(i) + (size_t)o->N1 * (size_t)(k))
/* interior work array with Fortran extents (1:a,1:b,1:c) */
#define R3(a, b, i, j, k) ((size_t)((i) - 1) + (size_t)(a) * ((size_t)((j) - 1) + (size_t)(b) * (size_t)((k) - 1)))

/* ------------------------------------------------------------------------ */
/* mesh: src/module_mpi_subdomain.f90:153-299 on one rank                    */
/* ------------------------------------------------------------------------ */
void orc_mesh_axis(int nm, double L, int uniform, double gamma, int periodic,
                   double *x, double *dx, double *dmx)
{
    int n = nm + 1; /* n?sub */
    for (int i = 0; i <= n; i++) { x[i] = 0.0; dx[i] = 0.0; dmx[i] = 0.0; }
    for (int i = 0; i <= n; i++) {
        /* global index ig = i (ista=1); formula uses (ig-1) */
        if (uniform == 1)
            x[i] = (double)(i - 1) * L / (double)nm + 0.0;
        else
            x[i] = L * 0.5 * (1.0 + tanh(0.5 * gamma * (2.0 * (double)(i - 1) / (double)nm - 1.0)) / tanh(gamma * 0.5)) + 0.0;
    }
    if (!periodic) x[0] = x[1];
    for (int i = 1; i <= nm; i++) dx[i] = x[i + 1] - x[i];
    /* single-rank exchange: periodic self-copy, walls MPI_PROC_NULL */
    if (periodic) { dx[0] = dx[nm]; dx[n] = dx[1]; }
    else { dx[0] = 0.0; dx[n] = 0.0; }
    for (int i = 1; i <= nm; i++) dmx[i] = 0.5 * (dx[i - 1] + dx[i]);
    if (periodic) { dmx[0] = dmx[nm]; dmx[n] = dmx[1]; }
    else dmx[n] = 0.5 * (dx[nm] + dx[n]);
}

orc *orc_create(int n1m, int n2m, int n3m,
                const double *dx1, const double *dx2, const double *dx3,
                const double *dmx1, const double *dmx2, const double *dmx3,
                const double *prm)
{
    orc *o = (orc *)calloc(1, sizeof(orc));
    o->n1m = n1m; o->n2m = n2m; o->n3m = n3m;
    o->n1sub = n1m + 1; o->n2sub = n2m + 1; o->n3sub = n3m + 1;
    o->n1msub = n1m; o->n2msub = n2m; o->n3msub = n3m;
    o->N1 = n1m + 2; o->N2 = n2m + 2; o->N3 = n3m + 2;
    /* pbc = (T,F,T): src/module_mpi_subdomain.f90:304-374 */
    o->i_indexS = 1; o->j_indexS = 2; o->k_indexS = 1;
    o->h1p = n1m / 2 + 1; o->h3p = n3m / 2 + 1;
    size_t a1 = (size_t)o->N1, a2 = (size_t)o->N2, a3 = (size_t)o->N3;
    o->dx1 = (double *)malloc(a1 * sizeof(double)); memcpy(o->dx1, dx1, a1 * sizeof(double));
    o->dmx1 = (double *)malloc(a1 * sizeof(double)); memcpy(o->dmx1, dmx1, a1 * sizeof(double));
    o->dx2 = (double *)malloc(a2 * sizeof(double)); memcpy(o->dx2, dx2, a2 * sizeof(double));
    o->dmx2 = (double *)malloc(a2 * sizeof(double)); memcpy(o->dmx2, dmx2, a2 * sizeof(double));
    o->dx3 = (double *)malloc(a3 * sizeof(double)); memcpy(o->dx3, dx3, a3 * sizeof(double));
    o->dmx3 = (double *)malloc(a3 * sizeof(double)); memcpy(o->dmx3, dmx3, a3 * sizeof(double));
    o->jC_BC = (int *)malloc(a2 * sizeof(int));
    o->jS_BC = (int *)malloc(a2 * sizeof(int));
    for (size_t j = 0; j < a2; j++) { o->jC_BC[j] = 1; o->jS_BC[j] = 1; }
    o->jC_BC[0] = 0; o->jS_BC[0] = 0; o->jS_BC[1] = 0;
    o->jC_BC[o->n2sub] = 0; o->jS_BC[o->n2sub] = 0;
    memcpy(o->prm, prm, OP_N * sizeof(double));
    size_t nf = a1 * a2 * a3;
    for (int f = 0; f < OF_NFIELD; f++) o->f[f] = (double *)calloc(nf, sizeof(double));
    for (int b = 0; b < OB_N; b++) o->bc[b] = (double *)calloc(a1 * a3, sizeof(double));
    o->dzk = (double *)calloc((size_t)n3m, sizeof(double));
    o->dxk2 = (double *)calloc((size_t)o->h1p, sizeof(double));
    return o;
}

void orc_destroy(orc *o)
{
    if (!o) return;
    cap_clear(o);
    free(o->dx1); free(o->dx2); free(o->dx3); free(o->dmx1); free(o->dmx2); free(o->dmx3);
    free(o->jC_BC); free(o->jS_BC);
    for (int f = 0; f < OF_NFIELD; f++) free(o->f[f]);
    for (int b = 0; b < OB_N; b++) free(o->bc[b]);
    free(o->dzk); free(o->dxk2);
    free(o);
}

double *orc_field(orc *o, int id) { return (id >= 0 && id < OF_NFIELD) ? o->f[id] : NULL; }
double *orc_bc(orc *o, int id) { return (id >= 0 && id < OB_N) ? o->bc[id] : NULL; }

void orc_wave_numbers(orc *o, double *dzk, double *dxk2)
{
    memcpy(dzk, o->dzk, sizeof(double) * (size_t)o->n3m);
    memcpy(dxk2, o->dxk2, sizeof(double) * (size_t)o->h1p);
}

/* ------------------------------------------------------------------------ */
/* PaScaL_TDMA/src/tdmas.f90                                                 */
/* ------------------------------------------------------------------------ */
/* tdmas.f90:27-51 */
void orc_tdma_single(double *a, double *b, double *c, double *d, int n1)
{
    (void)a;
    d[0] = d[0] / b[0];
    c[0] = c[0] / b[0];
    for (int i = 1; i < n1; i++) {
        double r = 1.0 / (b[i] - a[i] * c[i - 1]);
        d[i] = r * (d[i] - a[i] * d[i - 1]);
        c[i] = r * c[i];
    }
    for (int i = n1 - 2; i >= 0; i--) d[i] = d[i] - c[i] * d[i + 1];
}

/* tdmas.f90:61-100 (1-based e(2), e(n1) mapped to e[1], e[n1-1]) */
void orc_tdma_cycl_single(double *a, double *b, double *c, double *d, int n1)
{
    double *e = (double *)calloc((size_t)n1, sizeof(double));
    e[1] = -a[1];
    e[n1 - 1] = -c[n1 - 1];
    d[1] = d[1] / b[1];
    e[1] = e[1] / b[1];
    c[1] = c[1] / b[1];
    for (int i = 2; i < n1; i++) {
        double rr = 1.0 / (b[i] - a[i] * c[i - 1]);
        d[i] = rr * (d[i] - a[i] * d[i - 1]);
        e[i] = rr * (e[i] - a[i] * e[i - 1]);
        c[i] = rr * c[i];
    }
    for (int i = n1 - 2; i >= 1; i--) {
        d[i] = d[i] - c[i] * d[i + 1];
        e[i] = e[i] - c[i] * e[i + 1];
    }
    d[0] = (d[0] - a[0] * d[n1 - 1] - c[0] * d[1]) / (b[0] + a[0] * e[n1 - 1] + c[0] * e[1]);
    for (int i = 1; i < n1; i++) d[i] = d[i] + d[0] * e[i];
    free(e);
}

/* tdmas.f90:113-146; arrays (n1,n2) with n1 fastest */
void orc_tdma_many(double *a, double *b, double *c, double *d, int n1, int n2)
{
#define M2(x, i, j) x[(size_t)(i) + (size_t)n1 * (size_t)(j)]
    (void)a;
    for (int i = 0; i < n1; i++) {
        M2(d, i, 0) = M2(d, i, 0) / M2(b, i, 0);
        M2(c, i, 0) = M2(c, i, 0) / M2(b, i, 0);
    }
    for (int j = 1; j < n2; j++)
        for (int i = 0; i < n1; i++) {
            double r = 1.0 / (M2(b, i, j) - M2(a, i, j) * M2(c, i, j - 1));
            M2(d, i, j) = r * (M2(d, i, j) - M2(a, i, j) * M2(d, i, j - 1));
            M2(c, i, j) = r * M2(c, i, j);
        }
    for (int j = n2 - 2; j >= 0; j--)
        for (int i = 0; i < n1; i++)
            M2(d, i, j) = M2(d, i, j) - M2(c, i, j) * M2(d, i, j + 1);
}

/* tdmas.f90:159-212; note e(:,n2)=-c(:,n2) overwrites e(:,2)=-a(:,2) when n2=2 */
void orc_tdma_cycl_many(double *a, double *b, double *c, double *d, int n1, int n2)
{
    double *e = (double *)calloc((size_t)n1 * (size_t)n2, sizeof(double));
    for (int i = 0; i < n1; i++) {
        M2(e, i, 1) = -M2(a, i, 1);
        M2(e, i, n2 - 1) = -M2(c, i, n2 - 1);
    }
    for (int i = 0; i < n1; i++) {
        M2(d, i, 1) = M2(d, i, 1) / M2(b, i, 1);
        M2(e, i, 1) = M2(e, i, 1) / M2(b, i, 1);
        M2(c, i, 1) = M2(c, i, 1) / M2(b, i, 1);
    }
    for (int j = 2; j < n2; j++)
        for (int i = 0; i < n1; i++) {
            double rr = 1.0 / (M2(b, i, j) - M2(a, i, j) * M2(c, i, j - 1));
            M2(d, i, j) = rr * (M2(d, i, j) - M2(a, i, j) * M2(d, i, j - 1));
            M2(e, i, j) = rr * (M2(e, i, j) - M2(a, i, j) * M2(e, i, j - 1));
            M2(c, i, j) = rr * M2(c, i, j);
        }
    for (int j = n2 - 2; j >= 1; j--)
        for (int i = 0; i < n1; i++) {
            M2(d, i, j) = M2(d, i, j) - M2(c, i, j) * M2(d, i, j + 1);
            M2(e, i, j) = M2(e, i, j) - M2(c, i, j) * M2(e, i, j + 1);
        }
    for (int i = 0; i < n1; i++)
        M2(d, i, 0) = (M2(d, i, 0) - M2(a, i, 0) * M2(d, i, n2 - 1) - M2(c, i, 0) * M2(d, i, 1))
                    / (M2(b, i, 0) + M2(a, i, 0) * M2(e, i, n2 - 1) + M2(c, i, 0) * M2(e, i, 1));
    for (int j = 1; j < n2; j++)
        for (int i = 0; i < n1; i++)
            M2(d, i, j) = M2(d, i, j) + M2(d, i, 0) * M2(e, i, j);
    free(e);
#undef M2
}

/* ------------------------------------------------------------------------ */
/* PaScaL_TDMA_many_solve / _cycle: PaScaL_TDMA/src/pascal_tdma.f90:487-578   */
/* and :715-806.  The rows of every system are split over `nprocs` ranks by   */
/* para_range (PaScaL_TDMA/src/para_range.f90:26-40); each rank runs the      */
/* modified Thomas algorithm on its own rows, the reduced 2-rows-per-rank     */
/* systems are gathered in rank order (the Alltoallw of plan_many_create,     */
/* pascal_tdma.f90:243-266) and solved by tdma_many / tdma_cycl_many, and     */
/* each rank updates its rows.  nprocs = 1 is the single-rank solve.          */
/* ------------------------------------------------------------------------ */
#define P2(x, i, j) x[(size_t)(i) + (size_t)ld * (size_t)(j)]
/* one rank's rows j0 .. j0+m-1: elimination of both off-diagonals */
static void ptdma_rank_reduce(double *A, double *B, double *C, double *D, int n_sys, size_t ld, int j0, int m)
{
    for (int i = 0; i < n_sys; i++) {
        P2(A, i, j0) = P2(A, i, j0) / P2(B, i, j0);
        P2(D, i, j0) = P2(D, i, j0) / P2(B, i, j0);
        P2(C, i, j0) = P2(C, i, j0) / P2(B, i, j0);
        P2(A, i, j0 + 1) = P2(A, i, j0 + 1) / P2(B, i, j0 + 1);
        P2(D, i, j0 + 1) = P2(D, i, j0 + 1) / P2(B, i, j0 + 1);
        P2(C, i, j0 + 1) = P2(C, i, j0 + 1) / P2(B, i, j0 + 1);
    }
    for (int j = j0 + 2; j < j0 + m; j++)
        for (int i = 0; i < n_sys; i++) {
            double r = 1.0 / (P2(B, i, j) - P2(A, i, j) * P2(C, i, j - 1));
            P2(D, i, j) = r * (P2(D, i, j) - P2(A, i, j) * P2(D, i, j - 1));
            P2(C, i, j) = r * P2(C, i, j);
            P2(A, i, j) = -r * P2(A, i, j) * P2(A, i, j - 1);
        }
    for (int j = j0 + m - 3; j >= j0 + 1; j--)
        for (int i = 0; i < n_sys; i++) {
            P2(D, i, j) = P2(D, i, j) - P2(C, i, j) * P2(D, i, j + 1);
            P2(A, i, j) = P2(A, i, j) - P2(C, i, j) * P2(A, i, j + 1);
            P2(C, i, j) = -P2(C, i, j) * P2(C, i, j + 1);
        }
    for (int i = 0; i < n_sys; i++) {
        double r = 1.0 / (1.0 - P2(A, i, j0 + 1) * P2(C, i, j0));
        P2(D, i, j0) = r * (P2(D, i, j0) - P2(C, i, j0) * P2(D, i, j0 + 1));
        P2(A, i, j0) = r * P2(A, i, j0);
        P2(C, i, j0) = -r * P2(C, i, j0) * P2(C, i, j0 + 1);
    }
}

/* Reduced system of one line (N = 2*nprocs rows, unit diagonal at entry)
   by parallel cyclic reduction, the split solver's replacement of the serial
   tdma_many / tdma_cycl_many when there are two or more parts.  Step with
   stride s: every row i eliminates its couplings to rows i-s and i+s using
   those rows' current values (all rows updated from the same old values):
     k1 = a_i/b_{i-s}, k2 = c_i/b_{i+s},
     a_i' = -a_{i-s} k1, c_i' = -c_{i+s} k2,
     b_i' = b_i - c_{i-s} k1 - a_{i+s} k2, d_i' = d_i - d_{i-s} k1 - d_{i+s} k2.
   Non-cyclic: rows outside 0..N-1 are identity rows (a=c=d=0, b=1); strides
   1..N/2, then x_i = d_i/b_i.  Cyclic (N a power of two): indices mod N,
   strides 1..N/4, after which row i couples only to j = i+N/2:
     x_i = (d_i b_j - e_i d_j) / (b_i b_j - e_i e_j), e = a + c.
   The arrays hold the n_sys systems interleaved (system index fastest). */
static void reduced_pcr(double *ra, double *rb, double *rc, double *rd, int n_sys, int N, int cyclic)
{
    double *na = (double *)malloc(sizeof(double) * (size_t)N), *nb = (double *)malloc(sizeof(double) * (size_t)N);
    double *nc = (double *)malloc(sizeof(double) * (size_t)N), *nd = (double *)malloc(sizeof(double) * (size_t)N);
    double *a = (double *)malloc(sizeof(double) * (size_t)N), *b = (double *)malloc(sizeof(double) * (size_t)N);
    double *c = (double *)malloc(sizeof(double) * (size_t)N), *d = (double *)malloc(sizeof(double) * (size_t)N);
    for (int q = 0; q < n_sys; q++) {
        for (int i = 0; i < N; i++) {
            size_t o = (size_t)q + (size_t)n_sys * (size_t)i;
            a[i] = ra[o]; b[i] = rb[o]; c[i] = rc[o]; d[i] = rd[o];
        }
        const int lim = cyclic ? N / 2 : N;
        for (int s = 1; s < lim; s <<= 1) {
            for (int i = 0; i < N; i++) {
                int im = i - s, ip = i + s;
                double am = 0.0, bm = 1.0, cm = 0.0, dm = 0.0, ap = 0.0, bp = 1.0, cp = 0.0, dp = 0.0;
                if (cyclic) { im = (im + N) % N; ip = ip % N; }
                if (cyclic || im >= 0) { am = a[im]; bm = b[im]; cm = c[im]; dm = d[im]; }
                if (cyclic || ip < N) { ap = a[ip]; bp = b[ip]; cp = c[ip]; dp = d[ip]; }
                const double k1 = a[i] / bm, k2 = c[i] / bp;
                na[i] = -am * k1;
                nc[i] = -cp * k2;
                nb[i] = b[i] - cm * k1 - ap * k2;
                nd[i] = d[i] - dm * k1 - dp * k2;
            }
            memcpy(a, na, sizeof(double) * (size_t)N); memcpy(b, nb, sizeof(double) * (size_t)N);
            memcpy(c, nc, sizeof(double) * (size_t)N); memcpy(d, nd, sizeof(double) * (size_t)N);
        }
        for (int i = 0; i < N; i++) {
            double x;
            if (cyclic) {
                const int j = (i + N / 2) % N;
                const double ei = a[i] + c[i], ej = a[j] + c[j];
                x = (d[i] * b[j] - ei * d[j]) / (b[i] * b[j] - ei * ej);
            } else {
                x = d[i] / b[i];
            }
            nd[i] = x;
        }
        for (int i = 0; i < N; i++) rd[(size_t)q + (size_t)n_sys * (size_t)i] = nd[i];
    }
    free(na); free(nb); free(nc); free(nd); free(a); free(b); free(c); free(d);
}

static void ptdma_many_ranks(double *A, double *B, double *C, double *D,
                             int n_sys, size_t ld, int n_row, int nprocs, int cyclic, int pcr)
{
    const int nr = 2 * nprocs;
    int *j0 = (int *)malloc(sizeof(int) * (size_t)nprocs);
    int *mm = (int *)malloc(sizeof(int) * (size_t)nprocs);
    const int w1 = n_row / nprocs, w2 = n_row % nprocs;
    for (int p = 0; p < nprocs; p++) {
        j0[p] = p * w1 + (p < w2 ? p : w2);
        mm[p] = w1 + (p < w2 ? 1 : 0);
    }
    double *ra = (double *)malloc(sizeof(double) * (size_t)n_sys * nr);
    double *rb = (double *)malloc(sizeof(double) * (size_t)n_sys * nr);
    double *rc = (double *)malloc(sizeof(double) * (size_t)n_sys * nr);
    double *rd = (double *)malloc(sizeof(double) * (size_t)n_sys * nr);
    for (int p = 0; p < nprocs; p++) {
        const int a = j0[p], b = j0[p] + mm[p] - 1;
        ptdma_rank_reduce(A, B, C, D, n_sys, ld, a, mm[p]);
        for (int i = 0; i < n_sys; i++) {
            /* plan%A_rd(i,1:2) .. D_rd, placed at reduced rows 2p, 2p+1 */
            size_t q0 = (size_t)i + (size_t)n_sys * (size_t)(2 * p), q1 = q0 + (size_t)n_sys;
            ra[q0] = P2(A, i, a); ra[q1] = P2(A, i, b);
            rb[q0] = 1.0;         rb[q1] = 1.0;
            rc[q0] = P2(C, i, a); rc[q1] = P2(C, i, b);
            rd[q0] = P2(D, i, a); rd[q1] = P2(D, i, b);
        }
    }
    if (pcr) reduced_pcr(ra, rb, rc, rd, n_sys, nr, cyclic);
    else if (cyclic) orc_tdma_cycl_many(ra, rb, rc, rd, n_sys, nr);
    else orc_tdma_many(ra, rb, rc, rd, n_sys, nr);
    for (int p = 0; p < nprocs; p++) {
        const int a = j0[p], b = j0[p] + mm[p] - 1;
        for (int i = 0; i < n_sys; i++) {
            size_t q0 = (size_t)i + (size_t)n_sys * (size_t)(2 * p), q1 = q0 + (size_t)n_sys;
            P2(D, i, a) = rd[q0];
            P2(D, i, b) = rd[q1];
        }
        for (int j = a + 1; j < b; j++)
            for (int i = 0; i < n_sys; i++)
                P2(D, i, j) = P2(D, i, j) - P2(A, i, j) * P2(D, i, a) - P2(C, i, j) * P2(D, i, b);
    }
    free(ra); free(rb); free(rc); free(rd); free(j0); free(mm);
}
#undef P2

void orc_ptdma_many_solve_ranks(double *A, double *B, double *C, double *D, int n_sys, int n_row,
                                int nprocs, int cyclic)
{
    ptdma_many_ranks(A, B, C, D, n_sys, (size_t)n_sys, n_row, nprocs, cyclic, 0);
}

void orc_ptdma_many_solve_split(double *A, double *B, double *C, double *D, int n_sys, int n_row, int cyclic)
{
    const int P = orc_tdma_parts(n_row);
    ptdma_many_ranks(A, B, C, D, n_sys, (size_t)n_sys, n_row, P, cyclic, P > 1);
}

void orc_ptdma_many_solve(double *A, double *B, double *C, double *D, int n_sys, int n_row)
{
    ptdma_many_ranks(A, B, C, D, n_sys, (size_t)n_sys, n_row, 1, 0, 0);
}

void orc_ptdma_many_solve_cycle(double *A, double *B, double *C, double *D, int n_sys, int n_row)
{
    ptdma_many_ranks(A, B, C, D, n_sys, (size_t)n_sys, n_row, 1, 1, 0);
}

/* Parts per line of the split solve: lines of 2..512 rows are split into
   the power of two P >= ceil(n_row/16) parts (at most 16 rows each, at
   least 8 once there are two parts); longer lines are solved on one rank. */
int orc_tdma_parts(int n_row)
{
    if (n_row < 2 || n_row > 512) return 1;
    int P = 1;
    while (P * 16 < n_row) P <<= 1;
    return P;
}

void orc_set_tdma_split(orc *o, int split) { o->tdma_split = split; }

/* the step's line solves: single rank, or split as above */
static void lines_solve(orc *o, double *A, double *B, double *C, double *D, int n_sys, int n_row, int cyclic)
{
    struct orc_cap *r = NULL;
    size_t nb = (size_t)n_sys * (size_t)n_row * sizeof(double);
    if (o->cap_on && o->ncap < ORC_MAXCAP) {
        r = &o->cap[o->ncap++];
        r->n_sys = n_sys; r->n_row = n_row; r->cyclic = cyclic;
        r->a = malloc(nb); r->b = malloc(nb); r->c = malloc(nb); r->d = malloc(nb); r->x = malloc(nb);
        memcpy(r->a, A, nb); memcpy(r->b, B, nb); memcpy(r->c, C, nb); memcpy(r->d, D, nb);
    }
#ifdef _OPENMP
    if (n_sys >= 1024) {
        /* independent systems: chunks of them to the threads */
        const int P = o->tdma_split ? orc_tdma_parts(n_row) : 1;
        const int nch = 8 * omp_get_max_threads();
        const int w = (n_sys + nch - 1) / nch;
        OMP(omp parallel for schedule(dynamic))
        for (int q = 0; q < nch; q++) {
            const int s0 = q * w, len = s0 + w <= n_sys ? w : n_sys - s0;
            if (len > 0)
                ptdma_many_ranks(A + s0, B + s0, C + s0, D + s0, len, (size_t)n_sys, n_row, P, cyclic,
                                 o->tdma_split && P > 1);
        }
    } else
#endif
    if (o->tdma_split) orc_ptdma_many_solve_split(A, B, C, D, n_sys, n_row, cyclic);
    else ptdma_many_ranks(A, B, C, D, n_sys, (size_t)n_sys, n_row, 1, cyclic, 0);
    if (r) memcpy(r->x, D, nb);
}

static void cap_clear(orc *o)
{
    for (int q = 0; q < o->ncap; q++) {
        free(o->cap[q].a); free(o->cap[q].b); free(o->cap[q].c); free(o->cap[q].d); free(o->cap[q].x);
    }
    o->ncap = 0;
}

void orc_capture(orc *o, int on)
{
    cap_clear(o);
    o->cap_on = on;
}

int orc_capture_count(orc *o) { return o->ncap; }

int orc_capture_info(orc *o, int idx, int *n_sys, int *n_row, int *cyclic)
{
    if (idx < 0 || idx >= o->ncap) return -1;
    *n_sys = o->cap[idx].n_sys; *n_row = o->cap[idx].n_row; *cyclic = o->cap[idx].cyclic;
    return 0;
}

int orc_capture_copy(orc *o, int idx, double *a, double *b, double *c, double *d, double *x)
{
    if (idx < 0 || idx >= o->ncap) return -1;
    size_t nb = (size_t)o->cap[idx].n_sys * (size_t)o->cap[idx].n_row * sizeof(double);
    memcpy(a, o->cap[idx].a, nb); memcpy(b, o->cap[idx].b, nb); memcpy(c, o->cap[idx].c, nb);
    memcpy(d, o->cap[idx].d, nb); memcpy(x, o->cap[idx].x, nb);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* halo: src/module_mpi_subdomain.f90:758-791 on one rank (x, y, z order)    */
/* ------------------------------------------------------------------------ */
void orc_ghostcell_update(orc *o, double *f)
{
    /* x: periodic self-exchange of full (j,k) planes */
    for (int k = 0; k <= o->n3sub; k++)
        for (int j = 0; j <= o->n2sub; j++) {
            f[IX(0, j, k)] = f[IX(o->n1sub - 1, j, k)];
            f[IX(o->n1sub, j, k)] = f[IX(1, j, k)];
        }
    /* y: walls, MPI_PROC_NULL -> no-op */
    /* z: periodic self-exchange of full (i,j) planes */
    for (int j = 0; j <= o->n2sub; j++)
        for (int i = 0; i <= o->n1sub; i++) {
            f[IX(i, j, 0)] = f[IX(i, j, o->n3sub - 1)];
            f[IX(i, j, o->n3sub)] = f[IX(i, j, 1)];
        }
}

/* ------------------------------------------------------------------------ */
/* initial BC rows: module_solve_thermal.f90:88-110, momentum.f90:105-140    */
/* ------------------------------------------------------------------------ */
void orc_thermal_boundary(orc *o)
{
    double *T = o->f[OF_T];
    for (int k = 0; k <= o->n3sub; k++)
        for (int i = 0; i <= o->n1sub; i++)
            T[IX(i, 0, k)] = o->bc[OB_TBT][B2(i, k)];
    /* single rank: rank 0 of comm_1d_x2 takes the first branch only
       (if ... else if, module_solve_thermal.f90:102-110), so the top row of
       T is not written here. */
}

void orc_momentum_boundary(orc *o)
{
    double *U = o->f[OF_U], *V = o->f[OF_V], *W = o->f[OF_W];
    for (int k = 0; k <= o->n3sub; k++)
        for (int i = 0; i <= o->n1sub; i++) {
            U[IX(i, 0, k)] = o->bc[OB_UBT][B2(i, k)];
            W[IX(i, 0, k)] = o->bc[OB_WBT][B2(i, k)];
            V[IX(i, 0, k)] = o->bc[OB_VBT][B2(i, k)];
            V[IX(i, 1, k)] = o->bc[OB_VBT][B2(i, k)];
        }
}

/* ------------------------------------------------------------------------ */
/* thermal: src/module_solve_thermal.f90:116-156                             */
/* ------------------------------------------------------------------------ */
void orc_thermal_coeffi(orc *o)
{
    const double *p = o->prm;
    double *T = o->f[OF_T];
    OMP(omp parallel for)
    for (int k = 0; k <= o->n3sub; k++)
        for (int j = 0; j <= o->n2sub; j++)
            for (int i = 0; i <= o->n1sub; i++) {
                double tmp1 = p[OP_DELTAT] * T[IX(i, j, k)];
                double tmp2 = tmp1 * tmp1;
                double tmp3 = tmp2 * tmp1;
                double tmp4 = tmp3 * tmp1;
                double tmp5 = tmp4 * tmp1;
                o->f[OF_KPP][IX(i, j, k)] = p[OP_C10] * (1.0 + p[OP_C11] * tmp1 + p[OP_C12] * tmp2 + p[OP_C13] * tmp3 + p[OP_C14] * tmp4 + p[OP_C15] * tmp5) / p[OP_KPP0];
                o->f[OF_DKPP][IX(i, j, k)] = p[OP_C10] * (p[OP_C11] + 2.0 * p[OP_C12] * tmp1 + 3.0 * p[OP_C13] * tmp2 + 4.0 * p[OP_C14] * tmp3 + 5.0 * p[OP_C15] * tmp4) / p[OP_KPP0];
                o->f[OF_INVRHOCP][IX(i, j, k)] = p[OP_RHO0] * p[OP_CP0]
                    / (p[OP_A10] + p[OP_A10] * (p[OP_A11] * tmp1 + p[OP_A12] * tmp2 + p[OP_A13] * tmp3 + p[OP_A14] * tmp4 + p[OP_A15] * tmp5))
                    / (p[OP_B10] + p[OP_B10] * (p[OP_B11] * tmp1 + p[OP_B12] * tmp2 + p[OP_B13] * tmp3 + p[OP_B14] * tmp4 + p[OP_B15] * tmp5));
                double da = 1.0 + p[OP_A11] * tmp1 + p[OP_A12] * tmp2 + p[OP_A13] * tmp3;
                double db = 1.0 + p[OP_B11] * tmp1 + p[OP_B12] * tmp2 + p[OP_B13] * tmp3;
                o->f[OF_DINVRHOCP][IX(i, j, k)] = -(p[OP_B11] * (1.0 + 3.0 * p[OP_A12] * tmp2 + 4.0 * p[OP_A13] * tmp3)
                                                   + p[OP_A11] * (1.0 + 2.0 * p[OP_B11] * tmp1 + 3.0 * p[OP_B12] * tmp2 + 4.0 * p[OP_B13] * tmp3)
                                                   + tmp1 * (p[OP_B12] * (2.0 + 5.0 * p[OP_A13] * tmp3)
                                                             + p[OP_A12] * (2.0 + 4.0 * p[OP_B12] * tmp2 + 5.0 * p[OP_B13] * tmp3)
                                                             + 3.0 * tmp1 * (p[OP_A13] + p[OP_B13] + 2.0 * p[OP_A13] * p[OP_B13] * tmp3)))
                                                  / (da * da) / (db * db);
            }
}

/* src/module_solve_thermal.f90:175-532 */
void orc_thermal_solver(orc *o, double dt)
{
    const double Ct = o->prm[OP_CT];
    const int n1msub = o->n1msub, n2msub = o->n2msub, n3msub = o->n3msub;
    double *T = o->f[OF_T];
    const double *U = o->f[OF_U], *V = o->f[OF_V], *W = o->f[OF_W];
    const double *KPP = o->f[OF_KPP], *dKPP = o->f[OF_DKPP];
    const double *invRhoCp = o->f[OF_INVRHOCP], *dinvRhoCp = o->f[OF_DINVRHOCP];
    const double *dx1 = o->dx1, *dx2 = o->dx2, *dx3 = o->dx3;
    const double *dmx1 = o->dmx1, *dmx2 = o->dmx2, *dmx3 = o->dmx3;
    const double *TBCup = o->bc[OB_TUP], *TBCbt = o->bc[OB_TBT];
    const int *jC_BC = o->jC_BC;

    int i, j, k, im = 0, jm = 0, km = 0, ip = 0, jp = 0, kp = 0, jep = 0, jem = 0;
    double e1, e2, e3, e4, e5, e6;
    double dedx1, dedx2, dedy3, dedy4, dedz5, dedz6;
    double convect_e1, convect_e2, convect_e3, viscous_e1, viscous_e2, viscous_e3;
    double kPP1, kPP2, kPP3, kPP4, kPP5, kPP6, dkPP1, dkPP2, dkPP3, dkPP4, dkPP5, dkPP6;
    double ebc_up, ebc_down, eACK, eAPK, eAMK, eACI, eAPI, eAMI, eACJ, eAPJ, eAMJ;
    double invRhoCpCt_half, dinvRhoCpCt_half;

    size_t n = (size_t)n1msub * n2msub * n3msub;
    double *RHS = (double *)calloc(n, sizeof(double));
    double *am = (double *)calloc(n, sizeof(double));
    double *ac = (double *)calloc(n, sizeof(double));
    double *ap = (double *)calloc(n, sizeof(double));

    /* 1st stage: z (:236-361) */
    OMP(omp parallel for private(i, j, im, ip, jep, jem, e1, e2, e3, e4, e5, e6, dedx1, dedx2, dedy3, dedy4, dedz5, dedz6, convect_e1, convect_e2, convect_e3, viscous_e1, viscous_e2, viscous_e3, kPP1, kPP2, kPP3, kPP4, kPP5, kPP6, dkPP1, dkPP2, dkPP3, dkPP4, dkPP5, dkPP6, ebc_up, ebc_down, eACK, eAPK, eAMK, eACI, eAPI, eAMI, eACJ, eAPJ, eAMJ, invRhoCpCt_half, dinvRhoCpCt_half) lastprivate(jm, jp, km, kp))
    for (k = 1; k <= n3msub; k++) {
        km = k - 1; kp = k + 1;
        for (j = 1; j <= n2msub; j++) {
            jm = j - 1; jp = j + 1;
            jem = jC_BC[jm]; jep = jC_BC[jp];
            for (i = 1; i <= n1msub; i++) {
                ip = i + 1; im = i - 1;
                invRhoCpCt_half = 0.5 * Ct * invRhoCp[IX(i, j, k)];
                dinvRhoCpCt_half = 0.5 * Ct * dinvRhoCp[IX(i, j, k)];

                e1 = (dx1[i] * T[IX(im, j, k)] + dx1[im] * T[IX(i, j, k)]) * (0.5 / dmx1[i]);
                e2 = (dx1[ip] * T[IX(i, j, k)] + dx1[i] * T[IX(ip, j, k)]) * (0.5 / dmx1[ip]);
                e3 = PLANT(102, (dx2[jm] * T[IX(i, jm, k)] + dx2[j] * T[IX(i, j, k)]) * (0.5 / dmx2[j]),
                           (dx2[j] * T[IX(i, jm, k)] + dx2[jm] * T[IX(i, j, k)]) * (0.5 / dmx2[j]));
                e4 = (dx2[jp] * T[IX(i, j, k)] + dx2[j] * T[IX(i, jp, k)]) * (0.5 / dmx2[jp]);
                e5 = (dx3[k] * T[IX(i, j, km)] + dx3[km] * T[IX(i, j, k)]) * (0.5 / dmx3[k]);
                e6 = (dx3[kp] * T[IX(i, j, k)] + dx3[k] * T[IX(i, j, kp)]) * (0.5 / dmx3[kp]);
                dedx1 = (T[IX(i, j, k)] - T[IX(im, j, k)]) / dmx1[i];
                dedx2 = (T[IX(ip, j, k)] - T[IX(i, j, k)]) / dmx1[ip];
                dedy3 = (T[IX(i, j, k)] - T[IX(i, jm, k)]) / dmx2[j];
                dedy4 = (T[IX(i, jp, k)] - T[IX(i, j, k)]) / dmx2[jp];
                dedz5 = (T[IX(i, j, k)] - T[IX(i, j, km)]) / dmx3[k];
                dedz6 = (T[IX(i, j, kp)] - T[IX(i, j, k)]) / dmx3[kp];

                convect_e1 = (U[IX(i, j, k)] * dedx1 + U[IX(ip, j, k)] * dedx2) * 0.5;
                convect_e2 = (V[IX(i, j, k)] * dedy3 + V[IX(i, jp, k)] * dedy4) * 0.5;
                convect_e3 = (W[IX(i, j, k)] * dedz5 + W[IX(i, j, kp)] * dedz6) * 0.5;
                double rhs = -0.5 * (convect_e1 + convect_e2 + convect_e3);

                kPP1 = 0.5 / dmx1[i] * (dx1[i] * KPP[IX(im, j, k)] + dx1[im] * KPP[IX(i, j, k)]);
                kPP2 = 0.5 / dmx1[ip] * (dx1[ip] * KPP[IX(i, j, k)] + dx1[i] * KPP[IX(ip, j, k)]);
                kPP3 = 0.5 / dmx2[j] * (dx2[j] * KPP[IX(i, jm, k)] + dx2[jm] * KPP[IX(i, j, k)]);
                kPP4 = 0.5 / dmx2[jp] * (dx2[jp] * KPP[IX(i, j, k)] + dx2[j] * KPP[IX(i, jp, k)]);
                kPP5 = 0.5 / dmx3[k] * (dx3[k] * KPP[IX(i, j, km)] + dx3[km] * KPP[IX(i, j, k)]);
                kPP6 = 0.5 / dmx3[kp] * (dx3[kp] * KPP[IX(i, j, k)] + dx3[k] * KPP[IX(i, j, PLANT(101, km, kp))]);

                dkPP1 = 0.5 / dmx1[i] * (dx1[i] * dKPP[IX(im, j, k)] + dx1[im] * dKPP[IX(i, j, k)]);
                dkPP2 = 0.5 / dmx1[ip] * (dx1[ip] * dKPP[IX(i, j, k)] + dx1[i] * dKPP[IX(ip, j, k)]);
                dkPP3 = 0.5 / dmx2[j] * (dx2[j] * dKPP[IX(i, jm, k)] + dx2[jm] * dKPP[IX(i, j, k)]);
                dkPP4 = 0.5 / dmx2[jp] * (dx2[jp] * dKPP[IX(i, j, k)] + dx2[j] * dKPP[IX(i, jp, k)]);
                dkPP5 = 0.5 / dmx3[k] * (dx3[k] * dKPP[IX(i, j, km)] + dx3[km] * dKPP[IX(i, j, k)]);
                dkPP6 = 0.5 / dmx3[kp] * (dx3[kp] * dKPP[IX(i, j, k)] + dx3[k] * dKPP[IX(i, j, kp)]);

                viscous_e1 = (kPP2 * dedx2 - kPP1 * dedx1 - e2 * dkPP2 * dedx2 + e1 * dkPP1 * dedx1) / dx1[i];
                viscous_e2 = (kPP4 * dedy4 - kPP3 * dedy3 - e4 * dkPP4 * dedy4 + e3 * dkPP3 * dedy3) / dx2[j];
                viscous_e3 = (kPP6 * dedz6 - kPP5 * dedz5 - e6 * dkPP6 * dedz6 + e5 * dkPP5 * dedz5) / dx3[k];

                rhs = rhs
                    + invRhoCpCt_half * (viscous_e1 + viscous_e2 + viscous_e3)
                    - dinvRhoCpCt_half * T[IX(i, j, k)] * ((kPP2 * dedx2 - kPP1 * dedx1) / dx1[i]
                                                         + (kPP4 * dedy4 - kPP3 * dedy3) / dx2[j]
                                                         + (kPP6 * dedz6 - kPP5 * dedz5) / dx3[k]);

                ebc_up = -0.25 * V[IX(i, jp, k)] / dmx2[jp] * TBCup[B2(i, k)];
                ebc_down = PLANT(103, -0.25, 0.25) * V[IX(i, j, k)] / dmx2[j] * TBCbt[B2(i, k)];
                ebc_up = ebc_up
                       + invRhoCpCt_half / dx2[j] * kPP4 / dmx2[jp] * TBCup[B2(i, k)]
                       + invRhoCpCt_half / dx2[j] * dkPP4 * dedy4 * TBCup[B2(i, k)];
                ebc_down = ebc_down
                         + invRhoCpCt_half / dx2[j] * kPP3 / dmx2[j] * TBCbt[B2(i, k)]
                         - invRhoCpCt_half / dx2[j] * dkPP3 * dedy3 * TBCbt[B2(i, k)];
                rhs = rhs + (double)(1 - jem) * ebc_down + (double)(1 - jep) * ebc_up;

                eACK = invRhoCpCt_half / dx3[k] * (kPP6 / dmx3[kp] + kPP5 / dmx3[k]
                                                   - dx3[kp] * (0.5 / dmx3[kp]) * dkPP6 * dedz6
                                                   + dx3[km] * (0.5 / dmx3[k]) * dkPP5 * dedz5)
                     + (-0.25 * W[IX(i, j, kp)] / dmx3[kp] + 0.25 * W[IX(i, j, k)] / dmx3[k]);
                eAPK = -invRhoCpCt_half / dx3[k] * (kPP6 / dmx3[kp] + dx3[k] * (0.5 / dmx3[kp]) * dkPP6 * dedz6) + 0.25 * W[IX(i, j, kp)] / dmx3[kp];
                eAMK = -invRhoCpCt_half / dx3[k] * (kPP5 / dmx3[k] - dx3[k] * (0.5 / dmx3[k]) * dkPP5 * dedz5) - 0.25 * W[IX(i, j, k)] / dmx3[k];

                eACI = invRhoCpCt_half / dx1[i] * (kPP2 / dmx1[ip] + kPP1 / dmx1[i]
                                                   - dx1[ip] * (0.5 / dmx1[ip]) * dkPP2 * dedx2
                                                   + dx1[im] * (0.5 / dmx1[i]) * dkPP1 * dedx1)
                     + (0.25 * U[IX(i, j, k)] / dmx1[i] - 0.25 * U[IX(ip, j, k)] / dmx1[ip]);
                eAPI = -invRhoCpCt_half / dx1[i] * (kPP2 / dmx1[ip] + dx1[i] * (0.5 / dmx1[ip]) * dkPP2 * dedx2) + 0.25 * U[IX(ip, j, k)] / dmx1[ip];
                eAMI = -invRhoCpCt_half / dx1[i] * (kPP1 / dmx1[i] - dx1[i] * (0.5 / dmx1[i]) * dkPP1 * dedx1) - 0.25 * U[IX(i, j, k)] / dmx1[i];

                eACJ = invRhoCpCt_half / dx2[j] * (kPP4 / dmx2[jp] + kPP3 / dmx2[j]
                                                   - dx2[jp] * (0.5 / dmx2[jp]) * dkPP4 * dedy4 * (double)jep
                                                   + dx2[jm] * (0.5 / dmx2[j]) * dkPP3 * dedy3 * (double)jem)
                     + (0.25 * V[IX(i, j, k)] / dmx2[j] - 0.25 * V[IX(i, jp, k)] / dmx2[jp]);
                eAPJ = -invRhoCpCt_half / dx2[j] * (kPP4 / dmx2[jp] + dx2[j] * (0.5 / dmx2[jp]) * dkPP4 * dedy4) + 0.25 * V[IX(i, jp, k)] / dmx2[jp];
                eAMJ = -invRhoCpCt_half / dx2[j] * (kPP3 / dmx2[j] - dx2[j] * (0.5 / dmx2[j]) * dkPP3 * dedy3) - 0.25 * V[IX(i, j, k)] / dmx2[j];
                eAPJ = eAPJ * (double)jep;
                eAMJ = eAMJ * (double)jem;

                rhs = rhs
                    - (eAPK * T[IX(i, j, kp)] + eACK * T[IX(i, j, k)] + eAMK * T[IX(i, j, km)]
                       + eAPJ * T[IX(i, jp, k)] + eACJ * T[IX(i, j, k)] + eAMJ * T[IX(i, jm, k)]
                       + eAPI * T[IX(ip, j, k)] + eACI * T[IX(i, j, k)] + eAMI * T[IX(im, j, k)]
                       - dinvRhoCpCt_half * T[IX(i, j, k)] * ((kPP2 * dedx2 - kPP1 * dedx1) / dx1[i]
                                                            + (kPP4 * dedy4 - kPP3 * dedy3) / dx2[j]
                                                            + (kPP6 * dedz6 - kPP5 * dedz5) / dx3[k]));

                size_t q = R3(n1msub, n2msub, i, j, k);
                ac[q] = eACK * dt
                      - 0.5 * dinvRhoCpCt_half * ((kPP2 * dedx2 - kPP1 * dedx1) / dx1[i]
                                                  + (kPP4 * dedy4 - kPP3 * dedy3) / dx2[j]
                                                  + (kPP6 * dedz6 - kPP5 * dedz5) / dx3[k]) * dt
                      + 1.0;
                ap[q] = eAPK * dt;
                am[q] = eAMK * dt;
                RHS[q] = rhs * dt;
            }
        }
    }
    lines_solve(o, am, ac, ap, RHS, n1msub * n2msub, n3msub, 1);

    /* transpose (i,j,k) -> (j,k,i) (:371-379) */
    double *RHSI = (double *)calloc(n, sizeof(double));
    OMP(omp parallel for private(i))
    for (k = 1; k <= n3msub; k++)
        for (i = 1; i <= n1msub; i++)
            for (int jj = 1; jj <= n2msub; jj++)
                RHSI[R3(n2msub, n3msub, jj, k, i)] = RHS[R3(n1msub, n2msub, i, jj, k)];

    /* 2nd stage: x (:388-429); jm,jp,km,kp keep their stage-1 end values */
    OMP(omp parallel for private(j, i, ip, im, e1, e2, e3, e4, e5, e6, dedx1, dedx2, dedy3, dedy4, dedz5, dedz6, convect_e1, convect_e2, convect_e3, viscous_e1, viscous_e2, viscous_e3, kPP1, kPP2, kPP3, kPP4, kPP5, kPP6, dkPP1, dkPP2, dkPP3, dkPP4, dkPP5, dkPP6, ebc_up, ebc_down, eACK, eAPK, eAMK, eACI, eAPI, eAMI, eACJ, eAPJ, eAMJ, invRhoCpCt_half, dinvRhoCpCt_half))
    for (k = 1; k <= n3msub; k++)
        for (j = 1; j <= n2msub; j++)
            for (i = 1; i <= n1msub; i++) {
                ip = i + 1; im = i - 1;
                if (PLANT(104, 1, 0)) { jm = j - 1; jp = j + 1; km = k - 1; kp = k + 1; }
                invRhoCpCt_half = 0.5 * Ct * invRhoCp[IX(i, j, k)];
                kPP1 = 0.5 / dmx1[i] * (dx1[i] * KPP[IX(im, j, k)] + dx1[im] * KPP[IX(i, j, k)]);
                kPP2 = 0.5 / dmx1[ip] * (dx1[ip] * KPP[IX(i, j, k)] + dx1[i] * KPP[IX(ip, j, k)]);
                kPP3 = 0.5 / dmx2[j] * (dx2[j] * KPP[IX(i, jm, k)] + dx2[jm] * KPP[IX(i, j, k)]);
                kPP4 = 0.5 / dmx2[jp] * (dx2[jp] * KPP[IX(i, j, k)] + dx2[j] * KPP[IX(i, jp, k)]);
                kPP5 = 0.5 / dmx3[k] * (dx3[k] * KPP[IX(i, j, km)] + dx3[km] * KPP[IX(i, j, k)]);
                kPP6 = 0.5 / dmx3[kp] * (dx3[kp] * KPP[IX(i, j, k)] + dx3[k] * KPP[IX(i, j, kp)]);
                dkPP1 = 0.5 / dmx1[i] * (dx1[i] * dKPP[IX(im, j, k)] + dx1[im] * dKPP[IX(i, j, k)]);
                dkPP2 = 0.5 / dmx1[ip] * (dx1[ip] * dKPP[IX(i, j, k)] + dx1[i] * dKPP[IX(ip, j, k)]);
                dedx1 = (T[IX(i, j, k)] - T[IX(im, j, k)]) / dmx1[i];
                dedx2 = (T[IX(ip, j, k)] - T[IX(i, j, k)]) / dmx1[ip];
                dedy3 = (T[IX(i, j, k)] - T[IX(i, jm, k)]) / dmx2[j];
                dedy4 = (T[IX(i, jp, k)] - T[IX(i, j, k)]) / dmx2[jp];
                dedz5 = (T[IX(i, j, k)] - T[IX(i, j, km)]) / dmx3[k];
                dedz6 = (T[IX(i, j, kp)] - T[IX(i, j, k)]) / dmx3[kp];
                eACI = invRhoCpCt_half / dx1[i] * (kPP2 / dmx1[ip] + kPP1 / dmx1[i]
                                                   - dx1[ip] * (0.5 / dmx1[ip]) * dkPP2 * dedx2
                                                   + dx1[im] * (0.5 / dmx1[i]) * dkPP1 * dedx1)
                     + (0.25 * U[IX(i, j, k)] / dmx1[i] - 0.25 * U[IX(ip, j, k)] / dmx1[ip])
                     - 0.5 * PLANT(105, 0.5 * Ct * dinvRhoCp[IX(i, j, k)], invRhoCpCt_half) * ((kPP2 * dedx2 - kPP1 * dedx1) / dx1[i]
                                                + (kPP4 * dedy4 - kPP3 * dedy3) / dx2[j]
                                                + (kPP6 * dedz6 - kPP5 * dedz5) / dx3[k]);
                eAPI = -invRhoCpCt_half / dx1[i] * (kPP2 / dmx1[ip] + dx1[i] * (0.5 / dmx1[ip]) * dkPP2 * dedx2) + 0.25 * U[IX(ip, j, k)] / dmx1[ip];
                eAMI = -invRhoCpCt_half / dx1[i] * (kPP1 / dmx1[i] - dx1[i] * (0.5 / dmx1[i]) * dkPP1 * dedx1) - 0.25 * U[IX(i, j, k)] / dmx1[i];
                size_t q = R3(n2msub, n3msub, j, k, i);
                ac[q] = eACI * dt + 1.0;
                ap[q] = eAPI * dt;
                am[q] = eAMI * dt;
            }
    lines_solve(o, am, ac, ap, RHSI, n2msub * n3msub, n1msub, 1);

    /* transpose (j,k,i) -> (i,k,j) (:439-447) */
    OMP(omp parallel for private(k, i))
    for (j = 1; j <= n2msub; j++)
        for (k = 1; k <= n3msub; k++)
            for (i = 1; i <= n1msub; i++)
                RHS[R3(n1msub, n3msub, i, k, j)] = RHSI[R3(n2msub, n3msub, j, k, i)];

    /* 3rd stage: y (:456-507); km,kp keep their stage-1 end values */
    OMP(omp parallel for private(j, i, jm, jp, jem, jep, ip, im, e1, e2, e3, e4, e5, e6, dedx1, dedx2, dedy3, dedy4, dedz5, dedz6, convect_e1, convect_e2, convect_e3, viscous_e1, viscous_e2, viscous_e3, kPP1, kPP2, kPP3, kPP4, kPP5, kPP6, dkPP1, dkPP2, dkPP3, dkPP4, dkPP5, dkPP6, ebc_up, ebc_down, eACK, eAPK, eAMK, eACI, eAPI, eAMI, eACJ, eAPJ, eAMJ, invRhoCpCt_half, dinvRhoCpCt_half))
    for (k = 1; k <= n3msub; k++)
        for (j = 1; j <= n2msub; j++) {
            jm = j - 1; jp = j + 1;
            if (PLANT(106, 1, 0)) { km = k - 1; kp = k + 1; }
            jem = jC_BC[jm]; jep = jC_BC[jp];
            for (i = 1; i <= n1msub; i++) {
                ip = i + 1; im = i - 1;
                invRhoCpCt_half = 0.5 * Ct * invRhoCp[IX(i, j, k)];
                kPP1 = 0.5 / dmx1[i] * (dx1[i] * KPP[IX(im, j, k)] + dx1[im] * KPP[IX(i, j, k)]);
                kPP2 = 0.5 / dmx1[ip] * (dx1[ip] * KPP[IX(i, j, k)] + dx1[i] * KPP[IX(ip, j, k)]);
                kPP3 = 0.5 / dmx2[j] * (dx2[j] * KPP[IX(i, jm, k)] + dx2[jm] * KPP[IX(i, j, k)]);
                kPP4 = 0.5 / dmx2[jp] * (dx2[jp] * KPP[IX(i, j, k)] + dx2[j] * KPP[IX(i, jp, k)]);
                kPP5 = 0.5 / dmx3[k] * (dx3[k] * KPP[IX(i, j, km)] + dx3[km] * KPP[IX(i, j, k)]);
                kPP6 = 0.5 / dmx3[kp] * (dx3[kp] * KPP[IX(i, j, k)] + dx3[k] * KPP[IX(i, j, kp)]);
                dkPP3 = 0.5 / dmx2[j] * (dx2[j] * dKPP[IX(i, jm, k)] + dx2[jm] * dKPP[IX(i, j, k)]);
                dkPP4 = 0.5 / dmx2[jp] * (dx2[jp] * dKPP[IX(i, j, k)] + dx2[j] * dKPP[IX(i, jp, k)]);
                dedx1 = (T[IX(i, j, k)] - T[IX(im, j, k)]) / dmx1[i];
                dedx2 = (T[IX(ip, j, k)] - T[IX(i, j, k)]) / dmx1[ip];
                dedy3 = (T[IX(i, j, k)] - T[IX(i, jm, k)]) / dmx2[j];
                dedy4 = (T[IX(i, jp, k)] - T[IX(i, j, k)]) / dmx2[jp];
                dedz5 = (T[IX(i, j, k)] - T[IX(i, j, km)]) / dmx3[k];
                dedz6 = (T[IX(i, j, kp)] - T[IX(i, j, k)]) / dmx3[kp];
                eACJ = invRhoCpCt_half / dx2[j] * (kPP4 / dmx2[jp] + kPP3 / dmx2[j]
                                                   - dx2[jp] * (0.5 / dmx2[jp]) * dkPP4 * dedy4 * (double)jep
                                                   + dx2[jm] * (0.5 / dmx2[j]) * dkPP3 * dedy3 * (double)jem)
                     + (0.25 * V[IX(i, j, k)] / dmx2[j] - 0.25 * V[IX(i, jp, k)] / dmx2[jp])
                     - 0.5 * invRhoCpCt_half * ((kPP2 * dedx2 - kPP1 * dedx1) / dx1[i]
                                                + (kPP4 * dedy4 - kPP3 * dedy3) / dx2[j]
                                                + (kPP6 * dedz6 - kPP5 * dedz5) / dx3[k]);
                eAPJ = -invRhoCpCt_half / dx2[j] * (kPP4 / dmx2[jp] + dx2[j] * (0.5 / dmx2[jp]) * dkPP4 * dedy4) + 0.25 * V[IX(i, jp, k)] / dmx2[jp];
                eAMJ = -invRhoCpCt_half / dx2[j] * (kPP3 / dmx2[j] - dx2[j] * (0.5 / dmx2[j]) * dkPP3 * dedy3) - 0.25 * V[IX(i, j, k)] / dmx2[j];
                eAPJ = eAPJ * (double)jep;
                eAMJ = eAMJ * (double)jem;
                size_t q = R3(n1msub, n3msub, i, k, j);
                ac[q] = eACJ * dt + 1.0;
                ap[q] = eAPJ * dt;
                am[q] = eAMJ * dt;
            }
        }
    lines_solve(o, am, ac, ap, RHS, n3msub * n1msub, n2msub, 0);

    OMP(omp parallel for private(j, i))
    for (k = 1; k <= n3msub; k++)
        for (j = 1; j <= n2msub; j++)
            for (i = 1; i <= n1msub; i++)
                T[IX(i, j, k)] = T[IX(i, j, k)] + RHS[R3(n1msub, n3msub, i, k, j)];

    free(RHS); free(RHSI); free(am); free(ac); free(ap);
}

/* ------------------------------------------------------------------------ */
/* momentum: src/module_solve_momentum.f90                                   */
/* ------------------------------------------------------------------------ */
/* :145-173, including the Mu(:,j,k)/invRho(:,j,k) row broadcast inside the
   i loop: every i receives the value of i=n1sub. */
void orc_momentum_coeffi(orc *o)
{
    const double *p = o->prm;
    const double *T = o->f[OF_T];
    double *Mu = o->f[OF_MU], *invRho = o->f[OF_INVRHO];
    OMP(omp parallel for)
    for (int k = 0; k <= o->n3sub; k++)
        for (int j = 0; j <= o->n2sub; j++) {
            int i = o->n1sub; /* last iteration wins */
            double tmp1 = p[OP_DELTAT] * T[IX(i, j, k)];
            double tmp2 = tmp1 * tmp1;
            double tmp3 = tmp2 * tmp1;
            double tmp4 = tmp3 * tmp1;
            double tmp5 = tmp4 * tmp1;
            double mu = (p[OP_A10] + p[OP_A10] * (p[OP_A11] * tmp1 + p[OP_A12] * tmp2 + p[OP_A13] * tmp3 + p[OP_A14] * tmp4 + p[OP_A15] * tmp5))
                      * (p[OP_D10] + p[OP_D10] * (p[OP_D11] * tmp1 + p[OP_D12] * tmp2 + p[OP_D13] * tmp3 + p[OP_D14] * tmp4 + p[OP_D15] * tmp5)) / p[OP_MU0];
            double ir = p[OP_RHO0] / (p[OP_A10] + p[OP_A10] * (p[OP_A11] * tmp1 + p[OP_A12] * tmp2 + p[OP_A13] * tmp3 + p[OP_A14] * tmp4 + p[OP_A15] * tmp5));
            for (int ii = 0; ii <= o->n1sub; ii++) {
                Mu[IX(ii, j, k)] = mu;
                invRho[IX(ii, j, k)] = ir;
            }
        }
}

/* :178-208 */
void orc_momentum_pseudoupdateUVW(orc *o)
{
    double *U = o->f[OF_U], *V = o->f[OF_V], *W = o->f[OF_W];
    const double *dU = o->f[OF_DU], *dV = o->f[OF_DV], *dW = o->f[OF_DW];
    OMP(omp parallel for)
    for (int k = 1; k <= o->n3msub; k++) {
        for (int j = 2; j <= o->n2msub; j++)
            for (int i = 1; i <= o->n1msub; i++) {
                U[IX(i, j, k)] = U[IX(i, j, k)] + dU[IX(i, j, k)];
                V[IX(i, j, k)] = V[IX(i, j, k)] + dV[IX(i, j, k)];
                W[IX(i, j, k)] = W[IX(i, j, k)] + dW[IX(i, j, k)];
            }
        int j = 1;
        for (int i = 1; i <= o->n1msub; i++) {
            U[IX(i, j, k)] = U[IX(i, j, k)] + dU[IX(i, j, k)];
            V[IX(i, j, k)] = V[IX(i, j, k)] + dV[IX(i, j, k)] * (double)(2 - o->j_indexS);
            W[IX(i, j, k)] = W[IX(i, j, k)] + dW[IX(i, j, k)];
        }
    }
}

/* :225-563 */
void orc_momentum_solvedU(orc *o, double dt)
{
    const double *p = o->prm;
    const double Cmu = p[OP_CMU], Cmp = p[OP_CMP], Cmt = p[OP_CMT], DeltaT = p[OP_DELTAT], a12pera11 = p[OP_A12PERA11];
    const int n1msub = o->n1msub, n2msub = o->n2msub, n3msub = o->n3msub;
    const double *U = o->f[OF_U], *V = o->f[OF_V], *W = o->f[OF_W], *P = o->f[OF_P], *T = o->f[OF_T];
    const double *Mu = o->f[OF_MU], *invRho = o->f[OF_INVRHO];
    double *dU = o->f[OF_DU];
    const double *dx1 = o->dx1, *dx2 = o->dx2, *dx3 = o->dx3;
    const double *dmx1 = o->dmx1, *dmx2 = o->dmx2, *dmx3 = o->dmx3;
    const double *UBCup = o->bc[OB_UUP], *UBCbt = o->bc[OB_UBT], *VBCup = o->bc[OB_VUP], *VBCbt = o->bc[OB_VBT];
    const int *jC_BC = o->jC_BC;

    int i, j, k, im = 0, jm = 0, km = 0, ip = 0, jp = 0, kp = 0, jup = 0, jum = 0;
    double u1, u2, v3, v4, w5, w6, Tc;
    double dudx1, dudx2, dudy3, dudy4, dudz5, dudz6, dvdx3, dvdx4, dwdx5, dwdx6;
    double mua, mub, muc, mu3 = 0, mu4 = 0, mu5 = 0, mu6 = 0;
    double invRhoc, invRhocCmu_half, viscous_u1, viscous_u2, viscous_u3, viscous_u12, viscous_u13;
    double ubc_up, ubc_down;
    double mACI, mAPI, mAMI, mACJ, mAPJ, mAMJ, mACK, mAPK, mAMK;

    size_t n = (size_t)n1msub * n2msub * n3msub;
    double *RHS = (double *)calloc(n, sizeof(double));
    double *am = (double *)calloc(n, sizeof(double));
    double *ac = (double *)calloc(n, sizeof(double));
    double *ap = (double *)calloc(n, sizeof(double));

    OMP(omp parallel for private(i, j, im, ip, jum, jup, u1, u2, v3, v4, w5, w6, Tc, dudx1, dudx2, dudy3, dudy4, dudz5, dudz6, dvdx3, dvdx4, dwdx5, dwdx6, mua, mub, muc, mu3, mu4, mu5, mu6, invRhoc, invRhocCmu_half, viscous_u1, viscous_u2, viscous_u3, viscous_u12, viscous_u13, ubc_up, ubc_down, mACI, mAPI, mAMI, mACJ, mAPJ, mAMJ, mACK, mAPK, mAMK) lastprivate(jm, jp, km, kp))
    for (k = 1; k <= n3msub; k++) {
        km = k - 1; kp = k + 1;
        for (j = 1; j <= n2msub; j++) {
            jm = j - 1; jp = j + 1;
            jum = jC_BC[jm]; jup = jC_BC[jp];
            for (i = o->i_indexS; i <= n1msub; i++) {
                im = i - 1; ip = i + 1;
                u1 = 0.5 * (U[IX(im, j, k)] + U[IX(i, j, k)]);
                u2 = 0.5 * (U[IX(ip, j, k)] + U[IX(i, j, k)]);
                v3 = 0.5 * (dx1[im] * V[IX(i, j, k)] + dx1[i] * V[IX(im, j, k)]) / dmx1[i];
                v4 = 0.5 * (dx1[im] * V[IX(i, jp, k)] + dx1[i] * V[IX(im, jp, k)]) / dmx1[i];
                w5 = 0.5 * (dx1[im] * W[IX(i, j, k)] + dx1[i] * W[IX(im, j, k)]) / dmx1[i];
                w6 = 0.5 * (dx1[im] * W[IX(i, j, kp)] + dx1[i] * W[IX(im, j, kp)]) / dmx1[i];

                dudx1 = (U[IX(i, j, k)] - U[IX(im, j, k)]) / dx1[im];
                dudx2 = (U[IX(ip, j, k)] - U[IX(i, j, k)]) / dx1[i];
                dudy3 = (U[IX(i, j, k)] - U[IX(i, jm, k)]) / dmx2[j];
                dudy4 = (U[IX(i, jp, k)] - U[IX(i, j, k)]) / dmx2[jp];
                dudz5 = (U[IX(i, j, k)] - U[IX(i, j, km)]) / dmx3[k];
                dudz6 = (U[IX(i, j, kp)] - U[IX(i, j, k)]) / dmx3[kp];
                dvdx3 = (V[IX(i, j, k)] - V[IX(im, j, k)]) / dmx1[i];
                dvdx4 = (V[IX(i, jp, k)] - V[IX(im, jp, k)]) / dmx1[i];
                dwdx5 = (W[IX(i, j, k)] - W[IX(im, j, k)]) / dmx1[i];
                dwdx6 = (W[IX(i, j, kp)] - W[IX(im, j, kp)]) / dmx1[i];

                mua = 0.5 * (dx1[im] * Mu[IX(i, jm, k)] + dx1[i] * Mu[IX(im, jm, k)]) / dmx1[i];
                muc = 0.5 * (dx1[im] * Mu[IX(i, j, k)] + dx1[i] * Mu[IX(im, j, k)]) / dmx1[i];
                mub = 0.5 * (dx1[im] * Mu[IX(i, PLANT(201, jm, jp), k)] + dx1[i] * Mu[IX(im, PLANT(201, jm, jp), k)]) / dmx1[i];
                mu3 = 0.5 * (dx2[jm] * muc + dx2[j] * mua) / dmx2[j];
                mu4 = 0.5 * (dx2[jp] * muc + dx2[j] * mub) / dmx2[jp];
                mua = 0.5 * (dx1[im] * Mu[IX(i, j, km)] + dx1[i] * Mu[IX(im, j, km)]) / dmx1[i];
                muc = 0.5 * (dx1[im] * Mu[IX(i, j, k)] + dx1[i] * Mu[IX(im, j, k)]) / dmx1[i];
                mub = 0.5 * (dx1[im] * Mu[IX(i, j, kp)] + dx1[i] * Mu[IX(im, j, kp)]) / dmx1[i];
                mu5 = 0.5 * (dx3[km] * muc + dx3[k] * mua) / dmx3[k];
                mu6 = 0.5 * (dx3[kp] * muc + dx3[k] * mub) / dmx3[kp];

                invRhoc = 0.5 * (dx1[im] * invRho[IX(i, j, k)] + dx1[i] * invRho[IX(im, j, k)]) / dmx1[i];
                invRhocCmu_half = 0.5 * Cmu * invRhoc;

                viscous_u1 = 1.0 * (Mu[IX(i, j, k)] * dudx2 - Mu[IX(im, j, k)] * dudx1) / dmx1[i];
                viscous_u2 = 1.0 * (mu4 * dudy4 - mu3 * dudy3) / dx2[j];
                viscous_u3 = 1.0 * (mu6 * dudz6 - mu5 * dudz5) / dx3[k];
                viscous_u12 = 1.0 * (mu4 * dvdx4 - mu3 * dvdx3) / dx2[j];
                viscous_u13 = 1.0 * (mu6 * dwdx6 - mu5 * dwdx5) / dx3[k];

                double rhs = invRhocCmu_half * (2.0 * viscous_u1 + viscous_u2 + viscous_u3 + viscous_u12 + viscous_u13);
                rhs = rhs - Cmp * invRhoc * (P[IX(i, j, k)] - P[IX(im, j, k)]) / dmx1[i];
                Tc = 0.5 * (dx1[im] * T[IX(i, j, k)] + dx1[i] * T[IX(PLANT(202, ip, im), j, k)]) / dmx1[i];
                rhs = rhs + Cmt * (Tc + a12pera11 * (Tc * Tc) * DeltaT) * invRhoc;

                ubc_down = 0.25 * v3 / dmx2[j] * UBCbt[B2(i, k)] - 0.25 * dudy3 * (dx1[i] * VBCbt[B2(im, k)] + dx1[im] * VBCbt[B2(i, k)]) / dmx1[i] * 0.5;
                ubc_up = -0.25 * v4 / dmx2[jp] * UBCup[B2(i, k)] - 0.25 * dudy4 * (dx1[i] * VBCup[B2(im, k)] + dx1[im] * VBCup[B2(i, k)]) / dmx1[i] * 0.5;
                ubc_down = ubc_down + PLANT(203, -1.0, 1.0) * invRhocCmu_half / dx2[j] * mu3 / dmx2[j] * UBCbt[B2(i, k)];
                ubc_down = ubc_down - invRhocCmu_half / dx2[j] * mu3 / dmx1[i] * (VBCbt[B2(i, k)] - VBCbt[B2(im, k)]);
                ubc_up = ubc_up + invRhocCmu_half / dx2[j] * mu4 / dmx2[jp] * UBCup[B2(i, k)];
                ubc_up = ubc_up + invRhocCmu_half / dx2[j] * mu4 / dmx1[i] * (VBCup[B2(i, k)] - VBCup[B2(im, k)]);
                rhs = rhs + (1.0 - (double)jum) * ubc_down + (1.0 - (double)jup) * ubc_up;

                mACI = invRhocCmu_half / dmx1[i] * (Mu[IX(i, j, k)] / dx1[i] + Mu[IX(i, j, k)] / dx1[im]) * 2.0
                     + 0.25 * (dudx2 * 0.5 + dudx1 * 0.5 - u2 / dx1[i] + u1 / dx1[im]);
                mAPI = -invRhocCmu_half / dmx1[i] * Mu[IX(i, j, k)] / dx1[i] * 2.0 + 0.25 * (u2 / dx1[i] + dudx2 * 0.5);
                mAMI = -invRhocCmu_half / dmx1[i] * Mu[IX(i, j, k)] / dx1[im] * 2.0 + 0.25 * (-u1 / dx1[im] + dudx1 * 0.5);

                mACJ = invRhocCmu_half / dx2[j] * (mu4 / dmx2[jp] + mu3 / dmx2[j]) + 0.25 * (-v4 / dmx2[jp] + v3 / dmx2[j]);
                mAPJ = -invRhocCmu_half / dx2[j] * mu4 / dmx2[jp] + 0.25 * (v4 / dmx2[jp]);
                mAMJ = -invRhocCmu_half / dx2[j] * mu3 / dmx2[j] + 0.25 * (-v3 / dmx2[j]);
                mAPJ = mAPJ * (double)jup;
                mAMJ = mAMJ * (double)jum;

                mACK = invRhocCmu_half / dx3[k] * (mu6 / dmx3[kp] + mu5 / dmx3[k]) + 0.25 * (-w6 / dmx3[kp] + w5 / dmx3[k]);
                mAPK = -invRhocCmu_half / dx3[k] * mu6 / dmx3[kp] + 0.25 * (w6 / dmx3[kp]);
                mAMK = -invRhocCmu_half / dx3[k] * mu5 / dmx3[k] + 0.25 * (-w5 / dmx3[k]);

                rhs = rhs - (mAPI * U[IX(ip, j, k)] + mACI * U[IX(i, j, k)] + mAMI * U[IX(im, j, k)]
                             + mAPJ * U[IX(i, jp, k)] + mACJ * U[IX(i, j, k)] + mAMJ * U[IX(i, jm, k)]
                             + mAPK * U[IX(i, j, kp)] + mACK * U[IX(i, j, k)] + mAMK * U[IX(i, j, km)]);
                rhs = rhs - (0.25 * (v4 * (double)jup * dudy4 + v3 * (double)jum * dudy3)
                             - invRhocCmu_half / dx2[j] * (mu4 * dvdx4 * (double)jup - mu3 * dvdx3 * (double)jum));
                rhs = rhs - (0.25 * (w6 * dudz6 + w5 * dudz5)
                             - invRhocCmu_half / dx3[k] * (mu6 * dwdx6 - mu5 * dwdx5));

                size_t q = R3(n1msub, n2msub, i, j, k);
                ac[q] = mACK * dt + 1.0;
                ap[q] = mAPK * dt;
                am[q] = mAMK * dt;
                RHS[q] = rhs * dt;
            }
        }
    }
    lines_solve(o, am, ac, ap, RHS, n1msub * n2msub, n3msub, 1);

    double *RHSI = (double *)calloc(n, sizeof(double));
    OMP(omp parallel for private(i))
    for (k = 1; k <= n3msub; k++)
        for (int jj = 1; jj <= n2msub; jj++)
            for (i = o->i_indexS; i <= n1msub; i++)
                RHSI[R3(n2msub, n3msub, jj, k, i)] = RHS[R3(n1msub, n2msub, i, jj, k)];

    /* 2nd stage x (:443-472) */
    OMP(omp parallel for private(i, im, ip, jum, jup, u1, u2, v3, v4, w5, w6, Tc, dudx1, dudx2, dudy3, dudy4, dudz5, dudz6, dvdx3, dvdx4, dwdx5, dwdx6, mua, mub, muc, mu3, mu4, mu5, mu6, invRhoc, invRhocCmu_half, viscous_u1, viscous_u2, viscous_u3, viscous_u12, viscous_u13, ubc_up, ubc_down, mACI, mAPI, mAMI, mACJ, mAPJ, mAMJ, mACK, mAPK, mAMK))
    for (k = 1; k <= n3msub; k++)
        for (int jj = 1; jj <= n2msub; jj++)
            for (i = o->i_indexS; i <= n1msub; i++) {
                im = i - 1; ip = i + 1;
                u1 = 0.5 * (U[IX(im, jj, k)] + U[IX(i, jj, k)]);
                u2 = 0.5 * (U[IX(ip, jj, k)] + U[IX(i, jj, k)]);
                dudx1 = (U[IX(i, jj, k)] - U[IX(im, jj, k)]) / dx1[im];
                dudx2 = (U[IX(ip, jj, k)] - U[IX(i, jj, k)]) / dx1[i];
                invRhoc = 0.5 * (dx1[im] * invRho[IX(i, jj, k)] + dx1[i] * invRho[IX(im, jj, k)]) / dmx1[i];
                invRhocCmu_half = 0.5 * Cmu * invRhoc;
                mACI = invRhocCmu_half / dmx1[i] * (Mu[IX(i, jj, k)] / dx1[i] + Mu[IX(i, jj, k)] / dx1[im]) * 2.0
                     + 0.25 * (dudx2 * 0.5 + dudx1 * 0.5 - u2 / dx1[i] + u1 / dx1[im]);
                mAPI = -invRhocCmu_half / dmx1[i] * Mu[IX(i, jj, k)] / dx1[i] * 2.0 + 0.25 * (u2 / dx1[i] + dudx2 * 0.5);
                mAMI = -invRhocCmu_half / dmx1[i] * Mu[IX(i, jj, k)] / dx1[im] * 2.0 + 0.25 * (PLANT(204, 1.0, -1.0) * u1 / dx1[im] + dudx1 * 0.5);
                size_t q = R3(n2msub, n3msub, jj, k, i);
                ac[q] = mACI * dt + 1.0;
                ap[q] = mAPI * dt;
                am[q] = mAMI * dt;
            }
    lines_solve(o, am, ac, ap, RHSI, n2msub * n3msub, n1msub, 1);

    OMP(omp parallel for private(i, k))
    for (int jj = 1; jj <= n2msub; jj++)
        for (i = o->i_indexS; i <= n1msub; i++)
            for (k = 1; k <= n3msub; k++)
                RHS[R3(n1msub, n3msub, i, k, jj)] = RHSI[R3(n2msub, n3msub, jj, k, i)];

    /* 3rd stage y (:499-537) */
    OMP(omp parallel for private(j, i, jm, jp, im, ip, jum, jup, u1, u2, v3, v4, w5, w6, Tc, dudx1, dudx2, dudy3, dudy4, dudz5, dudz6, dvdx3, dvdx4, dwdx5, dwdx6, mua, mub, muc, mu3, mu4, mu5, mu6, invRhoc, invRhocCmu_half, viscous_u1, viscous_u2, viscous_u3, viscous_u12, viscous_u13, ubc_up, ubc_down, mACI, mAPI, mAMI, mACJ, mAPJ, mAMJ, mACK, mAPK, mAMK))
    for (k = 1; k <= n3msub; k++)
        for (j = 1; j <= n2msub; j++) {
            jm = j - 1; jp = j + 1;
            jum = jC_BC[jm]; jup = jC_BC[jp];
            for (i = o->i_indexS; i <= n1msub; i++) {
                im = i - 1; ip = i + 1;
                v3 = 0.5 * (dx1[im] * V[IX(i, j, k)] + dx1[i] * V[IX(im, j, k)]) / dmx1[i];
                v4 = 0.5 * (dx1[im] * V[IX(i, jp, k)] + dx1[i] * V[IX(im, jp, k)]) / dmx1[i];
                mua = 0.5 * (dx1[im] * Mu[IX(i, jm, k)] + dx1[i] * Mu[IX(im, jm, k)]) / dmx1[i];
                muc = 0.5 * (dx1[im] * Mu[IX(i, j, k)] + dx1[i] * Mu[IX(im, j, k)]) / dmx1[i];
                mub = 0.5 * (dx1[im] * Mu[IX(i, jp, k)] + dx1[i] * Mu[IX(im, jp, k)]) / dmx1[i];
                mu3 = 0.5 * (dx2[jm] * muc + dx2[j] * mua) / dmx2[j];
                mu4 = 0.5 * (dx2[jp] * muc + dx2[j] * PLANT(205, mua, mub)) / dmx2[jp];
                invRhoc = 0.5 * (dx1[im] * invRho[IX(i, j, k)] + dx1[i] * invRho[IX(im, j, k)]) / dmx1[i];
                invRhocCmu_half = 0.5 * Cmu * invRhoc;
                mACJ = invRhocCmu_half / dx2[j] * (mu4 / dmx2[jp] + mu3 / dmx2[j]) + 0.25 * (-v4 / dmx2[jp] + v3 / dmx2[j]);
                mAPJ = -invRhocCmu_half / dx2[j] * mu4 / dmx2[jp] + 0.25 * (v4 / dmx2[jp]);
                mAMJ = -invRhocCmu_half / dx2[j] * mu3 / dmx2[j] + 0.25 * (-v3 / dmx2[j]);
                mAPJ = mAPJ * (double)jup;
                mAMJ = mAMJ * (double)jum;
                size_t q = R3(n1msub, n3msub, i, k, j);
                ac[q] = mACJ * dt + 1.0;
                ap[q] = mAPJ * dt;
                am[q] = mAMJ * dt;
            }
        }
    lines_solve(o, am, ac, ap, RHS, n3msub * n1msub, n2msub, 0);

    OMP(omp parallel for private(j, i))
    for (k = 1; k <= n3msub; k++)
        for (j = 1; j <= n2msub; j++)
            for (i = 1; i <= n1msub; i++)
                dU[IX(i, j, k)] = RHS[R3(n1msub, n3msub, i, k, j)];

    free(RHS); free(RHSI); free(am); free(ac); free(ap);
}

/* :569-924 */
void orc_momentum_solvedV(orc *o, double dt)
{
    const double *p = o->prm;
    const double Cmu = p[OP_CMU], Cmp = p[OP_CMP];
    const int n1msub = o->n1msub, n2msub = o->n2msub, n3msub = o->n3msub, j_indexS = o->j_indexS;
    const double *U = o->f[OF_U], *V = o->f[OF_V], *W = o->f[OF_W], *P = o->f[OF_P];
    const double *dU = o->f[OF_DU];
    const double *Mu = o->f[OF_MU], *invRho = o->f[OF_INVRHO];
    double *dV = o->f[OF_DV];
    const double *dx1 = o->dx1, *dx2 = o->dx2, *dx3 = o->dx3;
    const double *dmx1 = o->dmx1, *dmx2 = o->dmx2, *dmx3 = o->dmx3;
    const double *VBCup = o->bc[OB_VUP], *VBCbt = o->bc[OB_VBT];
    const int *jS_BC = o->jS_BC;

    int i, j, k, im = 0, jm = 0, km = 0, ip = 0, jp = 0, kp = 0, jvp = 0, jvm = 0;
    double u1, u2, v3, v4, w5, w6;
    double dudy1, dudy2, dvdx1, dvdx2, dvdy3, dvdy4, dvdz5, dvdz6, dwdy5, dwdy6;
    double ddu1, ddu2, ddudy1, ddudy2;
    double mua, muc, mub, mu1 = 0, mu2 = 0, mu5, mu6;
    double invRhoc, invRhocCmu_half, viscous_v1, viscous_v2, viscous_v3, viscous_v21, viscous_v23;
    double vbc_up, vbc_down;
    double mACI, mAPI, mAMI, mACJ, mAPJ, mAMJ, mACK, mAPK, mAMK;

    size_t n = (size_t)n1msub * n2msub * n3msub;
    double *RHS = (double *)calloc(n, sizeof(double));
    double *am = (double *)calloc(n, sizeof(double));
    double *ac = (double *)calloc(n, sizeof(double));
    double *ap = (double *)calloc(n, sizeof(double));
    /* rows j=1 are never assembled (j starts at j_indexS); the reference
       solves them with uninitialised coefficients and discards them.  Give
       them an identity row so the discarded solves stay finite. */
    OMP(omp parallel for)
    for (size_t q = 0; q < n; q++) ac[q] = 1.0;

    OMP(omp parallel for private(i, j, im, ip, jvp, jvm, u1, u2, v3, v4, w5, w6, dudy1, dudy2, dvdx1, dvdx2, dvdy3, dvdy4, dvdz5, dvdz6, dwdy5, dwdy6, ddu1, ddu2, ddudy1, ddudy2, mua, muc, mub, mu5, mu6, invRhoc, invRhocCmu_half, viscous_v1, viscous_v2, viscous_v3, viscous_v21, viscous_v23, vbc_up, vbc_down, mACI, mAPI, mAMI, mACJ, mAPJ, mAMJ, mACK, mAPK, mAMK) lastprivate(jm, jp, km, kp, mu1, mu2))
    for (k = 1; k <= n3msub; k++) {
        kp = k + 1; km = k - 1;
        for (j = j_indexS; j <= n2msub; j++) {
            jp = j + 1; jm = j - 1;
            jvp = jS_BC[PLANT(211, jp, jm)]; jvm = jS_BC[PLANT(211, jm, jp)];
            for (i = 1; i <= n1msub; i++) {
                ip = i + 1; im = i - 1;
                v3 = 0.5 * (V[IX(i, jm, k)] + V[IX(i, j, k)]);
                v4 = 0.5 * (V[IX(i, j, k)] + V[IX(i, jp, k)]);
                u1 = (dx2[jm] * U[IX(i, j, k)] + dx2[j] * U[IX(i, jm, k)]) / dmx2[j] * 0.5;
                u2 = (dx2[jm] * U[IX(ip, j, k)] + dx2[j] * U[IX(ip, jm, k)]) / dmx2[j] * 0.5;
                w5 = (dx2[jm] * W[IX(i, j, k)] + dx2[j] * W[IX(i, jm, k)]) / dmx2[j] * 0.5;
                w6 = (dx2[jm] * W[IX(i, j, kp)] + dx2[j] * W[IX(i, jm, kp)]) / dmx2[j] * 0.5;
                ddu1 = (dx2[jm] * dU[IX(i, j, k)] + dx2[j] * dU[IX(i, PLANT(213, jp, jm), k)]) / dmx2[j] * 0.5;
                ddu2 = (dx2[jm] * dU[IX(ip, j, k)] + dx2[j] * dU[IX(ip, jm, k)]) / dmx2[j] * 0.5;

                dvdx1 = (V[IX(i, j, k)] - V[IX(im, j, k)]) / dmx1[i];
                dvdx2 = (V[IX(ip, j, k)] - V[IX(i, j, k)]) / dmx1[ip];
                dvdy3 = (V[IX(i, j, k)] - V[IX(i, jm, k)]) / dx2[jm];
                dvdy4 = (V[IX(i, jp, k)] - V[IX(i, j, k)]) / dx2[j];
                dvdz5 = (V[IX(i, j, k)] - V[IX(i, j, km)]) / dmx3[k];
                dvdz6 = (V[IX(i, j, kp)] - V[IX(i, j, k)]) / dmx3[kp];
                dudy1 = (U[IX(i, j, k)] - U[IX(i, jm, k)]) / dmx2[j];
                dudy2 = (U[IX(ip, j, k)] - U[IX(ip, jm, k)]) / dmx2[j];
                dwdy5 = (W[IX(i, j, k)] - W[IX(i, jm, k)]) / dmx2[j];
                dwdy6 = (W[IX(i, j, kp)] - W[IX(i, jm, kp)]) / dmx2[j];
                ddudy1 = (dU[IX(i, j, k)] - dU[IX(i, jm, k)]) / dmx2[j];
                ddudy2 = (dU[IX(ip, j, k)] - dU[IX(ip, jm, k)]) / dmx2[j];

                mua = 0.5 * (dx2[jm] * Mu[IX(im, j, k)] + dx2[j] * Mu[IX(im, jm, k)]) / dmx2[j];
                muc = 0.5 * (dx2[jm] * Mu[IX(i, j, k)] + dx2[j] * Mu[IX(i, jm, k)]) / dmx2[j];
                mub = 0.5 * (dx2[jm] * Mu[IX(ip, j, k)] + dx2[j] * Mu[IX(ip, jm, k)]) / dmx2[j];
                mu1 = 0.5 * (dx1[im] * muc + dx1[i] * mua) / dmx1[i];
                mu2 = 0.5 * (dx1[ip] * muc + dx1[i] * mub) / dmx1[ip];
                mua = 0.5 * (dx2[jm] * Mu[IX(i, j, km)] + dx2[j] * Mu[IX(i, jm, km)]) / dmx2[j];
                muc = 0.5 * (dx2[jm] * Mu[IX(i, j, k)] + dx2[j] * Mu[IX(i, jm, k)]) / dmx2[j];
                mub = 0.5 * (dx2[jm] * Mu[IX(i, j, kp)] + dx2[j] * Mu[IX(i, jm, kp)]) / dmx2[j];
                mu5 = 0.5 * (dx3[km] * muc + dx3[k] * mua) / dmx3[k];
                mu6 = 0.5 * (dx3[kp] * muc + dx3[k] * mub) / dmx3[kp];

                invRhoc = 0.5 * (dx2[jm] * invRho[IX(i, j, k)] + dx2[j] * invRho[IX(i, jm, k)]) / dmx2[j];
                invRhocCmu_half = 0.5 * Cmu * invRhoc;

                viscous_v1 = 1.0 * (mu2 * dvdx2 - mu1 * dvdx1) / dx1[i];
                viscous_v2 = 1.0 * (Mu[IX(i, j, k)] * dvdy4 - Mu[IX(i, jm, k)] * dvdy3) / dmx2[j];
                viscous_v3 = 1.0 * (mu6 * dvdz6 - mu5 * dvdz5) / dx3[k];
                viscous_v21 = 1.0 * (mu2 * dudy2 - mu1 * dudy1) / dx1[i];
                viscous_v23 = 1.0 * (mu6 * dwdy6 - mu5 * dwdy5) / dx3[k];

                double rhs = invRhocCmu_half * (viscous_v1 + 2.0 * viscous_v2 + viscous_v3 + viscous_v21 + viscous_v23);
                rhs = rhs - Cmp * invRhoc * (P[IX(i, j, k)] - P[IX(i, jm, k)]) / dmx2[j];

                vbc_down = (0.25 / dx2[jm] * v3 * VBCbt[B2(i, k)] - 0.25 * 0.5 * dvdy3 * VBCbt[B2(i, k)]);
                vbc_up = (-0.25 / dx2[j] * v4 * VBCup[B2(i, k)] - 0.25 * 0.5 * dvdy4 * VBCup[B2(i, k)]);
                vbc_down = vbc_down + invRhocCmu_half / dmx2[j] * Mu[IX(i, jm, k)] / dx2[jm] * VBCbt[B2(i, k)] * 2.0;
                vbc_up = vbc_up + invRhocCmu_half / dmx2[j] * Mu[IX(i, j, k)] / dx2[j] * VBCup[B2(i, k)] * 2.0;
                rhs = rhs + (1.0 - (double)jvm) * vbc_down + (1.0 - (double)jvp) * vbc_up;

                mACI = invRhocCmu_half / dx1[i] * (mu2 / dmx1[ip] + mu1 / dmx1[i]) + 0.25 * (-u2 / dmx1[ip] + u1 / dmx1[i]);
                mAPI = -invRhocCmu_half / dx1[i] * mu2 / dmx1[ip] + 0.25 * (u2 / dmx1[ip]);
                mAMI = -invRhocCmu_half / dx1[i] * mu1 / dmx1[i] + 0.25 * (-u1 / dmx1[i]);

                mACJ = invRhocCmu_half / dmx2[j] * (Mu[IX(i, j, k)] / dx2[j] + Mu[IX(i, jm, k)] / dx2[jm]) * 2.0
                     + 0.25 * (dvdy4 * 0.5 + dvdy3 * 0.5 - v4 / dx2[j] + v3 / dx2[jm]);
                mAPJ = -invRhocCmu_half / dmx2[j] * Mu[IX(i, j, k)] / dx2[j] * 2.0 + 0.25 * (v4 / dx2[j] + dvdy4 * 0.5);
                mAPJ = mAPJ * (double)jvp;
                mAMJ = -invRhocCmu_half / dmx2[j] * Mu[IX(i, jm, k)] / dx2[jm] * 2.0 + 0.25 * (-v3 / dx2[jm] + dvdy3 * 0.5);
                mAMJ = mAMJ * (double)jvm;

                mACK = invRhocCmu_half / dx3[k] * (mu6 / dmx3[kp] + mu5 / dmx3[k]) + 0.25 * (-w6 / dmx3[kp] + w5 / dmx3[k]);
                mAPK = -invRhocCmu_half / dx3[k] * mu6 / dmx3[kp] + 0.25 * (w6 / dmx3[kp]);
                mAMK = -invRhocCmu_half / dx3[k] * mu5 / dmx3[k] + 0.25 * (-w5 / dmx3[k]);

                rhs = rhs - (mAPI * V[IX(ip, j, k)] + mACI * V[IX(i, j, k)] + mAMI * V[IX(im, j, k)]
                             + mAPJ * V[IX(i, jp, k)] + mACJ * V[IX(i, j, k)] + mAMJ * V[IX(i, jm, k)]
                             + mAPK * V[IX(i, j, kp)] + mACK * V[IX(i, j, k)] + mAMK * V[IX(i, j, km)]);
                rhs = rhs - (0.25 * (u2 * dvdx2 + u1 * dvdx1) - invRhocCmu_half / dx1[i] * (mu2 * dudy2 - mu1 * dudy1));
                rhs = rhs - (0.25 * (w6 * dvdz6 + w5 * dvdz5) - invRhocCmu_half / dx3[k] * (mu6 * dwdy6 - mu5 * dwdy5));
                rhs = rhs - (0.25 * (ddu2 * dvdx2 + ddu1 * dvdx1) - invRhocCmu_half / dx1[i] * (mu2 * ddudy2 - mu1 * ddudy1));

                size_t q = R3(n1msub, n2msub, i, j, k);
                ac[q] = mACK * dt + 1.0;
                ap[q] = mAPK * dt;
                am[q] = mAMK * dt;
                RHS[q] = rhs * dt;
            }
        }
    }
    lines_solve(o, am, ac, ap, RHS, n1msub * n2msub, n3msub, 1);

    double *RHSI = (double *)calloc(n, sizeof(double));
    OMP(omp parallel for private(i))
    for (k = 1; k <= n3msub; k++)
        for (int jj = 1; jj <= n2msub; jj++)
            for (i = 1; i <= n1msub; i++)
                RHSI[R3(n2msub, n3msub, jj, k, i)] = RHS[R3(n1msub, n2msub, i, jj, k)];

    OMP(omp parallel for)
    for (size_t q = 0; q < n; q++) { am[q] = 0.0; ac[q] = 1.0; ap[q] = 0.0; }
    /* 2nd stage x (:802-833): mu1, mu2 are stale from stage 1's last point */
    OMP(omp parallel for private(j, i, jm, jp, im, ip, jvp, jvm, u1, u2, v3, v4, w5, w6, dudy1, dudy2, dvdx1, dvdx2, dvdy3, dvdy4, dvdz5, dvdz6, dwdy5, dwdy6, ddu1, ddu2, ddudy1, ddudy2, mua, muc, mub, mu5, mu6, invRhoc, invRhocCmu_half, viscous_v1, viscous_v2, viscous_v3, viscous_v21, viscous_v23, vbc_up, vbc_down, mACI, mAPI, mAMI, mACJ, mAPJ, mAMJ, mACK, mAPK, mAMK))
    for (k = 1; k <= n3msub; k++)
        for (j = j_indexS; j <= n2msub; j++) {
            jp = j + 1; jm = j - 1;
            jvp = jS_BC[jm]; jvm = jS_BC[jp];
            for (i = 1; i <= n1msub; i++) {
                ip = i + 1; im = i - 1;
                u1 = (dx2[jm] * U[IX(i, j, k)] + dx2[j] * U[IX(i, jm, k)]) / dmx2[j] * 0.5;
                u2 = (dx2[jm] * U[IX(ip, j, k)] + dx2[j] * U[IX(ip, jm, k)]) / dmx2[j] * 0.5;
                invRhoc = 0.5 * (dx2[jm] * invRho[IX(i, j, k)] + dx2[j] * invRho[IX(i, jm, k)]) / dmx2[j];
                invRhocCmu_half = 0.5 * Cmu * invRhoc;
                if (PLANT(212, 1, 0)) {
                    mua = 0.5 * (dx2[jm] * Mu[IX(im, j, k)] + dx2[j] * Mu[IX(im, jm, k)]) / dmx2[j];
                    muc = 0.5 * (dx2[jm] * Mu[IX(i, j, k)] + dx2[j] * Mu[IX(i, jm, k)]) / dmx2[j];
                    mub = 0.5 * (dx2[jm] * Mu[IX(ip, j, k)] + dx2[j] * Mu[IX(ip, jm, k)]) / dmx2[j];
                    mu1 = 0.5 * (dx1[im] * muc + dx1[i] * mua) / dmx1[i];
                    mu2 = 0.5 * (dx1[ip] * muc + dx1[i] * mub) / dmx1[ip];
                }
                mACI = invRhocCmu_half / dx1[i] * (mu2 / dmx1[ip] + mu1 / dmx1[i]) + 0.25 * (-u2 / dmx1[ip] + u1 / dmx1[i]);
                mAPI = -invRhocCmu_half / dx1[i] * mu2 / dmx1[ip] + 0.25 * (u2 / dmx1[ip]);
                mAMI = -invRhocCmu_half / dx1[i] * mu1 / dmx1[i] + 0.25 * (-u1 / dmx1[i]);
                size_t q = R3(n2msub, n3msub, j, k, i);
                ac[q] = mACI * dt + 1.0;
                ap[q] = mAPI * dt;
                am[q] = mAMI * dt;
            }
        }
    lines_solve(o, am, ac, ap, RHSI, n2msub * n3msub, n1msub, 1);

    int ny = o->n2sub - j_indexS; /* rows j=j_indexS..n2msub */
    size_t nyv = (size_t)n1msub * n3msub * ny;
    double *RHSJ = (double *)calloc(nyv, sizeof(double));
    double *amj = (double *)calloc(nyv, sizeof(double));
    double *acj = (double *)calloc(nyv, sizeof(double));
    double *apj = (double *)calloc(nyv, sizeof(double));
    OMP(omp parallel for private(j, i))
    for (k = 1; k <= n3msub; k++)
        for (j = j_indexS; j <= n2msub; j++)
            for (i = 1; i <= n1msub; i++)
                RHSJ[R3(n1msub, n3msub, i, k, j - j_indexS + 1)] = RHSI[R3(n2msub, n3msub, j, k, i)];

    /* 3rd stage y (:860-896) */
    OMP(omp parallel for private(j, i, jm, jp, im, ip, jvp, jvm, u1, u2, v3, v4, w5, w6, dudy1, dudy2, dvdx1, dvdx2, dvdy3, dvdy4, dvdz5, dvdz6, dwdy5, dwdy6, ddu1, ddu2, ddudy1, ddudy2, mua, muc, mub, mu5, mu6, invRhoc, invRhocCmu_half, viscous_v1, viscous_v2, viscous_v3, viscous_v21, viscous_v23, vbc_up, vbc_down, mACI, mAPI, mAMI, mACJ, mAPJ, mAMJ, mACK, mAPK, mAMK))
    for (k = 1; k <= n3msub; k++)
        for (j = j_indexS; j <= n2msub; j++) {
            jm = j - 1; jp = j + 1;
            jvp = jS_BC[PLANT(214, jp, jm)]; jvm = jS_BC[PLANT(214, jm, jp)];
            for (i = 1; i <= n1msub; i++) {
                ip = i + 1; im = i - 1;
                v3 = 0.5 * (V[IX(i, jm, k)] + V[IX(i, j, k)]);
                v4 = 0.5 * (V[IX(i, j, k)] + V[IX(i, jp, k)]);
                dvdy3 = (V[IX(i, j, k)] - V[IX(i, jm, k)]) / dx2[jm];
                dvdy4 = (V[IX(i, jp, k)] - V[IX(i, j, k)]) / dx2[j];
                invRhoc = 0.5 * (dx2[jm] * invRho[IX(i, j, k)] + dx2[j] * invRho[IX(i, jm, k)]) / dmx2[j];
                invRhocCmu_half = 0.5 * Cmu * invRhoc;
                mACJ = invRhocCmu_half / dmx2[j] * (Mu[IX(i, j, k)] / dx2[j] + Mu[IX(i, jm, k)] / dx2[jm]) * 2.0
                     + 0.25 * (dvdy4 * 0.5 + dvdy3 * 0.5 - v4 / dx2[j] + v3 / dx2[jm]);
                mAPJ = -invRhocCmu_half / dmx2[j] * Mu[IX(i, j, k)] / dx2[j] * 2.0 + 0.25 * (v4 / dx2[j] + dvdy4 * 0.5);
                mAPJ = mAPJ * (double)jvp;
                mAMJ = -invRhocCmu_half / dmx2[j] * Mu[IX(i, jm, k)] / dx2[jm] * 2.0 + 0.25 * (-v3 / dx2[jm] + dvdy3 * 0.5);
                mAMJ = mAMJ * (double)jvm;
                size_t q = R3(n1msub, n3msub, i, k, j - j_indexS + 1);
                acj[q] = mACJ * dt + 1.0;
                apj[q] = mAPJ * dt;
                amj[q] = mAMJ * dt;
            }
        }
    if (o->tdma_split && ny + 1 <= 512) {
        /* Split lines (not in the reference): the wall row j = 1, where dV is
           zero, leads each y line as an identity row, so that the 2^m - 1
           rows j = 2..n2m split into whole parts of 2^m rows.  Row 2's lower
           coefficient then multiplies the identity row's zero. */
        const size_t ns = (size_t)n3msub * n1msub, nb = sizeof(double) * ns;
        double *a2 = malloc(nb * (ny + 1)), *b2 = malloc(nb * (ny + 1));
        double *c2 = malloc(nb * (ny + 1)), *d2 = malloc(nb * (ny + 1));
        for (size_t s = 0; s < ns; s++) { a2[s] = 0.0; b2[s] = 1.0; c2[s] = 0.0; d2[s] = 0.0; }
        memcpy(a2 + ns, amj, nb * ny);
        memcpy(b2 + ns, acj, nb * ny);
        memcpy(c2 + ns, apj, nb * ny);
        memcpy(d2 + ns, RHSJ, nb * ny);
        lines_solve(o, a2, b2, c2, d2, ns, ny + 1, 0);
        memcpy(RHSJ, d2 + ns, nb * ny);
        free(a2); free(b2); free(c2); free(d2);
    } else {
        lines_solve(o, amj, acj, apj, RHSJ, n3msub * n1msub, ny, 0);
    }

    for (k = 0; k <= o->n3sub; k++)
        for (i = 0; i <= o->n1sub; i++) {
            dV[IX(i, 0, k)] = 0.0; dV[IX(i, 1, k)] = 0.0; dV[IX(i, o->n2sub, k)] = 0.0;
        }
    OMP(omp parallel for private(j, i))
    for (k = 1; k <= n3msub; k++)
        for (j = j_indexS; j <= n2msub; j++)
            for (i = 1; i <= n1msub; i++)
                dV[IX(i, j, k)] = RHSJ[R3(n1msub, n3msub, i, k, j - j_indexS + 1)];

    free(RHS); free(RHSI); free(am); free(ac); free(ap);
    free(RHSJ); free(amj); free(acj); free(apj);
}

/* :930-1272 */
void orc_momentum_solvedW(orc *o, double dt)
{
    const double *p = o->prm;
    const double Cmu = p[OP_CMU], Cmp = p[OP_CMP];
    const int n1msub = o->n1msub, n2msub = o->n2msub, n3msub = o->n3msub;
    const double *U = o->f[OF_U], *V = o->f[OF_V], *W = o->f[OF_W], *P = o->f[OF_P];
    const double *dU = o->f[OF_DU], *dV = o->f[OF_DV];
    const double *Mu = o->f[OF_MU], *invRho = o->f[OF_INVRHO];
    double *dW = o->f[OF_DW];
    const double *dx1 = o->dx1, *dx2 = o->dx2, *dx3 = o->dx3;
    const double *dmx1 = o->dmx1, *dmx2 = o->dmx2, *dmx3 = o->dmx3;
    const double *VBCup = o->bc[OB_VUP], *VBCbt = o->bc[OB_VBT], *WBCup = o->bc[OB_WUP], *WBCbt = o->bc[OB_WBT];
    const int *jC_BC = o->jC_BC;

    int i, j, k, im = 0, jm = 0, km = 0, ip = 0, jp = 0, kp = 0, jwp = 0, jwm = 0;
    double u1, u2, v3, v4, w6, w5;
    double dudz1, dudz2, dvdz3, dvdz4, dwdx1, dwdx2, dwdy3, dwdy4, dwdz5, dwdz6;
    double ddu1, ddu2, ddudz1, ddudz2, ddv3, ddv4, ddvdz3, ddvdz4;
    double mua, muc, mub, mu1 = 0, mu2 = 0, mu3 = 0, mu4 = 0, invRhoc, invRhocCmu_half;
    double viscous_w1, viscous_w2, viscous_w3, viscous_w31, viscous_w32;
    double zbc_up, zbc_down;
    double mACI, mAPI, mAMI, mACJ, mAPJ, mAMJ, mACK, mAPK, mAMK;

    size_t n = (size_t)n1msub * n2msub * n3msub;
    double *RHS = (double *)calloc(n, sizeof(double));
    double *am = (double *)calloc(n, sizeof(double));
    double *ac = (double *)calloc(n, sizeof(double));
    double *ap = (double *)calloc(n, sizeof(double));

    OMP(omp parallel for private(i, j, im, ip, jwp, jwm, u1, u2, v3, v4, w6, w5, dudz1, dudz2, dvdz3, dvdz4, dwdx1, dwdx2, dwdy3, dwdy4, dwdz5, dwdz6, ddu1, ddu2, ddudz1, ddudz2, ddv3, ddv4, ddvdz3, ddvdz4, mua, muc, mub, invRhoc, invRhocCmu_half, viscous_w1, viscous_w2, viscous_w3, viscous_w31, viscous_w32, zbc_up, zbc_down, mACI, mAPI, mAMI, mACJ, mAPJ, mAMJ, mACK, mAPK, mAMK) lastprivate(jm, jp, km, kp, mu1, mu2, mu3, mu4))
    for (k = o->k_indexS; k <= n3msub; k++) {
        kp = k + 1; km = k - 1;
        for (j = 1; j <= n2msub; j++) {
            jp = j + 1; jm = j - 1;
            jwp = jC_BC[jp]; jwm = jC_BC[jm];
            for (i = 1; i <= n1msub; i++) {
                ip = i + 1; im = i - 1;
                w5 = 0.5 * (W[IX(i, j, k)] + W[IX(i, j, km)]);
                w6 = 0.5 * (W[IX(i, j, k)] + W[IX(i, j, kp)]);
                u1 = (dx3[km] * U[IX(i, j, k)] + dx3[k] * U[IX(i, j, km)]) / dmx3[k] * 0.5;
                u2 = (dx3[km] * U[IX(ip, j, k)] + dx3[k] * U[IX(ip, j, km)]) / dmx3[k] * 0.5;
                v3 = (dx3[km] * V[IX(i, j, k)] + dx3[k] * V[IX(i, j, km)]) / dmx3[k] * 0.5;
                v4 = (dx3[km] * V[IX(i, jp, k)] + dx3[k] * V[IX(i, jp, km)]) / dmx3[k] * 0.5;

                dwdx1 = (W[IX(i, j, k)] - W[IX(im, j, k)]) / dmx1[i];
                dwdx2 = (W[IX(ip, j, k)] - W[IX(i, j, k)]) / dmx1[ip];
                dwdy3 = (W[IX(i, j, k)] - W[IX(i, jm, k)]) / dmx2[j];
                dwdy4 = (W[IX(i, jp, k)] - W[IX(i, j, k)]) / dmx2[jp];
                dwdz5 = (W[IX(i, j, k)] - W[IX(i, j, km)]) / dx3[km];
                dwdz6 = (W[IX(i, j, kp)] - W[IX(i, j, k)]) / dx3[k];
                dudz1 = (U[IX(i, j, k)] - U[IX(i, j, km)]) / dmx3[k];
                dudz2 = (U[IX(ip, j, k)] - U[IX(ip, j, km)]) / dmx3[k];
                dvdz3 = (V[IX(i, j, k)] - V[IX(i, j, km)]) / dmx3[k];
                dvdz4 = (V[IX(i, jp, k)] - V[IX(i, jp, km)]) / dmx3[k];
                ddudz1 = (dU[IX(i, j, k)] - dU[IX(i, j, km)]) / dmx3[k];
                ddudz2 = (dU[IX(ip, j, k)] - dU[IX(ip, j, km)]) / dmx3[k];
                ddvdz3 = (dV[IX(i, j, k)] - dV[IX(i, j, km)]) / dmx3[k];
                ddvdz4 = (dV[IX(i, jp, k)] - dV[IX(i, jp, km)]) / dmx3[k];
                ddu1 = (dx3[km] * dU[IX(i, j, k)] + dx3[k] * dU[IX(i, j, km)]) / dmx3[k] * 0.5;
                ddu2 = (dx3[km] * dU[IX(ip, j, k)] + dx3[k] * dU[IX(ip, j, km)]) / dmx3[k] * 0.5;
                ddv3 = (dx3[km] * dV[IX(i, j, k)] + dx3[k] * dV[IX(i, j, km)]) / dmx3[k] * 0.5 * (double)jwm;
                ddv4 = (dx3[km] * dV[IX(i, jp, k)] + dx3[k] * dV[IX(i, jp, km)]) / dmx3[k] * 0.5 * (double)jwp;

                mua = 0.5 * (dx3[km] * Mu[IX(im, j, k)] + dx3[k] * Mu[IX(im, j, km)]) / dmx3[k];
                muc = 0.5 * (dx3[km] * Mu[IX(i, j, k)] + dx3[k] * Mu[IX(i, j, km)]) / dmx3[k];
                mub = 0.5 * (dx3[km] * Mu[IX(ip, j, k)] + dx3[k] * Mu[IX(ip, j, km)]) / dmx3[k];
                mu1 = 0.5 * (dx1[im] * muc + dx1[i] * mua) / dmx1[i];
                mu2 = 0.5 * (dx1[ip] * muc + dx1[i] * mub) / dmx1[ip];
                mua = 0.5 * (dx3[km] * Mu[IX(i, jm, k)] + dx3[k] * Mu[IX(i, jm, km)]) / dmx3[k];
                muc = 0.5 * (dx3[km] * Mu[IX(i, j, k)] + dx3[k] * Mu[IX(i, j, km)]) / dmx3[k];
                mub = 0.5 * (dx3[km] * Mu[IX(i, jp, k)] + dx3[k] * Mu[IX(i, jp, km)]) / dmx3[k];
                mu3 = 0.5 * (dx2[jm] * muc + dx2[j] * mua) / dmx2[j];
                mu4 = 0.5 * (dx2[jp] * muc + dx2[j] * mub) / dmx2[jp];

                invRhoc = 0.5 * (dx3[km] * invRho[IX(i, j, k)] + dx3[k] * invRho[IX(i, j, km)]) / dmx3[k];
                invRhocCmu_half = 0.5 * Cmu * invRhoc;

                viscous_w1 = 1.0 * (mu2 * dwdx2 - mu1 * dwdx1) / dx1[i];
                viscous_w2 = 1.0 * (mu4 * dwdy4 - mu3 * dwdy3) / dx2[j];
                viscous_w3 = 1.0 * (Mu[IX(i, j, k)] * dwdz6 - Mu[IX(i, j, km)] * dwdz5) / dmx3[k];
                viscous_w31 = 1.0 * (mu2 * dudz2 - mu1 * dudz1) / dx1[i];
                viscous_w32 = 1.0 * (mu4 * dvdz4 - mu3 * dvdz3) / dx2[j];

                double rhs = invRhocCmu_half * (viscous_w1 + viscous_w2 + 2.0 * viscous_w3 + viscous_w31 + viscous_w32);
                rhs = rhs - Cmp * invRhoc * (P[IX(i, j, k)] - P[IX(i, j, km)]) / dmx3[k];

                zbc_down = 0.25 * v3 / dmx2[j] * WBCbt[B2(i, k)]
                         - 0.25 * dwdy3 * (dx3[km] * VBCbt[B2(i, k)] + dx3[k] * VBCbt[B2(i, km)]) / dmx3[k] * 0.5;
                zbc_up = -0.25 * v4 / dmx2[jp] * WBCup[B2(i, k)]
                       - 0.25 * dwdy4 * (dx3[km] * VBCup[B2(i, k)] + dx3[k] * VBCup[B2(i, km)]) / dmx3[k] * 0.5;
                zbc_down = zbc_down + invRhocCmu_half / dx2[j] * mu3 / dmx2[PLANT(221, j, jp)] * WBCbt[B2(i, k)];
                zbc_up = zbc_up + invRhocCmu_half / dx2[j] * mu4 / dmx2[j] * WBCup[B2(i, k)];
                zbc_down = zbc_down - invRhocCmu_half / dx2[j] * mu3 / dmx3[k] * (VBCbt[B2(i, k)] - VBCbt[B2(i, km)]);
                zbc_up = zbc_up + invRhocCmu_half / dx2[j] * mu4 / dmx3[k] * (VBCup[B2(i, k)] - VBCup[B2(i, km)]);
                rhs = rhs + (1.0 - (double)jwm) * zbc_down + (1.0 - (double)jwp) * zbc_up;

                mACI = invRhocCmu_half / dx1[i] * (mu2 / dmx1[ip] + mu1 / dmx1[i]) + 0.25 * (-u2 / dmx1[ip] + u1 / dmx1[i]);
                mAPI = -invRhocCmu_half / dx1[i] * mu2 / dmx1[ip] + 0.25 * (u2 / dmx1[ip]);
                mAMI = -invRhocCmu_half / dx1[i] * mu1 / dmx1[i] + 0.25 * (-u1 / dmx1[i]);

                mACJ = invRhocCmu_half / dx2[j] * (mu4 / dmx2[jp] + mu3 / dmx2[j]) + 0.25 * (-v4 / dmx2[jp] + v3 / dmx2[j]);
                mAPJ = -invRhocCmu_half / dx2[j] * mu4 / dmx2[jp] + 0.25 * (v4 / dmx2[jp]);
                mAPJ = mAPJ * (double)jwp;
                mAMJ = -invRhocCmu_half / dx2[j] * mu3 / dmx2[j] + 0.25 * (-v3 / dmx2[j]);
                mAMJ = mAMJ * (double)jwm;

                mACK = invRhocCmu_half / dmx3[k] * (Mu[IX(i, j, k)] / dx3[k] + Mu[IX(i, j, km)] / dx3[km]) * 2.0
                     + 0.25 * (-w6 / dx3[k] + w5 / dx3[km] + 0.5 * dwdz6 + 0.5 * dwdz5);
                mAPK = -invRhocCmu_half / dmx3[k] * Mu[IX(i, j, k)] / dx3[k] * 2.0 + 0.25 * (w6 / dx3[k] + 0.5 * dwdz6);
                mAMK = -invRhocCmu_half / dmx3[k] * Mu[IX(i, j, km)] / dx3[km] * 2.0 + 0.25 * (-w5 / dx3[km] + 0.5 * dwdz5);

                rhs = rhs - (mAPI * W[IX(ip, j, k)] + mACI * W[IX(i, j, k)] + mAMI * W[IX(im, j, k)]
                             + mAPJ * W[IX(i, jp, k)] + mACJ * W[IX(i, j, k)] + mAMJ * W[IX(i, jm, k)]
                             + mAPK * W[IX(i, j, kp)] + mACK * W[IX(i, j, k)] + mAMK * W[IX(i, j, km)]);
                rhs = rhs - (0.25 * (u2 * dwdx2 + u1 * dwdx1) - invRhocCmu_half / dx1[i] * (mu2 * dudz2 - mu1 * dudz1));
                rhs = rhs - (0.25 * (v3 * (double)jwm * dwdy3 + v4 * (double)jwp * dwdy4)
                             - invRhocCmu_half / dx2[j] * (mu4 * dvdz4 * (double)jwp - mu3 * dvdz3 * (double)jwm));
                rhs = rhs - (0.25 * (ddu2 * dwdx2 + ddu1 * dwdx1) - invRhocCmu_half / dx1[i] * (mu2 * ddudz2 - mu1 * ddudz1));
                rhs = rhs - (0.25 * (ddv3 * (double)jwm * dwdy3 + ddv4 * (double)jwp * dwdy4)
                             - invRhocCmu_half / dx2[j] * (mu4 * ddvdz4 * PLANT(224, 1.0, (double)jwp) - mu3 * ddvdz3 * (double)jwm));

                size_t q = R3(n1msub, n2msub, i, j, k);
                ac[q] = mACK * dt + 1.0;
                ap[q] = mAPK * dt;
                am[q] = mAMK * dt;
                RHS[q] = rhs * dt;
            }
        }
    }
    lines_solve(o, am, ac, ap, RHS, n1msub * n2msub, n3msub, 1);

    double *RHSI = (double *)calloc(n, sizeof(double));
    OMP(omp parallel for private(i))
    for (k = o->k_indexS; k <= n3msub; k++)
        for (int jj = 1; jj <= n2msub; jj++)
            for (i = 1; i <= n1msub; i++)
                RHSI[R3(n2msub, n3msub, jj, k, i)] = RHS[R3(n1msub, n2msub, i, jj, k)];

    /* 2nd stage x (:1156-1189): mu1, mu2 stale from stage 1 */
    OMP(omp parallel for private(i, j, jm, jp, im, ip, jwp, jwm, u1, u2, v3, v4, w6, w5, dudz1, dudz2, dvdz3, dvdz4, dwdx1, dwdx2, dwdy3, dwdy4, dwdz5, dwdz6, ddu1, ddu2, ddudz1, ddudz2, ddv3, ddv4, ddvdz3, ddvdz4, mua, muc, mub, invRhoc, invRhocCmu_half, viscous_w1, viscous_w2, viscous_w3, viscous_w31, viscous_w32, zbc_up, zbc_down, mACI, mAPI, mAMI, mACJ, mAPJ, mAMJ, mACK, mAPK, mAMK) lastprivate(km, kp))
    for (k = o->k_indexS; k <= n3msub; k++) {
        kp = k + 1; km = k - 1;
        for (j = 1; j <= n2msub; j++) {
            jp = j + 1; jm = j - 1;
            jwp = jC_BC[jp]; jwm = jC_BC[jm];
            for (i = 1; i <= n1msub; i++) {
                ip = i + 1; im = i - 1;
                u1 = (dx3[km] * U[IX(i, j, k)] + dx3[k] * U[IX(i, j, km)]) / dmx3[k] * 0.5;
                u2 = (dx3[km] * U[IX(ip, j, k)] + dx3[k] * U[IX(ip, j, km)]) / dmx3[k] * 0.5;
                invRhoc = 0.5 * (dx3[km] * invRho[IX(i, j, k)] + dx3[k] * invRho[IX(i, j, km)]) / dmx3[k];
                invRhocCmu_half = 0.5 * Cmu * invRhoc;
                if (PLANT(222, 1, 0)) {
                    mua = 0.5 * (dx3[km] * Mu[IX(im, j, k)] + dx3[k] * Mu[IX(im, j, km)]) / dmx3[k];
                    muc = 0.5 * (dx3[km] * Mu[IX(i, j, k)] + dx3[k] * Mu[IX(i, j, km)]) / dmx3[k];
                    mub = 0.5 * (dx3[km] * Mu[IX(ip, j, k)] + dx3[k] * Mu[IX(ip, j, km)]) / dmx3[k];
                    mu1 = 0.5 * (dx1[im] * muc + dx1[i] * mua) / dmx1[i];
                    mu2 = 0.5 * (dx1[ip] * muc + dx1[i] * mub) / dmx1[ip];
                }
                mACI = invRhocCmu_half / dx1[i] * (mu2 / dmx1[ip] + mu1 / dmx1[i]) + 0.25 * (-u2 / dmx1[ip] + u1 / dmx1[i]);
                mAPI = -invRhocCmu_half / dx1[i] * mu2 / dmx1[ip] + 0.25 * (u2 / dmx1[ip]);
                mAMI = -invRhocCmu_half / dx1[i] * mu1 / dmx1[i] + 0.25 * (-u1 / dmx1[i]);
                size_t q = R3(n2msub, n3msub, j, k, i);
                ac[q] = mACI * dt + 1.0;
                ap[q] = mAPI * dt;
                am[q] = mAMI * dt;
            }
        }
    }
    lines_solve(o, am, ac, ap, RHSI, n2msub * n3msub, n1msub, 1);

    OMP(omp parallel for private(j, i))
    for (k = 1; k <= n3msub; k++)
        for (j = 1; j <= n2msub; j++)
            for (i = 1; i <= n1msub; i++)
                RHS[R3(n1msub, n3msub, i, k, j)] = RHSI[R3(n2msub, n3msub, j, k, i)];

    /* 3rd stage y (:1216-1246): km stale (=n3msub-1), mu3, mu4 stale */
    OMP(omp parallel for private(j, i, jm, jp, im, ip, jwp, jwm, u1, u2, v3, v4, w6, w5, dudz1, dudz2, dvdz3, dvdz4, dwdx1, dwdx2, dwdy3, dwdy4, dwdz5, dwdz6, ddu1, ddu2, ddudz1, ddudz2, ddv3, ddv4, ddvdz3, ddvdz4, mua, muc, mub, invRhoc, invRhocCmu_half, viscous_w1, viscous_w2, viscous_w3, viscous_w31, viscous_w32, zbc_up, zbc_down, mACI, mAPI, mAMI, mACJ, mAPJ, mAMJ, mACK, mAPK, mAMK))
    for (k = o->k_indexS; k <= n3msub; k++)
        for (j = 1; j <= n2msub; j++) {
            jm = j - 1; jp = j + 1;
            if (PLANT(223, 1, 0)) km = k - 1;
            jwp = jC_BC[jp]; jwm = jC_BC[jm];
            for (i = 1; i <= n1msub; i++) {
                v3 = (dx3[km] * V[IX(i, j, k)] + dx3[k] * V[IX(i, j, km)]) / dmx3[k] * 0.5;
                v4 = (dx3[km] * V[IX(i, jp, k)] + dx3[k] * V[IX(i, jp, km)]) / dmx3[k] * 0.5;
                invRhoc = 0.5 * (dx3[km] * invRho[IX(i, j, k)] + dx3[k] * invRho[IX(i, j, km)]) / dmx3[k];
                invRhocCmu_half = 0.5 * Cmu * invRhoc;
                mACJ = invRhocCmu_half / dx2[j] * (mu4 / dmx2[jp] + mu3 / dmx2[j]) + 0.25 * (-v4 / dmx2[jp] + v3 / dmx2[j]);
                mAPJ = -invRhocCmu_half / dx2[j] * mu4 / dmx2[jp] + 0.25 * (v4 / dmx2[jp]);
                mAPJ = mAPJ * (double)jwp;
                mAMJ = -invRhocCmu_half / dx2[j] * mu3 / dmx2[j] + 0.25 * (-v3 / dmx2[j]);
                mAMJ = mAMJ * (double)jwm;
                size_t q = R3(n1msub, n3msub, i, k, j);
                ac[q] = mACJ * dt + 1.0;
                ap[q] = mAPJ * dt;
                am[q] = mAMJ * dt;
            }
        }
    lines_solve(o, am, ac, ap, RHS, n3msub * n1msub, n2msub, 0);

    OMP(omp parallel for private(j, i))
    for (k = o->k_indexS; k <= n3msub; k++)
        for (j = 1; j <= n2msub; j++)
            for (i = 1; i <= n1msub; i++)
                dW[IX(i, j, k)] = RHS[R3(n1msub, n3msub, i, k, j)];

    free(RHS); free(RHSI); free(am); free(ac); free(ap);
}

/* :1278-1356 */
void orc_momentum_blockLdV(orc *o, double dt)
{
    const double Cmu = o->prm[OP_CMU];
    const double *V = o->f[OF_V], *dW = o->f[OF_DW], *Mu = o->f[OF_MU], *invRho = o->f[OF_INVRHO];
    double *dV = o->f[OF_DV];
    const double *dx2 = o->dx2, *dx3 = o->dx3, *dmx2 = o->dmx2, *dmx3 = o->dmx3;
    OMP(omp parallel for)
    for (int k = 1; k <= o->n3msub; k++) {
        int kp = k + 1, km = k - 1;
        for (int j = o->j_indexS; j <= o->n2msub; j++) {
            int jm = j - 1;
            for (int i = 1; i <= o->n1msub; i++) {
                double dwm5 = (dx2[jm] * dW[IX(i, j, k)] + dx2[j] * dW[IX(i, jm, k)]) / dmx2[j] * 0.5;
                double dwm6 = (dx2[jm] * dW[IX(i, j, kp)] + dx2[j] * dW[IX(i, jm, kp)]) / dmx2[j] * 0.5;
                double ddwdy5 = (dW[IX(i, j, k)] - dW[IX(i, jm, k)]) / dmx2[j];
                double ddwdy6 = (dW[IX(i, j, kp)] - dW[IX(i, jm, kp)]) / dmx2[j];
                double dvdz5 = (V[IX(i, j, k)] - V[IX(i, j, km)]) / dmx3[k];
                double dvdz6 = (V[IX(i, j, kp)] - V[IX(i, j, k)]) / dmx3[kp];
                double invRhoc = 0.5 * (dx2[jm] * invRho[IX(i, j, k)] + dx2[j] * invRho[IX(i, jm, k)]) / dmx2[j];
                double mua = 0.5 * (dx2[jm] * Mu[IX(i, j, km)] + dx2[j] * Mu[IX(i, jm, km)]) / dmx2[j];
                double muc = 0.5 * (dx2[jm] * Mu[IX(i, j, k)] + dx2[j] * Mu[IX(i, jm, k)]) / dmx2[j];
                double mub = 0.5 * (dx2[jm] * Mu[IX(i, j, kp)] + dx2[j] * Mu[IX(i, jm, kp)]) / dmx2[j];
                double mu5 = 0.5 * (dx3[km] * muc + dx3[k] * mua) / dmx3[k];
                double mu6 = 0.5 * (dx3[kp] * muc + dx3[k] * mub) / dmx3[kp];
                dV[IX(i, j, k)] = dV[IX(i, j, k)]
                                - dt * (0.25 * (dwm5 * dvdz5 + dwm6 * dvdz6)
                                        - 0.5 * Cmu * invRhoc / dx3[k] * (mu6 * ddwdy6 - mu5 * ddwdy5));
            }
        }
    }
}

/* :1362-1460 */
void orc_momentum_blockLdU(orc *o, double dt)
{
    const double Cmu = o->prm[OP_CMU];
    const double *U = o->f[OF_U], *dV = o->f[OF_DV], *dW = o->f[OF_DW], *Mu = o->f[OF_MU], *invRho = o->f[OF_INVRHO];
    double *dU = o->f[OF_DU];
    const double *dx1 = o->dx1, *dx2 = o->dx2, *dx3 = o->dx3, *dmx1 = o->dmx1, *dmx2 = o->dmx2, *dmx3 = o->dmx3;
    const int *jC_BC = o->jC_BC;
    OMP(omp parallel for)
    for (int k = 1; k <= o->n3msub; k++) {
        int km = k - 1, kp = k + 1;
        for (int j = 1; j <= o->n2msub; j++) {
            int jm = j - 1, jp = j + 1;
            int jum = jC_BC[jm], jup = jC_BC[jp];
            for (int i = 1; i <= o->n1msub; i++) {
                int im = i - 1;
                double invRhoc = 0.5 * (dx1[im] * invRho[IX(i, j, k)] + dx1[i] * invRho[IX(im, j, k)]) / dmx1[i];
                double dvm3 = (dx1[i] * dV[IX(im, j, k)] + dx1[im] * dV[IX(i, j, k)]) / dmx1[i] * 0.5 * (double)jum;
                double dvm4 = (dx1[i] * dV[IX(im, jp, k)] + dx1[im] * dV[IX(i, jp, k)]) / dmx1[i] * 0.5 * (double)jup;
                double dudy3 = (U[IX(i, j, k)] - U[IX(i, jm, k)]) / dmx2[j];
                double dudy4 = (U[IX(i, jp, k)] - U[IX(i, j, k)]) / dmx2[jp];
                double dvmdx3 = (dV[IX(i, j, k)] - dV[IX(im, j, k)]) / dmx1[i] * (double)jum;
                double dvmdx4 = (dV[IX(i, jp, k)] - dV[IX(im, jp, k)]) / dmx1[i] * (double)jup;
                double mua = 0.5 * (dx1[im] * Mu[IX(i, jm, k)] + dx1[i] * Mu[IX(im, jm, k)]) / dmx1[i];
                double muc = 0.5 * (dx1[im] * Mu[IX(i, j, k)] + dx1[i] * Mu[IX(im, j, k)]) / dmx1[i];
                double mub = 0.5 * (dx1[im] * Mu[IX(i, jp, k)] + dx1[i] * Mu[IX(im, jp, k)]) / dmx1[i];
                double mu3 = 0.5 * (dx2[jm] * muc + dx2[j] * mua) / dmx2[j];
                double mu4 = 0.5 * (dx2[jp] * muc + dx2[j] * mub) / dmx2[jp];
                double dwm5 = (dx1[i] * dW[IX(im, j, k)] + dx1[im] * dW[IX(i, j, k)]) / dmx1[i] * 0.5;
                double dwm6 = (dx1[i] * dW[IX(im, j, kp)] + dx1[im] * dW[IX(i, j, kp)]) / dmx1[i] * 0.5;
                double dudz5 = (U[IX(i, j, k)] - U[IX(i, j, km)]) / dmx3[k];
                double dudz6 = (U[IX(i, j, kp)] - U[IX(i, j, k)]) / dmx3[kp];
                double dwmdx5 = (dW[IX(i, j, k)] - dW[IX(im, j, k)]) / dmx1[i];
                double dwmdx6 = (dW[IX(i, j, kp)] - dW[IX(im, j, kp)]) / dmx1[i];
                mua = 0.5 * (dx1[im] * Mu[IX(i, j, km)] + dx1[i] * Mu[IX(im, j, km)]) / dmx1[i];
                muc = 0.5 * (dx1[im] * Mu[IX(i, j, k)] + dx1[i] * Mu[IX(im, j, k)]) / dmx1[i];
                mub = 0.5 * (dx1[im] * Mu[IX(i, j, kp)] + dx1[i] * Mu[IX(im, j, kp)]) / dmx1[i];
                double mu5 = 0.5 * (dx3[km] * muc + dx3[k] * mua) / dmx3[k];
                double mu6 = 0.5 * (dx3[kp] * muc + dx3[k] * mub) / dmx3[kp];
                dU[IX(i, j, k)] = dU[IX(i, j, k)]
                                - dt * (0.25 * (dvm4 * dudy4 + dvm3 * dudy3) - 0.5 * Cmu * invRhoc / dx2[j] * (mu4 * dvmdx4 - mu3 * dvmdx3))
                                - dt * (0.25 * (dwm6 * dudz6 + dwm5 * dudz5) - 0.5 * Cmu * invRhoc / dx3[k] * (mu6 * dwmdx6 - mu5 * dwmdx5));
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* pressure: src/module_solve_pressure.f90                                   */
/* ------------------------------------------------------------------------ */
/* :93-129, single rank (h1psub_Ksub = h1p, ista = 1) */
void orc_pressure_wave_number(orc *o)
{
    const double PI = acos(-1.0);
    double L1 = o->prm[OP_L1], L3 = o->prm[OP_L3];
    double ddx = L1 / (double)o->n1m, ddz = L3 / (double)o->n3m;
    for (int k = 1; k <= o->n3m; k++) {
        int km = (k <= o->h3p) ? k - 1 : o->n3m - k + 1;
        o->dzk[k - 1] = 2.0 * (1.0 - cos(2.0 * PI * (double)km * ddz / L3)) / (ddz * ddz);
    }
    for (int i = 1; i <= o->h1p; i++) {
        int im = i - 1;
        o->dxk2[i - 1] = 2.0 * (1.0 - cos(2.0 * PI * (double)im * ddx / L1)) / (ddx * ddx);
    }
}

/* :140-231; prhs_out is (1:n1m,1:n2m,1:n3m) */
void orc_pressure_RHS(orc *o, double dt, double *PRHS)
{
    const double Cmp = o->prm[OP_CMP];
    const double *U = o->f[OF_U], *V = o->f[OF_V], *W = o->f[OF_W], *dPhat = o->f[OF_DPHAT], *invRho = o->f[OF_INVRHO];
    const double *VBCup = o->bc[OB_VUP], *VBCbt = o->bc[OB_VBT];
    const double *dx1 = o->dx1, *dx2 = o->dx2, *dx3 = o->dx3, *dmx1 = o->dmx1, *dmx2 = o->dmx2, *dmx3 = o->dmx3;
    const int *jC_BC = o->jC_BC;
    OMP(omp parallel for)
    for (int k = 1; k <= o->n3msub; k++) {
        int kp = k + 1, km = k - 1;
        for (int j = 1; j <= o->n2msub; j++) {
            int jp = j + 1, jm = j - 1;
            int jvm = jC_BC[jm], jvp = jC_BC[jp];
            for (int i = o->i_indexS; i <= o->n1msub; i++) {
                int ip = i + 1, im = i - 1;
                double DivUm = (U[IX(ip, j, k)] - U[IX(i, j, k)]) / dx1[i]
                             + ((double)jvp * V[IX(i, jp, k)] - (double)jvm * V[IX(i, j, k)]) / dx2[j]
                             + (W[IX(i, j, kp)] - W[IX(i, j, k)]) / dx3[k];
                double cbc = PLANT(303, -1.0, 1.0) * (1.0 - (double)jvm) * VBCbt[B2(i, k)] / dx2[j]
                           - (1.0 - (double)jvp) * VBCup[B2(i, k)] / dx2[j];
                double ddpdx1 = (dPhat[IX(i, j, k)] - dPhat[IX(im, j, k)]) / dmx1[i];
                double ddpdx2 = (dPhat[IX(ip, j, k)] - dPhat[IX(i, j, k)]) / dmx1[ip];
                double ddpdy3 = (dPhat[IX(i, j, k)] - dPhat[IX(i, jm, k)]) / dmx2[j];
                double ddpdy4 = (dPhat[IX(i, jp, k)] - dPhat[IX(i, j, k)]) / dmx2[jp];
                double ddpdz5 = (dPhat[IX(i, j, k)] - dPhat[IX(i, j, km)]) / dmx3[k];
                double ddpdz6 = (dPhat[IX(i, j, kp)] - dPhat[IX(i, j, k)]) / dmx3[kp];
                double invRho1 = 0.5 / dmx1[i] * (dx1[im] * invRho[IX(i, j, k)] + dx1[i] * invRho[IX(im, j, k)]);
                double invRho2 = 0.5 / dmx1[ip] * (dx1[ip] * invRho[IX(i, j, k)] + dx1[i] * invRho[IX(ip, j, k)]);
                double invRho3 = PLANT(302, 0.5 / dmx2[j] * (dx2[j] * invRho[IX(i, j, k)] + dx2[jm] * invRho[IX(i, jm, k)]),
                                       0.5 / dmx2[j] * (dx2[jm] * invRho[IX(i, j, k)] + dx2[j] * invRho[IX(i, jm, k)]));
                double invRho4 = 0.5 / dmx2[jp] * (dx2[jp] * invRho[IX(i, j, k)] + dx2[j] * invRho[IX(i, PLANT(301, jm, jp), k)]);
                double invRho5 = 0.5 / dmx3[k] * (dx3[km] * invRho[IX(i, j, k)] + dx3[k] * invRho[IX(i, j, km)]);
                double invRho6 = 0.5 / dmx3[kp] * (dx3[kp] * invRho[IX(i, j, k)] + dx3[k] * invRho[IX(i, j, kp)]);
                double ExtraTerm = ((1.0 - invRho2) * ddpdx2 - (1.0 - invRho1) * ddpdx1) / dx1[i]
                                 + ((1.0 - invRho4) * ddpdy4 - (1.0 - invRho3) * ddpdy3) / dx2[j]
                                 + ((1.0 - invRho6) * ddpdz6 - (1.0 - invRho5) * ddpdz5) / dx3[k];
                PRHS[R3(o->n1msub, o->n2msub, i, j, k)] = (DivUm - cbc) / dt / Cmp + ExtraTerm;
            }
        }
    }
}

/* radix-2 Stockham FFT (deviation D1) -------------------------------------- */
void orc_twiddles(int n, double *wr, double *wi)
{
    const double PI = acos(-1.0);
    for (int k = 0; k < n / 2; k++) {
        double a = 2.0 * PI * (double)k / (double)n;
        wr[k] = cos(a);
        wi[k] = -sin(a);
    }
}

/* sign=-1: X_k = sum x_n exp(-2 pi i k n / N); sign=+1: conjugate twiddles.
   Stage s (Ns = 2^s): butterfly j in [0,N/2): kk = j mod Ns,
   w = W_N^(kk*N/(2Ns)), a = x[j], b = w*x[j+N/2],
   y[(j/Ns)*2Ns + kk] = a+b, y[(j/Ns)*2Ns + kk + Ns] = a-b. */
void orc_fft(int n, double *re, double *im, int sign)
{
    double *wr = (double *)malloc(sizeof(double) * (size_t)(n / 2 + 1));
    double *wi = (double *)malloc(sizeof(double) * (size_t)(n / 2 + 1));
    double *bufa_r = (double *)malloc(sizeof(double) * (size_t)n);
    double *bufa_i = (double *)malloc(sizeof(double) * (size_t)n);
    double *bufb_r = (double *)malloc(sizeof(double) * (size_t)n);
    double *bufb_i = (double *)malloc(sizeof(double) * (size_t)n);
    orc_twiddles(n, wr, wi);
    memcpy(bufa_r, re, sizeof(double) * (size_t)n);
    memcpy(bufa_i, im, sizeof(double) * (size_t)n);
    double *xr = bufa_r, *xi = bufa_i, *yr = bufb_r, *yi = bufb_i;
    int half = n / 2;
    for (int Ns = 1; Ns < n; Ns *= 2) {
        for (int j = 0; j < half; j++) {
            int kk = j & (Ns - 1);
            int t = kk * (n / (2 * Ns));
            double twr = wr[t], twi = wi[t];
            double ar = xr[j], ai = xi[j];
            double br = xr[j + half], bi = xi[j + half];
            double tr, ti;
            if (sign < 0) { tr = br * twr - bi * twi; ti = br * twi + bi * twr; }
            else { tr = br * twr + bi * twi; ti = bi * twr - br * twi; }
            int oidx = (j / Ns) * 2 * Ns + kk;
            yr[oidx] = ar + tr; yi[oidx] = ai + ti;
            yr[oidx + Ns] = ar - tr; yi[oidx + Ns] = ai - ti;
        }
        double *sr = xr, *si = xi;
        xr = yr; xi = yi; yr = sr; yi = si;
    }
    memcpy(re, xr, sizeof(double) * (size_t)n);
    memcpy(im, xi, sizeof(double) * (size_t)n);
    free(bufa_r); free(bufa_i); free(bufb_r); free(bufb_i);
    free(wr); free(wi);
}

/* deterministic pairwise sum over n values zero-padded to a power of two (D2) */
double orc_tree_sum(const double *x, int n)
{
    int P = 1;
    while (P < n) P *= 2;
    double *b = (double *)malloc(sizeof(double) * (size_t)P);
    for (int i = 0; i < P; i++) b[i] = (i < n) ? x[i] : 0.0;
    for (int h = P / 2; h >= 1; h /= 2)
        for (int i = 0; i < h; i++) b[i] = b[i] + b[i + h];
    double s = b[0];
    free(b);
    return s;
}

/* :489-744 on one rank; prhs is (1:n1m,1:n2m,1:n3m) */
void orc_pressure_Poisson_FFT2(orc *o, const double *prhs)
{
    const int n1m = o->n1m, n2m = o->n2m, n3m = o->n3m, h1p = o->h1p, h3p = o->h3p;
    const double *dx2 = o->dx2, *dmx2 = o->dmx2;
    double *dP = o->f[OF_DP];
    size_t ns = (size_t)h1p * n2m * n3m;
    /* spectrum after x r2c: (1:h1p,1:n2m,1:n3m) */
    double *Sr = (double *)calloc(ns, sizeof(double));
    double *Si = (double *)calloc(ns, sizeof(double));
    const size_t nl = (size_t)(n1m > n3m ? n1m : n3m);  /* one line's buffer (per thread) */
    OMP(omp parallel)
    {
    double *lr = (double *)malloc(sizeof(double) * nl), *li = (double *)malloc(sizeof(double) * nl);
    OMP(omp for)
    for (int k = 1; k <= n3m; k++)
        for (int j = 1; j <= n2m; j++) {
            for (int i = 1; i <= n1m; i++) { lr[i - 1] = prhs[R3(n1m, n2m, i, j, k)]; li[i - 1] = 0.0; }
            orc_fft(n1m, lr, li, -1);
            for (int kx = 0; kx < h1p; kx++) {
                Sr[R3(h1p, n2m, kx + 1, j, k)] = lr[kx];
                Si[R3(h1p, n2m, kx + 1, j, k)] = li[kx];
            }
        }
    free(lr); free(li);
    }
    /* z c2c forward into (1:n3m,1:h1p,1:n2m) */
    double *Kr = (double *)calloc(ns, sizeof(double));
    double *Ki = (double *)calloc(ns, sizeof(double));
    OMP(omp parallel)
    {
    double *lr = (double *)malloc(sizeof(double) * nl), *li = (double *)malloc(sizeof(double) * nl);
    OMP(omp for)
    for (int j = 1; j <= n2m; j++)
        for (int kx = 1; kx <= h1p; kx++) {
            for (int k = 1; k <= n3m; k++) { lr[k - 1] = Sr[R3(h1p, n2m, kx, j, k)]; li[k - 1] = Si[R3(h1p, n2m, kx, j, k)]; }
            orc_fft(n3m, lr, li, -1);
            for (int kz = 1; kz <= n3m; kz++) {
                Kr[R3(n3m, h1p, kz, kx, j)] = lr[kz - 1];
                Ki[R3(n3m, h1p, kz, kx, j)] = li[kz - 1];
            }
        }
    free(lr); free(li);
    }
    /* tridiagonal systems in y (:599-640) */
    double *Am_r = (double *)malloc(ns * sizeof(double)), *Ac_r = (double *)malloc(ns * sizeof(double)), *Ap_r = (double *)malloc(ns * sizeof(double));
    double *Am_c = (double *)malloc(ns * sizeof(double)), *Ac_c = (double *)malloc(ns * sizeof(double)), *Ap_c = (double *)malloc(ns * sizeof(double));
    OMP(omp parallel for)
    for (int j = 1; j <= n2m; j++) {
        int jp = j + 1;
        double fft_amj = 1.0 / dx2[j] / dmx2[j];
        double fft_apj = 1.0 / dx2[j] / dmx2[jp];
        if (j == 1) fft_amj = 0.0;
        if (j == n2m) fft_apj = 0.0;
        double fft_acj = -fft_amj - fft_apj;
        for (int kx = 1; kx <= h1p; kx++)
            for (int kz = 1; kz <= n3m; kz++) {
                size_t q = R3(n3m, h1p, kz, kx, j);
                Am_r[q] = fft_amj; Ac_r[q] = fft_acj - o->dxk2[kx - 1] - o->dzk[kz - 1]; Ap_r[q] = fft_apj;
                Am_c[q] = fft_amj; Ac_c[q] = fft_acj - o->dxk2[kx - 1] - o->dzk[kz - 1]; Ap_c[q] = fft_apj;
            }
    }
    Am_r[0] = 0.0; Ac_r[0] = 1.0; Ap_r[0] = 0.0; Kr[0] = 0.0;
    lines_solve(o, Am_r, Ac_r, Ap_r, Kr, n3m * h1p, n2m, 0);
    lines_solve(o, Am_c, Ac_c, Ap_c, Ki, n3m * h1p, n2m, 0);
    for (int j = 1; j <= n2m; j++) {
        Ki[R3(n3m, h1p, 1, 1, j)] = 0.0;
        Ki[R3(n3m, h1p, h3p, 1, j)] = 0.0;
    }
    /* z c2c backward into (1:h1p,1:n2m,1:n3m) */
    OMP(omp parallel)
    {
    double *lr = (double *)malloc(sizeof(double) * nl), *li = (double *)malloc(sizeof(double) * nl);
    OMP(omp for)
    for (int j = 1; j <= n2m; j++)
        for (int kx = 1; kx <= h1p; kx++) {
            for (int kz = 1; kz <= n3m; kz++) { lr[kz - 1] = Kr[R3(n3m, h1p, kz, kx, j)]; li[kz - 1] = Ki[R3(n3m, h1p, kz, kx, j)]; }
            orc_fft(n3m, lr, li, +1);
            for (int k = 1; k <= n3m; k++) {
                Sr[R3(h1p, n2m, kx, j, k)] = lr[k - 1];
                Si[R3(h1p, n2m, kx, j, k)] = li[k - 1];
            }
        }
    free(lr); free(li);
    }
    /* x c2r, scale, tree sums */
    double *tmp = (double *)malloc(sizeof(double) * (size_t)n1m * n2m * n3m);
    double *planesum = (double *)malloc(sizeof(double) * (size_t)n3m);
    OMP(omp parallel)
    {
    double *lr = (double *)malloc(sizeof(double) * (size_t)n1m), *li = (double *)malloc(sizeof(double) * (size_t)n1m);
    double *linesum = (double *)malloc(sizeof(double) * (size_t)n2m);
    OMP(omp for)
    for (int k = 1; k <= n3m; k++) {
        for (int j = 1; j <= n2m; j++) {
            for (int kx = 0; kx < h1p; kx++) {
                lr[kx] = Sr[R3(h1p, n2m, kx + 1, j, k)];
                li[kx] = Si[R3(h1p, n2m, kx + 1, j, k)];
            }
            li[0] = 0.0;
            li[n1m / 2] = 0.0;
            for (int kx = 1; kx < n1m / 2; kx++) { lr[n1m - kx] = lr[kx]; li[n1m - kx] = -li[kx]; }
            orc_fft(n1m, lr, li, +1);
            for (int i = 1; i <= n1m; i++) {
                double v = lr[i - 1] / (double)n1m / (double)n3m;
                tmp[R3(n1m, n2m, i, j, k)] = v;
            }
            linesum[j - 1] = orc_tree_sum(&tmp[R3(n1m, n2m, 1, j, k)], n1m);
        }
        planesum[k - 1] = orc_tree_sum(linesum, n2m);
    }
    free(lr); free(li); free(linesum);
    }
    double AVER = orc_tree_sum(planesum, n3m);
    AVER = AVER / (double)n1m / (double)n2m / (double)n3m;
    OMP(omp parallel for)
    for (int k = 1; k <= n3m; k++)
        for (int j = 1; j <= n2m; j++)
            for (int i = 1; i <= n1m; i++)
                dP[IX(i, j, k)] = tmp[R3(n1m, n2m, i, j, k)] - AVER;
    /* wall extrapolation (:719-736), k=1..n3msub (D3) */
    {
        double s = dmx2[1] + dmx2[2];
        double Pbc_a = (s * s) / ((s * s) - (dmx2[1] * dmx2[1]));
        double Pbc_b = (dmx2[1] * dmx2[1]) / ((s * s) - (dmx2[1] * dmx2[1]));
        for (int k = 1; k <= o->n3msub; k++)
            for (int i = 1; i <= o->n1msub; i++)
                dP[IX(i, 0, k)] = Pbc_a * dP[IX(i, 1, k)] - Pbc_b * dP[IX(i, 2, k)];
        int n2sub = o->n2sub, n2msub = o->n2msub;
        double t = dmx2[n2sub] + dmx2[n2msub];
        Pbc_a = (t * t) / ((t * t) - (dmx2[n2sub] * dmx2[n2sub]));
        Pbc_b = (dmx2[n2sub] * dmx2[n2sub]) / ((t * t) - (dmx2[n2sub] * dmx2[n2sub]));
        for (int k = 1; k <= o->n3msub; k++)
            for (int i = 1; i <= o->n1msub; i++)
                dP[IX(i, n2sub, k)] = Pbc_a * dP[IX(i, n2msub, k)] - Pbc_b * dP[IX(i, n2sub - 2, k)];
    }
    free(Sr); free(Si); free(Kr); free(Ki);
    free(Am_r); free(Ac_r); free(Ap_r); free(Am_c); free(Ac_c); free(Ap_c);
    free(tmp); free(planesum);
}

/* :756-874 */
void orc_pressure_Projection(orc *o, double dt)
{
    const double Cmp = o->prm[OP_CMP];
    double *U = o->f[OF_U], *V = o->f[OF_V], *W = o->f[OF_W], *P = o->f[OF_P];
    double *dP = o->f[OF_DP], *dPhat = o->f[OF_DPHAT], *dP0 = o->f[OF_DP0];
    const double *invRho = o->f[OF_INVRHO];
    const double *dx1 = o->dx1, *dx2 = o->dx2, *dx3 = o->dx3, *dmx1 = o->dmx1, *dmx2 = o->dmx2, *dmx3 = o->dmx3;
    /* the gradients read dPhat as the loop nest leaves it (Q26); the
       planted site 315 reads a copy of the old dPhat instead */
    double *dPh_old = NULL;
    if (PLANT(315, 1, 0)) {
        size_t nb = (size_t)o->N1 * o->N2 * o->N3 * sizeof(double);
        dPh_old = (double *)malloc(nb);
        memcpy(dPh_old, dPhat, nb);
    }
    const double *dPr = dPh_old ? dPh_old : dPhat;
    for (int k = 1; k <= o->n3msub; k++) {
        int km = k - 1;
        for (int j = 2; j <= o->n2msub; j++) {
            int jm = j - 1;
            for (int i = 1; i <= o->n1msub; i++) {
                int im = i - 1;
                double invRhoc = 0.5 / dmx1[i] * (dx1[i] * invRho[IX(im, j, k)] + dx1[im] * invRho[IX(i, j, k)]);
                U[IX(i, j, k)] = U[IX(i, j, k)]
                               - dt * Cmp * ((dP[IX(i, j, k)] - dP[IX(im, j, k)]) / dmx1[i]
                                             + (invRhoc - 1.0) * (dPr[IX(i, j, k)] - dPr[IX(im, j, k)]) / dmx1[i]);
                invRhoc = 0.5 / dmx2[j] * (dx2[j] * invRho[IX(i, jm, k)] + dx2[jm] * invRho[IX(i, j, k)]);
                V[IX(i, j, k)] = V[IX(i, j, k)]
                               - dt * Cmp * ((dP[IX(i, j, k)] - dP[IX(i, jm, k)]) / dmx2[j]
                                             + (invRhoc - 1.0) * (dPr[IX(i, j, k)] - dPr[IX(i, jm, k)]) / dmx2[j]);
                invRhoc = 0.5 / dmx3[k] * (dx3[k] * invRho[IX(i, j, km)] + dx3[km] * invRho[IX(i, j, k)]);
                W[IX(i, j, k)] = W[IX(i, j, k)]
                               - dt * Cmp * ((dP[IX(i, j, k)] - dP[IX(i, j, km)]) / dmx3[k]
                                             + (invRhoc - 1.0) * (dPr[IX(i, j, k)] - dPr[IX(i, j, km)]) / dmx3[k]);
                P[IX(i, j, k)] = P[IX(i, j, k)] + dP[IX(i, j, k)];
                dP0[IX(i, j, k)] = (dPhat[IX(i, j, k)] + dP0[IX(i, j, k)]) / 2.0;
                dPhat[IX(i, j, k)] = 2.0 * dP[IX(i, j, k)] - PLANT(311, 2.0 * dP0[IX(i, j, k)] - dPhat[IX(i, j, k)], dP0[IX(i, j, k)]);
            }
        }
        {
            int j = 1, jm = 0;
            for (int i = 1; i <= o->n1msub; i++) {
                int im = i - 1;
                double invRhoc = 0.5 / dmx1[i] * (dx1[i] * invRho[IX(im, j, k)] + dx1[im] * invRho[IX(i, j, k)]);
                U[IX(i, j, k)] = U[IX(i, j, k)]
                               - dt * Cmp * ((dP[IX(i, j, k)] - dP[IX(im, j, k)]) / dmx1[i]
                                             + (invRhoc - 1.0) * (dPr[IX(i, j, k)] - dPr[IX(im, j, k)]) / dmx1[i]);
                invRhoc = 0.5 / dmx2[j] * (dx2[j] * invRho[IX(i, jm, k)] + dx2[jm] * invRho[IX(i, j, k)]);
                V[IX(i, j, k)] = V[IX(i, j, k)]
                               - dt * Cmp * ((dP[IX(i, j, k)] - dP[IX(i, jm, k)]) / dmx2[j]
                                             + (invRhoc - 1.0) * (dPr[IX(i, j, k)] - dPr[IX(i, jm, k)]) / dmx2[j]) * PLANT(312, 1.0, (double)(2 - o->j_indexS));
                invRhoc = 0.5 / dmx3[k] * (dx3[k] * invRho[IX(i, j, km)] + dx3[km] * invRho[IX(i, j, k)]);
                W[IX(i, j, k)] = W[IX(i, j, k)]
                               - dt * Cmp * ((dP[IX(i, j, k)] - dP[IX(i, j, km)]) / dmx3[k]
                                             + (invRhoc - 1.0) * (dPr[IX(i, j, k)] - dPr[IX(i, j, km)]) / dmx3[k]);
                P[IX(i, j, k)] = P[IX(i, j, k)] + dP[IX(i, j, k)];
                dP0[IX(i, j, k)] = (dPhat[IX(i, j, k)] + dP0[IX(i, j, k)]) / 2.0;
                dPhat[IX(i, j, k)] = 2.0 * dP[IX(i, j, k)] - PLANT(311, 2.0 * dP0[IX(i, j, k)] - dPhat[IX(i, j, k)], dP0[IX(i, j, k)]);
            }
        }
        {
            double s = dmx2[1] + dmx2[2];
            double Pbc_a = (s * s) / ((s * s) - (dmx2[1] * dmx2[1]));
            double Pbc_b = (dmx2[1] * dmx2[PLANT(313, 2, 1)]) / ((s * s) - (dmx2[1] * dmx2[1]));
            for (int i = 1; i <= o->n1sub - 1; i++)
                P[IX(i, 0, k)] = Pbc_a * P[IX(i, 1, k)] - Pbc_b * P[IX(i, 2, k)];
            int n2sub = o->n2sub, n2msub = o->n2msub;
            double t = dmx2[n2sub] + dmx2[n2msub];
            Pbc_a = (t * t) / ((t * t) - (dmx2[n2sub] * dmx2[n2sub]));
            Pbc_b = (dmx2[n2sub] * dmx2[n2sub]) / ((t * t) - (dmx2[n2sub] * dmx2[n2sub]));
            for (int i = 1; i <= o->n1sub - 1; i++)
                P[IX(i, n2sub, k)] = Pbc_a * P[IX(i, n2msub, k)] - Pbc_b * P[IX(i, n2sub - 2, k)];
        }
    }
    free(dPh_old);
}

/* ------------------------------------------------------------------------ */
/* post: src/module_post.f90:100-230                                         */
/* ------------------------------------------------------------------------ */
double orc_post_div(orc *o)
{
    const double *U = o->f[OF_U], *V = o->f[OF_V], *W = o->f[OF_W];
    double mx = -INFINITY;
    OMP(omp parallel for reduction(max : mx))
    for (int k = 1; k <= o->n3sub - 1; k++)
        for (int j = 1; j <= o->n2sub - 1; j++)
            for (int i = 1; i <= o->n1sub - 1; i++) {
                double d = (U[IX(i + 1, j, k)] - U[IX(i, j, k)]) / o->dx1[i]
                         + (V[IX(i, j + 1, k)] - V[IX(i, j, k)]) / o->dx2[j]
                         + (W[IX(i, j, k + 1)] - W[IX(i, j, k)]) / o->dx3[k];
                d = PLANT(323, fabs(d), d);
                if (d > mx) mx = d;
            }
    return mx;
}

double orc_post_cfl(orc *o, double dt, double *newdt)
{
    const double *U = o->f[OF_U], *V = o->f[OF_V], *W = o->f[OF_W];
    double mx = -INFINITY;
    OMP(omp parallel for reduction(max : mx))
    for (int k = 1; k <= o->n3sub - 1; k++)
        for (int j = 1; j <= o->n2sub - 1; j++)
            for (int i = 1; i <= o->n1sub - 1; i++) {
                double c = PLANT(322, (fabs(U[IX(i + 1, j, k)]) + fabs(U[IX(i, j, k)])) / 2.0,
                                 fabs(U[IX(i + 1, j, k)] + U[IX(i, j, k)]) / 2.0) / o->dx1[i]
                         + (fabs(V[IX(i, j + 1, k)] + V[IX(i, j, k)]) / 2.0) / PLANT(321, o->dmx2[j], o->dx2[j])
                         + (fabs(W[IX(i, j, k + 1)] + W[IX(i, j, k)]) / 2.0) / o->dx3[k];
                if (c > mx) mx = c;
            }
    (void)dt;
    double nd = o->prm[OP_MAXCFL] / mx;
    if (nd > o->prm[OP_DTMAX]) nd = o->prm[OP_DTMAX];
    *newdt = nd;
    return nd * mx; /* newCFL */
}

/* ------------------------------------------------------------------------ */
/* the time loop body: src/main.f90:131-203                                  */
/* ------------------------------------------------------------------------ */
void orc_step(orc *o, double dt)
{
    size_t n = (size_t)o->n1m * o->n2m * o->n3m;
    double *prhs = (double *)malloc(n * sizeof(double));
    orc_thermal_coeffi(o);
    orc_thermal_solver(o, dt);
    orc_ghostcell_update(o, o->f[OF_T]);
    orc_momentum_coeffi(o);
    orc_momentum_solvedU(o, dt);
    orc_ghostcell_update(o, o->f[OF_DU]);
    orc_momentum_solvedV(o, dt);
    orc_ghostcell_update(o, o->f[OF_DV]);
    orc_momentum_solvedW(o, dt);
    orc_ghostcell_update(o, o->f[OF_DW]);
    orc_momentum_blockLdV(o, dt);
    orc_ghostcell_update(o, o->f[OF_DV]);
    orc_momentum_blockLdU(o, dt);
    orc_momentum_pseudoupdateUVW(o);
    orc_ghostcell_update(o, o->f[OF_U]);
    orc_ghostcell_update(o, o->f[OF_V]);
    orc_ghostcell_update(o, o->f[OF_W]);
    orc_pressure_RHS(o, dt, prhs);
    orc_pressure_Poisson_FFT2(o, prhs);
    orc_ghostcell_update(o, o->f[OF_DP]);
    orc_pressure_Projection(o, dt);
    orc_ghostcell_update(o, o->f[OF_U]);
    orc_ghostcell_update(o, o->f[OF_V]);
    orc_ghostcell_update(o, o->f[OF_W]);
    orc_ghostcell_update(o, o->f[OF_P]);
    orc_ghostcell_update(o, o->f[OF_DPHAT]);
    free(prhs);
}

/* ------------------------------------------------------------------------ */
/* The PaScaL_TDMA heat-conduction example on one rank:                       */
/* PaScaL_TDMA/examples/solve_theta_plan_many.f90:30-203 (one time step) with */
/* the single-rank ghost-cell update of examples/mpi_subdomain.f90:219-249.  */
/* theta is (0:nx_sub,0:ny_sub,0:nz_sub), i fastest; bc3/bc4 (0:nx_sub,      */
/* 0:nz_sub); dmx/dmy/dmz (0:n_sub); jmbc/jpbc (0:ny_sub).                    */
/* ------------------------------------------------------------------------ */
void orc_heat_step(int nx_sub, int ny_sub, int nz_sub, double *theta, const double *bc3, const double *bc4,
                   const double *dmx, const double *dmy, const double *dmz, const int *jmbc, const int *jpbc,
                   double dx, double dy, double dz, double Ct, double dt)
{
    const size_t NX = (size_t)nx_sub + 1, NY = (size_t)ny_sub + 1;
#define TH(i, j, k) theta[(size_t)(i) + NX * ((size_t)(j) + NY * (size_t)(k))]
#define RH(i, j, k) rhs[(size_t)(i) + NX * ((size_t)(j) + NY * (size_t)(k))]
    double *rhs = (double *)calloc(NX * NY * ((size_t)nz_sub + 1), sizeof(double));
    /* right-hand side (:63-117) */
    for (int k = 1; k <= nz_sub - 1; k++) {
        int kp = k + 1, km = k - 1;
        for (int j = 1; j <= ny_sub - 1; j++) {
            int jp = j + 1, jm = j - 1;
            int jep = jpbc[j], jem = jmbc[j];
            for (int i = 1; i <= nx_sub - 1; i++) {
                int ip = i + 1, im = i - 1;
                double dedx1 = (TH(i, j, k) - TH(im, j, k)) / dmx[i];
                double dedx2 = (TH(ip, j, k) - TH(i, j, k)) / dmx[ip];
                double dedy3 = (TH(i, j, k) - TH(i, jm, k)) / dmy[j];
                double dedy4 = (TH(i, jp, k) - TH(i, j, k)) / dmy[jp];
                double dedz5 = (TH(i, j, k) - TH(i, j, km)) / dmz[k];
                double dedz6 = (TH(i, j, kp) - TH(i, j, k)) / dmz[kp];
                double viscous_e1 = 1.0 / dx * (dedx2 - dedx1);
                double viscous_e2 = 1.0 / dy * (dedy4 - dedy3);
                double viscous_e3 = 1.0 / dz * (dedz6 - dedz5);
                double viscous = 0.5 * Ct * (viscous_e1 + viscous_e2 + viscous_e3);
                double ebc_down = 0.5 * Ct / dy / dmy[j] * bc3[(size_t)i + NX * (size_t)k];
                double ebc_up = 0.5 * Ct / dy / dmy[jp] * bc4[(size_t)i + NX * (size_t)k];
                double ebc = (double)(1.0f - (float)jem) * ebc_down + (double)(1.0f - (float)jep) * ebc_up;
                double eAPI = -0.5 * Ct / dx / dmx[ip];
                double eAMI = -0.5 * Ct / dx / dmx[i];
                double eACI = 0.5 * Ct / dx * (1.0 / dmx[ip] + 1.0 / dmx[i]);
                double eAPK = -0.5 * Ct / dz / dmz[kp];
                double eAMK = -0.5 * Ct / dz / dmz[k];
                double eACK = 0.5 * Ct / dz * (1.0 / dmz[kp] + 1.0 / dmz[k]);
                double eAPJ = -0.5 * Ct / dy * (1.0 / dmy[jp]) * (double)jep;
                double eAMJ = -0.5 * Ct / dy * (1.0 / dmy[j]) * PLANT(402, 1.0, (double)jem);
                double eACJ = 0.5 * Ct / dy * (1.0 / dmy[jp] + 1.0 / dmy[j]);
                double eRHS = eAPK * TH(i, j, kp) + eACK * TH(i, j, k) + eAMK * TH(i, j, km)
                            + eAPJ * TH(i, jp, k) + eACJ * TH(i, j, k) + eAMJ * TH(i, jm, k)
                            + eAPI * TH(ip, j, k) + eACI * TH(i, j, k) + eAMI * TH(im, j, k);
                RH(i, j, k) = TH(i, j, k) / dt + viscous + ebc - (TH(i, j, k) / dt + eRHS);
            }
        }
    }
    const int nxm = nx_sub - 1, nym = ny_sub - 1, nzm = nz_sub - 1;
    size_t big = (size_t)(nxm > nym ? nxm : nym) * (size_t)(nym > nzm ? nym : nzm);
    if ((size_t)nxm * nzm > big) big = (size_t)nxm * nzm;
    double *ap = calloc(big, sizeof(double)), *am = calloc(big, sizeof(double));
    double *ac = calloc(big, sizeof(double)), *ad = calloc(big, sizeof(double));
    /* z: per plane j, systems i, rows k, cyclic (:122-149) */
    for (int j = 1; j <= nym; j++) {
        for (int k = 1; k <= nzm; k++) {
            int kp = k + 1;
            for (int i = 1; i <= nxm; i++) {
                size_t q = (size_t)(i - 1) + (size_t)nxm * (size_t)(k - 1);
                ap[q] = -0.5 * Ct / dz / dmz[kp] * dt;
                am[q] = -0.5 * Ct / dz / dmz[k] * dt;
                ac[q] = 0.5 * Ct / dz * (1.0 / dmz[kp] + 1.0 / dmz[k]) * dt + 1.0;
                ad[q] = RH(i, j, k) * PLANT(401, 1.0, dt);
            }
        }
        orc_ptdma_many_solve_cycle(am, ac, ap, ad, nxm, nzm);
        for (int k = 1; k <= nzm; k++)
            for (int i = 1; i <= nxm; i++) RH(i, j, k) = ad[(size_t)(i - 1) + (size_t)nxm * (size_t)(k - 1)];
    }
    /* y: per plane k, systems i, rows j (:158-186) */
    for (int k = 1; k <= nzm; k++) {
        for (int j = 1; j <= nym; j++) {
            int jp = j + 1;
            int jep = jpbc[j], jem = jmbc[j];
            for (int i = 1; i <= nxm; i++) {
                size_t q = (size_t)(i - 1) + (size_t)nxm * (size_t)(j - 1);
                ap[q] = -0.5 * Ct / dy / dmy[jp] * (double)jep * dt;
                am[q] = -0.5 * Ct / dy / dmy[j] * (double)jem * dt;
                ac[q] = 0.5 * Ct / dy * (1.0 / dmy[jp] + 1.0 / dmy[j]) * dt + 1.0;
                ad[q] = RH(i, j, k);
            }
        }
        orc_ptdma_many_solve(am, ac, ap, ad, nxm, nym);
        for (int j = 1; j <= nym; j++)
            for (int i = 1; i <= nxm; i++) RH(i, j, k) = ad[(size_t)(i - 1) + (size_t)nxm * (size_t)(j - 1)];
    }
    /* x: per plane k, systems j, rows i, cyclic; theta += solution (:191-218) */
    for (int k = 1; k <= nzm; k++) {
        for (int j = 1; j <= nym; j++)
            for (int i = 1; i <= nxm; i++) {
                int ip = i + 1;
                size_t q = (size_t)(j - 1) + (size_t)nym * (size_t)(i - 1);
                ap[q] = -0.5 * Ct / dx / dmx[ip] * dt;
                am[q] = -0.5 * Ct / dx / dmx[i] * dt;
                ac[q] = 0.5 * Ct / dx * (1.0 / dmx[ip] + 1.0 / dmx[i]) * dt + 1.0;
                ad[q] = RH(i, j, k) * PLANT(403, dt, 1.0);
            }
        orc_ptdma_many_solve_cycle(am, ac, ap, ad, nym, nxm);
        for (int j = 1; j <= nym; j++)
            for (int i = 1; i <= nxm; i++)
                TH(i, j, k) = TH(i, j, k) + ad[(size_t)(j - 1) + (size_t)nym * (size_t)(i - 1)];
    }
    /* ghost cells: x and z periodic on one rank, y walls untouched */
    for (int k = 0; k <= nz_sub; k++)
        for (int j = 0; j <= ny_sub; j++) {
            TH(0, j, k) = TH(nx_sub - 1, j, k);
            TH(nx_sub, j, k) = TH(1, j, k);
        }
    for (int j = 0; j <= ny_sub; j++)
        for (int i = 0; i <= nx_sub; i++) {
            TH(i, j, 0) = TH(i, j, nz_sub - 1);
            TH(i, j, nz_sub) = TH(i, j, 1);
        }
    free(rhs); free(ap); free(am); free(ac); free(ad);
#undef TH
#undef RH
}
