/*
 * tcs_oracle.h -- CPU restatement of the PaScaL_TCS MPM-STD time step on a
 * single rank (np1=np2=np3=1), used ONLY as test infrastructure: tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it as the
 * checker.  Nothing in the product (pascal_tcs_amd/, include/) includes,
 * links or calls it.
 *
 * Every routine restates the reference loop nest it cites, in the same
 * operation order (left-to-right sums, the reference's parenthesisation,
 * true FP64 divisions, no FMA contraction: build with -ffp-contract=off),
 * including the reference's stale-index quirks.  Deliberate deviations are
 * listed in the header comment of tcs_oracle.c.
 *
 * Layouts: full fields are (0:n1sub,0:n2sub,0:n3sub) FP64 with i fastest
 * (Fortran order), n?sub = n?m+1; wall tables are (0:n1sub,0:n3sub).
 */
#ifndef TCS_ORACLE_H
#define TCS_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* physical constants, same meaning as module global (src/module_global.f90:30-70) */
enum {
    OP_CT = 0, OP_CMU, OP_CMT, OP_CMP, OP_DELTAT,
    OP_A10, OP_A11, OP_A12, OP_A13, OP_A14, OP_A15,
    OP_B10, OP_B11, OP_B12, OP_B13, OP_B14, OP_B15,
    OP_C10, OP_C11, OP_C12, OP_C13, OP_C14, OP_C15,
    OP_D10, OP_D11, OP_D12, OP_D13, OP_D14, OP_D15,
    OP_KPP0, OP_RHO0, OP_CP0, OP_MU0, OP_A12PERA11,
    OP_L1, OP_L3, OP_MAXCFL, OP_DTMAX,
    OP_N
};

/* field ids (module variables of mpi_thermal / mpi_momentum / mpi_pressure) */
enum {
    OF_T = 0, OF_U, OF_V, OF_W, OF_P,
    OF_DU, OF_DV, OF_DW, OF_DP, OF_DPHAT, OF_DP0,
    OF_INVRHO, OF_MU, OF_KPP, OF_DKPP, OF_INVRHOCP, OF_DINVRHOCP,
    OF_NFIELD
};
/* wall tables (0:n1sub,0:n3sub) */
enum {
    OB_TUP = 0, OB_TBT, OB_UUP, OB_UBT, OB_VUP, OB_VBT, OB_WUP, OB_WBT, OB_N
};

typedef struct orc orc;

/* Test hooks.  orc_capture(o, 1) records, for every line solve a routine
   makes, its rows (a, b, c), right-hand side d and solution x; (o, 0) stops
   and clears.  Records are (n_row, n_sys) with the system index fastest,
   in the order of the solves (at most ORC_MAXCAP).  orc_set_plant(id)
   selects a planted mistake in the oracle built with -DORC_PLANTS
   (liboracle_plants.so); orc_plants_built() tells the builds apart. */
#define ORC_MAXCAP 32
void orc_capture(orc *o, int on);
int orc_capture_count(orc *o);
int orc_capture_info(orc *o, int idx, int *n_sys, int *n_row, int *cyclic);
int orc_capture_copy(orc *o, int idx, double *a, double *b, double *c, double *d, double *x);
void orc_set_plant(int id);
int orc_plants_built(void);

/* one axis of mpi_subdomain_mesh on a single rank; x,dx,dmx have nm+2 entries */
void orc_mesh_axis(int nm, double L, int uniform, double gamma, int periodic,
                   double *x, double *dx, double *dmx);

/* metrics (0:n?sub) and physical constants are copied in; pbc = (T,F,T) */
orc *orc_create(int n1m, int n2m, int n3m,
                const double *dx1, const double *dx2, const double *dx3,
                const double *dmx1, const double *dmx2, const double *dmx3,
                const double *prm);
void orc_destroy(orc *o);
double *orc_field(orc *o, int id);
double *orc_bc(orc *o, int id);
void orc_wave_numbers(orc *o, double *dzk, double *dxk2);

/* module routines (names of the reference) */
void orc_thermal_boundary(orc *o);
void orc_momentum_boundary(orc *o);
void orc_thermal_coeffi(orc *o);
void orc_thermal_solver(orc *o, double dt);
void orc_momentum_coeffi(orc *o);
void orc_momentum_solvedU(orc *o, double dt);
void orc_momentum_solvedV(orc *o, double dt);
void orc_momentum_solvedW(orc *o, double dt);
void orc_momentum_blockLdV(orc *o, double dt);
void orc_momentum_blockLdU(orc *o, double dt);
void orc_momentum_pseudoupdateUVW(orc *o);
void orc_pressure_wave_number(orc *o);
void orc_pressure_RHS(orc *o, double dt, double *prhs_out);
void orc_pressure_Poisson_FFT2(orc *o, const double *prhs);
void orc_pressure_Projection(orc *o, double dt);
void orc_ghostcell_update(orc *o, double *f);
double orc_post_div(orc *o);
double orc_post_cfl(orc *o, double dt, double *newdt);
/* the time loop body of src/main.f90:129-223 (without IO / blow-up exit) */
void orc_step(orc *o, double dt);

/* PaScaL_TDMA (single rank): arrays (n_sys,n_row), system index fastest */
void orc_ptdma_many_solve(double *A, double *B, double *C, double *D, int n_sys, int n_row);
void orc_ptdma_many_solve_cycle(double *A, double *B, double *C, double *D, int n_sys, int n_row);
/* the same solve with the rows of each system split over nprocs ranks
   (para_range), reduced systems of 2*nprocs rows; nprocs=1 is the above */
/* parts per line of the split line solves, and the switch that selects
   them for the step's sweeps (default 0: single-rank solves) */
int orc_tdma_parts(int n_row);
void orc_set_tdma_split(orc *o, int split);
void orc_ptdma_many_solve_ranks(double *A, double *B, double *C, double *D, int n_sys, int n_row,
                                int nprocs, int cyclic);
/* the split solve of the GPU kernels: orc_tdma_parts(n_row) ranks, reduced
   system by parallel cyclic reduction when there are two or more parts */
void orc_ptdma_many_solve_split(double *A, double *B, double *C, double *D, int n_sys, int n_row, int cyclic);
void orc_tdma_many(double *a, double *b, double *c, double *d, int n1, int n2);
void orc_tdma_cycl_many(double *a, double *b, double *c, double *d, int n1, int n2);
void orc_tdma_single(double *a, double *b, double *c, double *d, int n1);
void orc_tdma_cycl_single(double *a, double *b, double *c, double *d, int n1);

/* one step of the PaScaL_TDMA heat-conduction example on one rank
   (PaScaL_TDMA/examples/solve_theta_plan_many.f90:30-203 and the periodic
   ghost update of examples/mpi_subdomain.f90:219-249): theta
   (0:nx_sub,0:ny_sub,0:nz_sub) in place; wall tables (0:nx_sub,0:nz_sub);
   metrics (0:n_sub); jmbc/jpbc (0:ny_sub) */
void orc_heat_step(int nx_sub, int ny_sub, int nz_sub, double *theta, const double *bc3, const double *bc4,
                   const double *dmx, const double *dmy, const double *dmz, const int *jmbc, const int *jpbc,
                   double dx, double dy, double dz, double Ct, double dt);

/* FFT building blocks of the Poisson solve (radix-2 Stockham, power-of-two n) */
void orc_twiddles(int n, double *wr, double *wi);
void orc_fft(int n, double *re, double *im, int sign);
double orc_tree_sum(const double *x, int n);

#ifdef __cplusplus
}
#endif
#endif
