/*
 * pascal_tcs.h -- C-ABI of the MI355X-native PaScaL_TCS MPM-STD time step.
 *
 * One context owns the whole (single-rank) domain on one GPU: persistent
 * device copies of the module arrays of mpi_thermal / mpi_momentum /
 * mpi_pressure, the 1-D mesh metrics, scratch for the ADI and Poisson
 * solves, and one HIP stream.  Every entry point only enqueues work on that
 * stream and returns; the host waits only where a scalar comes back
 * (tcs_post_div, tcs_post_cfl, tcs_nusselt) or on an explicit download /
 * tcs_sync.  Plain pointers and sizes only; no torch or Fortran types.
 *
 * Error convention: every entry returns 0 on success, a negative TCS_E_*
 * code otherwise; tcs_last_error() returns a message for the last failure.
 * Entries are not re-entrant for one context (the reference is single-
 * threaded per rank; so is this).
 *
 * Array layouts are those of the reference's Fortran arrays:
 *   field  (0:n1sub,0:n2sub,0:n3sub) FP64, i fastest, n?sub = n?m+1
 *   wall   (0:n1sub,0:n3sub)          FP64, i fastest
 *   metric (0:n?sub)                  FP64
 *   TDMA   (1:n_sys,1:n_row)          FP64, system index fastest
 */
#ifndef PASCAL_TCS_H
#define PASCAL_TCS_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TCS_OK 0
#define TCS_E_ARG (-1)     /* bad argument / unsupported shape */
#define TCS_E_HIP (-2)     /* HIP runtime error */
#define TCS_E_STATE (-3)   /* call out of order */
#define TCS_E_NOMEM (-4)   /* device allocation failed */

/* Physical constants and switches of module global
   (src/module_global.f90:30-70, values set in :107-165). */
typedef struct tcs_params {
    int32_t n1m, n2m, n3m;        /* interior cells (namelist meshes) */
    int32_t pbc1, pbc2, pbc3;     /* periodicity; only (1,0,1) is supported */
    int32_t precision;            /* 0: FP64 in the reference's operation order;
                                     1: FP64 with contraction, reciprocals,
                                        Horner-form property polynomials and
                                        the mean of dP from the spectrum;
                                     2: mixed, FP32 stage rows, line solves and
                                        increments, FP64 state and pressure */
    int32_t tdma_split;           /* 0: every ADI line solved as on one rank;
                                     1: lines of up to 512 rows split into
                                     parts as PaScaL_TDMA splits them over
                                     ranks (see ptdma_many_solve_part_dev) */
    double Ct, Cmu, Cmt, Cmp;     /* :139-141 */
    double DeltaT;                /* namelist sim_parameter */
    double a10, a11, a12, a13, a14, a15;  /* density polynomial   :117 */
    double b10, b11, b12, b13, b14, b15;  /* heat capacity        :118 */
    double c10, c11, c12, c13, c14, c15;  /* conductivity         :119 */
    double d10, d11, d12, d13, d14, d15;  /* viscosity            :120 */
    double KPP0, Rho0, Cp0, Mu0, a12pera11;  /* :122-137 */
    double L1, L2, L3;            /* domain lengths :149-151 */
    double MaxCFL, dtMax;         /* :161-165 and namelist */
} tcs_params;

typedef struct tcs_ctx tcs_ctx;

/* Field ids: module variables of the reference solver modules.
   T: src/module_solve_thermal.f90:29; U,V,W,dU,dV,dW,P,invRho,Mu:
   src/module_solve_momentum.f90:31-44; dPhat,dP0,dP: module_solve_pressure.f90:32.
   invRho and Mu are stored as (0:n2sub,0:n3sub) tables because the
   reference broadcasts the i=n1sub value along every x-row
   (src/module_solve_momentum.f90:156-171); upload/download expand them. */
enum tcs_field {
    TCS_T = 0, TCS_U, TCS_V, TCS_W, TCS_P,
    TCS_DU, TCS_DV, TCS_DW, TCS_DP, TCS_DPHAT, TCS_DP0,
    TCS_INVRHO, TCS_MU,
    TCS_KPP, TCS_DKPP, TCS_INVRHOCP, TCS_DINVRHOCP,
    TCS_NFIELD
};
/* Wall tables {T,U,V,W}BC{up,bt}_sub: module_solve_thermal.f90:32,
   module_solve_momentum.f90:35-36. */
enum tcs_wall {
    TCS_TBCUP = 0, TCS_TBCBT, TCS_UBCUP, TCS_UBCBT, TCS_VBCUP, TCS_VBCBT,
    TCS_WBCUP, TCS_WBCBT, TCS_NWALL
};

/* ---- lifecycle --------------------------------------------------------- */
const char *tcs_last_error(void);
int tcs_device_count(int *count);

/* One axis of mpi_subdomain_mesh on one rank
   (src/module_mpi_subdomain.f90:153-299): x, dx, dmx each hold nm+2 values
   (0:n?sub).  uniform=1 uniform spacing, else tanh stretching by gamma. */
int tcs_mesh_axis(int32_t nm, double L, int32_t uniform, double gamma,
                  int32_t periodic, double *x, double *dx, double *dmx);

/* Replaces mpi_subdomain_make/_mesh/_indices + the *_allocation routines
   (src/module_mpi_subdomain.f90:98-374; module_solve_*.f90 allocation):
   copies the metrics (0:n?sub) to the device and allocates every module
   array once (all zero) plus solver scratch.  n1m and n3m must be powers of
   two (Poisson FFTs), n2m >= 3. */
int tcs_init(const tcs_params *prm,
             const double *dx1, const double *dx2, const double *dx3,
             const double *dmx1, const double *dmx2, const double *dmx3,
             tcs_ctx **out);
int tcs_finalize(tcs_ctx *ctx);
int tcs_sync(tcs_ctx *ctx);
/* the context's HIP stream (a hipStream_t) for callers that interoperate */
void *tcs_stream(tcs_ctx *ctx);

/* ---- module arrays ----------------------------------------------------- */
int tcs_field_upload(tcs_ctx *ctx, int32_t id, const double *host);
int tcs_field_download(tcs_ctx *ctx, int32_t id, double *host);
/* device address of a field; INVRHO/MU return the 2-D (j,k) table.  On the
   device a full field's rows are tcs_field_pitch elements apart (element
   (i,j,k) at dptr[i + pitch*(j + N2*k)], N2 = n2m+2; the pitch is N1 = n1m+2
   rounded up to a multiple of 16, so every row's interior starts on a
   128-byte line); the upload/download entries take the reference's
   unpadded layout.  Valid until the next tcs_step, which writes the updates
   of T, U, V, W, dPhat and dP0 into second buffers and exchanges them.
   tcs_post_div keeps the CFL maximum of the same pass for tcs_post_cfl
   until an entry point that may change U, V or W runs; a caller that
   writes a field through this pointer calls tcs_fields_touched before the
   next tcs_post_cfl.  In the mixed form (precision 2) the increments dU, dV,
   dW live in FP32: their FP64 arrays are filled from the FP32 values when
   the pointer is handed out (and by downloads), and writes through it reach
   the form after tcs_fields_touched. */
int tcs_field_device_ptr(tcs_ctx *ctx, int32_t id, double **dptr);
/* Declares that fields were written outside the entry points (through
   tcs_field_device_ptr): drops the CFL maximum kept by tcs_post_div and, in
   the mixed form, refreshes its FP32 copies and increments from the FP64
   arrays.  No counterpart in the reference (its fields are host arrays). */
int tcs_fields_touched(tcs_ctx *ctx);
/* Copies a field to (to_field = 0) or from (to_field = 1) a device buffer
   in the device layout (tcs_field_pitch * N2 * N3 doubles; INVRHO and MU:
   their N2 * N3 (j,k) tables), on the context's stream, without a host
   round trip.  No counterpart in the reference: device-resident states
   (checkpoints of an adjoint rollout). */
int tcs_field_copy_device(tcs_ctx *ctx, int32_t id, int32_t to_field, double *dev);
/* row pitch (elements) of the full fields on the device (no counterpart in
   the reference: the device layout of tcs_field_device_ptr) */
int tcs_field_pitch(tcs_ctx *ctx, int32_t *pitch);
int tcs_wall_upload(tcs_ctx *ctx, int32_t id, const double *host);
int tcs_wall_download(tcs_ctx *ctx, int32_t id, double *host);

/* ---- module routines (one entry per reference subroutine) --------------- */
/* mpi_thermal_boundary, src/module_solve_thermal.f90:88-111: TBC tables :=
   Tup/Tbt and the wall row of T.  On a single y-rank the reference's
   if/else-if writes only the bottom row (j=0). */
int tcs_thermal_boundary(tcs_ctx *ctx, double Tup, double Tbt);
/* mpi_momentum_boundary, src/module_solve_momentum.f90:105-140 (same rule) */
int tcs_momentum_boundary(tcs_ctx *ctx, double Uup, double Ubt, double Vup,
                          double Vbt, double Wup, double Wbt);
/* mpi_thermal_coeffi, src/module_solve_thermal.f90:116-156: materialises
   KPP, dKPP, invRhoCp, dinvRhoCp.  The thermal solver evaluates the same
   polynomials from T inline, so the time step does not need this call. */
int tcs_thermal_coeffi(tcs_ctx *ctx);
/* mpi_thermal_solver(U,V,W), src/module_solve_thermal.f90:175-532 */
int tcs_thermal_solver(tcs_ctx *ctx, double dt);
/* mpi_momentum_coeffi(T), src/module_solve_momentum.f90:145-173 */
int tcs_momentum_coeffi(tcs_ctx *ctx);
/* mpi_momentum_solvedU/V/W(T), src/module_solve_momentum.f90:225-563,
   :569-924, :930-1272 */
int tcs_momentum_solvedU(tcs_ctx *ctx, double dt);
int tcs_momentum_solvedV(tcs_ctx *ctx, double dt);
int tcs_momentum_solvedW(tcs_ctx *ctx, double dt);
/* mpi_momentum_blockLdV/blockLdU(T), :1278-1356, :1362-1460 */
int tcs_momentum_blockLdV(tcs_ctx *ctx, double dt);
int tcs_momentum_blockLdU(tcs_ctx *ctx, double dt);
/* mpi_momentum_pseudoupdateUVW, :178-208 */
int tcs_momentum_pseudoupdateUVW(tcs_ctx *ctx);
/* mpi_pressure_wave_number, src/module_solve_pressure.f90:93-129 (tcs_init
   already calls it; exposed for the drop-in module shell) */
int tcs_pressure_wave_number(tcs_ctx *ctx);
/* mpi_pressure_RHS(invRho,U,V,W,VBCup,VBCbt), :140-231; the RHS stays on
   the device for tcs_pressure_poisson_fft2 */
int tcs_pressure_rhs(tcs_ctx *ctx, double dt);
/* mpi_pressure_Poisson_FFT2(dx2,dmx2), :489-744: FFT x (r2c), FFT z, TDMA
   in y per wavenumber pair, inverse FFTs, scaling, mean removal, wall
   extrapolation of dP */
int tcs_pressure_poisson_fft2(tcs_ctx *ctx);
/* mpi_pressure_Projection(invRho,U,V,W,P,Jrank,Jnprocs), :756-874 */
int tcs_pressure_projection(tcs_ctx *ctx, double dt);
/* mpi_subdomain_ghostcell_update(X), src/module_mpi_subdomain.f90:758-791:
   x then y (walls: no-op) then z periodic plane copies on one rank */
int tcs_ghostcell_update(tcs_ctx *ctx, int32_t id);
/* mpi_Post_Div(U,V,W,maxDivU), src/module_post.f90:100-158 (signed max) */
int tcs_post_div(tcs_ctx *ctx, double *maxdiv);
/* mpi_Post_CFL(U,V,W,newCFL,newdt), src/module_post.f90:168-230.  Reuses
   the maximum of the preceding tcs_post_div pass when no entry point that
   may change U, V, W ran in between (see tcs_fields_touched). */
int tcs_post_cfl(tcs_ctx *ctx, double *newcfl, double *newdt);
/* The loop body of src/main.f90:131-203 (thermal, momentum, pressure,
   every halo update in main's order), fully on the device.  The same
   results as the routines called one by one; the last thermal sweep, the
   block corrections (with the pseudo-update) and the projection (with the
   dPhat/dP0 history) write their updates into second buffers of the fields,
   so the field device addresses change (tcs_field_device_ptr). */
int tcs_step(tcs_ctx *ctx, double dt);
/* Wall Nusselt numbers (no counterpart in the reference): mean over (i,k)
   of (TBCbt(i,k) - T(i,1,k))/dmx2(1) and (T(i,n2m,k) - TBCup(i,k))/dmx2(n2sub),
   times L2/(Tbt - Tup).  The wall arrays are the wall temperatures the
   thermal solver applies; on one y-rank the top ghost row of T is not set
   (src/module_solve_thermal.f90:102-110) and is not used here. */
int tcs_nusselt(tcs_ctx *ctx, double Tup, double Tbt, double *nu_bottom, double *nu_top);

/* ---- adjoint of the step (no counterpart in the reference) -------------- */
/* The discrete adjoint of tcs_step in the plain FP64 form (precision 0),
   for gradients of scalar functions of a rollout.  The context keeps
   adjoint fields for the step's state (T, U, V, W, P, dPhat, dP0; full
   (0:n1sub,0:n2sub,0:n3sub) layout) and for the wall arrays.
   tcs_adjoint_zero clears them.  tcs_adjoint_step(ctx, dt) takes the
   context's fields as the state before a step and the adjoint fields as
   the gradient with respect to the state after it; it runs the step (the
   fields end as after tcs_step) and leaves the gradient with respect to the
   state before it, adding the wall arrays' share to their adjoints.
   tcs_adjoint_copy moves an adjoint array from (to_adjoint = 1) or to
   (to_adjoint = 0) a device buffer of its size: kind 0 with a state field
   id, kind 1 with a wall id.  tcs_adjoint_add_nusselt adds weight times
   the gradient of tcs_nusselt's bottom value with respect to T and TBCbt.
   Returns TCS_E_ARG for precision != 0, and from tcs_adjoint_step for
   grids whose full fields hold 2^28 values or more (about 640^3). */
int tcs_adjoint_zero(tcs_ctx *ctx);
int tcs_adjoint_step(tcs_ctx *ctx, double dt);
int tcs_adjoint_copy(tcs_ctx *ctx, int32_t kind, int32_t id, int32_t to_adjoint, double *dev);
int tcs_adjoint_add_nusselt(tcs_ctx *ctx, double weight, double Tup, double Tbt);

/* ---- timers ----------------------------------------------------------- */
/* HIP-event timing per stage and per kernel on the context stream, the
   device counterpart of the reference's labelled timers
   (PaScaL_TDMA/src/timer.f90:14-118; labels src/main.f90:55-72).
   tcs_profile(ctx,1) clears and enables, (ctx,0) disables.  Entries are
   reported in first-seen order: stage labels ("thermal solve", ...) and
   kernel labels ("ptdma", "thermal_z", ...). */
int tcs_profile(tcs_ctx *ctx, int32_t enable);
int tcs_profile_count(tcs_ctx *ctx, int32_t *n);
int tcs_profile_entry(tcs_ctx *ctx, int32_t idx, char *name, int32_t name_len,
                      double *total_ms, int64_t *count);

/* ---- PaScaL_TDMA ----------------------------------------------------- */
/* PaScaL_TDMA_many_solve(plan,A,B,C,D,n_sys,n_row),
   PaScaL_TDMA/src/pascal_tdma.f90:487-578, and PaScaL_TDMA_many_solve_cycle,
   :715-806, on one rank (plan = nprocs 1; reduced 2-row solve by tdma_many /
   tdma_cycl_many, PaScaL_TDMA/src/tdmas.f90:113-212).  Device pointers to
   (n_sys,n_row) arrays; in place: D returns the solution, A and C return
   the reference's modified coefficients, B is read only.  Same operation
   order as the reference, so results are bit-identical to it.  stream may
   be NULL (the null stream).  n_row >= 2. */
int ptdma_many_solve_dev(double *A, const double *B, double *C, double *D,
                         int32_t n_sys, int32_t n_row, void *stream);
int ptdma_many_solve_cycle_dev(double *A, const double *B, double *C, double *D,
                               int32_t n_sys, int32_t n_row, void *stream);
/* Same contract with host arrays (copied in and out; synchronous). */
int ptdma_many_solve_host(double *A, double *B, double *C, double *D,
                          int32_t n_sys, int32_t n_row, int32_t cyclic);
/* No counterpart in the reference: ptdma_many_solve_host without any
   device allocation, for callers that cannot allocate (code under graph
   capture, harnesses that forbid hipMalloc, or tight loops such as a
   plane-by-plane solve, which the allocating entry would hipMalloc/hipFree
   on every call).  The arrays are staged in a static device array of four
   times 2^21 doubles (64 MiB, held by each GPU the library runs on until
   the process ends): n_sys*n_row must be <= 2097152, else TCS_E_ARG.
   Calls that overlap on one GPU (several host threads) race on that array.
   Same kernels and results as ptdma_many_solve_host; synchronous. */
int ptdma_many_solve_host_noalloc(double *A, double *B, double *C, double *D,
                                  int32_t n_sys, int32_t n_row, int32_t cyclic);
/* PaScaL_TDMA_single_solve / _single_solve_cycle(plan,A,B,C,D,n_row),
   PaScaL_TDMA/src/pascal_tdma.f90:301-382 and :393-475, on one rank: the
   per-rank modified Thomas of the many-system solve and the 2-row reduced
   solve by tdma_single / tdma_cycl_single (tdmas.f90:1-104), whose
   operations equal tdma_many / tdma_cycl_many on one system -- so this is
   the many-system solve with n_sys = 1, bit for bit.  Arrays (1:n_row), in
   place as in the reference; device (stream may be NULL) or host. */
int ptdma_single_solve_dev(double *A, const double *B, double *C, double *D,
                           int32_t n_row, int32_t cyclic, void *stream);
int ptdma_single_solve_host(double *A, double *B, double *C, double *D,
                            int32_t n_row, int32_t cyclic);
/* Axis solve on a (n1,n2,n3) array (i fastest) in place, systems along
   `axis` (0=x,1=y,2=z) -- the layout change the reference makes with host
   transposes before each PaScaL call (e.g. src/module_solve_thermal.f90:
   371-379) done by strided access instead.  Same arithmetic per system as
   ptdma_many_solve[_cycle]_dev. */
int ptdma_axis_solve_dev(double *A, const double *B, double *C, double *D,
                         int32_t n1, int32_t n2, int32_t n3, int32_t axis,
                         int32_t cyclic, void *stream);
/* PaScaL_TDMA_many_solve[_cycle] with every system's rows split over
   P ranks, P the power of two >= n_row/16 (rows per rank by para_range,
   PaScaL_TDMA/src/para_range.f90:26-40): per-rank modified Thomas
   (pascal_tdma.f90:501-536), the reduced system of 2P rows in rank order
   (pascal_tdma.f90:243-266), per-rank update (:567-576).  Each rank is one
   GPU thread.  The reduced system is solved by tdma_many / tdma_cycl_many
   when P = 1 (n_row <= 16: identical to the single-rank solve) and by
   parallel cyclic reduction across the P threads when P >= 2 (within
   rounding of the reference's serial reduced solve).  2 <= n_row <= 512.
   D returns the solution; A and C are read only here (the reference leaves
   its per-rank modified coefficients in them). */
int ptdma_many_solve_part_dev(const double *A, const double *B, const double *C, double *D,
                              int32_t n_sys, int32_t n_row, int32_t cyclic, void *stream);
/* Axis solve of ptdma_axis_solve_dev with the split lines above. */
int ptdma_axis_solve_part_dev(const double *A, const double *B, const double *C, double *D,
                              int32_t n1, int32_t n2, int32_t n3, int32_t axis,
                              int32_t cyclic, void *stream);

#ifdef __cplusplus
}
#endif
#endif
