/*
 * bn_amg_ext.h - extensions of libbnamg.so beyond the C_amg1 interface.
 * None of these has a counterpart in the reference; they add what a GPU
 * library needs at the boundary: explicit options (coarsening, V-cycle
 * precision, coarsest solver), device-pointer solves, a solve report,
 * hierarchy introspection for parity checks, the synthetic network
 * generator, the adjoint gradient and batched solves.
 */
#ifndef BN_AMG_EXT_H
#define BN_AMG_EXT_H

#include <stddef.h>
#include <stdint.h>

#include "amg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* error codes (negative, never a valid iteration count) */
#define BN_ERR_NO_DEVICE (-1000001)
#define BN_ERR_HIP (-1000002)
#define BN_ERR_ARG (-1000003)

/* V-cycle storage/arithmetic precision */
#define BN_PREC_FP64 0
#define BN_PREC_FP32 1

/* coarsest-level solver */
#define BN_COARSE_AUTO 0   /* dense when n_c <= BN_DENSE_MAX, else IC-PCG */
#define BN_COARSE_DENSE 1  /* explicit inverse of the coarsest operator   */
#define BN_COARSE_ICPCG 2  /* reference: IC-PCG to 1e-12, <= n_c steps    */
#define BN_DENSE_MAX 8192

/* fine-level IC(0) correction inside the preconditioner */
#define BN_IC_EXACT 0   /* level-scheduled forward/backward solves          */
#define BN_IC_JACOBI 1  /* ic_sweeps Jacobi sweeps per triangle (flexible)   */
#define BN_IC_NONE 2    /* no fine IC(0) step: the PCG preconditioner is     */
                        /* the plain V-cycle (amg_vcycle_precond)           */

/* strength test of the set-up: j strong in row i (j != i) iff
 *   C   (rs_classical.c:38-58):   -a_ij        >= theta * m_i
 *   MEX (rs_strength_mex.c:58-60,79-81): max(0, -a_ij) >= theta * m_i
 * with m_i = max(0, max_{j != i} -a_ij).  They differ only on positive
 * off-diagonals of rows whose off-diagonals are all >= 0 (m_i = 0): the
 * MEX rule makes every off-diagonal of such a row strong. */
#define BN_STRENGTH_C 0
#define BN_STRENGTH_MEX 1

/* amg_build_hierarchy (C_amg1/src/build_hierarchy.c:29-156) with a choice
 * of strength rule; amg_build_hierarchy is strength = BN_STRENGTH_C. */
void bn_build_hierarchy(amg_data *amg, const dcsr_mat *A, int strength);

/* outer stopping test */
#define BN_STOP_BNORM 0  /* sqrt(r'z / r0'z0) < tol (reference, amg_pcg.c:75) */
#define BN_STOP_RELRES 1 /* ||r||_2 / ||b||_2 < tol                          */

typedef struct {
  int coarsening;     /* COARSEN_RS / COARSEN_PMIS / COARSEN_HMIS          */
  double threshold;   /* strength threshold theta                         */
  int min_size;       /* coarsen while n >= min_size                      */
  int max_levels;     /* <= AMG_MAX_LEVELS                                */
  int nu;             /* smoothing sweeps per half cycle                  */
  int smoother;       /* SMOOTHER_*                                       */
  double omega;       /* damped-Jacobi weight                             */
  int precision;      /* BN_PREC_*                                        */
  int coarse_solver;  /* BN_COARSE_*                                      */
  int fine_ic;        /* BN_IC_*                                          */
  int ic_sweeps;      /* sweeps per triangle for BN_IC_JACOBI             */
  int max_colours;    /* colours above which a level smooths l1-Jacobi    */
  int flexible;       /* 1: flexible (Polak-Ribiere) beta                 */
  int stop;           /* BN_STOP_*                                        */
  int symmetrize;     /* 1: use (A + A^T)/2 as the reference wrapper does */
  int device;         /* HIP device ordinal, -1 = current                 */
  int verbose;        /* 1: hierarchy report on stderr                    */
  int graph;          /* 1: replay each PCG step as one captured hipGraph */
  int block_smoother; /* 1: every level smooths with chunked l1 block     */
                      /*    Gauss-Seidel (one launch per sweep) instead   */
                      /*    of multicolour GS / l1-Jacobi                 */
  int chunk_rows;     /* largest chunk of the block smoother (<= 8192;    */
                      /*    0: 8192)                                      */
  int strength;       /* BN_STRENGTH_*                                    */
  int gs_exact;       /* SMOOTHER_CF_GS: 0 runs the C-then-F multicolour  */
                      /*    sweep, 1 the reference's exact cf_order sweep */
                      /*    (level-scheduled over its dependency DAG).    */
                      /*    SMOOTHER_GS is always the exact natural-order */
                      /*    sweep (C_amg1/src/amg_cycle.c:21-80)          */
} bn_options;

typedef struct {
  int iterations;     /* iteration count, or -maxiter                      */
  double rel_bnorm;   /* final sqrt(r'z / r0'z0)                           */
  double relres;      /* final ||r|| / ||b|| from the recurrence           */
  double solve_ms;    /* device time from first to last kernel (hipEvent)  */
  double wall_ms;     /* host wall time of the call                        */
  double setup_ms;    /* host hierarchy + IC set-up                        */
  double upload_ms;   /* host analysis + device upload                     */
  int num_levels;
} bn_solve_info;

void bn_options_default(bn_options *o);

/* Create with explicit options (amg_solver_create = these defaults with
 * PMIS, CF-GS, FP64, symmetrise).  Returns NULL on bad arguments. */
void *bn_solver_create(int n, int nnz, const int *ia, const int *ja, const double *val,
                       const bn_options *opt);
/* Solve with device pointers (level-0 numbering = caller numbering);
 * runs on the solver's stream, synchronous on return. */
int bn_solver_solve_device(void *h, const double *d_b, double *d_x, double tol, int maxiter,
                           int print_level, bn_solve_info *info);
/* Solve with host pointers, also filling `info`. */
int bn_solver_solve_info(void *h, const double *b, double *x, double tol, int maxiter,
                         int print_level, bn_solve_info *info);
/* One preconditioner application z = B g (host pointers); which = 0
 * AMG+IC (amg_ichol_precond), 1 plain V-cycle (amg_vcycle_precond). */
int bn_solver_precond(void *h, const double *g, double *z, int which);
/* Device timing of the last solve etc. */
int bn_solver_last_info(void *h, bn_solve_info *info);
/* Change the V-cycle precision / IC mode / stop rule after creation. */
int bn_solver_set_mode(void *h, int precision, int fine_ic, int ic_sweeps, int flexible, int stop);

/* Hierarchy introspection (host copies, reference layout).
 * dims[0..7] = n, nnz(A), n_coarse, nnz(P), nnz(S), nnz(coarsest L),
 *              nnz(fine L), has_split;  what: 0 A, 1 P, 2 S */
int bn_solver_level_dims(void *h, int lev, int *dims);
int bn_solver_level_csr(void *h, int lev, int what, int *ia, int *ja, double *val);
int bn_solver_level_split(void *h, int lev, int *mis, int *cf_order, double *l1);
int bn_solver_ichol(void *h, int which, int *ia, int *ja, double *val);
/* device schedule: colours / levels used on each level (for reports) */
int bn_solver_schedule(void *h, int lev, int *ncolours, int *use_gs, int *lanes);
int bn_solver_ic_levels(void *h, int *fwd, int *bwd);

/* Time one kernel of the solve on the solver's stream with HIP events
 * (average over `reps` launches after one warm-up launch), and give its
 * algorithmic (compulsory) HBM bytes per launch.  which:
 *   0  fine-level CSR SpMV in FP64 (the outer PCG's q = A p)
 *   1  one full fine-level multicolour Gauss-Seidel sweep, V-cycle precision
 *   2  fine-level residual r = b - A x, V-cycle precision
 *   3  fine-level IC(0) forward + backward application (current IC mode)
 *   4  one preconditioner application (AMG + IC)
 *   5  plain FP64 copy of n values (16 n bytes): calibrates byte counters
 *   6  one Jacobi sweep over the fine IC(0) lower triangle, V-cycle
 *      precision (the kernel the Jacobi-swept IC repeats)
 * Returns 0 or a BN_ERR_* code. */
int bn_solver_time_kernel(void *h, int which, int reps, double *ms_avg, double *bytes);

/* Kernel census of one solve: the solve runs once (b, x as for
 * bn_solver_solve_device; no step graph) with a HIP event pair on the
 * solver's stream around every launch of its bandwidth kernels (sparse
 * products, residuals and transfers, Gauss-Seidel colours, IC(0) Jacobi
 * sweeps, the outer SpMV); afterwards every distinct launch shape is
 * launched `reps` times more, each after a 320 MB read (cold).  Per kernel
 * symbol (spelled as rocprofv3 prints it): launches, distinct shapes,
 * algorithmic bytes (each array touched once), device microseconds as the
 * solve ran, and as cold launches would take.  Up to max_out symbols;
 * *n_out = how many.  No counterpart in the reference (measurement). */
typedef struct {
  char name[80];
  int launches;
  int shapes;
  double bytes;
  double us_solve;
  double us_cold;
} bn_kernel_stat;
int bn_solver_kernel_census(void *h, const double *d_b, double *d_x, double tol, int maxiter, int reps,
                            bn_kernel_stat *out, int max_out, int *n_out);

/* Device memory helpers for callers that hold no HIP runtime of their
 * own (e.g. ctypes users): plain hipMalloc/hipMemcpy/hipFree on the
 * current device.  Return 0 or BN_ERR_HIP. */
int bn_dev_alloc(void **ptr, size_t bytes);
int bn_dev_free(void *ptr);
int bn_dev_copy(void *dst, const void *src, size_t bytes, int kind); /* 1 H2D, 2 D2H, 3 D2D */
int bn_dev_sync(void);

/* 1 if a HIP device is usable */
int bn_device_available(void);
/* message of the last error on this thread */
const char *bn_last_error(void);

/* ---- adjoint: lambda = A^{-1} dLdp with the same hierarchy, then the
 * edge-scatter gradient dL/dg_e = -(lam_u - lam_v)(p_u - p_v) and
 * dL/dlog r_e = 4 g_e dL/dg_e.  Host pointers; m segments with endpoints
 * eu/ev (unknown index or -1-d for Dirichlet node d, pressure pdir[d]). */
int bn_solver_adjoint_grad(void *h, const double *p, const double *dLdp, int m, const int *eu,
                           const int *ev, const double *g, const double *pdir, double tol,
                           int maxiter, double *lam, double *dLdg, double *dLdlogr);

/* The same with every array in device memory (no allocation, no host
 * copies).  d_lam is the initial guess on entry (zeros: cold start) and
 * the multipliers on return; nd = number of Dirichlet pressures in d_pdir.
 * Extension: C_amg1 has no gradient; it serves descent loops whose
 * pressures stay in HBM. */
int bn_solver_adjoint_grad_device(void *h, const double *d_p, const double *d_dLdp, int m,
                                  const int *d_eu, const int *d_ev, const double *d_g,
                                  const double *d_pdir, int nd, double tol, int maxiter,
                                  double *d_lam, double *d_dLdg, double *d_dLdlogr);

/* bn_solver_adjoint_grad (same arguments, host arrays, same results bit
 * for bit) without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (tight
 * descent or serving loops, code that forbids hipMalloc).  The solve runs
 * on the solver's own vectors (a step graph already captured for the
 * mode is replayed, none is captured) and the segment arrays pass through
 * a static device scratch of BN_SCRATCH_BYTES per GPU (about 32.5 MiB,
 * held until the process ends) in chunks, so m is unbounded; at most
 * BN_SCRATCH_DIRICHLET (65,536) Dirichlet pressures, else BN_ERR_ARG.
 * The solver must already be on the device (creation uploads it), and an
 * exact-order Gauss-Seidel smoother needs one ordinary solve first to
 * build its schedules.  Calls that overlap on one GPU share the scratch
 * and race: serialise them. */
int bn_solver_adjoint_grad_noalloc(void *h, const double *p, const double *dLdp, int m, const int *eu,
                                   const int *ev, const double *g, const double *pdir, double tol,
                                   int maxiter, double *lam, double *dLdg, double *dLdlogr);

/* ---- new solve-matrix values from device memory: nnz values on the
 * stored pattern, in its entry order (the creation order when the input
 * pattern was symmetric without duplicates; nnz must match).  The
 * hierarchy and the fine IC(0) are kept: the
 * two-matrix preconditioning of amg_solver_set_matrix
 * (C_amg1/src/amg_wrapper.c:168-199) without the host IC(0) and the
 * upload.  For finite-difference probes and descent steps between
 * hierarchy rebuilds; amg_solver_update rebuilds everything. */
int bn_solver_set_values_device(void *h, int nnz, const double *d_val);

/* ---- value-only rebuild of the hierarchy on the device: new solve-matrix
 * values as bn_solver_set_values_device takes them, then the interpolation
 * weights and Galerkin operators of every level recomputed on the stored
 * C/F split and patterns, bit-identical to amg_solver_update's host
 * rebuild (C_amg1/src/build_hierarchy.c:170-242, amg_wrapper.c:114-149);
 * the fine IC(0) and the coarsest solve are refreshed from the new values
 * on the host, and the host hierarchy follows.  Not for the block
 * smoother layout or a solver holding a separate solve matrix.  Extension:
 * the reference rebuilds on the host only. */
int bn_solver_rebuild_values_device(void *h, int nnz, const double *d_val);

/* ---- coarsest dense inverse: the solver's copy (row-major, n x n) into
 * out when out != NULL, its order in *n; BN_ERR_ARG if the coarsest
 * solve is not the dense inverse.  bn_dense_inverse is the host form of
 * the same inverse (A SPD, n = A->row; returns 1 if a pivot was clamped);
 * the solver forms it on the device entry for entry.  Extension: the
 * reference solves its coarsest level by an inner IC-PCG. */
int bn_solver_coarse_inverse(void *h, int *n, double *out);
int bn_dense_inverse(const dcsr_mat *A, double *inv);

/* ---- batched solves: k right-hand sides stored column-major (b[j*n+i]),
 * each solved to tol; iterations per column into its[k]. */
int bn_solver_solve_batch(void *h, int k, const double *b, double *x, double tol, int maxiter,
                          int *its, bn_solve_info *info);

/* Block PCG: the k <= 32 right-hand sides (b, k x n, one after another;
 * x the initial guesses, same layout, overwritten) are solved together.
 * Every sparse operation streams its matrix once for all columns (CSR
 * SpMM over columns interleaved by row); each column keeps its own PCG
 * scalars and stops on its own test (flexible beta, the solver's stop
 * mode and tol).  Needs the dense coarsest solve and a Jacobi-swept or
 * absent fine IC(0) (BN_IC_JACOBI / BN_IC_NONE); precision follows the
 * solver's mode.  its[j] = column j's iterations (-maxiter if it did not
 * converge); returns the largest count, negated if any column failed.
 * No counterpart in the reference (north-star batching form). */
int bn_solver_solve_block(void *h, int k, const double *b, double *x, double tol, int maxiter,
                          int *its, bn_solve_info *info);
/* Many independent systems in one solve: the solver was built on their
 * block-diagonal union, rows of system g are seg_ptr[g] .. seg_ptr[g+1]-1
 * (seg_ptr: nsys+1 host ints, 0 .. n).  One V-cycle serves all systems;
 * each keeps its own PCG scalars (flexible beta) and stop test.  b, x:
 * n-vectors (x the initial guess, overwritten); its[g] = system g's
 * iterations (-maxiter if it did not converge); returns the largest count,
 * negated if any system failed.  No counterpart in the reference
 * (north-star batching of uneven networks). */
int bn_solver_solve_segments(void *h, int nsys, const int *seg_ptr, const double *b, double *x,
                             double tol, int maxiter, int *its, bn_solve_info *info);
int bn_solver_solve_segments_device(void *h, int nsys, const int *seg_ptr, const double *d_b,
                                    double *d_x, double tol, int maxiter, int *its,
                                    bn_solve_info *info);
/* Same with device pointers. */
int bn_solver_solve_block_device(void *h, int k, const double *d_b, double *d_x, double tol,
                                 int maxiter, int *its, bn_solve_info *info);

/* ---- C-typed twin of the MEX smoother
 *   x = cf_gauss_seidel_mex(A, b, x0, cf_order, nu, direction)
 * (Matlab_amg/cf_gauss_seidel_mex.c:7,27-105), same argument list with C
 * types: A is the n x n CSR matrix (ia, ja, val; A symmetric, so MATLAB's
 * CSC arrays jc/ir/pr describe the same matrix), cf_order holds 1-based
 * row numbers as doubles (MATLAB's convention), direction 0 sweeps
 * cf_order(1) .. cf_order(n), 1 the reverse.  nu Gauss-Seidel sweeps start
 * from x0 and end in x (x may be x0): x_i = (b_i - sum_{j != i} a_ij x_j)
 * / a_ii with the row's first diagonal entry (1.0 if none).  The sweeps
 * run on the GPU level by level over the order's dependency DAG and read
 * exactly the values the sequential sweep reads.  Returns 0, BN_ERR_ARG
 * (bad sizes, direction, or a cf_order that is not a permutation of
 * 1..n; bn_last_error() then starts with the MEX error id it corresponds
 * to, "cf_gs:args" or "cf_gs:order"), BN_ERR_NO_DEVICE or BN_ERR_HIP. */
int cf_gauss_seidel(int n, const int *ia, const int *ja, const double *val, const double *b, const double *x0,
                    const double *cf_order, int nu, int direction, double *x);

/* The same sweeps on device arrays: the matrix (nnz entries), b and x
 * (the iterate, updated in place) and order (0-based row numbers, int)
 * are all device pointers.  The schedule is host analysis, so the pattern,
 * the values and the order are read back once per call. */
int bn_cf_gauss_seidel_device(int n, int nnz, const int *d_ia, const int *d_ja, const double *d_val,
                              const double *d_b, double *d_x, const int *d_order, int nu, int direction);

/* ---- options of the solvers behind the library-level entry points on a
 * caller's amg_data (amg_backslash, amg_fwdslash, amg_vcycle_precond,
 * amg_ichol_precond): coarsest solver, exact-order sweeps (gs_exact), ...
 * NULL restores bn_options_default.  The smoother, nu, omega and
 * coarsening always come from the amg_data itself.  Device mirrors built
 * under the previous options are dropped.  The reference has no options
 * beyond the amg_data fields. */
void bn_set_library_options(const bn_options *o);

#ifdef __cplusplus
}
#endif
#endif
