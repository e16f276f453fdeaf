/*
 * amg.h - C ABI of the MI355X (gfx950) AMG-preconditioned PCG library
 * libbnamg.so, a drop-in for the C_amg1 library (libamg.so).
 *
 * Types keep the layout of C_amg1/include/agg_lib.h:29-44 and
 * C_amg1/include/amg.h:19-69 (INT = int32, REAL = double, 0-based CSR) so
 * existing callers compile unchanged.  Hierarchy set-up runs on the host
 * (C, bit-identical C/F splits, patterns and values to the reference);
 * every per-iteration operation of a solve runs as hand-written gfx950 HIP
 * kernels.  There is no CPU solve path: when no GPU is usable, solve entry
 * points return BN_ERR_NO_DEVICE (solver handle API) or abort with a
 * message (entry points whose reference signature has no error return).
 */
#ifndef BN_AMG_H
#define BN_AMG_H

#ifdef __cplusplus
extern "C" {
#endif

#ifndef INT
#define INT int
#endif
#ifndef REAL
#define REAL double
#endif

/* 0-based CSR (C_amg1/include/agg_lib.h:37-44) */
typedef struct {
  INT row;
  INT col;
  INT nnz;
  INT *ia;
  INT *ja;
  REAL *val;
} dcsr_mat;

#define AMG_MAX_LEVELS 20 /* C_amg1/include/amg.h:19 */

/* smoother ids (C_amg1/include/amg.h:22-25) */
#define SMOOTHER_L1 0
#define SMOOTHER_GS 1
#define SMOOTHER_CF_GS 2
#define SMOOTHER_JACOBI 3

/* coarsening ids (C_amg1/include/amg.h:28-30) */
#define COARSEN_RS 0
#define COARSEN_PMIS 1
#define COARSEN_HMIS 2

/* lower-triangular IC(0) factor (C_amg1/include/amg.h:33-39) */
typedef struct {
  INT n;
  INT *ia;
  INT *ja;
  REAL *val;
  INT nnz;
} ichol_factor;

/* one level (C_amg1/include/amg.h:42-53) */
typedef struct {
  dcsr_mat A;
  dcsr_mat P;
  dcsr_mat A_filtered;
  INT *mis;
  INT *isolated;
  INT *cf_order;
  INT n;
  INT n_coarse;
  ichol_factor L_ichol;
  REAL *l1_diag;
} amg_level;

/* hierarchy + options (C_amg1/include/amg.h:56-69) */
typedef struct {
  amg_level levels[AMG_MAX_LEVELS];
  INT num_levels;
  REAL threshold;
  INT min_size;
  INT max_levels;
  INT maxit_prol;
  INT lump;
  INT nu;
  INT smoother;
  REAL omega_jacobi;
  INT coarsening;
} amg_data;

/* ---------------------------------------------------------------------
 * Opaque-handle solver API (replaces C_amg1/src/amg_wrapper.c:51-268).
 * Same argument meaning, ownership and return conventions:
 *   create copies and symmetrises (A + A^T)/2, builds the hierarchy with
 *   PMIS coarsening and the CF-GS smoother, computes IC(0) of the fine
 *   matrix and uploads everything to the current HIP device once;
 *   solve reads b, uses x as the initial guess and overwrites it, returns
 *   the iteration count or -maxiter on non-convergence.
 * ------------------------------------------------------------------- */
/* C_amg1/src/amg_wrapper.c:51 */
void *amg_solver_create(int n, int nnz, const int *ia, const int *ja,
                        const double *val, double threshold, int min_size,
                        int max_levels, int nu);
/* C_amg1/src/amg_wrapper.c:114 - same pattern, new values */
int amg_solver_update(void *handle, int n, int nnz, const int *ia,
                      const int *ja, const double *val);
/* C_amg1/src/amg_wrapper.c:168 - replace the solve matrix, keep hierarchy */
int amg_solver_set_matrix(void *handle, int n, int nnz, const int *ia,
                          const int *ja, const double *val);
/* C_amg1/src/amg_wrapper.c:214 */
int amg_solver_solve(void *handle, const double *b, double *x, double tol,
                     int maxiter, int print_level);
/* C_amg1/src/amg_wrapper.c:241 */
void amg_solver_set_smoother(void *handle, int smoother, double omega);
/* C_amg1/src/amg_wrapper.c:249 */
int amg_solver_size(void *handle);
/* C_amg1/src/amg_wrapper.c:255 */
int amg_solver_num_levels(void *handle);
/* C_amg1/src/amg_wrapper.c:261 */
void amg_solver_free(void *handle);

/* ---------------------------------------------------------------------
 * Library-level set-up entry points (host), C_amg1/include/amg.h:76-108.
 * Outputs are bit-identical to the reference routines.
 * ------------------------------------------------------------------- */
void amg_build_hierarchy(amg_data *amg, const dcsr_mat *A);    /* amg.h:76 */
void amg_rebuild_values(amg_data *amg, const dcsr_mat *A_new); /* amg.h:77 */
void amg_free(amg_data *amg);                                  /* amg.h:78 */
void ichol_compute(const dcsr_mat *A, ichol_factor *L);        /* amg.h:89 */
void ichol_free(ichol_factor *L);                              /* amg.h:91 */
void rs_strength(const dcsr_mat *A, REAL theta, dcsr_mat *S,
                 INT *max_lam);                                /* amg.h:99 */
void rs_coarsening(const dcsr_mat *S, INT max_lam, INT *cf,
                   INT *n_coarse);                             /* amg.h:100 */
void pmis_coarsening(const dcsr_mat *S, INT *cf, INT *n_coarse); /* amg.h:101 */
void hmis_coarsening(const dcsr_mat *S, INT *cf, INT *n_coarse); /* amg.h:102 */
void rs_standard_interpolation_sparsity(const dcsr_mat *S, const INT *cf,
                                        dcsr_mat *P);          /* amg.h:103 */
void rs_standard_interpolation_values(const dcsr_mat *A, const dcsr_mat *S,
                                      const INT *cf, dcsr_mat *P); /* :105 */
void rs_standard_interpolation(const dcsr_mat *A, const dcsr_mat *S,
                               const INT *cf, dcsr_mat *P);    /* amg.h:107 */
/* Galerkin triple product, C_amg1/include/agg_lib.h:300 */
void dcsr_rap(dcsr_mat *r, dcsr_mat *a, dcsr_mat *p, dcsr_mat *rap);

/* ---------------------------------------------------------------------
 * Library-level solve entry points, C_amg1/include/amg.h:81-96,111-115.
 * They take the reference's host structures; the library uploads the
 * hierarchy to the device on first use (cached per amg_data, dropped by
 * amg_free / amg_rebuild_values) and runs the work on the GPU.
 * ------------------------------------------------------------------- */
void amg_backslash(const amg_data *amg, INT lev, const REAL *b, REAL *x,
                   INT nu);                                    /* amg.h:81 */
void amg_fwdslash(const amg_data *amg, INT lev, const REAL *b, REAL *x,
                  INT nu);                                     /* amg.h:83 */
void amg_vcycle_precond(const amg_data *amg, const REAL *g, REAL *x); /* :85 */
void ichol_solve(const ichol_factor *L, const REAL *b, REAL *x); /* amg.h:90 */
void amg_ichol_precond(const amg_data *amg, const dcsr_mat *A,
                       const ichol_factor *L, const REAL *g,
                       REAL *x);                               /* amg.h:94 */
/* amg.h:111.  SpMV and BLAS-1 run on the device; `precond` is called with
 * host vectors each iteration (it may itself be amg_ichol_precond). */
INT amg_pcg(const dcsr_mat *A, const REAL *b, REAL *x,
            void (*precond)(const void *data, const REAL *g, REAL *Bg),
            const void *precond_data, REAL tol, INT maxiter,
            INT print_level);

#ifdef __cplusplus
}
#endif
#endif /* BN_AMG_H */
