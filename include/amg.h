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

/* the reference header brings these in for its callers
 * (C_amg1/include/agg_lib.h:20-24, amg.h:16-17) */
#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

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

/* 0-based COO (C_amg1/include/agg_lib.h:47-54) */
typedef struct {
  INT row;
  INT col;
  INT nnz;
  INT *icols;
  INT *irows;
  REAL *val;
} dcoo_mat;

/* ---------------------------------------------------------------------
 * Header-only CSR/COO helpers of C_amg1/include/agg_lib.h:57-242, same
 * semantics (callers such as test drivers use them unchanged).
 * ------------------------------------------------------------------- */
static inline void dcsr_mat_free(dcsr_mat *a) {
  if (!a) return;
  free(a->ia);
  free(a->ja);
  free(a->val);
  a->ia = a->ja = NULL;
  a->val = NULL;
  a->row = a->col = a->nnz = 0;
}

static inline void dcoo_mat_free(dcoo_mat *a) {
  if (!a) return;
  free(a->irows);
  free(a->icols);
  free(a->val);
  a->irows = a->icols = NULL;
  a->val = NULL;
  a->row = a->col = a->nnz = 0;
}

/* zero-filled arrays for an nrows x ncols matrix with nnz entries */
static inline void dcsr_mat_alloc(dcsr_mat *a, INT nrows, INT ncols, INT nnz) {
  a->row = nrows;
  a->col = ncols;
  a->nnz = nnz;
  a->ia = (INT *)calloc((size_t)nrows + 1, sizeof(INT));
  a->ja = (INT *)calloc((size_t)nnz, sizeof(INT));
  a->val = (REAL *)calloc((size_t)nnz, sizeof(REAL));
}

/* at = transpose of A with its rows taken in the order p (row i of the
 * permuted matrix is row p[i] of A; p NULL: natural order).  Rows of at
 * list their columns ascending.  at's arrays are pre-allocated; with
 * A->val or at->val NULL only the structure is formed. */
static inline void dcsr_transz(const dcsr_mat *a_in, INT *p, dcsr_mat *at_out) {
  const INT n = a_in->row, m = a_in->col;
  at_out->row = m;
  at_out->col = n;
  at_out->nnz = a_in->nnz;
  INT *ptr = at_out->ia;
  for (INT j = 0; j <= m; j++) ptr[j] = 0;
  for (INT k = 0; k < a_in->nnz; k++) ptr[a_in->ja[k] + 1]++;
  for (INT j = 0; j < m; j++) ptr[j + 1] += ptr[j];
  INT *fill = (INT *)calloc((size_t)(m > 0 ? m : 1), sizeof(INT));
  const int vals = a_in->val != NULL && at_out->val != NULL;
  for (INT i = 0; i < n; i++) {
    const INT src = p ? p[i] : i;
    for (INT k = a_in->ia[src]; k < a_in->ia[src + 1]; k++) {
      const INT j = a_in->ja[k];
      const INT at = ptr[j] + fill[j]++;
      at_out->ja[at] = i;
      if (vals) at_out->val[at] = a_in->val[k];
    }
  }
  free(fill);
}

/* c = a + b, allocated here.  Row i lists a's entries in a's order, then
 * b's columns not yet in the row; a b entry in a column already in the
 * row adds to it.  Returns -1 when the shapes differ, else 0. */
static inline INT dcsr_mat_add(dcsr_mat *a, dcsr_mat *b, dcsr_mat *c) {
  if (a->row != b->row || a->col != b->col) return -1;
  const INT n = a->row, m = a->col;
  INT *where = (INT *)malloc((size_t)(m > 0 ? m : 1) * sizeof(INT));
  REAL *acc = (REAL *)calloc((size_t)(m > 0 ? m : 1), sizeof(REAL));
  for (INT j = 0; j < m; j++) where[j] = -1;
  INT cnt = 0;
  for (INT i = 0; i < n; i++) {
    for (INT q = a->ia[i]; q < a->ia[i + 1]; q++) where[a->ja[q]] = i, cnt++;
    for (INT q = b->ia[i]; q < b->ia[i + 1]; q++)
      if (where[b->ja[q]] != i) where[b->ja[q]] = i, cnt++;
  }
  dcsr_mat_alloc(c, n, m, cnt);
  for (INT j = 0; j < m; j++) where[j] = -1;
  INT pos = 0;
  for (INT i = 0; i < n; i++) {
    const INT row0 = pos;
    c->ia[i] = pos;
    for (INT q = a->ia[i]; q < a->ia[i + 1]; q++) {
      const INT j = a->ja[q];
      where[j] = pos;
      c->ja[pos++] = j;
      acc[j] = a->val[q];
    }
    for (INT q = b->ia[i]; q < b->ia[i + 1]; q++) {
      const INT j = b->ja[q];
      if (where[j] >= row0) {
        acc[j] += b->val[q];
      } else {
        where[j] = pos;
        c->ja[pos++] = j;
        acc[j] = b->val[q];
      }
    }
    for (INT q = row0; q < pos; q++) {
      c->val[q] = acc[c->ja[q]];
      acc[c->ja[q]] = 0.0;
    }
  }
  c->ia[n] = pos;
  c->nnz = pos;
  free(where);
  free(acc);
  return 0;
}

/* ---------------------------------------------------------------------
 * Host-array helpers (C_amg1/include/agg_lib.h:251-310).  They work on
 * host memory on the CPU, as in the reference; each row's sum runs in the
 * reference's order.  The solve entry points below do not use them.
 * ------------------------------------------------------------------- */
void nullv(REAL *u, INT n);                                 /* agg_lib.h:251 */
void copyv(const REAL *u, REAL *v, INT n);                  /* agg_lib.h:254 */
void scpro(const REAL *u, const REAL *v, REAL *scpr, INT n); /* :257 */
void l2norm(const REAL *u, REAL *l2nr, INT n);              /* agg_lib.h:260 */
void uuplmv(REAL *u, const REAL *v, INT n, REAL smult);     /* agg_lib.h:263 */
void umuplv(REAL *u, const REAL *v, INT n, REAL smult);     /* agg_lib.h:266 */
void randv(REAL *u, INT n);                                 /* agg_lib.h:269 */
void dcsr_cp(const dcsr_mat *src, dcsr_mat *dst);           /* agg_lib.h:276 */
void abyvg(const INT *ia, const INT *ja, const REAL *a, const REAL *x, INT n,
           REAL *y);                                        /* agg_lib.h:284 */
void abyvcs(const INT *ia, const INT *ja, const REAL *a, const REAL *x, INT n,
            REAL *y);                                       /* agg_lib.h:288 */
void abyvam(const REAL *x, const INT *ia, const INT *ja, const REAL *a,
            const REAL *y, INT n, REAL *z);                 /* agg_lib.h:292 */
void vbya(const INT *ia, const INT *ja, const REAL *a, const REAL *x, INT n,
          INT m_dim, REAL *y);                              /* agg_lib.h:296 */
void dcoo2dcsr(dcoo_mat *acoo, dcsr_mat *a);                /* agg_lib.h:307 */
/* agg_lib.h:310: "nrows ncols nnz" then "row col value" lines, 1-based
 * indices detected; returns 0, or 1 with a message (entries outside the
 * stated shape are also rejected) */
INT dcsr_mat_dcoo(FILE *fp, dcsr_mat *a_out);

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
