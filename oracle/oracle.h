/*
 * oracle.h - CPU restatement of the C_amg1 AMG-preconditioned PCG path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code; only tests/, __graft_entry__.smoke() and the CPU-baseline
 * leg of bench.py use it, and only as the checker.
 *
 * Types and function names follow C_amg1/include/agg_lib.h:29-54 and
 * C_amg1/include/amg.h:19-115 so that the same harness (harness.c) drives
 * either this restatement or the compiled reference (oracle/_ref).
 */
#ifndef BN_ORACLE_H
#define BN_ORACLE_H

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define INT int
#define REAL double

typedef struct {
  INT row, col, nnz;
  INT *ia, *ja;
  REAL *val;
} dcsr_mat;

#define AMG_MAX_LEVELS 20
#define SMOOTHER_L1 0
#define SMOOTHER_GS 1
#define SMOOTHER_CF_GS 2
#define SMOOTHER_JACOBI 3
#define COARSEN_RS 0
#define COARSEN_PMIS 1
#define COARSEN_HMIS 2

typedef struct {
  INT n;
  INT *ia, *ja;
  REAL *val;
  INT nnz;
} ichol_factor;

typedef struct {
  dcsr_mat A, P, A_filtered;
  INT *mis, *isolated, *cf_order;
  INT n, n_coarse;
  ichol_factor L_ichol;
  REAL *l1_diag;
} amg_level;

typedef struct {
  amg_level levels[AMG_MAX_LEVELS];
  INT num_levels;
  REAL threshold;
  INT min_size, max_levels, maxit_prol, lump, nu, smoother;
  REAL omega_jacobi;
  INT coarsening;
} amg_data;

/* vector / sparse kernels */
void nullv(REAL *u, INT n);
void copyv(const REAL *u, REAL *v, INT n);
void scpro(const REAL *u, const REAL *v, REAL *s, INT n);
void l2norm(const REAL *u, REAL *s, INT n);
void uuplmv(REAL *u, const REAL *v, INT n, REAL a);
void umuplv(REAL *u, const REAL *v, INT n, REAL a);
void abyvg(const INT *ia, const INT *ja, const REAL *a, const REAL *x, INT n, REAL *y);
void abyvcs(const INT *ia, const INT *ja, const REAL *a, const REAL *x, INT n, REAL *y);
void abyvam(const REAL *x, const INT *ia, const INT *ja, const REAL *a,
            const REAL *y, INT n, REAL *z);
void vbya(const INT *ia, const INT *ja, const REAL *a, const REAL *x, INT n,
          INT m, REAL *y);
void dcsr_rap(dcsr_mat *r, dcsr_mat *a, dcsr_mat *p, dcsr_mat *rap);
void oracle_transpose(const dcsr_mat *a, dcsr_mat *at);
void oracle_add(const dcsr_mat *a, const dcsr_mat *b, dcsr_mat *c);
void oracle_symmetrize(const dcsr_mat *a, dcsr_mat *out);
void oracle_csr_free(dcsr_mat *a);

/* setup */
void rs_strength(const dcsr_mat *A, REAL theta, dcsr_mat *S, INT *max_lam);
void rs_coarsening(const dcsr_mat *S, INT max_lam, INT *cf, INT *n_coarse);
void pmis_coarsening(const dcsr_mat *S, INT *cf, INT *n_coarse);
void hmis_coarsening(const dcsr_mat *S, INT *cf, INT *n_coarse);
void rs_standard_interpolation_sparsity(const dcsr_mat *S, const INT *cf, dcsr_mat *P);
void rs_standard_interpolation_values(const dcsr_mat *A, const dcsr_mat *S,
                                      const INT *cf, dcsr_mat *P);
void rs_standard_interpolation(const dcsr_mat *A, const dcsr_mat *S,
                               const INT *cf, dcsr_mat *P);
void amg_build_hierarchy(amg_data *amg, const dcsr_mat *A);
void amg_rebuild_values(amg_data *amg, const dcsr_mat *A_new);
void amg_free(amg_data *amg);
void ichol_compute(const dcsr_mat *A, ichol_factor *L);
void ichol_solve(const ichol_factor *L, const REAL *b, REAL *x);
void ichol_free(ichol_factor *L);

/* solve */
void amg_backslash(const amg_data *amg, INT lev, const REAL *b, REAL *x, INT nu);
void amg_fwdslash(const amg_data *amg, INT lev, const REAL *b, REAL *x, INT nu);
void amg_vcycle_precond(const amg_data *amg, const REAL *g, REAL *x);
void amg_ichol_precond(const amg_data *amg, const dcsr_mat *A,
                       const ichol_factor *L, const REAL *g, REAL *x);
/* MEX smoother (Matlab_amg/cf_gauss_seidel_mex.c:27-105) */
void oracle_cf_gauss_seidel(INT n, const INT *ia, const INT *ja, const REAL *val, const REAL *b, const REAL *x0,
                            const REAL *cf_order, INT nu, INT direction, REAL *x);
INT amg_pcg(const dcsr_mat *A, const REAL *b, REAL *x,
            void (*precond)(const void *, const REAL *, REAL *),
            const void *data, REAL tol, INT maxiter, INT print_level);

#endif
