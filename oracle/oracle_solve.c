/*
 * oracle_solve.c - CPU restatement of the C_amg1 solve path: BLAS-1 and
 * CSR kernels, smoothers, backslash/fwdslash half cycles, the AMG+IC
 * preconditioner and PCG.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Loop orders
 * and operation orders follow the cited reference lines so the FP64
 * results are bit-identical to the reference built by gcc without FMA.
 */
#include "oracle.h"

/* C_amg1/src/vector_ops.c:17-72 */
void nullv(REAL *u, INT n) { for (INT i = 0; i < n; i++) u[i] = 0.0; }
void copyv(const REAL *u, REAL *v, INT n) { for (INT i = 0; i < n; i++) v[i] = u[i]; }
void scpro(const REAL *u, const REAL *v, REAL *s, INT n) {
  REAL acc = 0.0;
  for (INT i = 0; i < n; i++) acc += u[i] * v[i];
  *s = acc;
}
void l2norm(const REAL *u, REAL *s, INT n) { scpro(u, u, s, n); *s = sqrt(*s); }
void uuplmv(REAL *u, const REAL *v, INT n, REAL a) {
  for (INT i = 0; i < n; i++) u[i] = u[i] + v[i] * a;
}
void umuplv(REAL *u, const REAL *v, INT n, REAL a) {
  for (INT i = 0; i < n; i++) u[i] = a * u[i] + v[i];
}

/* C_amg1/src/sparse_ops.c:19-75: row sums in CSR order */
void abyvg(const INT *ia, const INT *ja, const REAL *a, const REAL *x, INT n, REAL *y) {
  for (INT i = 0; i < n; i++) {
    REAL s = 0.0;
    for (INT k = ia[i]; k < ia[i + 1]; k++) s += a[k] * x[ja[k]];
    y[i] = s;
  }
}
void abyvcs(const INT *ia, const INT *ja, const REAL *a, const REAL *x, INT n, REAL *y) {
  for (INT i = 0; i < n; i++) {
    REAL s = y[i];
    for (INT k = ia[i]; k < ia[i + 1]; k++) s += a[k] * x[ja[k]];
    y[i] = s;
  }
}
void abyvam(const REAL *x, const INT *ia, const INT *ja, const REAL *a,
            const REAL *y, INT n, REAL *z) {
  for (INT i = 0; i < n; i++) {
    REAL s = x[i];
    for (INT k = ia[i]; k < ia[i + 1]; k++) s -= a[k] * y[ja[k]];
    z[i] = s;
  }
}
/* y = A^T x as a scatter over rows in order (sparse_ops.c:61-75) */
void vbya(const INT *ia, const INT *ja, const REAL *a, const REAL *x, INT n,
          INT m, REAL *y) {
  for (INT j = 0; j < m; j++) y[j] = 0.0;
  for (INT i = 0; i < n; i++)
    for (INT k = ia[i]; k < ia[i + 1]; k++) y[ja[k]] += a[k] * x[i];
}

/* Forward L y = b then backward L^T x = y with L stored by rows
 * (C_amg1/src/ichol.c:121-165).  The forward sweep takes the diagonal from
 * the row (1.0 if absent); the backward sweep scatters column updates. */
void ichol_solve(const ichol_factor *L, const REAL *b, REAL *x) {
  INT n = L->n;
  REAL *y = (REAL *)calloc((size_t)(n > 0 ? n : 1), sizeof(REAL));
  REAL *dg = (REAL *)calloc((size_t)(n > 0 ? n : 1), sizeof(REAL));
  for (INT i = 0; i < n; i++) {
    REAL s = b[i], d = 1.0;
    for (INT k = L->ia[i]; k < L->ia[i + 1]; k++) {
      INT j = L->ja[k];
      if (j < i) s -= L->val[k] * y[j];
      else if (j == i) d = L->val[k];
    }
    y[i] = s / d;
  }
  for (INT i = 0; i < n; i++) x[i] = y[i];
  for (INT i = 0; i < n; i++)
    for (INT k = L->ia[i]; k < L->ia[i + 1]; k++)
      if (L->ja[k] == i) { dg[i] = L->val[k]; break; }
  for (INT i = n - 1; i >= 0; i--) {
    x[i] /= dg[i];
    for (INT k = L->ia[i]; k < L->ia[i + 1]; k++)
      if (L->ja[k] < i) x[L->ja[k]] -= L->val[k] * x[i];
  }
  free(y); free(dg);
}

/* One Gauss-Seidel update of row i: x_i = (b_i - sum_{j!=i} a_ij x_j)/a_ii,
 * diagonal 1.0 if absent (C_amg1/src/amg_cycle.c:21-80). */
static void gs_row(const dcsr_mat *A, INT i, const REAL *b, REAL *x) {
  REAL s = b[i], d = 1.0;
  for (INT k = A->ia[i]; k < A->ia[i + 1]; k++) {
    if (A->ja[k] == i) d = A->val[k];
    else s -= A->val[k] * x[A->ja[k]];
  }
  x[i] = s / d;
}

/* damped Jacobi (amg_cycle.c:85-102) and l1-Jacobi (amg_cycle.c:113-129):
 * the residual includes the diagonal term */
static void jacobi_like(const dcsr_mat *A, const REAL *l1, REAL omega, INT n,
                        const REAL *b, REAL *x, INT nu) {
  REAL *r = (REAL *)malloc((size_t)(n > 0 ? n : 1) * sizeof(REAL));
  for (INT s = 0; s < nu; s++) {
    for (INT i = 0; i < n; i++) {
      REAL t = b[i], d = 1.0;
      for (INT k = A->ia[i]; k < A->ia[i + 1]; k++) {
        if (A->ja[k] == i) d = A->val[k];
        t -= A->val[k] * x[A->ja[k]];
      }
      r[i] = l1 ? t : t / d;
    }
    for (INT i = 0; i < n; i++) {
      if (l1) x[i] += r[i] / l1[i];
      else x[i] += omega * r[i];
    }
  }
  free(r);
}

/* smoother dispatch (amg_cycle.c:137-175): GS variants run the given order
 * forwards (pre) or backwards (post); Jacobi variants are symmetric */
static void smooth(const amg_data *amg, INT lev, const REAL *b, REAL *x, INT nu, int backward) {
  const amg_level *L = &amg->levels[lev];
  INT n = L->n;
  switch (amg->smoother) {
    case SMOOTHER_GS:
    case SMOOTHER_CF_GS:
      for (INT s = 0; s < nu; s++)
        for (INT t = 0; t < n; t++) {
          INT ii = backward ? n - 1 - t : t;
          INT i = (amg->smoother == SMOOTHER_CF_GS) ? L->cf_order[ii] : ii;
          gs_row(&L->A, i, b, x);
        }
      break;
    case SMOOTHER_JACOBI:
      jacobi_like(&L->A, NULL, amg->omega_jacobi, n, b, x, nu);
      break;
    default:
      jacobi_like(&L->A, L->l1_diag, 0.0, n, b, x, nu);
      break;
  }
}

static void ichol_wrap(const void *data, const REAL *g, REAL *z) {
  ichol_solve((const ichol_factor *)data, g, z);
}

/* coarsest level: IC-preconditioned PCG to 1e-12 with at most n_c steps
 * (amg_cycle.c:186-190, :228-231) */
static void coarsest(const amg_level *L, const REAL *b, REAL *x) {
  amg_pcg(&L->A, b, x, ichol_wrap, &L->L_ichol, 1e-12, L->n, 0);
}

/* backslash: pre-smooth, residual, restrict, recurse from zero, prolong
 * (amg_cycle.c:180-217) */
void amg_backslash(const amg_data *amg, INT lev, const REAL *b, REAL *x, INT nu) {
  const amg_level *L = &amg->levels[lev];
  if (lev == amg->num_levels - 1) { coarsest(L, b, x); return; }
  INT n = L->n, nc = L->n_coarse;
  smooth(amg, lev, b, x, nu, 0);
  REAL *r = (REAL *)calloc((size_t)n, sizeof(REAL));
  REAL *rc = (REAL *)calloc((size_t)nc, sizeof(REAL));
  REAL *ec = (REAL *)calloc((size_t)nc, sizeof(REAL));
  abyvam(b, L->A.ia, L->A.ja, L->A.val, x, n, r);
  vbya(L->P.ia, L->P.ja, L->P.val, r, n, nc, rc);
  amg_backslash(amg, lev + 1, rc, ec, nu);
  abyvcs(L->P.ia, L->P.ja, L->P.val, ec, n, x);
  free(r); free(rc); free(ec);
}

/* fwdslash: residual, restrict, recurse from zero, prolong, post-smooth
 * (amg_cycle.c:222-258) */
void amg_fwdslash(const amg_data *amg, INT lev, const REAL *b, REAL *x, INT nu) {
  const amg_level *L = &amg->levels[lev];
  if (lev == amg->num_levels - 1) { coarsest(L, b, x); return; }
  INT n = L->n, nc = L->n_coarse;
  REAL *r = (REAL *)calloc((size_t)n, sizeof(REAL));
  REAL *rc = (REAL *)calloc((size_t)nc, sizeof(REAL));
  REAL *ec = (REAL *)calloc((size_t)nc, sizeof(REAL));
  abyvam(b, L->A.ia, L->A.ja, L->A.val, x, n, r);
  vbya(L->P.ia, L->P.ja, L->P.val, r, n, nc, rc);
  amg_fwdslash(amg, lev + 1, rc, ec, nu);
  abyvcs(L->P.ia, L->P.ja, L->P.val, ec, n, x);
  smooth(amg, lev, b, x, nu, 1);
  free(r); free(rc); free(ec);
}

/* x = backslash(g) + fwdslash(g - A backslash(g)) (amg_cycle.c:263-284) */
void amg_vcycle_precond(const amg_data *amg, const REAL *g, REAL *x) {
  INT n = amg->levels[0].n;
  const dcsr_mat *A = &amg->levels[0].A;
  nullv(x, n);
  amg_backslash(amg, 0, g, x, amg->nu);
  REAL *r = (REAL *)calloc((size_t)n, sizeof(REAL));
  REAL *c = (REAL *)calloc((size_t)n, sizeof(REAL));
  abyvam(g, A->ia, A->ja, A->val, x, n, r);
  amg_fwdslash(amg, 0, r, c, amg->nu);
  uuplmv(x, c, n, 1.0);
  free(r); free(c);
}

/* u1 = backslash(g); x = u1 + (LL^T)^{-1}(g - A u1);
 * x += fwdslash(g - A x)  (C_amg1/src/amg_ichol_precond.c:22-59) */
void amg_ichol_precond(const amg_data *amg, const dcsr_mat *A,
                       const ichol_factor *L, const REAL *g, REAL *x) {
  INT n = amg->levels[0].n, nu = amg->nu;
  REAL *u1 = (REAL *)calloc((size_t)n, sizeof(REAL));
  REAL *r1 = (REAL *)calloc((size_t)n, sizeof(REAL));
  REAL *ic = (REAL *)calloc((size_t)n, sizeof(REAL));
  REAL *r2 = (REAL *)calloc((size_t)n, sizeof(REAL));
  REAL *cr = (REAL *)calloc((size_t)n, sizeof(REAL));
  amg_backslash(amg, 0, g, u1, nu);
  abyvam(g, A->ia, A->ja, A->val, u1, n, r1);
  ichol_solve(L, r1, ic);
  copyv(u1, x, n);
  uuplmv(x, ic, n, 1.0);
  abyvam(g, A->ia, A->ja, A->val, x, n, r2);
  amg_fwdslash(amg, 0, r2, cr, nu);
  uuplmv(x, cr, n, 1.0);
  free(u1); free(r1); free(ic); free(r2); free(cr);
}

/* PCG with the preconditioned-norm stop sqrt(r^T z / r0^T z0) < tol;
 * returns the iteration count or -maxiter (C_amg1/src/amg_pcg.c:12-101). */
INT amg_pcg(const dcsr_mat *A, const REAL *b, REAL *x,
            void (*precond)(const void *, const REAL *, REAL *),
            const void *data, REAL tol, INT maxiter, INT print_level) {
  INT n = A->row;
  REAL *r = (REAL *)calloc((size_t)(n > 0 ? n : 1), sizeof(REAL));
  REAL *z = (REAL *)calloc((size_t)(n > 0 ? n : 1), sizeof(REAL));
  REAL *p = (REAL *)calloc((size_t)(n > 0 ? n : 1), sizeof(REAL));
  REAL *Ap = (REAL *)calloc((size_t)(n > 0 ? n : 1), sizeof(REAL));
  abyvam(b, A->ia, A->ja, A->val, x, n, r);
  REAL bnorm, rnorm, rz, rz0, rz_new, pAp;
  l2norm(b, &bnorm, n);
  if (bnorm == 0.0) bnorm = 1.0;
  precond(data, r, z);
  copyv(z, p, n);
  scpro(r, z, &rz, n);
  rz0 = rz;
  l2norm(r, &rnorm, n);
  if (print_level >= 1) fprintf(stderr, "  %4d  %8.2e  %8.2e\n", 0, sqrt(rz0), 1.0);
  INT it;
  for (it = 1; it <= maxiter; it++) {
    abyvg(A->ia, A->ja, A->val, p, n, Ap);
    scpro(p, Ap, &pAp, n);
    REAL alpha = rz / pAp;
    uuplmv(x, p, n, alpha);
    uuplmv(r, Ap, n, -alpha);
    l2norm(r, &rnorm, n);
    precond(data, r, z);
    scpro(r, z, &rz_new, n);
    REAL rel = sqrt(rz_new / rz0);
    if (print_level >= 1)
      fprintf(stderr, "  %4d  %8.2e  %8.2e  %6.4f\n", it, sqrt(rz_new), rel, rz_new / rz);
    if (rel < tol) break;
    REAL beta = rz_new / rz;
    rz = rz_new;
    umuplv(p, z, n, beta);
  }
  if (it > maxiter) {
    fprintf(stderr, "  PCG: no convergence after %d iterations, relres = %.2e\n",
            maxiter, rnorm / bnorm);
    it = -maxiter;
  }
  free(r); free(z); free(p); free(Ap);
  return it;
}

/* The MEX smoother x = cf_gauss_seidel_mex(A, b, x0, cf_order, nu,
 * direction) (Matlab_amg/cf_gauss_seidel_mex.c:27-105): x starts as a copy
 * of x0; each row's diagonal is its first diagonal entry, 1.0 if none,
 * taken once before the sweeps (:66-75); cf_order holds 1-based row numbers
 * as doubles (:83); direction 0 visits cf_order(1) .. cf_order(n), 1 the
 * reverse (:79-103); x_i = (b_i - sum_{j != i} a_ij x_j) / diag_i. */
void oracle_cf_gauss_seidel(INT n, const INT *ia, const INT *ja, const REAL *val, const REAL *b, const REAL *x0,
                            const REAL *cf_order, INT nu, INT direction, REAL *x) {
  REAL *dg = (REAL *)malloc((size_t)(n > 0 ? n : 1) * sizeof(REAL));
  memcpy(x, x0, (size_t)n * sizeof(REAL));
  for (INT i = 0; i < n; i++) {
    dg[i] = 1.0;
    for (INT k = ia[i]; k < ia[i + 1]; k++)
      if (ja[k] == i) {
        dg[i] = val[k];
        break;
      }
  }
  for (INT s = 0; s < nu; s++)
    for (INT t = 0; t < n; t++) {
      const INT ii = direction ? n - 1 - t : t;
      const INT i = (INT)(cf_order[ii] - 1);
      REAL acc = b[i];
      for (INT k = ia[i]; k < ia[i + 1]; k++)
        if (ja[k] != i) acc -= val[k] * x[ja[k]];
      x[i] = acc / dg[i];
    }
  free(dg);
}
