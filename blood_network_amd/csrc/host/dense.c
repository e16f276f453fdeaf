/*
 * dense.c - explicit inverse of the (small) coarsest operator, used by
 * the direct coarsest-level solve.  Left-looking Cholesky on a row-major
 * dense copy, then A^{-1} column by column with two column-oriented
 * triangular sweeps; rows/columns are spread over OpenMP threads.  The
 * device version (bn_dn_inverse, rebuild.hip) forms every entry with the
 * same operations in the same order, so the two inverses are identical.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "amg.h"
#include "csr.h"

/* Returns 0 on success, 1 if a pivot was <= 0 and had to be clamped
 * (the operator is then not numerically SPD). */
int bn_dense_inverse(const dcsr_mat *A, double *inv) {
  const int n = A->row;
  double *L = (double *)bn_xcalloc((size_t)n * (size_t)n, sizeof(double));
  for (int i = 0; i < n; i++)
    for (int k = A->ia[i]; k < A->ia[i + 1]; k++) L[(size_t)i * n + A->ja[k]] += A->val[k];
  int clamped = 0;
  for (int k = 0; k < n; k++) {
    double *Lk = L + (size_t)k * n;
    double d = Lk[k];
    for (int p = 0; p < k; p++) d -= Lk[p] * Lk[p];
    if (!(d > 0.0)) {
      d = 1e-300;
      clamped = 1;
    }
    const double lkk = sqrt(d);
    Lk[k] = lkk;
#pragma omp parallel for schedule(static)
    for (int i = k + 1; i < n; i++) {
      double *Li = L + (size_t)i * n;
      double s = Li[k];
      for (int p = 0; p < k; p++) s -= Li[p] * Lk[p];
      Li[k] = s / lkk;
    }
  }
  /* row-major L^T: row p of U is column p of L */
  double *U = (double *)bn_xmalloc((size_t)n * (size_t)n * sizeof(double));
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++)
    for (int p = 0; p < n; p++) U[(size_t)i * n + p] = (p >= i) ? L[(size_t)p * n + i] : 0.0;
  /* columns of A^{-1} (A^{-1} is symmetric), both sweeps column-oriented:
   * once y_p is final it is taken out of every later (forward) or earlier
   * (backward) entry.  Each entry of the forward sweep sees its terms in
   * the order of the row-oriented dot, p ascending; the backward sweep
   * takes them p descending.  The device inverse runs the same order. */
#pragma omp parallel
  {
    double *y = (double *)bn_xmalloc((size_t)n * sizeof(double));
#pragma omp for schedule(dynamic, 8)
    for (int c = 0; c < n; c++) {
      for (int i = 0; i < n; i++) y[i] = (i == c) ? 1.0 : 0.0;
      for (int p = c; p < n; p++) {
        const double yp = y[p] / L[(size_t)p * n + p];
        y[p] = yp;
        const double *Up = U + (size_t)p * n;
        for (int i = p + 1; i < n; i++) y[i] -= Up[i] * yp;
      }
      for (int p = n - 1; p >= 0; p--) {
        const double yp = y[p] / L[(size_t)p * n + p];
        y[p] = yp;
        const double *Lp = L + (size_t)p * n;
        for (int i = 0; i < p; i++) y[i] -= Lp[i] * yp;
      }
      for (int i = 0; i < n; i++) inv[(size_t)i * n + c] = y[i];
    }
    free(y);
  }
  free(U);
  free(L);
  return clamped;
}
