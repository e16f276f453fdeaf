/*
 * dense.c - explicit inverse of the (small) coarsest operator, used by
 * the direct coarsest-level solve.  Left-looking Cholesky on a row-major
 * dense copy, then A^{-1} column by column with two triangular solves;
 * rows/columns are spread over OpenMP threads.
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
  /* row-major L^T so both triangular sweeps read contiguous rows */
  double *U = (double *)bn_xmalloc((size_t)n * (size_t)n * sizeof(double));
#pragma omp parallel for schedule(static)
  for (int i = 0; i < n; i++)
    for (int p = 0; p < n; p++) U[(size_t)i * n + p] = (p >= i) ? L[(size_t)p * n + i] : 0.0;
  /* columns of A^{-1}: L y = e_c, L^T x = y  (A^{-1} is symmetric) */
#pragma omp parallel
  {
    double *y = (double *)bn_xmalloc((size_t)n * sizeof(double));
#pragma omp for schedule(dynamic, 8)
    for (int c = 0; c < n; c++) {
      for (int i = 0; i < c; i++) y[i] = 0.0;
      for (int i = c; i < n; i++) {
        const double *Li = L + (size_t)i * n;
        double s = (i == c) ? 1.0 : 0.0;
        for (int p = c; p < i; p++) s -= Li[p] * y[p];
        y[i] = s / Li[i];
      }
      for (int i = n - 1; i >= 0; i--) {
        const double *Ui = U + (size_t)i * n;
        double s = y[i];
        for (int p = i + 1; p < n; p++) s -= Ui[p] * y[p];
        y[i] = s / Ui[i];
      }
      for (int i = 0; i < n; i++) inv[(size_t)i * n + c] = y[i];
    }
    free(y);
  }
  free(U);
  free(L);
  return clamped;
}
