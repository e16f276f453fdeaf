/*
 * twins.c - C-typed twin of the MATLAB MEX smoother
 * x = cf_gauss_seidel_mex(A, b, x0, cf_order, nu, direction)
 * (Matlab_amg/cf_gauss_seidel_mex.c:7,27-105) on the GPU.
 *
 * The sequential sweep visits rows in cf_order (forward) or in its
 * reverse (backward); each row sets x_i = (b_i - sum_{j != i} a_ij x_j) /
 * a_ii with the values current at that moment.  Here the visiting order's
 * dependency DAG is cut into levels on the host (bn_sweep_levels: rows of
 * one level are uncoupled), and the device updates one level per launch,
 * so every row reads exactly the values the sequential sweep reads.  The
 * diagonal is the row's first diagonal entry, 1.0 when it has none, as
 * the MEX code takes it (cf_gauss_seidel_mex.c:66-75).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bn_solver.h"
#include "csr.h"
#include "schedule.h"

void bn_set_error(const char *fmt, ...);

static int valid_pattern(int n, int nnz, const int *ia, const int *ja) {
  if (n <= 0 || nnz < 0 || !ia || (nnz > 0 && !ja) || ia[0] != 0 || ia[n] != nnz) return 0;
  for (int i = 0; i < n; i++)
    if (ia[i + 1] < ia[i]) return 0;
  for (int k = 0; k < nnz; k++)
    if (ja[k] < 0 || ja[k] >= n) return 0;
  return 1;
}

/* nu sweeps of the host matrix A's order on device copies: d_ia, d_ja,
 * d_val hold A, d_b the right-hand side, d_x the iterate (in place) */
static int sweep_device(const dcsr_mat *A, const int *order, int nu, int backward, const int *d_ia,
                        const int *d_ja, const double *d_val, const double *d_b, double *d_x) {
  const int n = A->row;
  int *rows = (int *)bn_xmalloc((size_t)n * sizeof(int));
  int *ptr = NULL;
  const int nl = bn_sweep_levels(A, order, backward, rows, &ptr);
  if (nl < 0) {
    free(rows);
    bn_set_error("cf_gs:order: cf_order is not a permutation of 1..%d", n);
    return BN_ERR_ARG;
  }
  double *dinv = (double *)bn_xmalloc((size_t)n * sizeof(double));
  for (int i = 0; i < n; i++) {
    double d = 1.0;
    for (int k = A->ia[i]; k < A->ia[i + 1]; k++)
      if (A->ja[k] == i) {
        d = A->val[k];
        break;
      }
    dinv[i] = 1.0 / d;
  }
  int rc = 0;
  int *d_rows = NULL;
  double *d_dinv = NULL;
  hipStream_t st = NULL;
  if (hipMalloc((void **)&d_rows, (size_t)n * sizeof(int)) != hipSuccess ||
      hipMalloc((void **)&d_dinv, (size_t)n * sizeof(double)) != hipSuccess ||
      hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
    bn_set_error("cf_gauss_seidel: device allocation failed");
    rc = BN_ERR_HIP;
  }
  if (!rc && (hipMemcpyAsync(d_rows, rows, (size_t)n * sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess ||
              hipMemcpyAsync(d_dinv, dinv, (size_t)n * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess)) {
    bn_set_error("cf_gauss_seidel: copy to the device failed");
    rc = BN_ERR_HIP;
  }
  if (!rc) {
    const int lanes = bn_lanes_for(A->nnz, n);
    for (int sw = 0; sw < nu; sw++)
      for (int l = 0; l < nl; l++)
        bn_trsv_level('d', ptr[l + 1] - ptr[l], d_rows + ptr[l], d_ia, d_ja, d_val, d_dinv, d_b, d_x, lanes, NULL,
                      NULL, st);
    if (hipStreamSynchronize(st) != hipSuccess) {
      bn_set_error("cf_gauss_seidel: device failure during the sweeps");
      rc = BN_ERR_HIP;
    }
  }
  if (st) hipStreamDestroy(st);
  hipFree(d_rows);
  hipFree(d_dinv);
  free(rows);
  free(ptr);
  free(dinv);
  return rc;
}

int cf_gauss_seidel(int n, const int *ia, const int *ja, const double *val, const double *b, const double *x0,
                    const double *cf_order, int nu, int direction, double *x) {
  if (n <= 0 || !ia || !valid_pattern(n, ia[n], ia, ja) || !val || !b || !x0 || !cf_order || !x || nu < 0 ||
      (direction != 0 && direction != 1)) {
    bn_set_error("cf_gs:args: need an n x n CSR matrix, b, x0, cf_order, nu >= 0 and direction 0 or 1");
    return BN_ERR_ARG;
  }
  const int nnz = ia[n];
  int *order = (int *)bn_xmalloc((size_t)n * sizeof(int));
  for (int k = 0; k < n; k++) {
    const double v = cf_order[k];
    if (!(v >= 1.0 && v <= (double)n) || v != (double)(int)v) {
      free(order);
      bn_set_error("cf_gs:order: cf_order(%d) = %g is not a row number in 1..%d", k + 1, v, n);
      return BN_ERR_ARG;
    }
    order[k] = (int)v - 1; /* 1-based -> 0-based, cf_gauss_seidel_mex.c:83 */
  }
  if (x != x0) memcpy(x, x0, (size_t)n * sizeof(double));
  if (nu == 0) {
    free(order);
    return 0;
  }
  if (!bn_device_ok()) {
    free(order);
    bn_set_error("no HIP device available: libbnamg sweeps run only on the GPU");
    return BN_ERR_NO_DEVICE;
  }
  const dcsr_mat A = {n, n, nnz, (int *)ia, (int *)ja, (double *)val};
  int *d_ia = NULL, *d_ja = NULL;
  double *d_val = NULL, *d_b = NULL, *d_x = NULL;
  int rc = 0;
  if (hipMalloc((void **)&d_ia, ((size_t)n + 1) * sizeof(int)) != hipSuccess ||
      hipMalloc((void **)&d_ja, (size_t)(nnz > 0 ? nnz : 1) * sizeof(int)) != hipSuccess ||
      hipMalloc((void **)&d_val, (size_t)(nnz > 0 ? nnz : 1) * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&d_b, (size_t)n * sizeof(double)) != hipSuccess ||
      hipMalloc((void **)&d_x, (size_t)n * sizeof(double)) != hipSuccess) {
    bn_set_error("cf_gauss_seidel: device allocation failed");
    rc = BN_ERR_HIP;
  }
  if (!rc && (hipMemcpy(d_ia, ia, ((size_t)n + 1) * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
              (nnz > 0 && hipMemcpy(d_ja, ja, (size_t)nnz * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) ||
              (nnz > 0 && hipMemcpy(d_val, val, (size_t)nnz * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) ||
              hipMemcpy(d_b, b, (size_t)n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
              hipMemcpy(d_x, x, (size_t)n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)) {
    bn_set_error("cf_gauss_seidel: copy to the device failed");
    rc = BN_ERR_HIP;
  }
  if (!rc) rc = sweep_device(&A, order, nu, direction, d_ia, d_ja, d_val, d_b, d_x);
  if (!rc && hipMemcpy(x, d_x, (size_t)n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) {
    bn_set_error("cf_gauss_seidel: copy from the device failed");
    rc = BN_ERR_HIP;
  }
  hipFree(d_ia);
  hipFree(d_ja);
  hipFree(d_val);
  hipFree(d_b);
  hipFree(d_x);
  free(order);
  return rc;
}

int bn_cf_gauss_seidel_device(int n, int nnz, const int *d_ia, const int *d_ja, const double *d_val,
                              const double *d_b, double *d_x, const int *d_order, int nu, int direction) {
  if (n <= 0 || nnz < 0 || !d_ia || (nnz > 0 && (!d_ja || !d_val)) || !d_b || !d_x || !d_order || nu < 0 ||
      (direction != 0 && direction != 1)) {
    bn_set_error("cf_gs:args: need n > 0, device CSR arrays, b, x, order, nu >= 0 and direction 0 or 1");
    return BN_ERR_ARG;
  }
  if (nu == 0) return 0;
  if (!bn_device_ok()) {
    bn_set_error("no HIP device available: libbnamg sweeps run only on the GPU");
    return BN_ERR_NO_DEVICE;
  }
  /* the schedule is host analysis: the pattern, the diagonal and the
   * order come back once */
  dcsr_mat A;
  bn_csr_alloc(&A, n, n, nnz);
  int *order = (int *)bn_xmalloc((size_t)n * sizeof(int));
  int rc = 0;
  if (hipMemcpy(A.ia, d_ia, ((size_t)n + 1) * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
      (nnz > 0 && hipMemcpy(A.ja, d_ja, (size_t)nnz * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) ||
      (nnz > 0 && hipMemcpy(A.val, d_val, (size_t)nnz * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) ||
      hipMemcpy(order, d_order, (size_t)n * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) {
    bn_set_error("bn_cf_gauss_seidel_device: copy from the device failed");
    rc = BN_ERR_HIP;
  }
  if (!rc && !valid_pattern(n, nnz, A.ia, A.ja)) {
    bn_set_error("cf_gs:args: malformed CSR pattern (n=%d, nnz=%d)", n, nnz);
    rc = BN_ERR_ARG;
  }
  if (!rc) rc = sweep_device(&A, order, nu, direction, d_ia, d_ja, d_val, d_b, d_x);
  bn_csr_free(&A);
  free(order);
  return rc;
}
