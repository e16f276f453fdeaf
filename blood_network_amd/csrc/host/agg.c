/*
 * agg.c - the host-array helpers of the C_amg1 library ABI
 * (C_amg1/include/agg_lib.h:251-310, bodies in vector_ops.c, sparse_ops.c
 * and io.c): BLAS-1 on host vectors, CSR products on host arrays, and the
 * COO text reader.  Callers use them around a solve (test drivers compute
 * their final residual with abyvam, read matrices with dcsr_mat_dcoo); the
 * solve itself never calls them.  Each row's sum runs in the reference's
 * order, so results match it bit for bit; rows run in parallel.
 */
#include <limits.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "amg.h"

/* vector_ops.c:17-22 */
void nullv(REAL *u, INT n) {
  if (n > 0) memset(u, 0, (size_t)n * sizeof(REAL));
}

/* vector_ops.c:27-32 */
void copyv(const REAL *u, REAL *v, INT n) {
  if (n > 0) memmove(v, u, (size_t)n * sizeof(REAL));
}

/* vector_ops.c:37-43: one running sum in index order */
void scpro(const REAL *u, const REAL *v, REAL *scpr, INT n) {
  REAL s = 0.0;
  for (INT i = 0; i < n; i++) s += u[i] * v[i];
  *scpr = s;
}

/* vector_ops.c:48-52 */
void l2norm(const REAL *u, REAL *l2nr, INT n) {
  scpro(u, u, l2nr, n);
  *l2nr = sqrt(*l2nr);
}

/* vector_ops.c:57-62: u += smult v */
void uuplmv(REAL *u, const REAL *v, INT n, REAL smult) {
#pragma omp parallel for schedule(static) if (n > 65536)
  for (INT i = 0; i < n; i++) u[i] = u[i] + v[i] * smult;
}

/* vector_ops.c:67-72: u = smult u + v */
void umuplv(REAL *u, const REAL *v, INT n, REAL smult) {
#pragma omp parallel for schedule(static) if (n > 65536)
  for (INT i = 0; i < n; i++) u[i] = smult * u[i] + v[i];
}

/* vector_ops.c:77-82: the libc generator, so a caller that seeds it gets
 * the reference's vector */
void randv(REAL *u, INT n) {
  for (INT i = 0; i < n; i++) u[i] = (REAL)rand() / (REAL)RAND_MAX;
}

/* sparse_ops.c:80-87: dst pre-allocated */
void dcsr_cp(const dcsr_mat *src, dcsr_mat *dst) {
  dst->row = src->row;
  dst->col = src->col;
  dst->nnz = src->nnz;
  memcpy(dst->ia, src->ia, ((size_t)src->row + 1) * sizeof(INT));
  memcpy(dst->ja, src->ja, (size_t)src->nnz * sizeof(INT));
  memcpy(dst->val, src->val, (size_t)src->nnz * sizeof(REAL));
}

/* sparse_ops.c:19-28: y = A x */
void abyvg(const INT *ia, const INT *ja, const REAL *a, const REAL *x, INT n, REAL *y) {
#pragma omp parallel for schedule(static) if (n > 16384)
  for (INT i = 0; i < n; i++) {
    REAL s = 0.0;
    for (INT q = ia[i]; q < ia[i + 1]; q++) s += a[q] * x[ja[q]];
    y[i] = s;
  }
}

/* sparse_ops.c:33-42: y += A x (the sum starts from y_i) */
void abyvcs(const INT *ia, const INT *ja, const REAL *a, const REAL *x, INT n, REAL *y) {
#pragma omp parallel for schedule(static) if (n > 16384)
  for (INT i = 0; i < n; i++) {
    REAL s = y[i];
    for (INT q = ia[i]; q < ia[i + 1]; q++) s += a[q] * x[ja[q]];
    y[i] = s;
  }
}

/* sparse_ops.c:47-56: z = x - A y */
void abyvam(const REAL *x, const INT *ia, const INT *ja, const REAL *a, const REAL *y, INT n, REAL *z) {
#pragma omp parallel for schedule(static) if (n > 16384)
  for (INT i = 0; i < n; i++) {
    REAL s = x[i];
    for (INT q = ia[i]; q < ia[i + 1]; q++) s -= a[q] * y[ja[q]];
    z[i] = s;
  }
}

/* sparse_ops.c:61-75: y = A^T x (m_dim entries), scattered in row order */
void vbya(const INT *ia, const INT *ja, const REAL *a, const REAL *x, INT n, INT m_dim, REAL *y) {
  nullv(y, m_dim);
  for (INT i = 0; i < n; i++) {
    const REAL xi = x[i];
    for (INT q = ia[i]; q < ia[i + 1]; q++) y[ja[q]] += a[q] * xi;
  }
}

/* io.c:21-60: COO -> CSR by a counting sort on the row index; entries of
 * a row keep their order in the COO arrays (duplicates stay separate) */
void dcoo2dcsr(dcoo_mat *acoo, dcsr_mat *a) {
  const INT m = acoo->row, nnz = acoo->nnz;
  a->row = acoo->row;
  a->col = acoo->col;
  a->nnz = nnz;
  a->ia = (INT *)calloc((size_t)m + 1, sizeof(INT));
  a->ja = (INT *)calloc((size_t)(nnz > 0 ? nnz : 1), sizeof(INT));
  a->val = (REAL *)calloc((size_t)(nnz > 0 ? nnz : 1), sizeof(REAL));
  for (INT k = 0; k < nnz; k++) a->ia[acoo->irows[k] + 1]++;
  for (INT i = 0; i < m; i++) a->ia[i + 1] += a->ia[i];
  INT *next = (INT *)malloc(((size_t)m + 1) * sizeof(INT));
  memcpy(next, a->ia, ((size_t)m + 1) * sizeof(INT));
  for (INT k = 0; k < nnz; k++) {
    const INT at = next[acoo->irows[k]]++;
    a->ja[at] = acoo->icols[k];
    a->val[at] = acoo->val[k];
  }
  free(next);
}

/* io.c:71-115: "nrows ncols nnz" then nnz lines "row col value".  When the
 * smallest row or column index is 1 the indices are taken as 1-based.
 * Returns 0, or 1 with a message on stderr. */
INT dcsr_mat_dcoo(FILE *fp, dcsr_mat *a_out) {
  dcoo_mat c;
  memset(&c, 0, sizeof(c));
  if (fscanf(fp, "%d %d %d", &c.row, &c.col, &c.nnz) != 3 || c.row < 0 || c.col < 0 || c.nnz < 0) {
    fprintf(stderr, "Error reading matrix dimensions\n");
    return 1;
  }
  fprintf(stdout, "%%%%Degrees of freedom: %d, Nonzeros: %d\n", c.row, c.nnz);
  fflush(stdout);
  const size_t cap = (size_t)(c.nnz > 0 ? c.nnz : 1);
  c.irows = (INT *)calloc(cap, sizeof(INT));
  c.icols = (INT *)calloc(cap, sizeof(INT));
  c.val = (REAL *)calloc(cap, sizeof(REAL));
  if (!c.irows || !c.icols || !c.val) {
    fprintf(stderr, "Error allocating COO arrays\n");
    dcoo_mat_free(&c);
    return 1;
  }
  INT lo = INT_MAX;
  for (INT k = 0; k < c.nnz; k++) {
    if (fscanf(fp, "%d %d %lf", &c.irows[k], &c.icols[k], &c.val[k]) != 3) {
      fprintf(stderr, "Error reading matrix entry %d\n", k);
      dcoo_mat_free(&c);
      return 1;
    }
    if (c.irows[k] < lo) lo = c.irows[k];
    if (c.icols[k] < lo) lo = c.icols[k];
  }
  if (lo == 1)
    for (INT k = 0; k < c.nnz; k++) c.irows[k]--, c.icols[k]--;
  for (INT k = 0; k < c.nnz; k++)
    if (c.irows[k] < 0 || c.irows[k] >= c.row || c.icols[k] < 0 || c.icols[k] >= c.col) {
      fprintf(stderr, "Error: matrix entry %d (%d, %d) outside %d x %d\n", k, c.irows[k], c.icols[k], c.row,
              c.col);
      dcoo_mat_free(&c);
      return 1;
    }
  dcoo2dcsr(&c, a_out);
  dcoo_mat_free(&c);
  return 0;
}
