/*
 * csr.c - host CSR helpers for the hierarchy set-up.
 *
 * Row-independent work is split over OpenMP threads; every row is produced
 * by the same sequence of operations as a sequential pass, so the results
 * do not depend on the thread count.
 */
#include "csr.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

void *bn_xmalloc(size_t bytes) {
  void *p = malloc(bytes ? bytes : 1);
  if (!p) {
    fprintf(stderr, "bnamg: out of host memory (%zu bytes)\n", bytes);
    abort();
  }
  return p;
}

void *bn_xcalloc(size_t count, size_t size) {
  void *p = calloc(count ? count : 1, size ? size : 1);
  if (!p) {
    fprintf(stderr, "bnamg: out of host memory (%zu x %zu bytes)\n", count, size);
    abort();
  }
  return p;
}

int bn_num_threads(void) {
#ifdef _OPENMP
  return omp_get_max_threads();
#else
  return 1;
#endif
}

void bn_csr_alloc(dcsr_mat *a, INT rows, INT cols, INT nnz) {
  a->row = rows;
  a->col = cols;
  a->nnz = nnz;
  a->ia = (INT *)bn_xcalloc((size_t)rows + 1, sizeof(INT));
  a->ja = (INT *)bn_xmalloc((size_t)nnz * sizeof(INT));
  a->val = (REAL *)bn_xmalloc((size_t)nnz * sizeof(REAL));
}

void bn_csr_free(dcsr_mat *a) {
  if (!a) return;
  free(a->ia);
  free(a->ja);
  free(a->val);
  a->ia = NULL;
  a->ja = NULL;
  a->val = NULL;
  a->row = a->col = a->nnz = 0;
}

void bn_csr_copy(const dcsr_mat *src, dcsr_mat *dst) {
  bn_csr_alloc(dst, src->row, src->col, src->nnz);
  memcpy(dst->ia, src->ia, ((size_t)src->row + 1) * sizeof(INT));
  memcpy(dst->ja, src->ja, (size_t)src->nnz * sizeof(INT));
  memcpy(dst->val, src->val, (size_t)src->nnz * sizeof(REAL));
}

void bn_prefix(const INT *counts, INT n, INT *ptr) {
  long long acc = 0;
  ptr[0] = 0;
  for (INT i = 0; i < n; i++) {
    acc += counts[i];
    ptr[i + 1] = (INT)acc;
  }
}

/* Stable transpose (column j of a lists its rows ascending, with their
 * values).  Large matrices split the rows into contiguous ranges, one per
 * thread: per-range column counts, then each column's entries laid out
 * range by range, so every thread scatters its rows in order and the
 * result is the sequential one. */
void bn_csr_transpose(const dcsr_mat *a, dcsr_mat *at) {
#ifdef _OPENMP
  int nt = omp_get_max_threads();
  if (nt > 8) nt = 8;
  if (a->nnz >= (1 << 20) && nt > 1) {
    const INT nc = a->col, nr = a->row;
    bn_csr_alloc(at, nc, nr, a->nnz);
    INT *cnt = (INT *)bn_xcalloc((size_t)nt * ((size_t)nc + 1), sizeof(INT));
#pragma omp parallel num_threads(nt)
    {
      const int t = omp_get_thread_num();
      const INT r0 = (INT)((long long)nr * t / nt), r1 = (INT)((long long)nr * (t + 1) / nt);
      INT *c = cnt + (size_t)t * ((size_t)nc + 1);
      for (INT k = a->ia[r0]; k < a->ia[r1]; k++) c[a->ja[k]]++;
#pragma omp barrier
#pragma omp single
      {
        INT run = 0;
        for (INT j = 0; j < nc; j++) {
          at->ia[j] = run;
          for (int u = 0; u < nt; u++) run += cnt[(size_t)u * ((size_t)nc + 1) + j];
        }
        at->ia[nc] = run;
      }
#pragma omp for schedule(static)
      for (INT j = 0; j < nc; j++) {
        INT pos = at->ia[j];
        for (int u = 0; u < nt; u++) {
          INT *cu = cnt + (size_t)u * ((size_t)nc + 1);
          const INT m = cu[j];
          cu[j] = pos;
          pos += m;
        }
      }
      for (INT i = r0; i < r1; i++)
        for (INT k = a->ia[i]; k < a->ia[i + 1]; k++) {
          const INT q = c[a->ja[k]]++;
          at->ja[q] = i;
          if (a->val && at->val) at->val[q] = a->val[k];
        }
    }
    free(cnt);
    return;
  }
#endif
  bn_csr_alloc(at, a->col, a->row, a->nnz);
  INT *cnt = (INT *)bn_xcalloc((size_t)a->col + 1, sizeof(INT));
  for (INT k = 0; k < a->nnz; k++) cnt[a->ja[k]]++;
  bn_prefix(cnt, a->col, at->ia);
  memcpy(cnt, at->ia, (size_t)a->col * sizeof(INT));
  for (INT i = 0; i < a->row; i++) {
    for (INT k = a->ia[i]; k < a->ia[i + 1]; k++) {
      INT q = cnt[a->ja[k]]++;
      at->ja[q] = i;
      at->val[q] = a->val[k];
    }
  }
  free(cnt);
}

/* Union of two rows with A's order first (C_amg1/include/agg_lib.h:186-243
 * fixes this layout).  Two passes (count, fill), rows split over threads,
 * each thread stamping its own column marker with the row index. */
void bn_csr_add(const dcsr_mat *a, const dcsr_mat *b, dcsr_mat *c) {
  const INT n = a->row, m = a->col;
  INT *cnt = (INT *)bn_xmalloc((size_t)(n > 0 ? n : 1) * sizeof(INT));
  const int nt = bn_num_threads();
  INT *marks = (INT *)bn_xmalloc((size_t)nt * (size_t)(m > 0 ? m : 1) * sizeof(INT));
  INT *slots = (INT *)bn_xmalloc((size_t)nt * (size_t)(m > 0 ? m : 1) * sizeof(INT));
#pragma omp parallel
  {
#ifdef _OPENMP
    const int t = omp_get_thread_num();
#else
    const int t = 0;
#endif
    INT *mark = marks + (size_t)t * (size_t)(m > 0 ? m : 1);
    for (INT j = 0; j < m; j++) mark[j] = -1;
#pragma omp for schedule(static)
    for (INT i = 0; i < n; i++) {
      INT k = 0;
      for (INT q = a->ia[i]; q < a->ia[i + 1]; q++) {
        mark[a->ja[q]] = i;
        k++;
      }
      for (INT q = b->ia[i]; q < b->ia[i + 1]; q++)
        if (mark[b->ja[q]] != i) {
          mark[b->ja[q]] = i;
          k++;
        }
      cnt[i] = k;
    }
  }
  c->row = n;
  c->col = m;
  c->ia = (INT *)bn_xmalloc(((size_t)n + 1) * sizeof(INT));
  bn_prefix(cnt, n, c->ia);
  c->nnz = c->ia[n];
  c->ja = (INT *)bn_xmalloc((size_t)c->nnz * sizeof(INT));
  c->val = (REAL *)bn_xmalloc((size_t)c->nnz * sizeof(REAL));
#pragma omp parallel
  {
#ifdef _OPENMP
    const int t = omp_get_thread_num();
#else
    const int t = 0;
#endif
    INT *mark = marks + (size_t)t * (size_t)(m > 0 ? m : 1);
    INT *slot = slots + (size_t)t * (size_t)(m > 0 ? m : 1);
    for (INT j = 0; j < m; j++) mark[j] = -1;
#pragma omp for schedule(static)
    for (INT i = 0; i < n; i++) {
      INT p = c->ia[i];
      for (INT q = a->ia[i]; q < a->ia[i + 1]; q++) {
        INT col = a->ja[q];
        mark[col] = i;
        slot[col] = p;
        c->ja[p] = col;
        c->val[p] = a->val[q];
        p++;
      }
      for (INT q = b->ia[i]; q < b->ia[i + 1]; q++) {
        INT col = b->ja[q];
        if (mark[col] == i) {
          c->val[slot[col]] += b->val[q];
        } else {
          mark[col] = i;
          slot[col] = p;
          c->ja[p] = col;
          c->val[p] = b->val[q];
          p++;
        }
      }
    }
  }
  free(marks);
  free(slots);
  free(cnt);
}

void bn_csr_symmetrize(const dcsr_mat *a, dcsr_mat *out) {
  dcsr_mat at;
  bn_csr_transpose(a, &at);
  bn_csr_add(a, &at, out);
  bn_csr_free(&at);
#pragma omp parallel for schedule(static)
  for (INT k = 0; k < out->nnz; k++) out->val[k] *= 0.5;
}

void bn_csr_l1(const dcsr_mat *a, REAL *l1) {
#pragma omp parallel for schedule(static)
  for (INT i = 0; i < a->row; i++) {
    REAL s = 0.0;
    for (INT k = a->ia[i]; k < a->ia[i + 1]; k++) s += fabs(a->val[k]);
    l1[i] = s;
  }
}
