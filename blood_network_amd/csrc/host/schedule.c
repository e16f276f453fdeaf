/*
 * schedule.c - host analysis that turns sequential sweeps into parallel
 * schedules for the device:
 *   - multicolouring of the C set and of the F set of a level (C-then-F
 *     multicolour Gauss-Seidel), or of all points (plain GS): two colours
 *     by breadth-first parity where the couplings allow, greedy otherwise;
 *   - level sets of a triangular factor (dependency depth per row) for
 *     level-scheduled triangular solves;
 *   - level sets of a sequential Gauss-Seidel sweep in a given visiting
 *     order (natural or C/F order), for sweeps that reproduce it exactly.
 */
#include <stdlib.h>
#include <string.h>

#include "schedule.h"
#include "csr.h"

/* Greedy colouring of the nodes with cls[i] == want (visited ascending):
 * each takes the smallest colour unused by already coloured neighbours
 * of the same class.  Returns the number of colours; col[i] set for the
 * class members. */
static int greedy_class(const dcsr_mat *A, const int *cls, int want, int *col) {
  const int n = A->row;
  int cap = 64, ncol = 0;
  int *stamp = (int *)bn_xmalloc((size_t)cap * sizeof(int));
  for (int c = 0; c < cap; c++) stamp[c] = -1;
  for (int i = 0; i < n; i++) {
    if (cls && cls[i] != want) continue;
    for (int k = A->ia[i]; k < A->ia[i + 1]; k++) {
      const int j = A->ja[k];
      if (j == i || (cls && cls[j] != want) || col[j] < 0) continue;
      stamp[col[j]] = i;
    }
    int c = 0;
    while (c < cap && stamp[c] == i) c++;
    if (c == cap) {
      int nc = cap * 2;
      stamp = (int *)realloc(stamp, (size_t)nc * sizeof(int));
      for (int q = cap; q < nc; q++) stamp[q] = -1;
      cap = nc;
    }
    col[i] = c;
    if (c + 1 > ncol) ncol = c + 1;
  }
  free(stamp);
  return ncol;
}

/* Two colours (one when no two members couple) by breadth-first parity
 * when the couplings among the nodes with cls[i] == want have no odd
 * cycle: components are searched from their lowest node, neighbours in
 * row order.  A lattice with trees hung on it is such a graph, and its C
 * and F sets are too, where the greedy pass spends extra colours on the
 * rows it meets late.  Returns the colour count with col[i] set for the
 * class members, or 0 (col untouched for them) when there is an odd
 * cycle. */
static int bipartite_class(const dcsr_mat *A, const int *cls, int want, int *col) {
  const int n = A->row;
  int *q = (int *)bn_xmalloc((size_t)(n > 0 ? n : 1) * sizeof(int));
  int *par = (int *)bn_xmalloc((size_t)(n > 0 ? n : 1) * sizeof(int));
  for (int i = 0; i < n; i++) par[i] = -1;
  int ok = 1, ncol = 0;
  for (int s = 0; s < n && ok; s++) {
    if ((cls && cls[s] != want) || par[s] >= 0) continue;
    int h = 0, t = 0;
    q[t++] = s;
    par[s] = 0;
    if (ncol < 1) ncol = 1;
    while (h < t && ok) {
      const int i = q[h++];
      for (int k = A->ia[i]; k < A->ia[i + 1]; k++) {
        const int j = A->ja[k];
        if (j == i || (cls && cls[j] != want)) continue;
        if (par[j] < 0) {
          par[j] = par[i] ^ 1;
          ncol = 2;
          q[t++] = j;
        } else if (par[j] == par[i]) {
          ok = 0;
          break;
        }
      }
    }
  }
  if (ok)
    for (int i = 0; i < n; i++)
      if (!cls || cls[i] == want) col[i] = par[i];
  free(q);
  free(par);
  return ok ? ncol : 0;
}

/* the class's colours: two by parity when its couplings allow, else the
 * greedy pass */
static int colour_class(const dcsr_mat *A, const int *cls, int want, int *col) {
  const int nc = bipartite_class(A, cls, want, col);
  return nc > 0 ? nc : greedy_class(A, cls, want, col);
}

int bn_colour_order(const dcsr_mat *A, const int *mis, int split, int *perm, int **col_ptr) {
  const int n = A->row;
  int *col = (int *)bn_xmalloc((size_t)(n > 0 ? n : 1) * sizeof(int));
  for (int i = 0; i < n; i++) col[i] = -1;
  int *cls = NULL;
  int nclass = 1;
  if (split && mis) {
    cls = (int *)bn_xmalloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    for (int i = 0; i < n; i++) cls[i] = mis[i] ? 0 : 1; /* C first */
    nclass = 2;
  }
  int ncol_cls[2] = {0, 0};
  /* the classes colour independently (a node looks only at neighbours of
   * its own class): the F class as a task beside the C class */
  if (nclass == 2) {
#pragma omp task shared(ncol_cls, col) firstprivate(A, cls) if (n > 65536)
    ncol_cls[1] = colour_class(A, cls, 1, col);
    ncol_cls[0] = colour_class(A, cls, 0, col);
#pragma omp taskwait
  } else {
    ncol_cls[0] = colour_class(A, cls, 0, col);
  }
  const int ncol = ncol_cls[0] + ncol_cls[1];
  /* global colour id: C colours first, then F colours */
  int *cnt = (int *)bn_xcalloc((size_t)ncol + 1, sizeof(int));
  for (int i = 0; i < n; i++) {
    int g = col[i] + ((cls && cls[i] == 1) ? ncol_cls[0] : 0);
    col[i] = g;
    cnt[g + 1]++;
  }
  for (int c = 0; c < ncol; c++) cnt[c + 1] += cnt[c];
  int *ptr = (int *)bn_xmalloc(((size_t)ncol + 1) * sizeof(int));
  memcpy(ptr, cnt, ((size_t)ncol + 1) * sizeof(int));
  for (int i = 0; i < n; i++) perm[cnt[col[i]]++] = i; /* ascending within a colour */
  free(cnt);
  free(col);
  free(cls);
  *col_ptr = ptr;
  return ncol;
}

int bn_chunk_order(const dcsr_mat *A, const int *mis, int split, int nchunks, int *perm, int **chunk_ptr,
                   int **ccol_ptr, int **cbnd) {
  const int n = A->row;
  if (nchunks < 1) nchunks = 1;
  const long long S = n > 0 ? ((long long)n + nchunks - 1) / nchunks : 1;
  const int nk = n > 0 ? (int)(((long long)n + S - 1) / S) : 0;
  int *cp = (int *)bn_xmalloc(((size_t)nk + 1) * sizeof(int));
  for (int k = 0; k <= nk; k++) cp[k] = (int)((long long)k * S < n ? (long long)k * S : n);
  int *col = (int *)bn_xmalloc((size_t)(n > 0 ? n : 1) * sizeof(int));
  int *nck = (int *)bn_xmalloc((size_t)(nk > 0 ? nk : 1) * sizeof(int));
  const int use_split = split && mis;
  for (int i = 0; i < n; i++) col[i] = -1;
#pragma omp parallel
  {
    int cap = 64;
    int *stamp = (int *)bn_xmalloc((size_t)cap * sizeof(int));
    for (int c = 0; c < cap; c++) stamp[c] = -1;
#pragma omp for schedule(dynamic, 16)
    for (int k = 0; k < nk; k++) {
      const int lo = cp[k], hi = cp[k + 1];
      int nclass[2] = {0, 0};
      for (int cl = 0; cl < (use_split ? 2 : 1); cl++) {
        int ncol = 0;
        for (int q = lo; q < hi; q++) {
          if (use_split && (mis[q] ? 0 : 1) != cl) continue;
          for (int e = A->ia[q]; e < A->ia[q + 1]; e++) {
            const int j = A->ja[e];
            if (j < lo || j >= hi || j == q || col[j] < 0) continue;
            if (use_split && (mis[j] ? 0 : 1) != cl) continue;
            stamp[col[j]] = q;
          }
          int c = 0;
          while (c < cap && stamp[c] == q) c++;
          if (c == cap) {
            const int nc2 = cap * 2;
            stamp = (int *)realloc(stamp, (size_t)nc2 * sizeof(int));
            for (int t = cap; t < nc2; t++) stamp[t] = -1;
            cap = nc2;
          }
          col[q] = c;
          if (c + 1 > ncol) ncol = c + 1;
        }
        nclass[cl] = ncol;
      }
      if (use_split)
        for (int q = lo; q < hi; q++)
          if (!mis[q]) col[q] += nclass[0];
      nck[k] = nclass[0] + nclass[1];
    }
    free(stamp);
  }
  int *cc = (int *)bn_xmalloc(((size_t)nk + 1) * sizeof(int));
  cc[0] = 0;
  for (int k = 0; k < nk; k++) cc[k + 1] = cc[k] + nck[k] + 1;
  int *bnd = (int *)bn_xmalloc((size_t)(cc[nk] > 0 ? cc[nk] : 1) * sizeof(int));
#pragma omp parallel for schedule(dynamic, 16)
  for (int k = 0; k < nk; k++) {
    const int lo = cp[k], hi = cp[k + 1], nc = nck[k];
    int *b = bnd + cc[k];
    for (int c = 0; c <= nc; c++) b[c] = 0;
    for (int q = lo; q < hi; q++) b[col[q] + 1]++;
    b[0] = lo;
    for (int c = 0; c < nc; c++) b[c + 1] += b[c];
    /* b[c] is now the first slot of colour c; fill, then restore */
    for (int q = lo; q < hi; q++) perm[b[col[q]]++] = q;
    for (int c = nc; c > 0; c--) b[c] = b[c - 1];
    b[0] = lo;
  }
  free(col);
  free(nck);
  *chunk_ptr = cp;
  *ccol_ptr = cc;
  *cbnd = bnd;
  return nk;
}

int bn_level_sets(const int *ia, const int *ja, int n, int lower, int *rows, int **lev_ptr) {
  int *depth = (int *)bn_xmalloc((size_t)(n > 0 ? n : 1) * sizeof(int));
  int nlev = 0;
  if (lower) {
    for (int i = 0; i < n; i++) {
      int d = 0;
      for (int k = ia[i]; k < ia[i + 1]; k++) {
        const int j = ja[k];
        if (j < i && depth[j] + 1 > d) d = depth[j] + 1;
      }
      depth[i] = d;
      if (d + 1 > nlev) nlev = d + 1;
    }
  } else {
    for (int i = n - 1; i >= 0; i--) {
      int d = 0;
      for (int k = ia[i]; k < ia[i + 1]; k++) {
        const int j = ja[k];
        if (j > i && depth[j] + 1 > d) d = depth[j] + 1;
      }
      depth[i] = d;
      if (d + 1 > nlev) nlev = d + 1;
    }
  }
  int *cnt = (int *)bn_xcalloc((size_t)nlev + 1, sizeof(int));
  for (int i = 0; i < n; i++) cnt[depth[i] + 1]++;
  for (int l = 0; l < nlev; l++) cnt[l + 1] += cnt[l];
  int *ptr = (int *)bn_xmalloc(((size_t)nlev + 1) * sizeof(int));
  memcpy(ptr, cnt, ((size_t)nlev + 1) * sizeof(int));
  for (int i = 0; i < n; i++) rows[cnt[depth[i]]++] = i;
  free(cnt);
  free(depth);
  *lev_ptr = ptr;
  return nlev;
}

int bn_sweep_levels(const dcsr_mat *A, const int *order, int backward, int *rows, int **lev_ptr) {
  const int n = A->row;
  *lev_ptr = NULL;
  /* pos[i] = when the sweep visits row i */
  int *pos = (int *)bn_xmalloc((size_t)(n > 0 ? n : 1) * sizeof(int));
  for (int i = 0; i < n; i++) pos[i] = -1;
  for (int k = 0; k < n; k++) {
    const int i = order ? order[k] : k;
    if (i < 0 || i >= n || pos[i] >= 0) {
      free(pos);
      return -1;
    }
    pos[i] = backward ? n - 1 - k : k;
  }
  /* couplings from either triangle: row i of A and of A^T */
  dcsr_mat At;
  bn_csr_transpose(A, &At);
  int *depth = (int *)bn_xmalloc((size_t)(n > 0 ? n : 1) * sizeof(int));
  int nlev = 0;
  for (int t = 0; t < n; t++) {
    const int k = backward ? n - 1 - t : t;
    const int i = order ? order[k] : k;
    const int pi = pos[i];
    int d = 0;
    for (int q = A->ia[i]; q < A->ia[i + 1]; q++) {
      const int j = A->ja[q];
      if (j != i && pos[j] < pi && depth[j] + 1 > d) d = depth[j] + 1;
    }
    for (int q = At.ia[i]; q < At.ia[i + 1]; q++) {
      const int j = At.ja[q];
      if (j != i && pos[j] < pi && depth[j] + 1 > d) d = depth[j] + 1;
    }
    depth[i] = d;
    if (d + 1 > nlev) nlev = d + 1;
  }
  bn_csr_free(&At);
  free(pos);
  int *cnt = (int *)bn_xcalloc((size_t)nlev + 1, sizeof(int));
  for (int i = 0; i < n; i++) cnt[depth[i] + 1]++;
  for (int l = 0; l < nlev; l++) cnt[l + 1] += cnt[l];
  int *ptr = (int *)bn_xmalloc(((size_t)nlev + 1) * sizeof(int));
  memcpy(ptr, cnt, ((size_t)nlev + 1) * sizeof(int));
  for (int i = 0; i < n; i++) rows[cnt[depth[i]]++] = i;
  free(cnt);
  free(depth);
  *lev_ptr = ptr;
  return nlev;
}

void bn_csr_offdiag_rows(const dcsr_mat *A, const int *perm_r, dcsr_mat *B) {
  const int n = A->row;
  int *cnt = (int *)bn_xmalloc((size_t)(n > 0 ? n : 1) * sizeof(int));
#pragma omp parallel for schedule(static) if (A->nnz > 65536)
  for (int k = 0; k < n; k++) {
    const int i = perm_r ? perm_r[k] : k;
    int c = 0;
    for (int q = A->ia[i]; q < A->ia[i + 1]; q++) c += (A->ja[q] != i);
    cnt[k] = c;
  }
  int *ptr = (int *)bn_xmalloc(((size_t)n + 1) * sizeof(int));
  bn_prefix(cnt, n, ptr);
  free(cnt);
  bn_csr_alloc(B, A->row, A->col, ptr[n]);
  memcpy(B->ia, ptr, ((size_t)n + 1) * sizeof(int));
  free(ptr);
#pragma omp parallel for schedule(static) if (A->nnz > 65536)
  for (int k = 0; k < n; k++) {
    const int i = perm_r ? perm_r[k] : k;
    int p = B->ia[k];
    for (int q = A->ia[i]; q < A->ia[i + 1]; q++)
      if (A->ja[q] != i) {
        B->ja[p] = A->ja[q];
        B->val[p] = A->val[q];
        p++;
      }
  }
}

void bn_csr_permute(const dcsr_mat *A, const int *perm_r, const int *inv_c, dcsr_mat *B) {
  const int n = A->row;
  bn_csr_alloc(B, A->row, A->col, A->nnz);
  int *cnt = (int *)bn_xmalloc((size_t)(n > 0 ? n : 1) * sizeof(int));
  for (int k = 0; k < n; k++) {
    const int i = perm_r ? perm_r[k] : k;
    cnt[k] = A->ia[i + 1] - A->ia[i];
  }
  bn_prefix(cnt, n, B->ia);
  free(cnt);
#pragma omp parallel for schedule(static)
  for (int k = 0; k < n; k++) {
    const int i = perm_r ? perm_r[k] : k;
    int p = B->ia[k];
    for (int q = A->ia[i]; q < A->ia[i + 1]; q++, p++) {
      B->ja[p] = inv_c ? inv_c[A->ja[q]] : A->ja[q];
      B->val[p] = A->val[q];
    }
  }
}

int bn_row_blocks(const int *ia, int n, int max_nnz, int max_rows, int *blocks) {
  int nb = 0, r = 0;
  blocks[0] = 0;
  while (r < n) {
    if (ia[r + 1] - ia[r] > max_nnz) {
      blocks[++nb] = ++r;
      continue;
    }
    const int r0 = r;
    long long acc = 0;
    while (r < n && r - r0 < max_rows && acc + (ia[r + 1] - ia[r]) <= max_nnz) {
      acc += ia[r + 1] - ia[r];
      r++;
    }
    blocks[++nb] = r;
  }
  return nb;
}
