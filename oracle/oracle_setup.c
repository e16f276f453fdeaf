/*
 * oracle_setup.c - CPU restatement of the C_amg1 hierarchy setup.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Each function
 * follows the reference routine cited beside it statement by statement in
 * its arithmetic and its ordering, so integer outputs (C/F split, sparsity
 * patterns, row orders) and FP64 values are bit-identical to the reference
 * when both are compiled by gcc without FMA contraction.
 */
#include "oracle.h"

void oracle_csr_free(dcsr_mat *a) {
  free(a->ia); free(a->ja); free(a->val);
  a->ia = NULL; a->ja = NULL; a->val = NULL;
  a->row = a->col = a->nnz = 0;
}

static void csr_alloc(dcsr_mat *a, INT r, INT c, INT nnz) {
  a->row = r; a->col = c; a->nnz = nnz;
  a->ia = (INT *)calloc((size_t)r + 1, sizeof(INT));
  a->ja = (INT *)calloc(nnz > 0 ? (size_t)nnz : 1, sizeof(INT));
  a->val = (REAL *)calloc(nnz > 0 ? (size_t)nnz : 1, sizeof(REAL));
}

/* Transpose by counting sort on columns; rows of A^T list source rows in
 * ascending order.  Follows C_amg1/include/agg_lib.h:120-183 (p == NULL). */
void oracle_transpose(const dcsr_mat *a, dcsr_mat *at) {
  csr_alloc(at, a->col, a->row, a->nnz);
  for (INT k = 0; k < a->nnz; k++) at->ia[a->ja[k] + 1]++;
  for (INT i = 0; i < a->col; i++) at->ia[i + 1] += at->ia[i];
  INT *next = (INT *)malloc(((size_t)a->col + 1) * sizeof(INT));
  memcpy(next, at->ia, ((size_t)a->col + 1) * sizeof(INT));
  for (INT i = 0; i < a->row; i++)
    for (INT k = a->ia[i]; k < a->ia[i + 1]; k++) {
      INT q = next[a->ja[k]]++;
      at->ja[q] = i;
      at->val[q] = a->val[k];
    }
  free(next);
}

/* C = A + B: row i lists A's columns in A's order, then B's new columns in
 * B's order; shared entries are a + b.  Follows agg_lib.h:186-243. */
void oracle_add(const dcsr_mat *a, const dcsr_mat *b, dcsr_mat *c) {
  INT n = a->row, m = a->col;
  INT *where = (INT *)malloc((size_t)(m > 0 ? m : 1) * sizeof(INT));
  for (INT j = 0; j < m; j++) where[j] = -1;
  INT cap = a->nnz + b->nnz;
  csr_alloc(c, n, m, cap);
  INT pos = 0;
  for (INT i = 0; i < n; i++) {
    INT start = pos;
    c->ia[i] = start;
    for (INT k = a->ia[i]; k < a->ia[i + 1]; k++) {
      where[a->ja[k]] = pos;
      c->ja[pos] = a->ja[k];
      c->val[pos] = a->val[k];
      pos++;
    }
    for (INT k = b->ia[i]; k < b->ia[i + 1]; k++) {
      INT col = b->ja[k];
      if (where[col] >= start) {
        c->val[where[col]] += b->val[k];
      } else {
        where[col] = pos;
        c->ja[pos] = col;
        c->val[pos] = b->val[k];
        pos++;
      }
    }
  }
  c->ia[n] = pos;
  c->nnz = pos;
  free(where);
}

/* A_sym = (A + A^T) / 2 as in C_amg1/test_amg.c:52-63 and
 * C_amg1/src/amg_wrapper.c:65-76, C_amg1/src/build_hierarchy.c:118-130. */
void oracle_symmetrize(const dcsr_mat *a, dcsr_mat *out) {
  dcsr_mat at;
  oracle_transpose(a, &at);
  oracle_add(a, &at, out);
  for (INT k = 0; k < out->nnz; k++) out->val[k] *= 0.5;
  oracle_csr_free(&at);
}

/* Strength of connection: j strong for i iff j != i and
 * -a_ij >= theta * max(0, max_{k!=i} -a_ik).  max_lam = longest row of A
 * (diagonal included), at least 1.  C_amg1/src/rs_classical.c:30-88. */
void rs_strength(const dcsr_mat *A, REAL theta, dcsr_mat *S, INT *max_lam) {
  INT n = A->row, longest = 0;
  REAL *thr = (REAL *)malloc((size_t)(n > 0 ? n : 1) * sizeof(REAL));
  for (INT i = 0; i < n; i++) {
    INT len = A->ia[i + 1] - A->ia[i];
    if (len > longest) longest = len;
    REAL mx = 0.0;
    for (INT k = A->ia[i]; k < A->ia[i + 1]; k++)
      if (A->ja[k] != i && -(A->val[k]) > mx) mx = -(A->val[k]);
    thr[i] = theta * mx;
  }
  *max_lam = longest < 1 ? 1 : longest;
  INT cnt = 0;
  for (INT i = 0; i < n; i++)
    for (INT k = A->ia[i]; k < A->ia[i + 1]; k++)
      if (A->ja[k] != i && -(A->val[k]) >= thr[i]) cnt++;
  csr_alloc(S, n, n, cnt);
  INT p = 0;
  for (INT i = 0; i < n; i++) {
    for (INT k = A->ia[i]; k < A->ia[i + 1]; k++)
      if (A->ja[k] != i && -(A->val[k]) >= thr[i]) {
        S->ja[p] = A->ja[k];
        S->val[p] = A->val[k];
        p++;
      }
    S->ia[i + 1] = p;
  }
  free(thr);
}

/* ---- doubly linked bucket lists keyed by lambda (head = most recent) ---- */
typedef struct { INT *head, *next, *prev; } buckets;

static void bk_unlink(buckets *b, INT key, INT i) {
  if (b->prev[i] >= 0) b->next[b->prev[i]] = b->next[i];
  else b->head[key] = b->next[i];
  if (b->next[i] >= 0) b->prev[b->next[i]] = b->prev[i];
}

static void bk_push(buckets *b, INT key, INT i) {
  b->next[i] = b->head[key];
  b->prev[i] = -1;
  if (b->head[key] >= 0) b->prev[b->head[key]] = i;
  b->head[key] = i;
}

/* Greedy first pass shared by RS (rs_classical.c:146-205) and HMIS phase 1
 * (rs_classical.c:435-486): take the head of the highest non-empty bucket
 * as C, make its undecided influencees F, and raise lambda of the
 * undecided strong dependencies of each new F (moving them to the head of
 * their new bucket).  `limit` >= 0 stops after that many decisions (HMIS);
 * the RS pass runs until every bucket is empty. */
static void greedy_pass(const dcsr_mat *S, const dcsr_mat *ST, INT *lambda,
                        INT nbuckets, INT top, INT *cf, INT limit) {
  INT n = S->row;
  buckets b;
  b.head = (INT *)malloc((size_t)nbuckets * sizeof(INT));
  b.next = (INT *)malloc((size_t)(n > 0 ? n : 1) * sizeof(INT));
  b.prev = (INT *)malloc((size_t)(n > 0 ? n : 1) * sizeof(INT));
  for (INT k = 0; k < nbuckets; k++) b.head[k] = -1;
  for (INT i = n - 1; i >= 0; i--) bk_push(&b, lambda[i], i);
  INT decided = 0;
  while (limit < 0 || decided < limit) {
    while (top >= 0 && b.head[top] < 0) top--;
    if (top < 0) break;
    INT c = b.head[top];
    bk_unlink(&b, lambda[c], c);
    cf[c] = 1;
    decided++;
    for (INT q = ST->ia[c]; q < ST->ia[c + 1]; q++) {
      INT j = ST->ja[q];
      if (cf[j] != 0) continue;
      bk_unlink(&b, lambda[j], j);
      cf[j] = -1;
      decided++;
      for (INT r = S->ia[j]; r < S->ia[j + 1]; r++) {
        INT k = S->ja[r];
        if (cf[k] != 0) continue;
        bk_unlink(&b, lambda[k], k);
        lambda[k]++;
        bk_push(&b, lambda[k], k);
        if (lambda[k] > top) top = lambda[k];
      }
    }
  }
  free(b.head); free(b.next); free(b.prev);
}

/* Promote every F point with no strong C neighbour; count C points.
 * rs_classical.c:211-225 (same block at :356-370 and :596-610). */
static void second_pass(const dcsr_mat *S, INT *cf, INT *n_coarse) {
  INT n = S->row, nc = 0;
  for (INT i = 0; i < n; i++) {
    if (cf[i] != -1) continue;
    INT found = 0;
    for (INT k = S->ia[i]; k < S->ia[i + 1] && !found; k++)
      found = (cf[S->ja[k]] == 1);
    if (!found) cf[i] = 1;
  }
  for (INT i = 0; i < n; i++) nc += (cf[i] == 1);
  *n_coarse = nc;
}

/* Classical RS split.  lambda[0] is left at 0 (the initialisation loop of
 * rs_classical.c:119-123 starts at i = 1); 2*max_lam+1 buckets, search
 * starts at max_lam (rs_classical.c:126,146). */
void rs_coarsening(const dcsr_mat *S, INT max_lam, INT *cf, INT *n_coarse) {
  INT n = S->row;
  dcsr_mat ST;
  oracle_transpose(S, &ST);
  INT *lambda = (INT *)calloc((size_t)(n > 0 ? n : 1), sizeof(INT));
  for (INT i = 0; i < n; i++) cf[i] = 0;
  for (INT i = 1; i < n; i++) lambda[i] = ST.ia[i + 1] - ST.ia[i];
  INT nb = 2 * max_lam + 1;
  /* the reference sizes its buckets at 2*max_lam+1; extra room here only
   * guards against inputs where lambda would outgrow that */
  INT cap = nb + 2 * S->nnz + 2;
  greedy_pass(S, &ST, lambda, cap, max_lam, cf, -1);
  free(lambda);
  oracle_csr_free(&ST);
  second_pass(S, cf, n_coarse);
}

/* G = S u S^T without self loops: S's row order, then S^T's new entries
 * (rs_classical.c:256-290, :500-532). */
static void sym_graph(const dcsr_mat *S, const dcsr_mat *ST, INT **gia, INT **gja) {
  INT n = S->row;
  INT *mark = (INT *)malloc((size_t)(n > 0 ? n : 1) * sizeof(INT));
  INT *ia = (INT *)calloc((size_t)n + 1, sizeof(INT));
  for (INT i = 0; i < n; i++) mark[i] = -1;
  for (int pass = 0; pass < 2; pass++) {
    INT *ja = pass ? (INT *)malloc((size_t)(ia[n] > 0 ? ia[n] : 1) * sizeof(INT)) : NULL;
    for (INT i = 0; i < n; i++) mark[i] = -1;
    for (INT i = 0; i < n; i++) {
      INT p = pass ? ia[i] : 0;
      const dcsr_mat *src[2] = {S, ST};
      for (int s = 0; s < 2; s++)
        for (INT k = src[s]->ia[i]; k < src[s]->ia[i + 1]; k++) {
          INT j = src[s]->ja[k];
          if (j == i || mark[j] == i) continue;
          mark[j] = i;
          if (pass) ja[p] = j;
          p++;
        }
      if (!pass) ia[i + 1] = p;
    }
    if (!pass) for (INT i = 0; i < n; i++) ia[i + 1] += ia[i];
    else *gja = ja;
  }
  *gia = ia;
  free(mark);
}

/* Jacobi-style independent-set rounds on G (rs_classical.c:305-349,
 * :549-586): an undecided node becomes C unless an undecided neighbour has
 * a strictly larger weight; then undecided nodes next to a C become F.
 * A round that selects nothing turns every undecided node into C. */
static void mis_rounds(INT n, const INT *gia, const INT *gja, const REAL *w,
                       INT *cf, INT undecided) {
  INT *sel = (INT *)malloc((size_t)(n > 0 ? n : 1) * sizeof(INT));
  while (undecided > 0) {
    INT ns = 0;
    for (INT i = 0; i < n; i++) {
      if (cf[i] != 0) continue;
      INT best = 1;
      for (INT k = gia[i]; k < gia[i + 1] && best; k++)
        if (cf[gja[k]] == 0 && w[gja[k]] > w[i]) best = 0;
      if (best) sel[ns++] = i;
    }
    if (ns == 0) {
      for (INT i = 0; i < n; i++)
        if (cf[i] == 0) { cf[i] = 1; undecided--; }
      break;
    }
    for (INT k = 0; k < ns; k++) { cf[sel[k]] = 1; undecided--; }
    for (INT i = 0; i < n; i++) {
      if (cf[i] != 0) continue;
      for (INT k = gia[i]; k < gia[i + 1]; k++)
        if (cf[gja[k]] == 1) { cf[i] = -1; undecided--; break; }
    }
  }
  free(sel);
}

/* PMIS: weights lambda_i + rand()/RAND_MAX after srand(12345), drawn in
 * index order (rs_classical.c:246-371). */
void pmis_coarsening(const dcsr_mat *S, INT *cf, INT *n_coarse) {
  INT n = S->row;
  dcsr_mat ST;
  oracle_transpose(S, &ST);
  INT *gia, *gja;
  sym_graph(S, &ST, &gia, &gja);
  REAL *w = (REAL *)malloc((size_t)(n > 0 ? n : 1) * sizeof(REAL));
  srand(12345);
  for (INT i = 0; i < n; i++) {
    w[i] = (REAL)(ST.ia[i + 1] - ST.ia[i]) + (REAL)rand() / (REAL)RAND_MAX;
    cf[i] = 0;
  }
  oracle_csr_free(&ST);
  mis_rounds(n, gia, gja, w, cf, n);
  free(w); free(gia); free(gja);
  second_pass(S, cf, n_coarse);
}

/* HMIS: the RS greedy pass with every lambda initialised (no lambda[0]
 * quirk), max_lam + longest S row + 1 buckets, search from the largest
 * lambda, stopping after n decisions; then PMIS rounds on whatever is
 * still undecided, weights drawn only for those nodes
 * (rs_classical.c:390-611). */
void hmis_coarsening(const dcsr_mat *S, INT *cf, INT *n_coarse) {
  INT n = S->row;
  dcsr_mat ST;
  oracle_transpose(S, &ST);
  INT *lambda = (INT *)calloc((size_t)(n > 0 ? n : 1), sizeof(INT));
  INT max_lam = 0, max_row = 0;
  for (INT i = 0; i < n; i++) {
    lambda[i] = ST.ia[i + 1] - ST.ia[i];
    cf[i] = 0;
    if (lambda[i] > max_lam) max_lam = lambda[i];
    if (S->ia[i + 1] - S->ia[i] > max_row) max_row = S->ia[i + 1] - S->ia[i];
  }
  INT cap = max_lam + max_row + 1 + 2 * S->nnz + 2;
  greedy_pass(S, &ST, lambda, cap, max_lam, cf, n);
  free(lambda);
  INT undecided = 0;
  for (INT i = 0; i < n; i++) undecided += (cf[i] == 0);
  if (undecided > 0) {
    INT *gia, *gja;
    sym_graph(S, &ST, &gia, &gja);
    REAL *w = (REAL *)malloc((size_t)n * sizeof(REAL));
    srand(12345);
    for (INT i = 0; i < n; i++)
      w[i] = (cf[i] == 0)
                 ? (REAL)(ST.ia[i + 1] - ST.ia[i]) + (REAL)rand() / (REAL)RAND_MAX
                 : -1.0;
    mis_rounds(n, gia, gja, w, cf, undecided);
    free(w); free(gia); free(gja);
  }
  oracle_csr_free(&ST);
  second_pass(S, cf, n_coarse);
}

/* P pattern: C rows inject their coarse index; F rows list their strong C
 * neighbours in S-row order (rs_classical.c:643-694). */
void rs_standard_interpolation_sparsity(const dcsr_mat *S, const INT *cf, dcsr_mat *P) {
  INT n = S->row, nc = 0;
  INT *cidx = (INT *)malloc((size_t)(n > 0 ? n : 1) * sizeof(INT));
  for (INT i = 0; i < n; i++) cidx[i] = (cf[i] == 1) ? nc++ : -1;
  INT cnt = 0;
  for (INT i = 0; i < n; i++) {
    if (cf[i] == 1) { cnt++; continue; }
    for (INT k = S->ia[i]; k < S->ia[i + 1]; k++) cnt += (cf[S->ja[k]] == 1);
  }
  csr_alloc(P, n, nc, cnt);
  INT p = 0;
  for (INT i = 0; i < n; i++) {
    if (cf[i] == 1) {
      P->ja[p++] = cidx[i];
    } else {
      for (INT k = S->ia[i]; k < S->ia[i + 1]; k++)
        if (cf[S->ja[k]] == 1) P->ja[p++] = cidx[S->ja[k]];
    }
    P->ia[i + 1] = p;
  }
  free(cidx);
}

/* Direct interpolation weights (rs_classical.c:703-805):
 *   w_ij = -(a_ij + sum_{k in D_i^s} a_ik a_kj / sum_{l in C_i} a_kl) / a^_ii,
 *   a^_ii = a_ii + sum_weak a_im (+ a_ik of strong F neighbours with
 *   |denominator| < 1e-15); no scaling when |a^_ii| <= 1e-15. */
void rs_standard_interpolation_values(const dcsr_mat *A, const dcsr_mat *S,
                                      const INT *cf, dcsr_mat *P) {
  INT n = A->row;
  INT *strong_of = (INT *)malloc((size_t)(n > 0 ? n : 1) * sizeof(INT));
  INT *slot = (INT *)malloc((size_t)(n > 0 ? n : 1) * sizeof(INT));
  for (INT i = 0; i < n; i++) { strong_of[i] = -1; slot[i] = -1; }
  for (INT i = 0; i < n; i++) {
    if (cf[i] == 1) { P->val[P->ia[i]] = 1.0; continue; }
    for (INT k = S->ia[i]; k < S->ia[i + 1]; k++) strong_of[S->ja[k]] = i;
    INT p = P->ia[i];
    for (INT k = S->ia[i]; k < S->ia[i + 1]; k++) {
      INT j = S->ja[k];
      if (cf[j] == 1) { P->val[p] = 0.0; slot[j] = p; p++; }
    }
    REAL diag = 0.0, weak = 0.0;
    for (INT k = A->ia[i]; k < A->ia[i + 1]; k++) {
      INT j = A->ja[k];
      REAL a = A->val[k];
      if (j == i) diag = a;
      else if (strong_of[j] == i && cf[j] == 1) P->val[slot[j]] += a;
      else if (strong_of[j] == i && cf[j] == -1) { /* strong F: below */ }
      else weak += a;
    }
    REAL dhat = diag + weak;
    for (INT k = A->ia[i]; k < A->ia[i + 1]; k++) {
      INT f = A->ja[k];
      if (f == i || strong_of[f] != i || cf[f] != -1) continue;
      REAL aik = A->val[k], den = 0.0;
      for (INT q = A->ia[f]; q < A->ia[f + 1]; q++)
        if (cf[A->ja[q]] == 1 && strong_of[A->ja[q]] == i) den += A->val[q];
      if (fabs(den) < 1e-15) {
        dhat += aik;
      } else {
        for (INT q = A->ia[f]; q < A->ia[f + 1]; q++)
          if (cf[A->ja[q]] == 1 && strong_of[A->ja[q]] == i)
            P->val[slot[A->ja[q]]] += aik * A->val[q] / den;
      }
    }
    if (fabs(dhat) > 1e-15)
      for (INT q = P->ia[i]; q < P->ia[i + 1]; q++) P->val[q] /= -dhat;
    for (INT k = S->ia[i]; k < S->ia[i + 1]; k++)
      if (cf[S->ja[k]] == 1) slot[S->ja[k]] = -1;
  }
  free(strong_of); free(slot);
}

void rs_standard_interpolation(const dcsr_mat *A, const dcsr_mat *S,
                               const INT *cf, dcsr_mat *P) {
  rs_standard_interpolation_sparsity(S, cf, P);
  rs_standard_interpolation_values(A, S, cf, P);
}

/* Galerkin R*A*P with the SMMP row layout of C_amg1/src/sparse_ops.c:102-213:
 * each coarse row starts with its diagonal (value 0.0 + contributions),
 * then columns in order of first discovery; a fine row i2 reached twice in
 * one coarse row adds to known slots without re-discovering. */
void dcsr_rap(dcsr_mat *R, dcsr_mat *A, dcsr_mat *P, dcsr_mat *RAP) {
  INT nc = R->row, nf = A->row;
  INT *col_slot = (INT *)malloc((size_t)(nc > 0 ? nc : 1) * sizeof(INT));
  INT *row_seen = (INT *)malloc((size_t)(nf > 0 ? nf : 1) * sizeof(INT));
  INT *ia = (INT *)calloc((size_t)nc + 1, sizeof(INT));
  INT *ja = NULL;
  REAL *va = NULL;
  for (int pass = 0; pass < 2; pass++) {
    for (INT i = 0; i < nc; i++) col_slot[i] = -1;
    for (INT i = 0; i < nf; i++) row_seen[i] = -1;
    INT cnt = 0;
    for (INT ic = 0; ic < nc; ic++) {
      INT beg = cnt;
      col_slot[ic] = cnt;
      if (pass) { ja[cnt] = ic; va[cnt] = 0.0; }
      cnt++;
      for (INT q1 = R->ia[ic]; q1 < R->ia[ic + 1]; q1++) {
        INT i1 = R->ja[q1];
        REAL r = pass ? R->val[q1] : 0.0;
        for (INT q2 = A->ia[i1]; q2 < A->ia[i1 + 1]; q2++) {
          INT i2 = A->ja[q2];
          REAL ra = pass ? r * A->val[q2] : 0.0;
          int fresh = (row_seen[i2] != ic);
          row_seen[i2] = ic;
          for (INT q3 = P->ia[i2]; q3 < P->ia[i2 + 1]; q3++) {
            INT i3 = P->ja[q3];
            if (fresh && col_slot[i3] < beg) {
              col_slot[i3] = cnt;
              if (pass) { ja[cnt] = i3; va[cnt] = ra * P->val[q3]; }
              cnt++;
            } else if (pass) {
              va[col_slot[i3]] += ra * P->val[q3];
            }
          }
        }
      }
      if (!pass) ia[ic] = beg;
    }
    if (!pass) {
      ia[nc] = cnt;
      ja = (INT *)calloc(cnt > 0 ? (size_t)cnt : 1, sizeof(INT));
      va = (REAL *)calloc(cnt > 0 ? (size_t)cnt : 1, sizeof(REAL));
    }
  }
  RAP->row = nc; RAP->col = nc; RAP->nnz = ia[nc];
  RAP->ia = ia; RAP->ja = ja; RAP->val = va;
  free(col_slot); free(row_seen);
}

/* IC(0) on the lower triangle of A in its stored row order
 * (C_amg1/src/ichol.c:12-118).  The inner sums stop at the first column
 * >= the pivot column, so rows that start with their diagonal (every
 * Galerkin level) give L(i,k) = a_ik / sqrt(a_kk), L(i,i) = sqrt(a_ii).
 * Non-positive pivots are replaced by 1e-14 before the square root. */
void ichol_compute(const dcsr_mat *A, ichol_factor *L) {
  INT n = A->row, cnt = 0;
  for (INT i = 0; i < n; i++)
    for (INT k = A->ia[i]; k < A->ia[i + 1]; k++) cnt += (A->ja[k] <= i);
  L->n = n; L->nnz = cnt;
  L->ia = (INT *)calloc((size_t)n + 1, sizeof(INT));
  L->ja = (INT *)calloc(cnt > 0 ? (size_t)cnt : 1, sizeof(INT));
  L->val = (REAL *)calloc(cnt > 0 ? (size_t)cnt : 1, sizeof(REAL));
  INT p = 0;
  for (INT i = 0; i < n; i++) {
    L->ia[i] = p;
    for (INT k = A->ia[i]; k < A->ia[i + 1]; k++)
      if (A->ja[k] <= i) { L->ja[p] = A->ja[k]; L->val[p] = A->val[k]; p++; }
  }
  L->ia[n] = p;
  REAL *w = (REAL *)calloc((size_t)(n > 0 ? n : 1), sizeof(REAL));
  INT *wrow = (INT *)malloc((size_t)(n > 0 ? n : 1) * sizeof(INT));
  INT *dpos = (INT *)malloc((size_t)(n > 0 ? n : 1) * sizeof(INT));
  for (INT i = 0; i < n; i++) {
    wrow[i] = -1;
    dpos[i] = -1;
    for (INT k = L->ia[i]; k < L->ia[i + 1]; k++)
      if (L->ja[k] == i) { dpos[i] = k; break; }
  }
  for (INT i = 0; i < n; i++) {
    for (INT k = L->ia[i]; k < L->ia[i + 1]; k++) {
      w[L->ja[k]] = L->val[k];
      wrow[L->ja[k]] = i;
    }
    for (INT q = L->ia[i]; q < L->ia[i + 1]; q++) {
      INT k = L->ja[q];
      if (k >= i) continue;
      REAL s = w[k];
      for (INT t = L->ia[k]; t < L->ia[k + 1]; t++) {
        INT j = L->ja[t];
        if (j >= k) break;
        if (wrow[j] == i) s -= w[j] * L->val[t];
      }
      w[k] = s / L->val[dpos[k]];
    }
    REAL d = w[i];
    for (INT q = L->ia[i]; q < L->ia[i + 1]; q++) {
      INT j = L->ja[q];
      if (j >= i) break;
      d -= w[j] * w[j];
    }
    if (d <= 0.0) d = 1e-14;
    w[i] = sqrt(d);
    for (INT q = L->ia[i]; q < L->ia[i + 1]; q++) L->val[q] = w[L->ja[q]];
    for (INT q = L->ia[i]; q < L->ia[i + 1]; q++) {
      w[L->ja[q]] = 0.0;
      wrow[L->ja[q]] = -1;
    }
  }
  free(w); free(wrow); free(dpos);
}

void ichol_free(ichol_factor *L) {
  free(L->ia); free(L->ja); free(L->val);
  L->ia = NULL; L->ja = NULL; L->val = NULL;
  L->n = 0; L->nnz = 0;
}

/* l1_i = sum_j |a_ij| (C_amg1/src/build_hierarchy.c:13-23) */
static REAL *l1_rows(const dcsr_mat *A) {
  REAL *d = (REAL *)malloc((size_t)(A->row > 0 ? A->row : 1) * sizeof(REAL));
  for (INT i = 0; i < A->row; i++) {
    REAL s = 0.0;
    for (INT k = A->ia[i]; k < A->ia[i + 1]; k++) s += fabs(A->val[k]);
    d[i] = s;
  }
  return d;
}

/* Galerkin step shared by build and rebuild: A_c = sym(P^T A P)
 * (build_hierarchy.c:108-130, :202-224). */
static void galerkin(const amg_level *cur, dcsr_mat *Ac) {
  dcsr_mat Pt, rap;
  oracle_transpose(&cur->P, &Pt);
  dcsr_rap(&Pt, (dcsr_mat *)&cur->A, (dcsr_mat *)&cur->P, &rap);
  oracle_csr_free(&Pt);
  oracle_symmetrize(&rap, Ac);
  oracle_csr_free(&rap);
}

/* Level loop of C_amg1/src/build_hierarchy.c:29-156: coarsen while the
 * current level has >= min_size rows and fewer than max_levels (<= 20)
 * levels exist; stop when the split yields 0 or n C points; cf_order lists
 * C points ascending then F points ascending; IC(0) of the coarsest A. */
void amg_build_hierarchy(amg_data *amg, const dcsr_mat *A) {
  INT max_levels = amg->max_levels > AMG_MAX_LEVELS ? AMG_MAX_LEVELS : amg->max_levels;
  amg_level *l0 = &amg->levels[0];
  memset(l0, 0, sizeof(*l0));
  l0->A.row = A->row; l0->A.col = A->col; l0->A.nnz = A->nnz;
  l0->A.ia = (INT *)malloc(((size_t)A->row + 1) * sizeof(INT));
  l0->A.ja = (INT *)malloc((size_t)(A->nnz > 0 ? A->nnz : 1) * sizeof(INT));
  l0->A.val = (REAL *)malloc((size_t)(A->nnz > 0 ? A->nnz : 1) * sizeof(REAL));
  memcpy(l0->A.ia, A->ia, ((size_t)A->row + 1) * sizeof(INT));
  memcpy(l0->A.ja, A->ja, (size_t)A->nnz * sizeof(INT));
  memcpy(l0->A.val, A->val, (size_t)A->nnz * sizeof(REAL));
  l0->n = A->row;
  l0->l1_diag = l1_rows(&l0->A);
  INT nlev = 1;
  while (amg->levels[nlev - 1].A.row >= amg->min_size && nlev < max_levels) {
    amg_level *cur = &amg->levels[nlev - 1];
    INT nf = cur->A.row, ml = 0, nc = 0;
    rs_strength(&cur->A, amg->threshold, &cur->A_filtered, &ml);
    INT *cf = (INT *)calloc((size_t)(nf > 0 ? nf : 1), sizeof(INT));
    if (amg->coarsening == COARSEN_PMIS) pmis_coarsening(&cur->A_filtered, cf, &nc);
    else if (amg->coarsening == COARSEN_HMIS) hmis_coarsening(&cur->A_filtered, cf, &nc);
    else rs_coarsening(&cur->A_filtered, ml, cf, &nc);
    cur->n_coarse = nc;
    if (nc == 0 || nc >= nf) {
      /* the reference keeps the strength matrix allocated here */
      free(cf);
      break;
    }
    cur->mis = (INT *)calloc((size_t)nf, sizeof(INT));
    cur->isolated = (INT *)calloc((size_t)nf, sizeof(INT));
    cur->cf_order = (INT *)malloc((size_t)nf * sizeof(INT));
    INT q = 0;
    for (INT i = 0; i < nf; i++) {
      cur->mis[i] = (cf[i] == 1);
      if (cf[i] == 1) cur->cf_order[q++] = i;
    }
    for (INT i = 0; i < nf; i++) if (cf[i] != 1) cur->cf_order[q++] = i;
    rs_standard_interpolation(&cur->A, &cur->A_filtered, cf, &cur->P);
    free(cf);
    amg_level *nx = &amg->levels[nlev];
    memset(nx, 0, sizeof(*nx));
    galerkin(cur, &nx->A);
    nx->n = nc;
    nx->l1_diag = l1_rows(&nx->A);
    nlev++;
  }
  amg->num_levels = nlev;
  ichol_compute(&amg->levels[nlev - 1].A, &amg->levels[nlev - 1].L_ichol);
}

/* Value-only rebuild on the stored C/F split and P pattern
 * (C_amg1/src/build_hierarchy.c:170-242). */
void amg_rebuild_values(amg_data *amg, const dcsr_mat *A_new) {
  INT nlev = amg->num_levels;
  ichol_free(&amg->levels[nlev - 1].L_ichol);
  amg_level *l0 = &amg->levels[0];
  oracle_csr_free(&l0->A);
  l0->A.row = A_new->row; l0->A.col = A_new->col; l0->A.nnz = A_new->nnz;
  l0->A.ia = (INT *)malloc(((size_t)A_new->row + 1) * sizeof(INT));
  l0->A.ja = (INT *)malloc((size_t)(A_new->nnz > 0 ? A_new->nnz : 1) * sizeof(INT));
  l0->A.val = (REAL *)malloc((size_t)(A_new->nnz > 0 ? A_new->nnz : 1) * sizeof(REAL));
  memcpy(l0->A.ia, A_new->ia, ((size_t)A_new->row + 1) * sizeof(INT));
  memcpy(l0->A.ja, A_new->ja, (size_t)A_new->nnz * sizeof(INT));
  memcpy(l0->A.val, A_new->val, (size_t)A_new->nnz * sizeof(REAL));
  free(l0->l1_diag);
  l0->l1_diag = l1_rows(&l0->A);
  for (INT k = 0; k < nlev - 1; k++) {
    amg_level *cur = &amg->levels[k];
    INT *cf = (INT *)malloc((size_t)cur->n * sizeof(INT));
    for (INT i = 0; i < cur->n; i++) cf[i] = cur->mis[i] ? 1 : -1;
    rs_standard_interpolation_values(&cur->A, &cur->A_filtered, cf, &cur->P);
    free(cf);
    amg_level *nx = &amg->levels[k + 1];
    oracle_csr_free(&nx->A);
    galerkin(cur, &nx->A);
    free(nx->l1_diag);
    nx->l1_diag = l1_rows(&nx->A);
  }
  ichol_compute(&amg->levels[nlev - 1].A, &amg->levels[nlev - 1].L_ichol);
}

void amg_free(amg_data *amg) {
  for (INT k = 0; k < amg->num_levels; k++) {
    amg_level *l = &amg->levels[k];
    oracle_csr_free(&l->A);
    if (k < amg->num_levels - 1) {
      oracle_csr_free(&l->P);
      oracle_csr_free(&l->A_filtered);
    }
    free(l->mis); free(l->isolated); free(l->cf_order); free(l->l1_diag);
    l->mis = l->isolated = l->cf_order = NULL;
    l->l1_diag = NULL;
    if (k == amg->num_levels - 1) ichol_free(&l->L_ichol);
  }
  amg->num_levels = 0;
}
