/*
 * setup.c - host construction of the classical Ruge-Stueben hierarchy.
 *
 * Outputs (C/F split, S, P, Galerkin operators, IC(0) factors) are bit-
 * identical to the C_amg1 routines cited on each function: the split and
 * every row's entry order and FP operation order are the same.  Work that
 * is independent per row (strength, PMIS rounds, interpolation, RAP rows,
 * l1 norms) runs on OpenMP threads with thread-private markers; the greedy
 * RS/HMIS pass, the second pass and IC(0) are inherently sequential.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "amg.h"
#include "bn_amg_ext.h"
#include "csr.h"

void bn_cache_drop(const void *p);

#define CPT 1
#define FPT (-1)

static inline int thread_id(void) {
#ifdef _OPENMP
  return omp_get_thread_num();
#else
  return 0;
#endif
}

/* ---------------------------------------------------------------------
 * Strength: j in S_i iff j != i and -a_ij >= theta * max(0, max_k -a_ik);
 * S keeps A's values and order.  max_lam = longest A row (>= 1).
 * C_amg1/src/rs_classical.c:30-88.
 * ------------------------------------------------------------------- */
static void strength_rule(const dcsr_mat *A, REAL theta, int clamp, dcsr_mat *S, INT *max_lam);

void rs_strength(const dcsr_mat *A, REAL theta, dcsr_mat *S, INT *max_lam) {
  strength_rule(A, theta, 0, S, max_lam);
}

/* clamp = 1: the MATLAB MEX rule, max(0, -a_ij) >= theta * m_i
 * (Matlab_amg/rs_strength_mex.c:58-60,79-81) */
static inline int is_strong(REAL a, REAL t, int clamp) {
  const REAL v = -a;
  return clamp ? ((v < 0.0 ? 0.0 : v) >= t) : (v >= t);
}

static void strength_rule(const dcsr_mat *A, REAL theta, int clamp, dcsr_mat *S, INT *max_lam) {
  const INT n = A->row;
  REAL *thr = (REAL *)bn_xmalloc((size_t)n * sizeof(REAL));
  INT *cnt = (INT *)bn_xmalloc((size_t)n * sizeof(INT));
  INT longest = 0;
#pragma omp parallel for schedule(static) reduction(max : longest)
  for (INT i = 0; i < n; i++) {
    REAL mx = 0.0;
    INT len = A->ia[i + 1] - A->ia[i];
    if (len > longest) longest = len;
    for (INT k = A->ia[i]; k < A->ia[i + 1]; k++) {
      if (A->ja[k] == i) continue;
      REAL v = -A->val[k];
      if (v > mx) mx = v;
    }
    REAL t = theta * mx;
    INT c = 0;
    for (INT k = A->ia[i]; k < A->ia[i + 1]; k++)
      if (A->ja[k] != i && is_strong(A->val[k], t, clamp)) c++;
    thr[i] = t;
    cnt[i] = c;
  }
  *max_lam = longest < 1 ? 1 : longest;
  S->row = n;
  S->col = n;
  S->ia = (INT *)bn_xmalloc(((size_t)n + 1) * sizeof(INT));
  bn_prefix(cnt, n, S->ia);
  S->nnz = S->ia[n];
  S->ja = (INT *)bn_xmalloc((size_t)S->nnz * sizeof(INT));
  S->val = (REAL *)bn_xmalloc((size_t)S->nnz * sizeof(REAL));
#pragma omp parallel for schedule(static)
  for (INT i = 0; i < n; i++) {
    INT p = S->ia[i];
    for (INT k = A->ia[i]; k < A->ia[i + 1]; k++)
      if (A->ja[k] != i && is_strong(A->val[k], thr[i], clamp)) {
        S->ja[p] = A->ja[k];
        S->val[p] = A->val[k];
        p++;
      }
  }
  free(thr);
  free(cnt);
}

/* ---------------------------------------------------------------------
 * Greedy max-lambda selection with bucket lists (the head of a bucket is
 * the node pushed last).  Used by RS (C_amg1/src/rs_classical.c:137-205)
 * and HMIS phase 1 (rs_classical.c:427-486).
 * ------------------------------------------------------------------- */
typedef struct {
  INT *head, *next, *prev;
} lambda_buckets;

static inline void lb_remove(lambda_buckets *b, INT key, INT i) {
  INT p = b->prev[i], q = b->next[i];
  if (p >= 0) b->next[p] = q; else b->head[key] = q;
  if (q >= 0) b->prev[q] = p;
}

static inline void lb_push(lambda_buckets *b, INT key, INT i) {
  INT h = b->head[key];
  b->next[i] = h;
  b->prev[i] = -1;
  if (h >= 0) b->prev[h] = i;
  b->head[key] = i;
}

/* decisions_cap < 0: run until all buckets are empty (RS); otherwise stop
 * once that many nodes are decided (HMIS). */
static void greedy_select(const dcsr_mat *S, const dcsr_mat *ST, INT *lambda,
                          INT nbuckets, INT top, INT *cf, long long decisions_cap) {
  const INT n = S->row;
  lambda_buckets b;
  b.head = (INT *)bn_xmalloc((size_t)nbuckets * sizeof(INT));
  b.next = (INT *)bn_xmalloc((size_t)n * sizeof(INT));
  b.prev = (INT *)bn_xmalloc((size_t)n * sizeof(INT));
  for (INT k = 0; k < nbuckets; k++) b.head[k] = -1;
  for (INT i = n - 1; i >= 0; i--) lb_push(&b, lambda[i], i);
  long long decided = 0;
  for (;;) {
    if (decisions_cap >= 0 && decided >= decisions_cap) break;
    while (top >= 0 && b.head[top] < 0) top--;
    if (top < 0) break;
    const INT c = b.head[top];
    lb_remove(&b, lambda[c], c);
    cf[c] = CPT;
    decided++;
    /* the pass is bound by memory latency on scattered rows: fetch the
     * next neighbours' state ahead (same operations, same order) */
    const INT q0 = ST->ia[c], q1 = ST->ia[c + 1];
    for (INT q = q0; q < q1 && q < q0 + 4; q++) {
      __builtin_prefetch(&cf[ST->ja[q]]);
      __builtin_prefetch(&S->ia[ST->ja[q]]);
    }
    for (INT q = q0; q < q1; q++) {
      const INT f = ST->ja[q];
      if (q + 4 < q1) {
        __builtin_prefetch(&cf[ST->ja[q + 4]]);
        __builtin_prefetch(&S->ia[ST->ja[q + 4]]);
      }
      if (cf[f] != 0) continue;
      lb_remove(&b, lambda[f], f);
      cf[f] = FPT;
      decided++;
      const INT r0 = S->ia[f], r1 = S->ia[f + 1];
      for (INT r = r0; r < r1; r++) {
        const INT k = S->ja[r];
        __builtin_prefetch(&lambda[k]);
        __builtin_prefetch(&b.next[k]);
        __builtin_prefetch(&b.prev[k]);
      }
      for (INT r = r0; r < r1; r++) {
        const INT k = S->ja[r];
        if (cf[k] != 0) continue;
        lb_remove(&b, lambda[k], k);
        const INT nl = ++lambda[k];
        lb_push(&b, nl, k);
        if (nl > top) top = nl;
      }
    }
  }
  free(b.head);
  free(b.next);
  free(b.prev);
}

/* F points without a strong C neighbour become C, in index order and in
 * place (rs_classical.c:211-225). */
static INT promote_and_count(const dcsr_mat *S, INT *cf) {
  const INT n = S->row;
  for (INT i = 0; i < n; i++) {
    if (cf[i] != FPT) continue;
    int has_c = 0;
    for (INT k = S->ia[i]; k < S->ia[i + 1]; k++)
      if (cf[S->ja[k]] == CPT) { has_c = 1; break; }
    if (!has_c) cf[i] = CPT;
  }
  INT nc = 0;
#pragma omp parallel for schedule(static) reduction(+ : nc)
  for (INT i = 0; i < n; i++) nc += (cf[i] == CPT);
  return nc;
}

/* C_amg1/src/rs_classical.c:108-229; lambda[0] stays 0 (the reference's
 * initialisation loop starts at i = 1, :119-123). */
void rs_coarsening(const dcsr_mat *S, INT max_lam, INT *cf, INT *n_coarse) {
  const INT n = S->row;
  dcsr_mat ST;
  bn_csr_transpose(S, &ST);
  INT *lambda = (INT *)bn_xcalloc((size_t)n, sizeof(INT));
  for (INT i = 0; i < n; i++) cf[i] = 0;
  for (INT i = 1; i < n; i++) lambda[i] = ST.ia[i + 1] - ST.ia[i];
  /* 2*max_lam+1 buckets in the reference (:126).  A node's lambda starts
   * at |S^T_k| and grows by one per new F point f with k in S_f, i.e. f in
   * S^T_k, so it never exceeds 2 max_k |S^T_k|; the top pointer starts at
   * max_lam */
  INT max_t = 0;
  for (INT i = 0; i < n; i++)
    if (ST.ia[i + 1] - ST.ia[i] > max_t) max_t = ST.ia[i + 1] - ST.ia[i];
  INT nb = (2 * max_t > max_lam ? 2 * max_t : max_lam) + 2;
  greedy_select(S, &ST, lambda, nb, max_lam, cf, -1);
  free(lambda);
  bn_csr_free(&ST);
  *n_coarse = promote_and_count(S, cf);
}

/* symmetric strength graph G = S u S^T, S's order first
 * (rs_classical.c:256-290) */
static void strength_graph(const dcsr_mat *S, const dcsr_mat *ST, INT **gia_out, INT **gja_out) {
  const INT n = S->row;
  const int nt = bn_num_threads();
  INT *marks = (INT *)bn_xmalloc((size_t)nt * (size_t)n * sizeof(INT));
  INT *cnt = (INT *)bn_xmalloc((size_t)n * sizeof(INT));
  INT *gia = (INT *)bn_xmalloc(((size_t)n + 1) * sizeof(INT));
  INT *gja = NULL;
  for (int pass = 0; pass < 2; pass++) {
#pragma omp parallel
    {
      INT *mark = marks + (size_t)thread_id() * (size_t)n;
      for (INT j = 0; j < n; j++) mark[j] = -1;
#pragma omp for schedule(static)
      for (INT i = 0; i < n; i++) {
        INT p = pass ? gia[i] : 0;
        for (INT k = S->ia[i]; k < S->ia[i + 1]; k++) {
          INT j = S->ja[k];
          if (j == i || mark[j] == i) continue;
          mark[j] = i;
          if (pass) gja[p] = j;
          p++;
        }
        for (INT k = ST->ia[i]; k < ST->ia[i + 1]; k++) {
          INT j = ST->ja[k];
          if (j == i || mark[j] == i) continue;
          mark[j] = i;
          if (pass) gja[p] = j;
          p++;
        }
        if (!pass) cnt[i] = p;
      }
    }
    if (!pass) {
      bn_prefix(cnt, n, gia);
      gja = (INT *)bn_xmalloc((size_t)gia[n] * sizeof(INT));
    }
  }
  free(marks);
  free(cnt);
  *gia_out = gia;
  *gja_out = gja;
}

/* Independent-set rounds (rs_classical.c:305-349): the select phase only
 * reads the state of the previous round and the F phase only reacts to C
 * marks, so both are order-free and run in parallel. */
static void independent_set_rounds(INT n, const INT *gia, const INT *gja, const REAL *w,
                                   INT *cf, INT undecided) {
  unsigned char *pick = (unsigned char *)bn_xmalloc((size_t)n);
  while (undecided > 0) {
    INT picked = 0;
#pragma omp parallel for schedule(static) reduction(+ : picked)
    for (INT i = 0; i < n; i++) {
      pick[i] = 0;
      if (cf[i] != 0) continue;
      int top = 1;
      const REAL wi = w[i];
      for (INT k = gia[i]; k < gia[i + 1]; k++) {
        const INT j = gja[k];
        if (cf[j] == 0 && w[j] > wi) { top = 0; break; }
      }
      if (top) { pick[i] = 1; picked++; }
    }
    if (picked == 0) {
      for (INT i = 0; i < n; i++)
        if (cf[i] == 0) cf[i] = CPT;
      break;
    }
#pragma omp parallel for schedule(static)
    for (INT i = 0; i < n; i++)
      if (pick[i]) cf[i] = CPT;
    undecided -= picked;
    INT to_f = 0;
#pragma omp parallel for schedule(static) reduction(+ : to_f)
    for (INT i = 0; i < n; i++) {
      if (cf[i] != 0) continue;
      for (INT k = gia[i]; k < gia[i + 1]; k++)
        if (cf[gja[k]] == CPT) { cf[i] = FPT; to_f++; break; }
    }
    undecided -= to_f;
  }
  free(pick);
}

/* The PMIS/HMIS tie-break draws: the sequence rand() gives after
 * srand(12345) (C_amg1/src/rs_classical.c:295-298,537-541), taken from a
 * private glibc random_r state of the same type (TYPE_3, 128-byte state,
 * which is what srand/rand use), so the draws match the reference's on the
 * same C library while concurrent set-ups on other threads neither disturb
 * them nor the caller's own rand() sequence. */
typedef struct {
  struct random_data rd;
  char state[128];
} weight_rng;

static void weight_rng_init(weight_rng *g) {
  memset(g, 0, sizeof(*g));
  initstate_r(12345u, g->state, sizeof(g->state), &g->rd);
}

static REAL weight_draw(weight_rng *g) {
  int32_t v = 0;
  random_r(&g->rd, &v);
  return (REAL)v / (REAL)RAND_MAX;
}

/* PMIS, C_amg1/src/rs_classical.c:246-371.  Weights lambda_i +
 * rand()/RAND_MAX after srand(12345), drawn in index order, so the split
 * matches the reference on the same C library. */
/* the per-row set-up stages on the device (setup_dev.hip), entry for entry
 * the host routines' results; absent from builds without the device code */
typedef struct bn_dsetup bn_dsetup;
bn_dsetup *bn_dsetup_begin(const dcsr_mat *A) __attribute__((weak));
void bn_dsetup_end(bn_dsetup *c) __attribute__((weak));
int bn_dsetup_strength(bn_dsetup *c, double theta, int clamp, dcsr_mat *S, int *colcount) __attribute__((weak));
int bn_dsetup_rounds(bn_dsetup *c, const double *w, int *cf, int undecided) __attribute__((weak));
int bn_dsetup_interp(bn_dsetup *c, const int *cf, dcsr_mat *P) __attribute__((weak));

/* PMIS with the rounds on the device: the weights are the host's (column
 * counts from the device, draws in index order); 0, or -1 to run the host
 * routine instead */
static int pmis_device(bn_dsetup *dc, const INT *colcnt, const dcsr_mat *S, INT *cf, INT *n_coarse) {
  const INT n = S->row;
  REAL *w = (REAL *)bn_xmalloc((size_t)(n > 0 ? n : 1) * sizeof(REAL));
  weight_rng g;
  weight_rng_init(&g);
  for (INT i = 0; i < n; i++) {
    w[i] = (REAL)colcnt[i] + weight_draw(&g);
    cf[i] = 0;
  }
  const int rc = bn_dsetup_rounds(dc, w, cf, n);
  free(w);
  if (rc) return -1;
  *n_coarse = promote_and_count(S, cf);
  return 0;
}

void pmis_coarsening(const dcsr_mat *S, INT *cf, INT *n_coarse) {
  const INT n = S->row;
  dcsr_mat ST;
  bn_csr_transpose(S, &ST);
  INT *gia, *gja;
  strength_graph(S, &ST, &gia, &gja);
  REAL *w = (REAL *)bn_xmalloc((size_t)n * sizeof(REAL));
  weight_rng g;
  weight_rng_init(&g);
  for (INT i = 0; i < n; i++) {
    w[i] = (REAL)(ST.ia[i + 1] - ST.ia[i]) + weight_draw(&g);
    cf[i] = 0;
  }
  bn_csr_free(&ST);
  independent_set_rounds(n, gia, gja, w, cf, n);
  free(w);
  free(gia);
  free(gja);
  *n_coarse = promote_and_count(S, cf);
}

/* HMIS, C_amg1/src/rs_classical.c:390-611: greedy pass with every lambda
 * initialised, capped at n decisions, then PMIS rounds on any undecided
 * rest (weights drawn for those nodes only). */
void hmis_coarsening(const dcsr_mat *S, INT *cf, INT *n_coarse) {
  const INT n = S->row;
  dcsr_mat ST;
  bn_csr_transpose(S, &ST);
  INT *lambda = (INT *)bn_xmalloc((size_t)n * sizeof(INT));
  INT max_lam = 0, max_row = 0;
  for (INT i = 0; i < n; i++) {
    lambda[i] = ST.ia[i + 1] - ST.ia[i];
    cf[i] = 0;
    if (lambda[i] > max_lam) max_lam = lambda[i];
    INT rl = S->ia[i + 1] - S->ia[i];
    if (rl > max_row) max_row = rl;
  }
  /* lambda <= 2 max_lam (see rs_coarsening) */
  INT nb = 2 * max_lam + max_row + 2;
  greedy_select(S, &ST, lambda, nb, max_lam, cf, (long long)n);
  free(lambda);
  INT undecided = 0;
  for (INT i = 0; i < n; i++) undecided += (cf[i] == 0);
  if (undecided > 0) {
    INT *gia, *gja;
    strength_graph(S, &ST, &gia, &gja);
    REAL *w = (REAL *)bn_xmalloc((size_t)n * sizeof(REAL));
    weight_rng g;
    weight_rng_init(&g);
    for (INT i = 0; i < n; i++)
      w[i] = (cf[i] == 0) ? (REAL)(ST.ia[i + 1] - ST.ia[i]) + weight_draw(&g) : -1.0;
    independent_set_rounds(n, gia, gja, w, cf, undecided);
    free(w);
    free(gia);
    free(gja);
  }
  bn_csr_free(&ST);
  *n_coarse = promote_and_count(S, cf);
}

/* ---------------------------------------------------------------------
 * Direct interpolation (C_amg1/src/rs_classical.c:643-819)
 * ------------------------------------------------------------------- */
void rs_standard_interpolation_sparsity(const dcsr_mat *S, const INT *cf, dcsr_mat *P) {
  const INT n = S->row;
  INT *cidx = (INT *)bn_xmalloc((size_t)n * sizeof(INT));
  INT *cnt = (INT *)bn_xmalloc((size_t)n * sizeof(INT));
  INT nc = 0;
  for (INT i = 0; i < n; i++) cidx[i] = (cf[i] == CPT) ? nc++ : -1;
#pragma omp parallel for schedule(static)
  for (INT i = 0; i < n; i++) {
    if (cf[i] == CPT) { cnt[i] = 1; continue; }
    INT c = 0;
    for (INT k = S->ia[i]; k < S->ia[i + 1]; k++) c += (cf[S->ja[k]] == CPT);
    cnt[i] = c;
  }
  P->row = n;
  P->col = nc;
  P->ia = (INT *)bn_xmalloc(((size_t)n + 1) * sizeof(INT));
  bn_prefix(cnt, n, P->ia);
  P->nnz = P->ia[n];
  P->ja = (INT *)bn_xmalloc((size_t)P->nnz * sizeof(INT));
  P->val = (REAL *)bn_xcalloc((size_t)P->nnz, sizeof(REAL));
#pragma omp parallel for schedule(static)
  for (INT i = 0; i < n; i++) {
    INT p = P->ia[i];
    if (cf[i] == CPT) {
      P->ja[p] = cidx[i];
    } else {
      for (INT k = S->ia[i]; k < S->ia[i + 1]; k++)
        if (cf[S->ja[k]] == CPT) P->ja[p++] = cidx[S->ja[k]];
    }
  }
  free(cidx);
  free(cnt);
}

/* w_ij = -(a_ij + sum_{k in D_i^s} a_ik a_kj / sum_{l in C_i} a_kl) / a^_ii
 * with the reference's lumping (|den| < 1e-15 -> a^_ii) and scaling guard
 * (|a^_ii| > 1e-15); rs_classical.c:703-805.  Rows are independent. */
void rs_standard_interpolation_values(const dcsr_mat *A, const dcsr_mat *S,
                                      const INT *cf, dcsr_mat *P) {
  const INT n = A->row;
  const int nt = bn_num_threads();
  INT maxrow = 1;
  for (INT i = 0; i < n; i++)
    if (A->ia[i + 1] - A->ia[i] > maxrow) maxrow = A->ia[i + 1] - A->ia[i];
  unsigned *tag_all = (unsigned *)bn_xmalloc((size_t)nt * (size_t)n * sizeof(unsigned));
  INT *slot_all = (INT *)bn_xmalloc((size_t)nt * (size_t)n * sizeof(INT));
  INT *tslot_all = (INT *)bn_xmalloc((size_t)nt * (size_t)maxrow * sizeof(INT));
  REAL *tval_all = (REAL *)bn_xmalloc((size_t)nt * (size_t)maxrow * sizeof(REAL));
  /* about 64 chunks per thread: small dense levels balance their heavy
   * rows, large levels keep the scheduling traffic low */
  INT chunk = n / (64 * (INT)nt);
  if (chunk < 64) chunk = 64;
  if (chunk > 1024) chunk = 1024;
#pragma omp parallel
  {
    const size_t off = (size_t)thread_id() * (size_t)n;
    /* tag[j] = 2i + 1: j in S_i and a C point; 2i: j in S_i, not C
     * (unsigned: below the unset value 0xffffffff for every int row i) */
    unsigned *tag = tag_all + off;
    INT *slot = slot_all + off;
    INT *tslot = tslot_all + (size_t)thread_id() * (size_t)maxrow;
    REAL *tval = tval_all + (size_t)thread_id() * (size_t)maxrow;
    for (INT j = 0; j < n; j++) tag[j] = 0xffffffffu;
#pragma omp for schedule(dynamic, chunk)
    for (INT i = 0; i < n; i++) {
      if (cf[i] == CPT) {
        P->val[P->ia[i]] = 1.0;
        continue;
      }
      const unsigned own = 2u * (unsigned)i, cown = own + 1u;
      {
        INT p = P->ia[i];
        for (INT k = S->ia[i]; k < S->ia[i + 1]; k++) {
          const INT j = S->ja[k];
          if (cf[j] == CPT) {
            P->val[p] = 0.0;
            tag[j] = cown;
            slot[j] = p++;
          } else {
            tag[j] = own;
          }
        }
      }
      REAL diag = 0.0, weak = 0.0;
      for (INT k = A->ia[i]; k < A->ia[i + 1]; k++) {
        const INT j = A->ja[k];
        const REAL a = A->val[k];
        if (j == i) diag = a;
        else if (tag[j] == cown) P->val[slot[j]] += a;
        else if (tag[j] == own && cf[j] == FPT) continue;
        else weak += a;
      }
      REAL dhat = diag + weak;
      for (INT k = A->ia[i]; k < A->ia[i + 1]; k++) {
        const INT f = A->ja[k];
        if (f == i || tag[f] != own || cf[f] != FPT) continue;
        const REAL aik = A->val[k];
        /* row f's entries in C_i^s, in row order: summed into the
         * denominator, then (the same entries, same order) distributed */
        REAL den = 0.0;
        INT nb = 0;
        for (INT q = A->ia[f]; q < A->ia[f + 1]; q++) {
          const INT m = A->ja[q];
          if (tag[m] == cown) {
            den += A->val[q];
            tslot[nb] = slot[m];
            tval[nb++] = A->val[q];
          }
        }
        if (fabs(den) < 1e-15) {
          dhat += aik;
        } else {
          for (INT t = 0; t < nb; t++) P->val[tslot[t]] += aik * tval[t] / den;
        }
      }
      if (fabs(dhat) > 1e-15)
        for (INT q = P->ia[i]; q < P->ia[i + 1]; q++) P->val[q] /= -dhat;
    }
  }
  free(tag_all);
  free(slot_all);
  free(tslot_all);
  free(tval_all);
}

void rs_standard_interpolation(const dcsr_mat *A, const dcsr_mat *S, const INT *cf, dcsr_mat *P) {
  rs_standard_interpolation_sparsity(S, cf, P);
  rs_standard_interpolation_values(A, S, cf, P);
}

/* ---------------------------------------------------------------------
 * Galerkin product R*A*P, row layout of C_amg1/src/sparse_ops.c:102-213
 * (diagonal first, then columns in order of first discovery; a fine row
 * met again in the same coarse row only accumulates).  Coarse rows are
 * independent: count pass, prefix sum, fill pass, each thread with its
 * own column-slot and fine-row markers.
 * ------------------------------------------------------------------- */
void dcsr_rap(dcsr_mat *R, dcsr_mat *A, dcsr_mat *P, dcsr_mat *RAP) {
  const INT nc = R->row, nf = A->row;
  const int nt = bn_num_threads();
  INT *cs_all = (INT *)bn_xmalloc((size_t)nt * (size_t)nc * sizeof(INT));
  INT *rs_all = (INT *)bn_xmalloc((size_t)nt * (size_t)nf * sizeof(INT));
  INT *cnt = (INT *)bn_xmalloc((size_t)nc * sizeof(INT));
  INT *ia = (INT *)bn_xmalloc(((size_t)nc + 1) * sizeof(INT));
  INT *ja = NULL;
  REAL *va = NULL;
  for (int pass = 0; pass < 2; pass++) {
#pragma omp parallel
    {
      const int t = thread_id();
      INT *cstamp = cs_all + (size_t)t * (size_t)nc; /* row that owns the slot */
      INT *rseen = rs_all + (size_t)t * (size_t)nf;
      INT *cslot = NULL;
      for (INT j = 0; j < nc; j++) cstamp[j] = -1;
      for (INT j = 0; j < nf; j++) rseen[j] = -1;
      if (pass) cslot = (INT *)bn_xmalloc((size_t)nc * sizeof(INT));
#pragma omp for schedule(dynamic, 256)
      for (INT ic = 0; ic < nc; ic++) {
        INT p = pass ? ia[ic] : 0;
        cstamp[ic] = ic;
        if (pass) {
          cslot[ic] = p;
          ja[p] = ic;
          va[p] = 0.0;
        }
        p++;
        for (INT q1 = R->ia[ic]; q1 < R->ia[ic + 1]; q1++) {
          const INT i1 = R->ja[q1];
          const REAL r = pass ? R->val[q1] : 0.0;
          for (INT q2 = A->ia[i1]; q2 < A->ia[i1 + 1]; q2++) {
            const INT i2 = A->ja[q2];
            const int fresh = (rseen[i2] != ic);
            rseen[i2] = ic;
            if (!pass) {
              if (!fresh) continue;
              for (INT q3 = P->ia[i2]; q3 < P->ia[i2 + 1]; q3++) {
                const INT i3 = P->ja[q3];
                if (cstamp[i3] != ic) { cstamp[i3] = ic; p++; }
              }
              continue;
            }
            const REAL ra = r * A->val[q2];
            for (INT q3 = P->ia[i2]; q3 < P->ia[i2 + 1]; q3++) {
              const INT i3 = P->ja[q3];
              if (fresh && cstamp[i3] != ic) {
                cstamp[i3] = ic;
                cslot[i3] = p;
                ja[p] = i3;
                va[p] = ra * P->val[q3];
                p++;
              } else {
                va[cslot[i3]] += ra * P->val[q3];
              }
            }
          }
        }
        if (!pass) cnt[ic] = p;
      }
      free(cslot);
    }
    if (!pass) {
      bn_prefix(cnt, nc, ia);
      ja = (INT *)bn_xmalloc((size_t)ia[nc] * sizeof(INT));
      va = (REAL *)bn_xmalloc((size_t)ia[nc] * sizeof(REAL));
    }
  }
  free(cs_all);
  free(rs_all);
  free(cnt);
  RAP->row = nc;
  RAP->col = nc;
  RAP->nnz = ia[nc];
  RAP->ia = ia;
  RAP->ja = ja;
  RAP->val = va;
}

/* ---------------------------------------------------------------------
 * IC(0) in stored row order (C_amg1/src/ichol.c:12-118): the partial sums
 * stop at the first column >= the pivot, which for diagonal-first rows
 * yields L(i,k) = a_ik / sqrt(a_kk); non-positive pivots become 1e-14.
 * ------------------------------------------------------------------- */
void ichol_compute(const dcsr_mat *A, ichol_factor *L) {
  const INT n = A->row;
  bn_cache_drop(L); /* a device mirror of a factor held in *L is stale now */
  INT *cnt = (INT *)bn_xmalloc((size_t)n * sizeof(INT));
#pragma omp parallel for schedule(static)
  for (INT i = 0; i < n; i++) {
    INT c = 0;
    for (INT k = A->ia[i]; k < A->ia[i + 1]; k++) c += (A->ja[k] <= i);
    cnt[i] = c;
  }
  L->n = n;
  L->ia = (INT *)bn_xmalloc(((size_t)n + 1) * sizeof(INT));
  bn_prefix(cnt, n, L->ia);
  free(cnt);
  L->nnz = L->ia[n];
  L->ja = (INT *)bn_xmalloc((size_t)L->nnz * sizeof(INT));
  L->val = (REAL *)bn_xmalloc((size_t)L->nnz * sizeof(REAL));
  INT *dpos = (INT *)bn_xmalloc((size_t)n * sizeof(INT));
#pragma omp parallel for schedule(static)
  for (INT i = 0; i < n; i++) {
    INT p = L->ia[i];
    dpos[i] = -1;
    for (INT k = A->ia[i]; k < A->ia[i + 1]; k++)
      if (A->ja[k] <= i) {
        if (A->ja[k] == i && dpos[i] < 0) dpos[i] = p;
        L->ja[p] = A->ja[k];
        L->val[p] = A->val[k];
        p++;
      }
  }
  REAL *w = (REAL *)bn_xcalloc((size_t)n, sizeof(REAL));
  INT *wrow = (INT *)bn_xmalloc((size_t)n * sizeof(INT));
  for (INT i = 0; i < n; i++) wrow[i] = -1;
  for (INT i = 0; i < n; i++) {
    const INT lo = L->ia[i], hi = L->ia[i + 1];
    for (INT q = lo; q < hi; q++) {
      w[L->ja[q]] = L->val[q];
      wrow[L->ja[q]] = i;
    }
    for (INT q = lo; q < hi; q++) {
      const INT k = L->ja[q];
      if (k >= i) continue;
      REAL s = w[k];
      for (INT t = L->ia[k]; t < L->ia[k + 1]; t++) {
        const INT j = L->ja[t];
        if (j >= k) break;
        if (wrow[j] == i) s -= w[j] * L->val[t];
      }
      w[k] = s / L->val[dpos[k]];
    }
    REAL d = w[i];
    for (INT q = lo; q < hi; q++) {
      const INT j = L->ja[q];
      if (j >= i) break;
      d -= w[j] * w[j];
    }
    if (d <= 0.0) d = 1e-14;
    w[i] = sqrt(d);
    for (INT q = lo; q < hi; q++) L->val[q] = w[L->ja[q]];
    for (INT q = lo; q < hi; q++) {
      w[L->ja[q]] = 0.0;
      wrow[L->ja[q]] = -1;
    }
  }
  free(w);
  free(wrow);
  free(dpos);
}

void ichol_free(ichol_factor *L) {
  if (!L) return;
  bn_cache_drop(L);
  free(L->ia);
  free(L->ja);
  free(L->val);
  L->ia = NULL;
  L->ja = NULL;
  L->val = NULL;
  L->n = 0;
  L->nnz = 0;
}

/* ---------------------------------------------------------------------
 * Hierarchy (C_amg1/src/build_hierarchy.c:29-260)
 * ------------------------------------------------------------------- */
/* the same product on the device (galerkin.hip), identical entry for
 * entry; absent from builds without the device code, and it declines (-1)
 * when no device is present */
int bn_galerkin_device(const dcsr_mat *A, const dcsr_mat *P, const dcsr_mat *R, dcsr_mat *Ac)
    __attribute__((weak));

/* A device context for the per-row stages of a level.  BN_DEVICE_SETUP
 * unset (automatic): PMIS levels of at least 4,000,000 operator entries,
 * where the device rounds more than pay for the transfers (at 1.07M
 * unknowns, level 0: 71 ms of host strength, split and interpolation on
 * 16 threads against 37 ms); the other coarsenings keep the per-row
 * stages on the host, whose threads form strength and interpolation in
 * about the time the operator takes to cross to the device.  "all" (or
 * "1"): every level of at least BN_DEVICE_SETUP_MIN entries (default 0),
 * whatever the coarsening; "off" (or "0"): none.  No device: none. */
static bn_dsetup *device_setup_begin(const dcsr_mat *A, int coarsening) {
  if (!bn_dsetup_begin || !bn_dsetup_end || !bn_dsetup_strength || !bn_dsetup_rounds || !bn_dsetup_interp)
    return NULL;
  const char *mode = getenv("BN_DEVICE_SETUP");
  long long min_nnz = 4000000;
  if (mode && *mode) {
    if (!strcmp(mode, "off") || !strcmp(mode, "0")) return NULL;
    if (strcmp(mode, "all") && strcmp(mode, "1")) return NULL;
    const char *mn = getenv("BN_DEVICE_SETUP_MIN");
    min_nnz = (mn && *mn) ? atoll(mn) : 0;
  } else if (coarsening != COARSEN_PMIS) {
    return NULL;
  }
  if (A->nnz < min_nnz) return NULL;
  return bn_dsetup_begin(A);
}

static void dsetup_end(bn_dsetup *dc) {
  if (dc) bn_dsetup_end(dc);
}

static void galerkin_level(const amg_level *cur, dcsr_mat *Ac) {
  dcsr_mat Pt, rap;
  bn_csr_transpose(&cur->P, &Pt);
  if (bn_galerkin_device && bn_galerkin_device(&cur->A, &cur->P, &Pt, Ac) == 0) {
    bn_csr_free(&Pt);
    return;
  }
  dcsr_rap(&Pt, (dcsr_mat *)&cur->A, (dcsr_mat *)&cur->P, &rap);
  bn_csr_free(&Pt);
  bn_csr_symmetrize(&rap, Ac);
  bn_csr_free(&rap);
}

void amg_build_hierarchy(amg_data *amg, const dcsr_mat *A) { bn_build_hierarchy(amg, A, 0); }

static double setup_ms(void) { return 1e3 * omp_get_wtime(); }

/* amg_build_hierarchy with a choice of strength rule (BN_STRENGTH_*) */
void bn_build_hierarchy(amg_data *amg, const dcsr_mat *A, int strength) {
  bn_cache_drop(amg); /* a device mirror of a hierarchy held in *amg is stale now */
  const char *lg = getenv("BN_SETUP_LOG");
  const int log = lg && *lg && *lg != '0';
  INT max_levels = amg->max_levels;
  if (max_levels > AMG_MAX_LEVELS) max_levels = AMG_MAX_LEVELS;
  amg_level *l0 = &amg->levels[0];
  memset(l0, 0, sizeof(*l0));
  bn_csr_copy(A, &l0->A);
  l0->n = A->row;
  l0->l1_diag = (REAL *)bn_xmalloc((size_t)A->row * sizeof(REAL));
  bn_csr_l1(&l0->A, l0->l1_diag);
  INT nlev = 1;
  while (amg->levels[nlev - 1].A.row >= amg->min_size && nlev < max_levels) {
    amg_level *cur = &amg->levels[nlev - 1];
    const INT nf = cur->A.row;
    INT ml = 0, nc = 0;
    double t[5];
    t[0] = setup_ms();
    /* large levels: strength, PMIS rounds and interpolation on the device */
    bn_dsetup *dc = device_setup_begin(&cur->A, amg->coarsening);
    INT *colcnt = (dc && amg->coarsening == COARSEN_PMIS) ? (INT *)bn_xmalloc((size_t)nf * sizeof(INT)) : NULL;
    if (dc && bn_dsetup_strength(dc, amg->threshold, strength == 1, &cur->A_filtered, colcnt) == 0) {
      INT longest = 0;
      for (INT i = 0; i < nf; i++)
        if (cur->A.ia[i + 1] - cur->A.ia[i] > longest) longest = cur->A.ia[i + 1] - cur->A.ia[i];
      ml = longest < 1 ? 1 : longest;
    } else {
      dsetup_end(dc);
      dc = NULL;
      free(colcnt);
      colcnt = NULL;
      strength_rule(&cur->A, amg->threshold, strength == 1, &cur->A_filtered, &ml);
    }
    t[1] = setup_ms();
    INT *cf = (INT *)bn_xcalloc((size_t)nf, sizeof(INT));
    if (amg->coarsening == COARSEN_PMIS) {
      if (!colcnt || pmis_device(dc, colcnt, &cur->A_filtered, cf, &nc)) pmis_coarsening(&cur->A_filtered, cf, &nc);
    } else if (amg->coarsening == COARSEN_HMIS) {
      hmis_coarsening(&cur->A_filtered, cf, &nc);
    } else {
      rs_coarsening(&cur->A_filtered, ml, cf, &nc);
    }
    free(colcnt);
    t[2] = setup_ms();
    cur->n_coarse = nc;
    if (nc == 0 || nc >= nf) {
      /* stalled (build_hierarchy.c:78-83): this level becomes coarsest */
      dsetup_end(dc);
      bn_csr_free(&cur->A_filtered);
      free(cf);
      break;
    }
    cur->mis = (INT *)bn_xmalloc((size_t)nf * sizeof(INT));
    cur->isolated = (INT *)bn_xcalloc((size_t)nf, sizeof(INT));
    cur->cf_order = (INT *)bn_xmalloc((size_t)nf * sizeof(INT));
    INT q = 0;
    for (INT i = 0; i < nf; i++) {
      cur->mis[i] = (cf[i] == CPT);
      if (cf[i] == CPT) cur->cf_order[q++] = i;
    }
    for (INT i = 0; i < nf; i++)
      if (cf[i] != CPT) cur->cf_order[q++] = i;
    if (!dc || bn_dsetup_interp(dc, cf, &cur->P)) rs_standard_interpolation(&cur->A, &cur->A_filtered, cf, &cur->P);
    const int on_device = dc != NULL;
    dsetup_end(dc);
    free(cf);
    t[3] = setup_ms();
    amg_level *nx = &amg->levels[nlev];
    memset(nx, 0, sizeof(*nx));
    galerkin_level(cur, &nx->A);
    t[4] = setup_ms();
    if (log)
      fprintf(stderr, "[setup] level %d (%d rows, %d nnz%s): strength %.1f ms, split %.1f ms, interpolation %.1f ms, "
              "Galerkin %.1f ms -> %d rows, %d nnz\n", nlev - 1, nf, cur->A.nnz, on_device ? ", device stages" : "",
              t[1] - t[0], t[2] - t[1], t[3] - t[2], t[4] - t[3], nc, nx->A.nnz);
    nx->n = nc;
    nx->l1_diag = (REAL *)bn_xmalloc((size_t)nc * sizeof(REAL));
    bn_csr_l1(&nx->A, nx->l1_diag);
    nlev++;
  }
  amg->num_levels = nlev;
  amg->levels[nlev - 1].n_coarse = 0;
  ichol_compute(&amg->levels[nlev - 1].A, &amg->levels[nlev - 1].L_ichol);
}

void amg_rebuild_values(amg_data *amg, const dcsr_mat *A_new) {
  const INT nlev = amg->num_levels;
  bn_cache_drop(amg);
  ichol_free(&amg->levels[nlev - 1].L_ichol);
  amg_level *l0 = &amg->levels[0];
  bn_csr_free(&l0->A);
  bn_csr_copy(A_new, &l0->A);
  bn_csr_l1(&l0->A, l0->l1_diag);
  for (INT k = 0; k < nlev - 1; k++) {
    amg_level *cur = &amg->levels[k];
    INT *cf = (INT *)bn_xmalloc((size_t)cur->n * sizeof(INT));
    for (INT i = 0; i < cur->n; i++) cf[i] = cur->mis[i] ? CPT : FPT;
    rs_standard_interpolation_values(&cur->A, &cur->A_filtered, cf, &cur->P);
    free(cf);
    amg_level *nx = &amg->levels[k + 1];
    bn_csr_free(&nx->A);
    galerkin_level(cur, &nx->A);
    bn_csr_l1(&nx->A, nx->l1_diag);
  }
  ichol_compute(&amg->levels[nlev - 1].A, &amg->levels[nlev - 1].L_ichol);
}

void amg_free(amg_data *amg) {
  bn_cache_drop(amg);
  for (INT k = 0; k < amg->num_levels; k++) {
    amg_level *l = &amg->levels[k];
    bn_csr_free(&l->A);
    if (k < amg->num_levels - 1) {
      bn_csr_free(&l->P);
      bn_csr_free(&l->A_filtered);
    }
    free(l->mis);
    free(l->isolated);
    free(l->cf_order);
    free(l->l1_diag);
    l->mis = l->isolated = l->cf_order = NULL;
    l->l1_diag = NULL;
    if (k == amg->num_levels - 1) ichol_free(&l->L_ichol);
  }
  amg->num_levels = 0;
}
