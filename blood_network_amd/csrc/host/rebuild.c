/*
 * rebuild.c - value-only rebuild of the hierarchy on the device.
 *
 * amg_solver_update (C_amg1/src/amg_wrapper.c:114-149) rebuilds the
 * values of a hierarchy on its stored C/F split and patterns on the host
 * (C_amg1/src/build_hierarchy.c:170-242) and uploads it again.  Here the
 * new fine values come from device memory and the level operators and
 * interpolation weights are recomputed by the kernels of rebuild.hip, in
 * the host's order and without contraction, so they are bit-identical to
 * the host rebuild.  Every derived device array (FP32 copies, diagonal and
 * l1 inverses, the colour-ordered GS rows, R = P^T) is refreshed in place
 * through index maps built once from the host layout, so captured graphs
 * stay valid.  What stays on the host: the fine IC(0) and the coarsest
 * solve (dense inverse or IC(0)), whose inputs are downloaded; the host
 * hierarchy is brought up to date as well, so a later re-upload gives the
 * same state.
 */
#include <limits.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "bn_solver.h"
#include "csr.h"
#include "schedule.h"

static double rb_ms(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return 1e3 * (double)t.tv_sec + 1e-6 * (double)t.tv_nsec;
}

void bn_set_error(const char *fmt, ...);
void bn_track_device(struct bn_solver *s, void *p);
int bn_ensure_device(struct bn_solver *s);
int bn_dense_inverse(const dcsr_mat *A, double *inv);

typedef struct {
  /* standard interpolation (every level but the last): per A entry class
   * (0 strong, 1 weak, 2 diagonal), C points, the (F row, strong F
   * neighbour) pairs of each row with the A entry of each and the row-f
   * entries its den sums, and the terms of each P entry */
  unsigned char *cls;
  int *cpt, *rp, *pair_k, *doff, *dq;
  int *prow, *poff, *pa, *pq;
  int npairs;
  int *rmap; /* R entry -> P entry */
  /* Galerkin product of this level: per coarse entry triple lists (pull),
   * or one thread per coarse row over the sorted coarse rows below */
  int pull, nlong;
  int *toff, *t1, *t2, *t3, *elist;
  int *c_col, *c_pos, *tpos; /* sorted row copy and transpose positions (levels >= 1) */
  int *acmap;                /* colour-ordered GS row entry -> level entry */
} rb_level;

typedef struct {
  rb_level lev[AMG_MAX_LEVELS];
  int *tpos0;    /* fine level: transpose positions (input symmetrisation) */
  double *work;  /* product values / gathered input, max nnz over levels */
  double *iw;    /* interpolation scratch: den per pair, dhat per row */
  double *dn_c, *dn_r; /* coarsest inverse: the factor column- and row-major */
  int *ic_lmap, *ic_umap; /* fine IC(0) on the device: factor entry -> operator
                           * entry, transposed row entry -> factor entry */
  double *l1;              /* l1 row norms of every level (host set-up's l1_diag) */
  long long l1off[AMG_MAX_LEVELS + 1];
  int maxnnz;
  long long maxiw;
} rb_plan;

void bn_rebuild_release(struct bn_solver *s) {
  free(s->rb); /* the device arrays go with the solver's allocations */
  s->rb = NULL;
}

static void *rb_dev(struct bn_solver *s, size_t bytes) {
  void *p = NULL;
  if (hipMalloc(&p, bytes ? bytes : 8) != hipSuccess) return NULL;
  bn_track_device(s, p);
  return p;
}

static int *rb_upload_int(struct bn_solver *s, const int *h, size_t n) {
  int *d = (int *)rb_dev(s, n * sizeof(int));
  if (d && n && hipMemcpy(d, h, n * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) return NULL;
  return d;
}

typedef struct {
  int col, val;
} pair_t;

static int cmp_pair(const void *a, const void *b) {
  const pair_t *x = (const pair_t *)a, *y = (const pair_t *)b;
  return (x->col > y->col) - (x->col < y->col);
}

static void sort_row(pair_t *p, int len) {
  if (len > 24) {
    qsort(p, (size_t)len, sizeof(pair_t), cmp_pair);
    return;
  }
  for (int i = 1; i < len; i++) {
    pair_t t = p[i];
    int j = i - 1;
    while (j >= 0 && p[j].col > t.col) p[j + 1] = p[j], j--;
    p[j + 1] = t;
  }
}

/* per-row sorted copy of a CSR pattern: col[] sorted within each row, pos[]
 * the original position of each entry */
static void sorted_rows(const dcsr_mat *A, int *col, int *pos) {
#pragma omp parallel for schedule(dynamic, 1024)
  for (int i = 0; i < A->row; i++) {
    const int lo = A->ia[i], len = A->ia[i + 1] - lo;
    pair_t *p = (pair_t *)malloc((size_t)(len > 0 ? len : 1) * sizeof(pair_t));
    for (int k = 0; k < len; k++) p[k].col = A->ja[lo + k], p[k].val = lo + k;
    sort_row(p, len);
    for (int k = 0; k < len; k++) col[lo + k] = p[k].col, pos[lo + k] = p[k].val;
    free(p);
  }
}

static int find_sorted(const int *col, int lo, int hi, int c) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (col[mid] == c) return mid;
    if (col[mid] < c) lo = mid + 1;
    else hi = mid;
  }
  return -1;
}

/* position of entry (j, i) for every entry (i, j); -1 if a matrix is not
 * structurally symmetric */
static int transpose_positions(const dcsr_mat *A, const int *col, const int *pos, int *tp) {
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (int i = 0; i < A->row; i++)
    for (int q = A->ia[i]; q < A->ia[i + 1]; q++) {
      const int j = A->ja[q];
      const int at = find_sorted(col, A->ia[j], A->ia[j + 1], i);
      if (at < 0) bad = 1;
      else tp[q] = pos[at];
    }
  return bad ? -1 : 0;
}

static unsigned char *rb_upload_bytes(struct bn_solver *s, const unsigned char *h, size_t n) {
  unsigned char *d = (unsigned char *)rb_dev(s, n);
  if (d && n && hipMemcpy(d, h, n, hipMemcpyHostToDevice) != hipSuccess) return NULL;
  return d;
}

static int *xcalloc_int(size_t n) {
  int *p = (int *)bn_xmalloc((n > 0 ? n : 1) * sizeof(int));
  memset(p, 0, (n > 0 ? n : 1) * sizeof(int));
  return p;
}

/* exclusive prefix sums of cnt[0..n) into off[0..n]; -1 past INT_MAX */
static int scan_counts(const int *cnt, int n, int *off) {
  long long t = 0;
  for (int i = 0; i < n; i++) {
    off[i] = (int)t;
    t += cnt[i];
    if (t > INT_MAX) return -1;
  }
  off[n] = (int)t;
  return 0;
}

/* strong neighbours of F row i in mk: the P slot of a strong C point (slots
 * in the order the interpolation's sparsity pass assigns them), -2 for a
 * strong F point; returns the slot past the last */
static int mark_strong(const amg_level *L, int i, int *mk) {
  const dcsr_mat *S = &L->A_filtered;
  int slot = L->P.ia[i];
  for (int k = S->ia[i]; k < S->ia[i + 1]; k++) {
    const int j = S->ja[k];
    mk[j] = L->mis[j] ? slot++ : -2;
  }
  return slot;
}

static void unmark(const amg_level *L, int i, int *mk) {
  const dcsr_mat *S = &L->A_filtered;
  for (int k = S->ia[i]; k < S->ia[i + 1]; k++) mk[S->ja[k]] = -1;
}

/* Contribution lists of the standard interpolation of one level
 * (C_amg1/src/rs_classical.c, rs_standard_interpolation_values): for F
 * row i, the A entries (a_ij with j a strong C point) that start each
 * weight, the strong F neighbours f in row order with the entries of row
 * f that their den sums (columns that are strong C points of i), and for
 * each weight the terms a_if a_fj / den that follow, in the order the
 * host adds them. */
static int plan_interp(struct bn_solver *s, const amg_level *L, rb_level *r, long long *iw) {
  const dcsr_mat *A = &L->A, *P = &L->P;
  const int n = A->row, np = P->nnz;
  int *npr = xcalloc_int((size_t)n + 1), *ndr = xcalloc_int((size_t)n + 1), *pcnt = xcalloc_int((size_t)np + 1);
  unsigned char *cls = (unsigned char *)bn_xmalloc((size_t)(A->nnz > 0 ? A->nnz : 1));
  int *cpt = xcalloc_int((size_t)n);
  int bad = 0, maxp = 1, rc = 0;
  for (int i = 0; i < n; i++)
    if (P->ia[i + 1] - P->ia[i] > maxp) maxp = P->ia[i + 1] - P->ia[i];
#pragma omp parallel reduction(| : bad)
  {
    int *mk = (int *)bn_xmalloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    for (int j = 0; j < n; j++) mk[j] = -1;
#pragma omp for schedule(dynamic, 256)
    for (int i = 0; i < n; i++) {
      cpt[i] = L->mis[i] ? 1 : 0;
      if (cpt[i]) {
        for (int k = A->ia[i]; k < A->ia[i + 1]; k++) cls[k] = 1;
        continue;
      }
      if (mark_strong(L, i, mk) != P->ia[i + 1]) {
        bad = 1;
        unmark(L, i, mk);
        continue;
      }
      for (int k = A->ia[i]; k < A->ia[i + 1]; k++) {
        const int j = A->ja[k];
        if (j == i) {
          cls[k] = 2;
          continue;
        }
        const int m = mk[j];
        cls[k] = m == -1 ? 1 : 0;
        if (m >= 0) {
          pcnt[m]++;
        } else if (m == -2) {
          npr[i]++;
          for (int q = A->ia[j]; q < A->ia[j + 1]; q++) {
            const int m2 = mk[A->ja[q]];
            if (m2 >= 0) ndr[i]++, pcnt[m2]++;
          }
        }
      }
      unmark(L, i, mk);
    }
    free(mk);
  }
  int *rp = xcalloc_int((size_t)n + 1), *dbase = xcalloc_int((size_t)n + 1), *poff = xcalloc_int((size_t)np + 1);
  if (bad) {
    bn_set_error("interpolation pattern does not match the strong C points");
    rc = BN_ERR_ARG;
  } else if (scan_counts(npr, n, rp) || scan_counts(ndr, n, dbase) || scan_counts(pcnt, np, poff)) {
    bn_set_error("interpolation lists exceed 2^31 entries");
    rc = BN_ERR_ARG;
  }
  const int npairs = rc ? 0 : rp[n], nd = rc ? 0 : dbase[n], nt = rc ? 0 : poff[np];
  int *pair_k = xcalloc_int((size_t)npairs), *doff = xcalloc_int((size_t)npairs + 1), *dq = xcalloc_int((size_t)nd);
  int *prow = xcalloc_int((size_t)np), *pa = xcalloc_int((size_t)nt), *pq = xcalloc_int((size_t)nt);
  if (!rc) {
#pragma omp parallel
    {
      int *mk = (int *)bn_xmalloc((size_t)(n > 0 ? n : 1) * sizeof(int));
      int *cur = (int *)bn_xmalloc((size_t)maxp * sizeof(int));
      for (int j = 0; j < n; j++) mk[j] = -1;
#pragma omp for schedule(dynamic, 256)
      for (int i = 0; i < n; i++) {
        const int p0 = P->ia[i];
        for (int p = p0; p < P->ia[i + 1]; p++) prow[p] = i;
        if (cpt[i]) continue;
        mark_strong(L, i, mk);
        for (int p = p0; p < P->ia[i + 1]; p++) cur[p - p0] = poff[p];
        /* the direct entries first: the host adds them in a pass of its own */
        for (int k = A->ia[i]; k < A->ia[i + 1]; k++) {
          const int j = A->ja[k];
          if (j == i || mk[j] < 0) continue;
          const int t = cur[mk[j] - p0]++;
          pa[t] = k;
          pq[t] = -1;
        }
        int pair = rp[i], dpos = dbase[i];
        for (int k = A->ia[i]; k < A->ia[i + 1]; k++) {
          const int j = A->ja[k];
          if (j == i || mk[j] != -2) continue;
          pair_k[pair] = k;
          doff[pair] = dpos;
          for (int q = A->ia[j]; q < A->ia[j + 1]; q++) {
            const int m2 = mk[A->ja[q]];
            if (m2 < 0) continue;
            dq[dpos++] = q;
            const int t = cur[m2 - p0]++;
            pa[t] = pair;
            pq[t] = q;
          }
          pair++;
        }
        unmark(L, i, mk);
      }
      free(cur);
      free(mk);
    }
    doff[npairs] = nd;
    r->npairs = npairs;
    r->cls = rb_upload_bytes(s, cls, (size_t)A->nnz);
    r->cpt = rb_upload_int(s, cpt, (size_t)n);
    r->rp = rb_upload_int(s, rp, (size_t)n + 1);
    r->pair_k = rb_upload_int(s, pair_k, (size_t)npairs);
    r->doff = rb_upload_int(s, doff, (size_t)npairs + 1);
    r->dq = rb_upload_int(s, dq, (size_t)nd);
    r->prow = rb_upload_int(s, prow, (size_t)np);
    r->poff = rb_upload_int(s, poff, (size_t)np + 1);
    r->pa = rb_upload_int(s, pa, (size_t)nt);
    r->pq = rb_upload_int(s, pq, (size_t)nt);
    if ((long long)npairs + n > *iw) *iw = (long long)npairs + n;
  }
  free(npr);
  free(ndr);
  free(pcnt);
  free(cls);
  free(cpt);
  free(rp);
  free(dbase);
  free(poff);
  free(pair_k);
  free(doff);
  free(dq);
  free(prow);
  free(pa);
  free(pq);
  return rc;
}

/* Per coarse entry lists of the Galerkin product R A P (the host's
 * row-by-row product: R row, then A row, then P row), for a level whose
 * busiest coarse row has more than thresh terms (thresh < 0: never); a
 * lighter level keeps one thread per coarse row. */
static int plan_rap(struct bn_solver *s, const dcsr_mat *R, const dcsr_mat *A, const dcsr_mat *P, const dcsr_mat *C,
                    rb_level *r, long long thresh) {
  r->pull = 0;
  if (thresh < 0) return 0;
  const int nf = A->row, nc = C->row, nnz = C->nnz;
  long long *ap = (long long *)bn_xmalloc((size_t)(nf > 0 ? nf : 1) * sizeof(long long));
#pragma omp parallel for schedule(static)
  for (int i = 0; i < nf; i++) {
    long long t = 0;
    for (int q = A->ia[i]; q < A->ia[i + 1]; q++) t += P->ia[A->ja[q] + 1] - P->ia[A->ja[q]];
    ap[i] = t;
  }
  long long mx = 0, tot = 0;
#pragma omp parallel for schedule(static) reduction(max : mx) reduction(+ : tot)
  for (int I = 0; I < nc; I++) {
    long long t = 0;
    for (int q = R->ia[I]; q < R->ia[I + 1]; q++) t += ap[R->ja[q]];
    if (t > mx) mx = t;
    tot += t;
  }
  free(ap);
  if (mx <= thresh || tot > INT_MAX) return 0;
  int *cnt = xcalloc_int((size_t)nnz + 1), *toff = xcalloc_int((size_t)nnz + 1);
  for (int pass = 0; pass < 2; pass++) {
    if (pass == 1) {
      scan_counts(cnt, nnz, toff); /* below tot <= INT_MAX */
      memcpy(cnt, toff, (size_t)nnz * sizeof(int));
    }
    const int T = toff[nnz];
    int *t1 = NULL, *t2 = NULL, *t3 = NULL;
    if (pass == 1) t1 = xcalloc_int((size_t)T), t2 = xcalloc_int((size_t)T), t3 = xcalloc_int((size_t)T);
#pragma omp parallel
    {
      int *mk = (int *)bn_xmalloc((size_t)(nc > 0 ? nc : 1) * sizeof(int));
      for (int j = 0; j < nc; j++) mk[j] = -1;
#pragma omp for schedule(dynamic, 16)
      for (int I = 0; I < nc; I++) {
        for (int q = C->ia[I]; q < C->ia[I + 1]; q++) mk[C->ja[q]] = q;
        for (int q1 = R->ia[I]; q1 < R->ia[I + 1]; q1++) {
          const int i1 = R->ja[q1];
          for (int q2 = A->ia[i1]; q2 < A->ia[i1 + 1]; q2++) {
            const int i2 = A->ja[q2];
            for (int q3 = P->ia[i2]; q3 < P->ia[i2 + 1]; q3++) {
              const int m = mk[P->ja[q3]];
              if (m < 0) continue;
              if (pass == 0) {
                cnt[m]++;
              } else {
                const int t = cnt[m]++;
                t1[t] = q1, t2[t] = q2, t3[t] = q3;
              }
            }
          }
        }
        for (int q = C->ia[I]; q < C->ia[I + 1]; q++) mk[C->ja[q]] = -1;
      }
      free(mk);
    }
    if (pass == 1) {
      int nlong = 0;
      for (int e = 0; e < nnz; e++) nlong += toff[e + 1] - toff[e] > BN_RB_LONG;
      int *el = xcalloc_int((size_t)nlong);
      nlong = 0;
      for (int e = 0; e < nnz; e++)
        if (toff[e + 1] - toff[e] > BN_RB_LONG) el[nlong++] = e;
      r->toff = rb_upload_int(s, toff, (size_t)nnz + 1);
      r->t1 = rb_upload_int(s, t1, (size_t)T);
      r->t2 = rb_upload_int(s, t2, (size_t)T);
      r->t3 = rb_upload_int(s, t3, (size_t)T);
      r->elist = rb_upload_int(s, el, (size_t)nlong);
      r->nlong = nlong;
      r->pull = 1;
      free(el);
      free(t1);
      free(t2);
      free(t3);
    }
  }
  free(cnt);
  free(toff);
  return 0;
}

/* The fine IC(0) is formed on the device when the device factor exists
 * and every factor row is strictly ascending with its diagonal last (the
 * host factor is the operator's lower part in stored order); otherwise
 * the host forms it.  Maps: factor entry -> operator entry, and the
 * transposed rows' entries -> factor entry (as the upload transposed). */
static int plan_ic(struct bn_solver *s, rb_plan *pl) {
  const ichol_factor *F = &s->L_own;
  const dcsr_mat *A = s->A;
  if (!s->fine_ic.lo.ia || F->n != A->row || s->fine_ic.lo.nnz != F->nnz) return 0;
  for (int i = 0; i < F->n; i++) {
    const int lo = F->ia[i], hi = F->ia[i + 1];
    if (hi <= lo || F->ja[hi - 1] != i) return 0;
    for (int q = lo + 1; q < hi; q++)
      if (F->ja[q] <= F->ja[q - 1]) return 0;
  }
  int *lmap = (int *)bn_xmalloc((size_t)(F->nnz > 0 ? F->nnz : 1) * sizeof(int));
  int bad = 0;
  for (int i = 0; i < A->row && !bad; i++) {
    int p = F->ia[i];
    for (int k = A->ia[i]; k < A->ia[i + 1]; k++)
      if (A->ja[k] <= i) {
        if (p >= F->ia[i + 1] || F->ja[p] != A->ja[k]) bad = 1;
        else lmap[p++] = k;
      }
    if (p != F->ia[i + 1]) bad = 1;
  }
  if (!bad) {
    dcsr_mat Li = {F->n, F->n, F->nnz, F->ia, F->ja, NULL}, Ui;
    Li.val = (double *)bn_xmalloc((size_t)(F->nnz > 0 ? F->nnz : 1) * sizeof(double));
    for (int k = 0; k < F->nnz; k++) Li.val[k] = (double)k;
    bn_csr_transpose(&Li, &Ui);
    int *umap = (int *)bn_xmalloc((size_t)(Ui.nnz > 0 ? Ui.nnz : 1) * sizeof(int));
    for (int k = 0; k < Ui.nnz; k++) umap[k] = (int)Ui.val[k];
    pl->ic_lmap = rb_upload_int(s, lmap, (size_t)F->nnz);
    pl->ic_umap = rb_upload_int(s, umap, (size_t)Ui.nnz);
    free(umap);
    bn_csr_free(&Ui);
    free(Li.val);
  }
  free(lmap);
  if (!bad && (!pl->ic_lmap || !pl->ic_umap)) {
    bn_set_error("device memory for the rebuild plan");
    return BN_ERR_HIP;
  }
  return 0;
}

static int plan_build(struct bn_solver *s) {
  rb_plan *pl = (rb_plan *)calloc(1, sizeof(rb_plan));
  const amg_data *amg = s->amg;
  const int nlev = s->nlev;
  int rc = 0;
  /* product lists for levels whose busiest coarse row has more terms than
   * this (BN_RB_PULL; negative: never) */
  long long thresh = 2048;
  const char *env = getenv("BN_RB_PULL");
  if (env && *env) thresh = atoll(env);
  for (int l = 0; l < nlev && !rc; l++) {
    const amg_level *L = &amg->levels[l];
    const dcsr_mat *A = &L->A;
    rb_level *r = &pl->lev[l];
    if (A->nnz > pl->maxnnz) pl->maxnnz = A->nnz;
    int *col = (int *)bn_xmalloc((size_t)(A->nnz > 0 ? A->nnz : 1) * sizeof(int));
    int *pos = (int *)bn_xmalloc((size_t)(A->nnz > 0 ? A->nnz : 1) * sizeof(int));
    int *tp = (int *)bn_xmalloc((size_t)(A->nnz > 0 ? A->nnz : 1) * sizeof(int));
    sorted_rows(A, col, pos);
    if (transpose_positions(A, col, pos, tp)) {
      bn_set_error("level %d operator is not structurally symmetric", l);
      rc = BN_ERR_ARG;
    }
    if (!rc && l == 0) pl->tpos0 = rb_upload_int(s, tp, (size_t)A->nnz);
    if (!rc && l > 0) {
      if (!pl->lev[l - 1].pull) {
        r->c_col = rb_upload_int(s, col, (size_t)A->nnz);
        r->c_pos = rb_upload_int(s, pos, (size_t)A->nnz);
      }
      r->tpos = rb_upload_int(s, tp, (size_t)A->nnz);
    }
    free(col);
    free(pos);
    free(tp);
    /* colour-ordered GS rows: relabel a matrix whose values are the slot
     * numbers exactly as the upload relabelled the values */
    const bn_dlevel *d = &s->lev[l];
    if (!rc && d->use_gs) {
      int *cperm = (int *)bn_xmalloc((size_t)(d->n > 0 ? d->n : 1) * sizeof(int));
      if (hipMemcpy(cperm, d->cmap_d, (size_t)d->n * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) rc = BN_ERR_HIP;
      dcsr_mat I = *A, Ic;
      I.val = (double *)bn_xmalloc((size_t)(A->nnz > 0 ? A->nnz : 1) * sizeof(double));
      for (int k = 0; k < A->nnz; k++) I.val[k] = (double)k;
      bn_csr_offdiag_rows(&I, cperm, &Ic); /* the upload's off-diagonal copy */
      int *m = (int *)bn_xmalloc((size_t)(A->nnz > 0 ? A->nnz : 1) * sizeof(int));
      for (int k = 0; k < Ic.nnz; k++) m[k] = (int)Ic.val[k];
      if (!rc) r->acmap = rb_upload_int(s, m, (size_t)Ic.nnz);
      free(m);
      bn_csr_free(&Ic);
      free(I.val);
      free(cperm);
    }
    if (rc || l == nlev - 1) continue;
    if (!rc) rc = plan_interp(s, L, r, &pl->maxiw);
    /* R = P^T as the upload built it (stable transpose): entry -> P entry,
     * and the product lists on that R */
    if (!rc) {
      const dcsr_mat *P = &L->P;
      dcsr_mat Pi = *P, Ri;
      Pi.val = (double *)bn_xmalloc((size_t)(P->nnz > 0 ? P->nnz : 1) * sizeof(double));
      for (int k = 0; k < P->nnz; k++) Pi.val[k] = (double)k;
      bn_csr_transpose(&Pi, &Ri);
      int *m = (int *)bn_xmalloc((size_t)(Ri.nnz > 0 ? Ri.nnz : 1) * sizeof(int));
      for (int k = 0; k < Ri.nnz; k++) m[k] = (int)Ri.val[k];
      r->rmap = rb_upload_int(s, m, (size_t)Ri.nnz);
      free(m);
      rc = plan_rap(s, &Ri, A, P, &amg->levels[l + 1].A, r, thresh);
      bn_csr_free(&Ri);
      free(Pi.val);
    }
  }
  if (!rc) pl->work = (double *)rb_dev(s, (size_t)(pl->maxnnz > 0 ? pl->maxnnz : 1) * sizeof(double));
  if (!rc) pl->iw = (double *)rb_dev(s, (size_t)(pl->maxiw > 0 ? pl->maxiw : 1) * sizeof(double));
  for (int l = 0; l < nlev && !rc; l++) {
    const rb_level *r = &pl->lev[l];
    const int last = (l == nlev - 1);
    if ((l == 0 && !pl->tpos0) || (l > 0 && !r->tpos) || (l > 0 && !pl->lev[l - 1].pull && (!r->c_col || !r->c_pos)) ||
        (s->lev[l].use_gs && !r->acmap) ||
        (!last && (!r->cls || !r->cpt || !r->rp || !r->pair_k || !r->doff || !r->dq || !r->prow || !r->poff ||
                   !r->pa || !r->pq || !r->rmap)) ||
        (r->pull && (!r->toff || !r->t1 || !r->t2 || !r->t3 || !r->elist)))
      rc = BN_ERR_HIP;
  }
  if (!rc && s->coarse_dense && s->lev[nlev - 1].n > BN_DENSE_MAX) {
    bn_set_error("rebuild on the device: a dense coarsest solve of %d rows is beyond its %d-row limit",
                 s->lev[nlev - 1].n, BN_DENSE_MAX);
    rc = BN_ERR_ARG;
  }
  if (!rc && s->coarse_dense) {
    const size_t nc = (size_t)s->lev[nlev - 1].n;
    pl->dn_c = (double *)rb_dev(s, nc * nc * sizeof(double));
    pl->dn_r = (double *)rb_dev(s, nc * nc * sizeof(double));
    if (!pl->dn_c || !pl->dn_r) rc = BN_ERR_HIP;
  }
  if (!rc) {
    for (int l = 0; l < nlev; l++) pl->l1off[l + 1] = pl->l1off[l] + s->lev[l].n;
    pl->l1 = (double *)rb_dev(s, (size_t)(pl->l1off[nlev] > 0 ? pl->l1off[nlev] : 1) * sizeof(double));
  }
  if (!rc && (!pl->work || !pl->iw || !pl->l1)) rc = BN_ERR_HIP;
  if (!rc) rc = plan_ic(s, pl);
  if (rc) {
    free(pl);
    if (rc == BN_ERR_HIP) bn_set_error("device memory for the rebuild plan");
    return rc;
  }
  const char *lg = getenv("BN_RB_LOG");
  if (lg && *lg && *lg != '0')
    for (int l = 0; l + 1 < nlev; l++) {
      const rb_level *r = &pl->lev[l];
      int t = 0;
      if (r->pull && hipMemcpy(&t, r->toff + amg->levels[l + 1].A.nnz, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
        t = -1;
      fprintf(stderr, "[rebuild plan] level %d: %d strong F pairs; product %s", l, r->npairs,
              r->pull ? "per-entry lists" : "per coarse row");
      if (r->pull) fprintf(stderr, " (%d terms, %d entries over %d terms)", t, r->nlong, BN_RB_LONG);
      fprintf(stderr, "\n");
    }
  s->rb = pl;
  return 0;
}

/* host factor values of an IC(0) factor into its device copy (lower rows,
 * their transpose and the inverse diagonal, as the upload forms them) */
static int refresh_tri(struct bn_solver *s, const ichol_factor *F, bn_dtri *t) {
  const int n = F->n;
  dcsr_mat Lh = {n, n, F->nnz, F->ia, F->ja, F->val}, Uh;
  bn_csr_transpose(&Lh, &Uh);
  double *dinv = (double *)bn_xmalloc((size_t)(n > 0 ? n : 1) * sizeof(double));
  for (int i = 0; i < n; i++) {
    double d = 1.0;
    for (int k = Lh.ia[i]; k < Lh.ia[i + 1]; k++)
      if (Lh.ja[k] == i) {
        d = Lh.val[k];
        break;
      }
    dinv[i] = 1.0 / d;
  }
  const size_t nz = (size_t)F->nnz;
  float *fs = (float *)bn_xmalloc((nz > (size_t)n ? nz : (size_t)n) * sizeof(float) + 4);
  hipError_t e = hipMemcpy(t->lo.vd, Lh.val, nz * sizeof(double), hipMemcpyHostToDevice);
  for (size_t k = 0; k < nz; k++) fs[k] = (float)Lh.val[k];
  if (e == hipSuccess) e = hipMemcpy(t->lo.vs, fs, nz * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(t->up.vd, Uh.val, nz * sizeof(double), hipMemcpyHostToDevice);
  for (size_t k = 0; k < nz; k++) fs[k] = (float)Uh.val[k];
  if (e == hipSuccess) e = hipMemcpy(t->up.vs, fs, nz * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(t->dinv_d, dinv, (size_t)n * sizeof(double), hipMemcpyHostToDevice);
  for (int i = 0; i < n; i++) fs[i] = (float)dinv[i];
  if (e == hipSuccess) e = hipMemcpy(t->dinv_s, fs, (size_t)n * sizeof(float), hipMemcpyHostToDevice);
  free(fs);
  free(dinv);
  bn_csr_free(&Uh);
  return e == hipSuccess ? 0 : BN_ERR_HIP;
}

/* New values of the solve matrix (device memory, nnz entries in the
 * stored order, as bn_solver_set_values_device takes them), and a
 * value-only rebuild of the whole hierarchy from them on the device: the
 * same split and patterns, interpolation weights and Galerkin operators
 * recomputed bit for bit as amg_rebuild_values would, fine IC(0) and
 * coarsest solve refreshed. */
int bn_solver_rebuild_values_device(void *h, int nnz_in, const double *d_val) {
  struct bn_solver *s = (struct bn_solver *)h;
  if (!s || !d_val || s->borrowed) return BN_ERR_ARG;
  if (nnz_in != s->A->nnz) {
    bn_set_error("rebuild: %d values given, the stored solve matrix has %d", nnz_in, s->A->nnz);
    return BN_ERR_ARG;
  }
  int rc = bn_ensure_device(s);
  if (rc) return rc;
  const int nlev = s->nlev;
  amg_data *amg = s->amg;
  for (int l = 0; l < nlev; l++)
    if (s->lev[l].block) {
      bn_set_error("rebuild on the device does not support the block smoother layout");
      return BN_ERR_ARG;
    }
  if (amg->levels[0].A.nnz != s->A->nnz ||
      (!s->asolve_shared && (s->Asolve.ia != s->lev[0].A.ia || s->Asolve.ja != s->lev[0].A.ja))) {
    bn_set_error("rebuild: the solver holds a solve matrix apart from its hierarchy");
    return BN_ERR_ARG;
  }
  const char *lg = getenv("BN_RB_LOG");
  const int log = lg && *lg && *lg != '0';
  double tm[6];
  tm[0] = rb_ms();
  if (!s->vperm_state && (rc = bn_value_map(s))) return rc;
  if (!s->rb && (rc = plan_build(s))) return rc;
  rb_plan *pl = (rb_plan *)s->rb;
  tm[1] = rb_ms();
  hipStream_t st = s->stream;
  const int n0 = s->lev[0].n, nz0 = s->lev[0].A.nnz;

  /* fine values in stored order, symmetrised as the handle stores them */
  const double *in = d_val;
  if (s->vperm_state == 2) {
    bn_permute(nz0, s->vperm_d, d_val, pl->work, 0, st);
    in = pl->work;
  }
  if (s->opt.symmetrize) {
    if (in == pl->work) {
      /* symmetrise from a copy: the product scratch doubles as input */
      if (hipMemcpyAsync(s->lev[0].A.vd, in, (size_t)nz0 * sizeof(double), hipMemcpyDeviceToDevice, st) !=
          hipSuccess)
        return BN_ERR_HIP;
      bn_rb_symmetrize(nz0, pl->tpos0, s->lev[0].A.vd, pl->work, 0, st);
      if (hipMemcpyAsync(s->lev[0].A.vd, pl->work, (size_t)nz0 * sizeof(double), hipMemcpyDeviceToDevice, st) !=
          hipSuccess)
        return BN_ERR_HIP;
    } else {
      bn_rb_symmetrize(nz0, pl->tpos0, in, s->lev[0].A.vd, 0, st);
    }
  } else if (hipMemcpyAsync(s->lev[0].A.vd, in, (size_t)nz0 * sizeof(double), hipMemcpyDeviceToDevice, st) !=
             hipSuccess) {
    return BN_ERR_HIP;
  }

  /* level by level: interpolation weights, R = P^T, Galerkin product,
   * symmetrisation */
  for (int l = 0; l < nlev - 1; l++) {
    bn_dlevel *d = &s->lev[l], *c = &s->lev[l + 1];
    const rb_level *r = &pl->lev[l], *rc1 = &pl->lev[l + 1];
    double *den = pl->iw, *dhat = pl->iw + r->npairs;
    bn_rb_den(r->npairs, r->doff, r->dq, d->A.vd, den, st);
    bn_rb_dhat(d->n, d->A.ia, d->A.vd, r->cls, r->cpt, r->rp, r->pair_k, den, dhat, st);
    bn_rb_pval(d->P.nnz, r->prow, r->cpt, d->P.ia, r->poff, r->pa, r->pq, r->pair_k, d->A.vd, den, dhat, d->P.vd,
               st);
    bn_rb_to_float(d->P.nnz, d->P.vd, d->P.vs, st);
    bn_rb_gather(d->R.nnz, r->rmap, d->P.vd, d->R.vd, d->R.vs, st);
    if (r->pull)
      bn_rb_rap_pull(c->A.nnz, r->toff, r->t1, r->t2, r->t3, r->nlong, r->elist, d->R.vd, d->A.vd, d->P.vd,
                     pl->work, st);
    else
      bn_rb_rap(c->n, d->R.ia, d->R.ja, d->R.vd, d->A.ia, d->A.ja, d->A.vd, d->P.ia, d->P.ja, d->P.vd, c->A.ia,
                rc1->c_col, rc1->c_pos, pl->work, st);
    bn_rb_symmetrize(c->A.nnz, rc1->tpos, pl->work, c->A.vd, 1, st);
  }
  /* derived per-level arrays, and the l1 row norms of the host set-up */
  for (int l = 0; l < nlev; l++) {
    bn_dlevel *d = &s->lev[l];
    bn_rb_to_float(d->A.nnz, d->A.vd, d->A.vs, st);
    bn_rb_row_scalars(d->n, d->A.ia, d->A.ja, d->A.vd, d->dinv_d, d->dinv_s, d->l1inv_d, d->l1inv_s,
                      pl->l1 + pl->l1off[l], st);
    if (d->use_gs) {
      bn_rb_gather(d->Ac.nnz, pl->lev[l].acmap, d->A.vd, d->Ac.vd, d->Ac.vs, st);
      /* 1/a_ii in colour order: the level's own 1/a_ii gathered by the row map */
      bn_rb_gather(d->n, d->cmap_d, d->dinv_d, d->dinvc_d, d->dinvc_s, st);
    }
  }
  if (!s->asolve_shared) {
    if (hipMemcpyAsync(s->Asolve.vd, s->lev[0].A.vd, (size_t)nz0 * sizeof(double), hipMemcpyDeviceToDevice, st) !=
        hipSuccess)
      return BN_ERR_HIP;
    if (s->Asolve.vs != s->lev[0].A.vs) bn_rb_to_float(nz0, s->Asolve.vd, s->Asolve.vs, st);
  }
  hipEvent_t ev;
  if (hipEventCreateWithFlags(&ev, hipEventDisableTiming) != hipSuccess) return BN_ERR_HIP;
  hipError_t e = hipEventRecord(ev, st);

  /* then, on the device, the fine IC(0) (ichol_compute's entries) and the
   * coarsest inverse (bn_dense_inverse's entries) */
  bn_dtri *t = &s->fine_ic;
  if (pl->ic_lmap) {
    const int nzl = t->lo.nnz;
    bn_rb_gather(nzl, pl->ic_lmap, s->lev[0].A.vd, t->lo.vd, NULL, st);
    for (int lv = 0; lv < t->nlev_f; lv++)
      bn_ic0_rows(t->ptr_f[lv + 1] - t->ptr_f[lv], t->rows_f + t->ptr_f[lv], t->lo.ia, t->lo.ja, t->lo.vd, st);
    bn_rb_to_float(nzl, t->lo.vd, t->lo.vs, st);
    bn_rb_gather(t->up.nnz, pl->ic_umap, t->lo.vd, t->up.vd, t->up.vs, st);
    bn_rb_row_scalars(t->n, t->lo.ia, t->lo.ja, t->lo.vd, t->dinv_d, t->dinv_s, NULL, NULL, NULL, st);
  }
  const bn_dlevel *dc = &s->lev[nlev - 1];
  if (s->coarse_dense &&
      bn_dn_inverse(dc->n, dc->A.ia, dc->A.ja, dc->A.vd, pl->dn_c, pl->dn_r, s->cinv_d, s->cinv_s, NULL, st))
    e = hipErrorInvalidValue;

  /* meanwhile the host hierarchy follows: level and interpolation values,
   * l1 norms */
  if (e == hipSuccess) e = hipEventSynchronize(ev);
  hipEventDestroy(ev);
  if (e != hipSuccess) {
    hipStreamSynchronize(st);
    bn_set_error("device failure in the value-only rebuild");
    return BN_ERR_HIP;
  }
  tm[2] = rb_ms();
  for (int l = 0; l < nlev && e == hipSuccess; l++) {
    amg_level *L = &amg->levels[l];
    e = hipMemcpy(L->A.val, s->lev[l].A.vd, (size_t)L->A.nnz * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess && l < nlev - 1)
      e = hipMemcpy(L->P.val, s->lev[l].P.vd, (size_t)L->P.nnz * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess && L->l1_diag)
      e = hipMemcpy(L->l1_diag, pl->l1 + pl->l1off[l], (size_t)L->A.row * sizeof(double), hipMemcpyDeviceToHost);
  }
  if (e != hipSuccess) {
    hipStreamSynchronize(st);
    return BN_ERR_HIP;
  }
  memcpy(s->A_own.val, amg->levels[0].A.val, (size_t)nz0 * sizeof(double));
  (void)n0;
  tm[3] = rb_ms();

  /* host-side remainder: the coarsest IC(0) factor (the coarsest solve
   * when it is not the dense inverse), the fine IC(0) where the device
   * does not form it */
  amg_level *C = &amg->levels[nlev - 1];
  ichol_free(&C->L_ichol);
  ichol_compute(&C->A, &C->L_ichol);
  if (!pl->ic_lmap) {
    ichol_free(&s->L_own);
    ichol_compute(s->A, &s->L_own);
  }
  if (hipStreamSynchronize(st) != hipSuccess) {
    bn_set_error("device failure in the fine IC(0) or the coarsest inverse");
    return BN_ERR_HIP;
  }
  tm[4] = rb_ms();
  if (pl->ic_lmap) {
    if (hipMemcpy(s->L_own.val, t->lo.vd, (size_t)t->lo.nnz * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
      return BN_ERR_HIP;
  } else if (t->lo.ia && (rc = refresh_tri(s, &s->L_own, t))) {
    return rc;
  }
  if (!s->coarse_dense && (rc = refresh_tri(s, &C->L_ichol, &s->coarse_ic))) return rc;
  /* the exact solves' ELL copies follow their factors' new values */
  bn_tri_ell_refresh(s, &s->fine_ic);
  if (!s->coarse_dense) bn_tri_ell_refresh(s, &s->coarse_ic);
  tm[5] = rb_ms();
  if (log)
    fprintf(stderr,
            "[rebuild] plan %.1f ms, device levels %.1f ms, host copy of the hierarchy %.1f ms, fine IC(0) (%s) and "
            "coarsest %s (%d rows) %.1f ms more, factor copies %.1f ms\n",
            tm[1] - tm[0], tm[2] - tm[1], tm[3] - tm[2], pl->ic_lmap ? "device" : "host",
            s->coarse_dense ? "inverse" : "IC(0)", C->A.row, tm[4] - tm[3], tm[5] - tm[4]);
  return 0;
}

/* The coarsest inverse of a host CSR matrix formed on the device into
 * cinv_d / cinv_s (entries of bn_dense_inverse); the matrix copy and the
 * dense factor scratch are released before returning.  *clamped as
 * bn_dense_inverse returns it. */
int bn_coarse_inverse_device(struct bn_solver *s, const dcsr_mat *A, double *cinv_d, float *cinv_s, int *clamped) {
  const int n = A->row, nnz = A->nnz;
  const size_t nn = (size_t)n * (size_t)n;
  double *Lc = NULL, *Lr = NULL, *va = NULL;
  int *ia = NULL, *ja = NULL, *flag = NULL, zero = 0, rc = 0;
  hipError_t e = hipMalloc((void **)&Lc, nn * sizeof(double) + 8);
  if (e == hipSuccess) e = hipMalloc((void **)&Lr, nn * sizeof(double) + 8);
  if (e == hipSuccess) e = hipMalloc((void **)&ia, ((size_t)n + 1) * sizeof(int));
  if (e == hipSuccess) e = hipMalloc((void **)&ja, (size_t)(nnz > 0 ? nnz : 1) * sizeof(int));
  if (e == hipSuccess) e = hipMalloc((void **)&va, (size_t)(nnz > 0 ? nnz : 1) * sizeof(double));
  if (e == hipSuccess) e = hipMalloc((void **)&flag, sizeof(int));
  if (e == hipSuccess) e = hipMemcpy(ia, A->ia, ((size_t)n + 1) * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess && nnz) e = hipMemcpy(ja, A->ja, (size_t)nnz * sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess && nnz) e = hipMemcpy(va, A->val, (size_t)nnz * sizeof(double), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(flag, &zero, sizeof(int), hipMemcpyHostToDevice);
  if (e == hipSuccess && bn_dn_inverse(n, ia, ja, va, Lc, Lr, cinv_d, cinv_s, flag, s->stream)) rc = BN_ERR_ARG;
  if (e == hipSuccess && !rc) e = hipStreamSynchronize(s->stream);
  if (e == hipSuccess && !rc) e = hipMemcpy(&zero, flag, sizeof(int), hipMemcpyDeviceToHost);
  if (clamped) *clamped = zero;
  hipFree(Lc);
  hipFree(Lr);
  hipFree(ia);
  hipFree(ja);
  hipFree(va);
  hipFree(flag);
  if (e != hipSuccess) {
    bn_set_error("device failure in the coarsest inverse (%d rows)", n);
    return BN_ERR_HIP;
  }
  if (rc) bn_set_error("coarsest inverse: %d rows is beyond the dense limit", n);
  return rc;
}

