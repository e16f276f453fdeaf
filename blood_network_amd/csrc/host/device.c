/*
 * device.c - upload of the host hierarchy to the GPU.
 *
 * Levels keep the hierarchy's numbering, whose neighbouring rows share x
 * cache lines in every operator of the level.  A level swept by the
 * multicolour Gauss-Seidel smoother also gets a copy of its rows in colour
 * order (C colours, then F colours) without the diagonal, so each colour's
 * kernel streams one contiguous CSR segment and scatters into the level
 * vectors.  (Renumbering such a level in colour order instead measured
 * slower: its residual, restriction and interpolation lose the locality of
 * their x gathers and the sweep gains little.)  Only the chunked layout
 * of the block smoother relabels levels, P and R = P^T on both sides and
 * the solve matrix and its IC(0) factor in level-0 numbering.  Operators
 * are stored in FP64 and FP32 so the V-cycle precision can be switched
 * without re-uploading.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "bn_solver.h"
#include "csr.h"
#include "schedule.h"

int bn_dense_inverse(const dcsr_mat *A, double *inv);
void bn_set_error(const char *fmt, ...);
void bn_track_device(struct bn_solver *s, void *p);
void bn_track_host(struct bn_solver *s, void *p);

static int g_device_state = -1; /* -1 unknown, 0 none, 1 usable */

int bn_device_ok(void) {
  if (g_device_state < 0) {
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    g_device_state = (e == hipSuccess && n > 0) ? 1 : 0;
  }
  return g_device_state;
}

#define CK(call)                                                                   \
  do {                                                                             \
    hipError_t e_ = (call);                                                        \
    if (e_ != hipSuccess) {                                                        \
      bn_set_error("HIP error %d (%s) at %s:%d", (int)e_, hipGetErrorString(e_),   \
                   __FILE__, __LINE__);                                            \
      return BN_ERR_HIP;                                                           \
    }                                                                              \
  } while (0)

static void *dmalloc(struct bn_solver *s, size_t bytes) {
  void *p = NULL;
  if (hipMalloc(&p, bytes ? bytes : 8) != hipSuccess) return NULL;
  bn_track_device(s, p);
  return p;
}

/* Threads per CSR row: the smallest power of two that covers the mean
 * row in one round of four strides (about four entries per lane, so a
 * wave keeps several rows' loads in flight), raised until the launch has
 * about 2^17 threads (small coarse operators otherwise leave CUs idle),
 * at most one wave64 per row. */
static int lanes_for(long long nnz, long long rows) {
  double avg = rows > 0 ? (double)nnz / (double)rows : 1.0;
  int l = 1;
  while (4 * l < avg && l < 64) l *= 2;
  while (l < 64 && (long long)l * rows < (1LL << 17) && l < avg) l *= 2;
  return l;
}

int bn_lanes_for(long long nnz, long long rows) { return lanes_for(nnz, rows); }

static int cmp_desc_ll(const void *a, const void *b) {
  const long long x = *(const long long *)a, y = *(const long long *)b;
  return (x < y) - (x > y);
}

static int is_identity(const int *perm, int n) {
  for (int i = 0; i < n; i++)
    if (perm[i] != i) return 0;
  return 1;
}

static int upload_csr(struct bn_solver *s, const dcsr_mat *m, bn_dcsr *d) {
  d->rows = m->row;
  d->cols = m->col;
  d->nnz = m->nnz;
  d->lanes = d->lanes_s = lanes_for(m->nnz, m->row);
  d->deep = d->deep_s = 0;
  d->ia = (int *)dmalloc(s, ((size_t)m->row + 1) * sizeof(int));
  d->ja = (int *)dmalloc(s, (size_t)m->nnz * sizeof(int));
  d->vd = (double *)dmalloc(s, (size_t)m->nnz * sizeof(double));
  d->vs = (float *)dmalloc(s, (size_t)m->nnz * sizeof(float));
  if (!d->ia || !d->ja || !d->vd || !d->vs) {
    bn_set_error("device out of memory uploading a %d x %d CSR (%d nnz)", m->row, m->col, m->nnz);
    return BN_ERR_HIP;
  }
  /* row-length bins: rows over 16 strides of the lane group go to the
   * block-per-row kernel (coarse Galerkin rows are heavy-tailed).  One
   * list, longest rows first, serves every lanes choice as a prefix. */
  d->long_rows = NULL;
  d->seg_ptr = d->seg_row = d->seg_ptr_h = NULL;
  d->seg_part = NULL;
  d->seg_cnt = NULL;
  d->nmulti_all = 0;
  int nl = 0;
  for (int i = 0; i < m->row; i++) nl += (m->ia[i + 1] - m->ia[i] > bn_long_len(1));
  for (int k = 0; k < 7; k++) d->nlong_for[k] = 0;
  if (nl > 0) {
    long long *key = (long long *)bn_xmalloc((size_t)nl * sizeof(long long));
    int q = 0;
    for (int i = 0; i < m->row; i++) {
      const int len = m->ia[i + 1] - m->ia[i];
      if (len > bn_long_len(1)) key[q++] = ((long long)len << 32) | (long long)(0x7fffffff - i);
    }
    qsort(key, (size_t)nl, sizeof(long long), cmp_desc_ll);
    int *lr = (int *)bn_xmalloc((size_t)nl * sizeof(int));
    int *sp = (int *)bn_xmalloc(((size_t)nl + 1) * sizeof(int));
    sp[0] = 0;
    d->nmulti_all = 0;
    for (int t = 0; t < nl; t++) {
      lr[t] = 0x7fffffff - (int)(key[t] & 0xffffffffLL);
      const int len = (int)(key[t] >> 32);
      for (int k = 0; k < 7; k++)
        if (len > bn_long_len(1 << k)) d->nlong_for[k] = t + 1;
      sp[t + 1] = sp[t] + (len + BN_SEG_NNZ - 1) / BN_SEG_NNZ;
      if (len > BN_SEG_NNZ) d->nmulti_all = t + 1;
    }
    free(key);
    int *sr = (int *)bn_xmalloc((size_t)sp[nl] * sizeof(int));
    for (int t = 0; t < nl; t++)
      for (int q = sp[t]; q < sp[t + 1]; q++) sr[q] = t;
    d->long_rows = (int *)dmalloc(s, (size_t)nl * sizeof(int));
    d->seg_ptr = (int *)dmalloc(s, ((size_t)nl + 1) * sizeof(int));
    d->seg_row = (int *)dmalloc(s, (size_t)sp[nl] * sizeof(int));
    d->seg_part = (double *)dmalloc(s, (size_t)(sp[d->nmulti_all] > 0 ? sp[d->nmulti_all] : 1) * sizeof(double));
    d->seg_cnt = (int *)dmalloc(s, (size_t)(d->nmulti_all > 0 ? d->nmulti_all : 1) * sizeof(int));
    if (!d->long_rows || !d->seg_ptr || !d->seg_row || !d->seg_part || !d->seg_cnt) {
      free(lr);
      free(sp);
      free(sr);
      return BN_ERR_HIP;
    }
    CK(hipMemcpy(d->long_rows, lr, (size_t)nl * sizeof(int), hipMemcpyHostToDevice));
    CK(hipMemcpy(d->seg_ptr, sp, ((size_t)nl + 1) * sizeof(int), hipMemcpyHostToDevice));
    CK(hipMemcpy(d->seg_row, sr, (size_t)sp[nl] * sizeof(int), hipMemcpyHostToDevice));
    bn_seg_reset(sp[d->nmulti_all], d->seg_part, d->nmulti_all, d->seg_cnt, s->stream);
    free(lr);
    free(sr);
    d->seg_ptr_h = sp; /* kept for the per-lanes segment counts */
    bn_track_host(s, sp);
  }
  /* row blocks of <= BN_ROWBLOCK_NNZ entries for the adaptive kernel */
  {
    int *bk = (int *)bn_xmalloc(((size_t)m->row + 1) * sizeof(int));
    d->nblk = bn_row_blocks(m->ia, m->row, BN_ROWBLOCK_NNZ, BN_ROWBLOCK_ROWS, bk);
    d->blk = (int *)dmalloc(s, ((size_t)d->nblk + 1) * sizeof(int));
    d->blkq = (int *)dmalloc(s, ((size_t)d->nblk + 1) * sizeof(int));
    d->bpart = (double *)dmalloc(s, ((size_t)d->nblk + 1) * sizeof(double));
    if (!d->blk || !d->blkq || !d->bpart) {
      free(bk);
      return BN_ERR_HIP;
    }
    CK(hipMemcpy(d->blk, bk, ((size_t)d->nblk + 1) * sizeof(int), hipMemcpyHostToDevice));
    for (int k = 0; k <= d->nblk; k++) bk[k] = m->ia[bk[k]];
    CK(hipMemcpy(d->blkq, bk, ((size_t)d->nblk + 1) * sizeof(int), hipMemcpyHostToDevice));
    free(bk);
    d->adapt = d->adapt_s = 0;
  }
  CK(hipMemcpy(d->ia, m->ia, ((size_t)m->row + 1) * sizeof(int), hipMemcpyHostToDevice));
  CK(hipMemcpy(d->ja, m->ja, (size_t)m->nnz * sizeof(int), hipMemcpyHostToDevice));
  CK(hipMemcpy(d->vd, m->val, (size_t)m->nnz * sizeof(double), hipMemcpyHostToDevice));
  /* the FP32 copy is rounded on the device (same round-to-nearest cast) */
  bn_rb_to_float(m->nnz, d->vd, d->vs, s->stream);
  CK(hipStreamSynchronize(s->stream));
  return 0;
}

static int upload_vec2(struct bn_solver *s, const double *h, int n, double **vd, float **vs) {
  *vd = (double *)dmalloc(s, (size_t)n * sizeof(double));
  *vs = (float *)dmalloc(s, (size_t)n * sizeof(float));
  if (!*vd || !*vs) {
    bn_set_error("device out of memory (vector of %d)", n);
    return BN_ERR_HIP;
  }
  float *tmp = (float *)bn_xmalloc((size_t)n * sizeof(float));
  for (int i = 0; i < n; i++) tmp[i] = (float)h[i];
  CK(hipMemcpy(*vd, h, (size_t)n * sizeof(double), hipMemcpyHostToDevice));
  CK(hipMemcpy(*vs, tmp, (size_t)n * sizeof(float), hipMemcpyHostToDevice));
  free(tmp);
  return 0;
}

/* diagonal inverse as the GS row update uses it: 1/a_ii, 1.0 if absent
 * (C_amg1/src/amg_cycle.c:57); l1 inverse 1/sum|a_ij| */
static void diag_inverses(const dcsr_mat *A, double *dinv, double *l1inv) {
#pragma omp parallel for schedule(static) if (A->nnz > 65536)
  for (int i = 0; i < A->row; i++) {
    double d = 1.0, l1 = 0.0;
    for (int k = A->ia[i]; k < A->ia[i + 1]; k++) {
      if (A->ja[k] == i) d = A->val[k];
      l1 += fabs(A->val[k]);
    }
    dinv[i] = 1.0 / d;
    if (l1inv) l1inv[i] = l1 > 0.0 ? 1.0 / l1 : 1.0;
  }
}

/* Upload a triangular factor L (rows hold columns <= i, diagonal
 * included) given in original numbering, relabelled by perm/inv. */
static int upload_tri(struct bn_solver *s, const ichol_factor *F, const int *perm, const int *inv,
                      bn_dtri *t) {
  const int n = F->n;
  dcsr_mat Lh = {n, n, F->nnz, F->ia, F->ja, F->val};
  dcsr_mat Uh;
  bn_csr_transpose(&Lh, &Uh);
  int *rows_f = (int *)bn_xmalloc((size_t)n * sizeof(int));
  int *rows_b = (int *)bn_xmalloc((size_t)n * sizeof(int));
  t->nlev_f = bn_level_sets(Lh.ia, Lh.ja, n, 1, rows_f, &t->ptr_f);
  t->nlev_b = bn_level_sets(Uh.ia, Uh.ja, n, 0, rows_b, &t->ptr_b);
  double *dinv = (double *)bn_xmalloc((size_t)n * sizeof(double));
  for (int i = 0; i < n; i++) {
    double d = 1.0;
    for (int k = Lh.ia[i]; k < Lh.ia[i + 1]; k++)
      if (Lh.ja[k] == i) { d = Lh.val[k]; break; }
    dinv[i] = 1.0 / d;
  }
  dcsr_mat Lp, Up;
  double *dinv_p = (double *)bn_xmalloc((size_t)n * sizeof(double));
  if (perm) {
    bn_csr_permute(&Lh, perm, inv, &Lp);
    bn_csr_permute(&Uh, perm, inv, &Up);
    for (int k = 0; k < n; k++) dinv_p[k] = dinv[perm[k]];
    for (int k = 0; k < n; k++) {
      rows_f[k] = inv[rows_f[k]];
      rows_b[k] = inv[rows_b[k]];
    }
  } else {
    Lp = Lh; /* read only: no copy */
    Up = Uh;
    memcpy(dinv_p, dinv, (size_t)n * sizeof(double));
  }
  int rc = upload_csr(s, &Lp, &t->lo);
  if (!rc) rc = upload_csr(s, &Up, &t->up);
  if (!rc) rc = upload_vec2(s, dinv_p, n, &t->dinv_d, &t->dinv_s);
  if (!rc) {
    t->rows_f = (int *)dmalloc(s, (size_t)n * sizeof(int));
    t->rows_b = (int *)dmalloc(s, (size_t)n * sizeof(int));
    if (!t->rows_f || !t->rows_b) rc = BN_ERR_HIP;
  }
  if (!rc) {
    CK(hipMemcpy(t->rows_f, rows_f, (size_t)n * sizeof(int), hipMemcpyHostToDevice));
    CK(hipMemcpy(t->rows_b, rows_b, (size_t)n * sizeof(int), hipMemcpyHostToDevice));
  }
  t->n = n;
  if (perm) {
    bn_csr_free(&Lp);
    bn_csr_free(&Up);
  }
  bn_csr_free(&Uh);
  free(dinv);
  free(dinv_p);
  free(rows_f);
  free(rows_b);
  return rc;
}

/* Autotuner: time the operator's SpMV for every lanes-per-row choice
 * (1..64; from 4 lanes up also with 8 and 16 load rounds) whose long-row
 * bin stays small, and the row-block kernel, in both precisions, and
 * keep the fastest.
 * Every timed launch starts cold: a read of a buffer larger than the
 * last-level cache first evicts the operator, as the rest of a V-cycle
 * does between two uses of a coarse operator (timing repeated launches
 * would serve operators below the cache size from cache and favour
 * kernels that are slow from HBM).  x and y are scratch device buffers
 * of at least cols / rows doubles; flush (nflush doubles) may be NULL. */
#define BN_TUNE_REPS 5
typedef struct {
  const double *flush;
  int nflush;
  hipEvent_t ev[2 * BN_TUNE_REPS + 2];
} bn_tune_env;

/* total device time of reps cold launches (one warm-up launch first);
 * every launch has its own event pair, read after one synchronisation */
static float time_launches(struct bn_solver *s, const bn_tune_env *env, bn_dcsr *d, char vt, int cand,
                           void *x, void *y, int reps) {
  const void *val = vt == 's' ? (const void *)d->vs : (const void *)d->vd;
  for (int r = -1; r < reps; r++) {
    if (env->flush) bn_dot_partial(env->nflush, env->flush, env->flush, NULL, NULL, s->part1, NULL, s->stream);
    if (r >= 0) hipEventRecord(env->ev[2 * r], s->stream);
    if (cand > 0)
      bn_dcsr_apply(d, vt, BN_OP_SET, cand, x, y, NULL, s->stream);
    else
      bn_csr_apply_blocks(vt, BN_OP_SET, d->nblk, d->blk, d->blkq, d->ia, d->ja, val, x, y, NULL, s->stream);
    if (r >= 0) hipEventRecord(env->ev[2 * r + 1], s->stream);
  }
  hipEventSynchronize(env->ev[2 * reps - 1]);
  float total = 0.0f;
  for (int r = 0; r < reps; r++) {
    float ms = 0.0f;
    hipEventElapsedTime(&ms, env->ev[2 * r], env->ev[2 * r + 1]);
    total += ms;
  }
  return total;
}

/* Lanes per row of a GS level's colour kernels: each option times whole
 * cold forward sweeps (all colours back to back, as the V-cycle runs
 * them), FP64 and FP32 */
static void tune_gs(struct bn_solver *s, const bn_tune_env *env, bn_dlevel *d, void *x, void *b, float *best_ms) {
  enum { NOPT = 6 };
  const int opts[NOPT] = {1, 2, 4, 8, 16, BN_GS_BLOCKS};
  /* BN_GS_LANES pins the colour kernels' lanes per row (1, 2, 4, 8 or
   * 16, or 1024 for the row-block kernel): one value for every level, or a
   * comma-separated list by level (levels past the list, or with 0, are
   * timed) */
  const char *pin = getenv("BN_GS_LANES");
  if (pin && pin[0]) {
    const int lev = (int)(d - s->lev);
    int l = atoi(pin);
    if (strchr(pin, ',')) {
      const char *c = pin;
      for (int k = 0; k < lev && c; k++) {
        c = strchr(c, ',');
        if (c) c++;
      }
      l = c ? atoi(c) : 0;
    }
    for (int o = 0; o < NOPT; o++)
      if (opts[o] == l) {
        d->gs_lanes = d->gs_lanes_s = l;
        best_ms[0] = best_ms[1] = -1.0f; /* pinned: not timed */
        return;
      }
  }
  for (int p = 0; p < 2; p++) {
    const char vt = p ? 's' : 'd';
    int best_l = 0;
    float best = 1e30f, cand_ms[NOPT];
    /* two passes over the candidates, each keeping its better pass, so
     * that a drift in clocks or cache state across the candidate order
     * does not decide between two close choices */
    for (int o10 = 0; o10 < 2 * NOPT; o10++) {
      const int o = o10 % NOPT, pass = o10 / NOPT;
      const int lanes = opts[o];
      for (int r = -1; r < BN_TUNE_REPS; r++) {
        if (env->flush) bn_dot_partial(env->nflush, env->flush, env->flush, NULL, NULL, s->part1, NULL, s->stream);
        if (r >= 0) hipEventRecord(env->ev[2 * r], s->stream);
        for (int c = 0; c < d->ncol; c++) bn_gs_colour(s, d, vt, c, lanes, b, x, NULL, NULL);
        if (r >= 0) hipEventRecord(env->ev[2 * r + 1], s->stream);
      }
      hipEventSynchronize(env->ev[2 * BN_TUNE_REPS - 1]);
      float total = 0.0f;
      for (int r = 0; r < BN_TUNE_REPS; r++) {
        float ms = 0.0f;
        hipEventElapsedTime(&ms, env->ev[2 * r], env->ev[2 * r + 1]);
        total += ms;
      }
      const char *lg = getenv("BN_TUNE_LOG");
      if (lg && lg[0] == '2')
        fprintf(stderr, "  gs %c rows=%d colours=%d lanes=%d pass %d: %.1f us per cold sweep\n", vt, d->n, d->ncol,
                lanes, pass, 1e3 * total / BN_TUNE_REPS);
      if (pass == 0 || total < cand_ms[o]) cand_ms[o] = total;
    }
    for (int o = 0; o < NOPT; o++)
      if (cand_ms[o] < best) best = cand_ms[o], best_l = opts[o];
    const char *lg = getenv("BN_TUNE_LOG");
    if (lg && (lg[0] == '1' || lg[0] == '2'))
      fprintf(stderr, "tune gs %c rows=%d colours=%d: lanes=%d (%.1f us per cold sweep)\n", vt, d->n, d->ncol, best_l,
              1e3 * best / BN_TUNE_REPS);
    if (p) d->gs_lanes_s = best_l;
    else d->gs_lanes = best_l;
    best_ms[p] = best;
  }
}

/* the one-launch sweep of a GS level (bn_gs_walk) with 1, 2, 4 or 8 lanes
 * per row, timed like the colour launches (whole cold forward sweeps from a
 * nonzero start, the input copy included); kept where it beats their best
 * time.  BN_GS_WALK=0 keeps a launch per colour; BN_GS_WALK=1/2/4/8 pins
 * the lanes of the one-launch sweep on every level that has one. */
static void tune_gs_walk(struct bn_solver *s, const bn_tune_env *env, bn_dlevel *d, void *x, void *b,
                         const float *best_ms) {
  d->gs_walk = d->gs_walk_s = 0;
  if (!d->gs_jx || !s->walk_ok) return;
  const char *pin = getenv("BN_GS_WALK");
  if (pin && pin[0]) {
    const int l = atoi(pin);
    if (l == 1 || l == 2 || l == 4 || l == 8) d->gs_walk = d->gs_walk_s = l;
    return;
  }
  if (best_ms[0] < 0.0f) return; /* colour lanes pinned: keep the colour launches */
  const int opts[4] = {1, 2, 4, 8};
  double *ctl = s->scal + BN_S_WALKERR;
  for (int p = 0; p < 2; p++) {
    const char vt = p ? 's' : 'd';
    const void *dinvc = vt == 's' ? (const void *)d->dinvc_s : (const void *)d->dinvc_d;
    const void *val = vt == 's' ? (const void *)d->Ac.vs : (const void *)d->Ac.vd;
    int best_l = 0;
    float best = best_ms[p];
    for (int o = 0; o < 4; o++) {
      bn_fill('d', 1, ctl, 0.0, s->stream);
      for (int r = -1; r < BN_TUNE_REPS; r++) {
        if (env->flush) bn_dot_partial(env->nflush, env->flush, env->flush, NULL, NULL, s->part1, NULL, s->stream);
        if (r >= 0) hipEventRecord(env->ev[2 * r], s->stream);
        bn_gs_prep(vt, d->n, x, d->gs_xold, s->stream);
        bn_gs_walk(vt, opts[o], d->gs_wnp, d->ncol, 0, d->gs_wsteps, d->Ac.ia, d->gs_jx, val, d->cmap_d, dinvc, b,
                   x, d->gs_xold, 0, NULL, NULL, ctl, s->walk_limit, s->stream);
        if (r >= 0) hipEventRecord(env->ev[2 * r + 1], s->stream);
      }
      hipEventSynchronize(env->ev[2 * BN_TUNE_REPS - 1]);
      double flag = 0.0;
      hipMemcpy(&flag, ctl, sizeof(double), hipMemcpyDeviceToHost);
      float total = 0.0f;
      for (int r = 0; r < BN_TUNE_REPS; r++) {
        float ms = 0.0f;
        hipEventElapsedTime(&ms, env->ev[2 * r], env->ev[2 * r + 1]);
        total += ms;
      }
      const char *lg = getenv("BN_TUNE_LOG");
      if (lg && lg[0] == '2')
        fprintf(stderr, "  gs walk %c rows=%d colours=%d lanes=%d: %.1f us per cold sweep%s\n", vt, d->n, d->ncol,
                opts[o], 1e3 * total / BN_TUNE_REPS, flag != 0.0 ? " (gave up)" : "");
      if (flag != 0.0) {
        bn_fill('d', 1, ctl, 0.0, s->stream);
        best_l = 0;
        break; /* a wait gave up: the colour launches serve */
      }
      if (total < best) best = total, best_l = opts[o];
    }
    const char *lg = getenv("BN_TUNE_LOG");
    if (lg && (lg[0] == '1' || lg[0] == '2'))
      fprintf(stderr, "tune gs walk %c rows=%d colours=%d: lanes=%d (%.1f us per cold sweep; colour launches %.1f)\n",
              vt, d->n, d->ncol, best_l, 1e3 * best / BN_TUNE_REPS, 1e3 * best_ms[p] / BN_TUNE_REPS);
    if (p) d->gs_walk_s = best_l;
    else d->gs_walk = best_l;
  }
  hipStreamSynchronize(s->stream);
}

/* Without the autotuner the colour kernels run Ac's SpMV lanes and no
 * sweep runs in one launch, unless BN_GS_LANES / BN_GS_WALK pin them (as
 * with the tuner; one value for every level) */
static void gs_pins(struct bn_solver *s) {
  const char *pl = getenv("BN_GS_LANES"), *pw = getenv("BN_GS_WALK");
  const int l = pl ? atoi(pl) : 0, w = pw ? atoi(pw) : 0;
  for (int k = 0; k < s->nlev; k++) {
    bn_dlevel *d = &s->lev[k];
    if (!d->use_gs || d->block) continue;
    if (l == 1 || l == 2 || l == 4 || l == 8 || l == 16 || l == BN_GS_BLOCKS) d->gs_lanes = d->gs_lanes_s = l;
    d->gs_walk = d->gs_walk_s = (d->gs_jx && s->walk_ok && (w == 1 || w == 2 || w == 4 || w == 8)) ? w : 0;
  }
}

static void autotune_csr(struct bn_solver *s, const bn_tune_env *env, bn_dcsr *d, void *x, void *y,
                         int allow_blocks, int allow_deep) {
  if (!d->ia || d->rows <= 0 || d->nnz <= 0) return;
  const int base = d->lanes;
  int cand[24], nc = 0;
  const int opts[7] = {1, 2, 4, 8, 16, 32, 64};
  /* lanes far beyond the mean row length only idle, unless the operator
   * is small enough that more lanes are what fills the device */
  int np2 = 1;
  while (np2 < 64 && (long long)np2 * d->rows < (long long)d->nnz) np2 *= 2;
  for (int k = 0; k < 7; k++) {
    int l = opts[k];
    if (l < 1 || l > 64) continue;
    if (l > 4 && l > 2 * np2 && (long long)l * d->rows > (1LL << 20)) continue;
    int dup = 0;
    for (int q = 0; q < nc; q++) dup |= (cand[q] == l);
    if (!dup) cand[nc++] = l;
  }
  if (allow_deep) {
    const int n0 = nc;
    for (int k = 0; k < n0; k++)
      if (cand[k] >= 4) {
        cand[nc++] = cand[k] | BN_LANES_DEEP8;
        cand[nc++] = cand[k] | BN_LANES_DEEP16;
      }
  }
  /* a lanes choice that sends more than 1/64 of the rows to the long
   * bin is never competitive: drop it (keeps at least one candidate) */
  {
    int keep = 0;
    for (int k = 0; k < nc; k++)
      if ((long long)bn_nlong(d, cand[k]) * 64 <= d->rows || (cand[k] & BN_LANES_MASK) == 64) cand[keep++] = cand[k];
    if (keep > 0) nc = keep;
  }
  if (allow_blocks && d->blk) cand[nc++] = 0; /* 0: the row-block kernel */
  for (int pv = 0; pv < 2; pv++) {
    const char vt = pv ? 's' : 'd';
    float best = 1e30f, best_lanes_ms = 1e30f, blocks_ms = 1e30f;
    int bl = base, use_blocks = 0;
    /* a few cold launches each; more for small operators, whose times
     * are short and noisier */
    const double opb = (double)d->nnz * (4.0 + (pv ? 4.0 : 8.0)) + (double)d->rows * 12.0;
    int reps = (int)(1e8 / (opb > 1.0 ? opb : 1.0));
    if (reps < 3) reps = 3;
    if (reps > BN_TUNE_REPS) reps = BN_TUNE_REPS;
    for (int k = 0; k < nc; k++) {
      const float ms = time_launches(s, env, d, vt, cand[k], x, y, reps);
      const char *lgc = getenv("BN_TUNE_LOG");
      if (lgc && lgc[0] == '2')
        fprintf(stderr, "  cand %c rows=%d nnz=%d lanes=%d rounds=%d blocks=%d: %.1f us\n", vt, d->rows, d->nnz,
                cand[k] & BN_LANES_MASK, cand[k] >> 8, cand[k] == 0, 1e3 * ms / reps);
      if (cand[k] > 0 && ms < best_lanes_ms) {
        best_lanes_ms = ms;
        bl = cand[k];
      }
      if (cand[k] == 0) blocks_ms = ms;
    }
    /* the row-block kernel unless a lanes kernel beats it by more than 3%:
     * timing noise between close candidates then does not flip the choice
     * from one run to the next */
    use_blocks = blocks_ms <= 1.03f * best_lanes_ms;
    best = use_blocks ? blocks_ms : best_lanes_ms;
    const int lanes = bl & BN_LANES_MASK, deep = bl & ~BN_LANES_MASK;
    const char *lg = getenv("BN_TUNE_LOG");
    if (lg && (lg[0] == '1' || lg[0] == '2'))
      fprintf(stderr, "tune %c rows=%d nnz=%d: %s lanes=%d rounds=%s  (%.1f us per cold launch)\n", vt, d->rows,
              d->nnz, use_blocks ? "row-block" : "lanes", lanes,
              deep == BN_LANES_DEEP16 ? "16" : deep == BN_LANES_DEEP8 ? "8" : "default", 1e3 * best / reps);
    if (pv) {
      d->lanes_s = lanes;
      d->deep_s = deep;
      d->adapt_s = use_blocks;
    } else {
      d->lanes = lanes;
      d->deep = deep;
      d->adapt = use_blocks;
    }
  }
}

/* BN_APPLY in the environment pins the level-operator kernel choice:
 * "lanes" (lanes-per-row kernel only), "blocks" (row-block kernel for
 * every level operator and the outer PCG's), anything else: timed choice. */
static int apply_mode(void) {
  const char *e = getenv("BN_APPLY");
  if (!e) return 0;
  if (!strcmp(e, "lanes")) return 1;
  if (!strcmp(e, "blocks")) return 2;
  return 0;
}


static void pin_blocks(struct bn_solver *s) {
  for (int l = 0; l < s->nlev; l++) {
    bn_dcsr *ops[3] = {&s->lev[l].A, &s->lev[l].P, &s->lev[l].R};
    for (int k = 0; k < 3; k++)
      if (ops[k]->blk) ops[k]->adapt = ops[k]->adapt_s = 1;
  }
  /* and the outer PCG's operator */
  if (s->asolve_shared) s->Asolve = s->lev[0].A;
  else if (s->Asolve.blk) s->Asolve.adapt = 1;
}

/* the tuned operators in a fixed order (NULL entries skipped) */
static int tune_list(struct bn_solver *s, bn_dcsr **ops, int *flags) {
  int n = 0;
  for (int l = 0; l < s->nlev; l++) {
    const int allow = apply_mode() == 0;
    ops[n] = &s->lev[l].A, flags[n++] = allow ? 3 : 2;
    ops[n] = &s->lev[l].P, flags[n++] = allow ? 3 : 2;
    ops[n] = &s->lev[l].R, flags[n++] = allow ? 3 : 2;
    ops[n] = &s->lev[l].Ac, flags[n++] = 0;
  }
  /* the outer PCG's q = A p: a lanes kernel with p'q fused, or the row
   * blocks followed by the dot */
  ops[n] = s->asolve_shared ? NULL : &s->Asolve, flags[n++] = apply_mode() == 0 ? 3 : 2;
  ops[n] = &s->fine_ic.lo, flags[n++] = 0;
  ops[n] = &s->fine_ic.up, flags[n++] = 0;
  ops[n] = &s->coarse_ic.lo, flags[n++] = 0;
  ops[n] = &s->coarse_ic.up, flags[n++] = 0;
  return n;
}

#define BN_TUNE_FIELDS 8
static void tune_save(const bn_dcsr *d, int *t) {
  t[0] = d->rows, t[1] = d->nnz, t[2] = d->lanes, t[3] = d->lanes_s;
  t[4] = d->deep, t[5] = d->deep_s, t[6] = d->adapt, t[7] = d->adapt_s;
}

/* BN_TUNE_FILE names a file of tuned kernel choices: read when it exists
 * and matches this device mirror's operators (count and every shape),
 * written after tuning otherwise.  A run under a profiler, whose overhead
 * distorts the timings the tuner compares, then runs the choices of an
 * unprofiled run. */
static int tune_file_load(struct bn_solver *s, int nops, bn_dcsr *const *ops) {
  const char *path = getenv("BN_TUNE_FILE");
  if (!path || !*path) return 0;
  FILE *f = fopen(path, "r");
  if (!f) return 0;
  int nf = 0, nl = 0;
  int ok = fscanf(f, "%d %d", &nf, &nl) == 2 && nf == nops && nl == s->nlev;
  int *t = (int *)bn_xmalloc((size_t)(nops > 0 ? nops : 1) * BN_TUNE_FIELDS * sizeof(int));
  int g[AMG_MAX_LEVELS][4];
  for (int k = 0; ok && k < nops * BN_TUNE_FIELDS; k++) ok = fscanf(f, "%d", &t[k]) == 1;
  for (int l = 0; ok && l < nl; l++) ok = fscanf(f, "%d %d %d %d", &g[l][0], &g[l][1], &g[l][2], &g[l][3]) == 4;
  fclose(f);
  for (int k = 0; ok && k < nops; k++)
    if (ops[k] && (t[BN_TUNE_FIELDS * k] != ops[k]->rows || t[BN_TUNE_FIELDS * k + 1] != ops[k]->nnz)) ok = 0;
  if (!ok) {
    free(t);
    return 0;
  }
  free(s->tune);
  s->tune = t;
  s->ntune = nops;
  for (int l = 0; l < nl; l++)
    for (int q = 0; q < 4; q++) s->tune_gs[l][q] = g[l][q];
  return 1;
}

static void tune_file_save(const struct bn_solver *s) {
  const char *path = getenv("BN_TUNE_FILE");
  if (!path || !*path) return;
  FILE *f = fopen(path, "w");
  if (!f) return;
  fprintf(f, "%d %d\n", s->ntune, s->nlev);
  for (int k = 0; k < s->ntune; k++) {
    for (int q = 0; q < BN_TUNE_FIELDS; q++) fprintf(f, "%d ", s->tune[BN_TUNE_FIELDS * k + q]);
    fprintf(f, "\n");
  }
  for (int l = 0; l < s->nlev; l++)
    fprintf(f, "%d %d %d %d\n", s->tune_gs[l][0], s->tune_gs[l][1], s->tune_gs[l][2], s->tune_gs[l][3]);
  fclose(f);
}

static void autotune_all(struct bn_solver *s) {
  bn_dcsr *ops[4 * AMG_MAX_LEVELS + 8];
  int flags[4 * AMG_MAX_LEVELS + 8];
  const int nops = tune_list(s, ops, flags);
  /* a rebuild on the same pattern (amg_solver_update) keeps the choices */
  int reuse = s->tune && s->ntune == nops;
  for (int k = 0; reuse && k < nops; k++) {
    const int *t = s->tune + BN_TUNE_FIELDS * k;
    if (ops[k] && (t[0] != ops[k]->rows || t[1] != ops[k]->nnz)) reuse = 0;
  }
  const int loaded = !reuse && tune_file_load(s, nops, ops);
  if (loaded) reuse = 1;
  if (reuse) {
    for (int k = 0; k < nops; k++) {
      const int *t = s->tune + BN_TUNE_FIELDS * k;
      if (!ops[k] || !ops[k]->ia) continue;
      ops[k]->lanes = t[2], ops[k]->lanes_s = t[3], ops[k]->deep = t[4], ops[k]->deep_s = t[5];
      ops[k]->adapt = t[6], ops[k]->adapt_s = t[7];
    }
    for (int l = 0; l < s->nlev; l++) {
      s->lev[l].gs_lanes = s->tune_gs[l][0];
      s->lev[l].gs_lanes_s = s->tune_gs[l][1];
      s->lev[l].gs_walk = s->lev[l].gs_jx ? s->tune_gs[l][2] : 0;
      s->lev[l].gs_walk_s = s->lev[l].gs_jx ? s->tune_gs[l][3] : 0;
    }
  } else {
    void *x = s->x, *y = s->q;
    bn_fill('d', s->lev[0].n, x, 1.0, s->stream);
    /* eviction buffer: 320 MB, above the 256 MB last-level cache */
    bn_tune_env env;
    memset(&env, 0, sizeof(env));
    env.nflush = 40 << 20;
    void *fb = NULL;
    if (hipMalloc(&fb, (size_t)env.nflush * sizeof(double)) == hipSuccess) {
      bn_fill('d', env.nflush, fb, 0.0, s->stream);
      env.flush = (const double *)fb;
    }
    int evok = 1;
    for (int k = 0; k < 2 * BN_TUNE_REPS + 2; k++) evok &= hipEventCreate(&env.ev[k]) == hipSuccess;
    for (int k = 0; evok && k < nops; k++)
      if (ops[k]) autotune_csr(s, &env, ops[k], x, y, flags[k] & 1, (flags[k] & 2) != 0);
    for (int l = 0; evok && l < s->nlev; l++) {
      bn_dlevel *d = &s->lev[l];
      if (d->use_gs && !d->block && d->ncol > 0 && d->Ac.ia) {
        float best[2];
        tune_gs(s, &env, d, x, y, best);
        tune_gs_walk(s, &env, d, x, y, best);
      }
      s->tune_gs[l][0] = d->gs_lanes;
      s->tune_gs[l][1] = d->gs_lanes_s;
      s->tune_gs[l][2] = d->gs_walk;
      s->tune_gs[l][3] = d->gs_walk_s;
    }
    hipStreamSynchronize(s->stream);
    for (int k = 0; k < 2 * BN_TUNE_REPS + 2; k++)
      if (env.ev[k]) hipEventDestroy(env.ev[k]);
    if (fb) hipFree(fb);
    free(s->tune);
    s->tune = (int *)bn_xmalloc((size_t)nops * BN_TUNE_FIELDS * sizeof(int));
    s->ntune = nops;
    for (int k = 0; k < nops; k++) {
      if (ops[k]) tune_save(ops[k], s->tune + BN_TUNE_FIELDS * k);
      else memset(s->tune + BN_TUNE_FIELDS * k, 0, BN_TUNE_FIELDS * sizeof(int));
    }
    tune_file_save(s);
  }
  if (s->asolve_shared) s->Asolve = s->lev[0].A;
  hipStreamSynchronize(s->stream);
}

static int alloc_vec(struct bn_solver *s, void **p, int n) {
  *p = dmalloc(s, (size_t)(n > 0 ? n : 1) * sizeof(double));
  if (!*p) {
    bn_set_error("device out of memory (work vector of %d)", n);
    return BN_ERR_HIP;
  }
  return 0;
}

static double now_ms(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return 1e3 * (double)t.tv_sec + 1e-6 * (double)t.tv_nsec;
}

static int cmp_int(const void *a, const void *b) {
  const int x = *(const int *)a, y = *(const int *)b;
  return (x > y) - (x < y);
}

static void sort_levels(int nlev, const int *ptr, int *rows) {
#pragma omp parallel for schedule(dynamic, 16)
  for (int l = 0; l < nlev; l++) qsort(rows + ptr[l], (size_t)(ptr[l + 1] - ptr[l]), sizeof(int), cmp_int);
}

/* One direction's ELL copy from the CSR pattern (host copies ia, ja) and
 * the level schedule (ptr, rows).  Values come later, gathered on the
 * device from the CSR copy (bn_tri_ell_refresh). */
static int ell_build_dir(struct bn_solver *s, int nlev, const int *ptr, const int *rows, const int *ia,
                         const int *ja, int n, int np, bn_tell *e) {
  e->w = (int *)bn_xmalloc((size_t)(nlev > 0 ? nlev : 1) * sizeof(int));
  e->q = (int *)bn_xmalloc((size_t)(nlev > 0 ? nlev : 1) * sizeof(int));
  bn_track_host(s, e->w);
  bn_track_host(s, e->q);
  long long slots = 0;
  for (int l = 0; l < nlev; l++) {
    int w = 0;
    for (int k = ptr[l]; k < ptr[l + 1]; k++) {
      const int len = ia[rows[k] + 1] - ia[rows[k]];
      if (len > w) w = len;
    }
    if (w > BN_ELL_MAX_W) w = 0; /* this level keeps the CSR rows */
    e->w[l] = w;
    e->q[l] = (int)slots;
    slots += (long long)w * (ptr[l + 1] - ptr[l]);
    if (slots > 0x7fffffffLL) return -1;
  }
  e->slots = slots;
  int *src = (int *)bn_xmalloc((size_t)(slots > 0 ? slots : 1) * sizeof(int));
  int *col = (int *)bn_xmalloc((size_t)(slots > 0 ? slots : 1) * sizeof(int));
#pragma omp parallel for schedule(dynamic, 4)
  for (int l = 0; l < nlev; l++) {
    const int w = e->w[l], cnt = ptr[l + 1] - ptr[l];
    for (int k = 0; k < cnt && w > 0; k++) {
      const int r = rows[ptr[l] + k], q0 = ia[r], len = ia[r + 1] - q0;
      for (int t = 0; t < w; t++) {
        const long long at = (long long)e->q[l] + (long long)t * cnt + k;
        src[at] = t < len ? q0 + t : -1;
        col[at] = t < len ? ja[q0 + t] : r;
      }
    }
  }
  /* the one-launch solve's steps (bn_trsv_parts): part p holds rows
   * [p n / np, (p + 1) n / np); in each run of ELL levels, its steps are its
   * rows of each level, contiguous in the level's ascending row list */
  e->np = np;
  e->nrun = 0;
  e->run_lo = (int *)bn_xmalloc((size_t)(nlev > 0 ? nlev : 1) * sizeof(int));
  e->run_hi = (int *)bn_xmalloc((size_t)(nlev > 0 ? nlev : 1) * sizeof(int));
  bn_track_host(s, e->run_lo);
  bn_track_host(s, e->run_hi);
  for (int l = 0; l < nlev;) {
    if (e->w[l] == 0) {
      l++;
      continue;
    }
    int l1 = l;
    while (l1 < nlev && e->w[l1] > 0) l1++;
    e->run_lo[e->nrun] = l;
    e->run_hi[e->nrun] = l1;
    e->nrun++;
    l = l1;
  }
  /* kb[l * (np + 1) + p]: part p's first entry in level l's row list */
  int *kb = (int *)bn_xmalloc((size_t)(nlev > 0 ? nlev : 1) * (size_t)(np + 1) * sizeof(int));
  long long nsteps = 0;
  for (int l = 0; l < nlev; l++) {
    const int cnt = ptr[l + 1] - ptr[l];
    int k = 0;
    for (int p = 0; p <= np; p++) {
      const long long lo = (long long)p * n / np;
      while (k < cnt && rows[ptr[l] + k] < lo) k++;
      kb[(size_t)l * (np + 1) + p] = p == np ? cnt : k;
    }
    if (e->w[l] > 0)
      for (int p = 0; p < np; p++) nsteps += kb[(size_t)l * (np + 1) + p + 1] > kb[(size_t)l * (np + 1) + p];
  }
  int *sp = (int *)bn_xmalloc((size_t)(e->nrun > 0 ? e->nrun : 1) * (size_t)(np + 1) * sizeof(int));
  int *st = (int *)bn_xmalloc((size_t)(nsteps > 0 ? nsteps : 1) * 4 * sizeof(int));
  long long at = 0;
  for (int r = 0; r < e->nrun; r++) {
    for (int p = 0; p < np; p++) {
      sp[(size_t)r * (np + 1) + p] = (int)at;
      for (int l = e->run_lo[r]; l < e->run_hi[r]; l++) {
        const int k0 = kb[(size_t)l * (np + 1) + p], k1 = kb[(size_t)l * (np + 1) + p + 1];
        if (k1 == k0) continue;
        st[4 * at] = ptr[l] + k0;
        st[4 * at + 1] = k1 - k0;
        st[4 * at + 2] = (ptr[l + 1] - ptr[l]) * 16 + e->w[l];
        st[4 * at + 3] = e->q[l] + k0;
        at++;
      }
    }
    sp[(size_t)r * (np + 1) + np] = (int)at;
  }
  free(kb);
  int rc = 0;
  e->src = (int *)dmalloc(s, (size_t)(slots > 0 ? slots : 1) * sizeof(int));
  e->ja = (int *)dmalloc(s, (size_t)(slots > 0 ? slots : 1) * sizeof(int));
  e->vd = (double *)dmalloc(s, (size_t)(slots > 0 ? slots : 1) * sizeof(double));
  e->vs = (float *)dmalloc(s, (size_t)(slots > 0 ? slots : 1) * sizeof(float));
  e->sp = (int *)dmalloc(s, (size_t)(e->nrun > 0 ? e->nrun : 1) * (size_t)(np + 1) * sizeof(int));
  e->steps = (int *)dmalloc(s, (size_t)(nsteps > 0 ? nsteps : 1) * 4 * sizeof(int));
  if (!e->src || !e->ja || !e->vd || !e->vs || !e->sp || !e->steps) rc = -1;
  if (!rc && slots > 0 &&
      (hipMemcpy(e->src, src, (size_t)slots * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
       hipMemcpy(e->ja, col, (size_t)slots * sizeof(int), hipMemcpyHostToDevice) != hipSuccess))
    rc = -1;
  if (!rc && e->nrun > 0 &&
      (hipMemcpy(e->sp, sp, (size_t)e->nrun * (np + 1) * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
       hipMemcpy(e->steps, st, (size_t)nsteps * 4 * sizeof(int), hipMemcpyHostToDevice) != hipSuccess))
    rc = -1;
  free(sp);
  free(st);
  free(src);
  free(col);
  return rc;
}

/* BN_TRI_ELL=0 keeps every exact triangular level on the CSR rows */
int bn_tri_ell_build(struct bn_solver *s, bn_dtri *t) {
  if (t->ell || t->ell_tried || !t->lo.ia || t->n <= 0) return 0;
  /* never inside a stream capture (the build copies synchronously) */
  hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
  if (hipStreamIsCapturing(s->stream, &cs) != hipSuccess || cs != hipStreamCaptureStatusNone) return 0;
  t->ell_tried = 1;
  const char *env = getenv("BN_TRI_ELL");
  if (env && env[0] == '0') return 0;
  const int n = t->n, nz = t->lo.nnz;
  int *lia = (int *)bn_xmalloc(((size_t)n + 1) * sizeof(int)), *uia = (int *)bn_xmalloc(((size_t)n + 1) * sizeof(int));
  int *lja = (int *)bn_xmalloc((size_t)(nz > 0 ? nz : 1) * sizeof(int));
  int *uja = (int *)bn_xmalloc((size_t)(nz > 0 ? nz : 1) * sizeof(int));
  int *rf = (int *)bn_xmalloc((size_t)n * sizeof(int)), *rb = (int *)bn_xmalloc((size_t)n * sizeof(int));
  int rc = 0;
  if (hipMemcpy(lia, t->lo.ia, ((size_t)n + 1) * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(uia, t->up.ia, ((size_t)n + 1) * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(lja, t->lo.ja, (size_t)nz * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(uja, t->up.ja, (size_t)nz * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(rf, t->rows_f, (size_t)n * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(rb, t->rows_b, (size_t)n * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
    rc = -1;
  /* rows ascending within each level (a relabelled factor's level lists
   * come in the original order): the one-launch solve's parts are then
   * contiguous in every level; rows of one level are independent, so the
   * order changes no value */
  if (!rc) {
    sort_levels(t->nlev_f, t->ptr_f, rf);
    sort_levels(t->nlev_b, t->ptr_b, rb);
    if (hipMemcpy(t->rows_f, rf, (size_t)n * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(t->rows_b, rb, (size_t)n * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
      rc = -1;
  }
  /* parts of the one-launch solve: BN_TRI_PARTS, else two per compute
   * unit; never more than fit on the device at once (both precisions' kernels,
   * so every part is resident), nor more than one per 1,024 rows */
  int fit = bn_trsv_parts_per_cu('d');
  const int fit_s = bn_trsv_parts_per_cu('s');
  if (fit_s < fit) fit = fit_s;
  if (fit > 2) fit = 2;
  int np = fit * s->walk_blocks;
  const char *penv = getenv("BN_TRI_PARTS");
  if (penv && atoi(penv) > 0) np = atoi(penv);
  if (np > fit * s->walk_blocks) np = fit * s->walk_blocks;
  if (np > (n + 1023) / 1024) np = (n + 1023) / 1024;
  if (np < 1) {
    np = 1;
    s->walk_ok = 0; /* nothing known to fit: keep a launch per level */
  }
  if (!rc) rc = ell_build_dir(s, t->nlev_f, t->ptr_f, rf, lia, lja, n, np, &t->ef);
  if (!rc) rc = ell_build_dir(s, t->nlev_b, t->ptr_b, rb, uia, uja, n, np, &t->eb);
  if (!rc) {
    t->ef.dd = (double *)dmalloc(s, (size_t)n * sizeof(double));
    t->ef.ds = (float *)dmalloc(s, (size_t)n * sizeof(float));
    t->eb.dd = (double *)dmalloc(s, (size_t)n * sizeof(double));
    t->eb.ds = (float *)dmalloc(s, (size_t)n * sizeof(float));
    if (!t->ef.dd || !t->ef.ds || !t->eb.dd || !t->eb.ds) rc = -1;
  }
  free(lia);
  free(uia);
  free(lja);
  free(uja);
  free(rf);
  free(rb);
  if (rc) {
    (void)hipGetLastError();
    return 0; /* the CSR rows serve: same results */
  }
  t->ell = 1;
  bn_tri_ell_refresh(s, t);
  return hipStreamSynchronize(s->stream) == hipSuccess ? 0 : BN_ERR_HIP;
}

/* the ELL values and level-ordered 1/t_ii from the device CSR copy (after
 * every change of the factor's values) */
void bn_tri_ell_refresh(struct bn_solver *s, bn_dtri *t) {
  if (!t->ell) return;
  bn_ell_gather(t->ef.slots, t->ef.src, t->lo.vd, t->ef.vd, t->ef.vs, s->stream);
  bn_ell_gather(t->eb.slots, t->eb.src, t->up.vd, t->eb.vd, t->eb.vs, s->stream);
  bn_rb_gather(t->n, t->rows_f, t->dinv_d, t->ef.dd, t->ef.ds, s->stream);
  bn_rb_gather(t->n, t->rows_b, t->dinv_d, t->eb.dd, t->eb.ds, s->stream);
}

int bn_tri_upload(struct bn_solver *s, const ichol_factor *F, bn_dtri *t) {
  return upload_tri(s, F, NULL, NULL, t);
}

int bn_csr_upload(struct bn_solver *s, const dcsr_mat *m, bn_dcsr *d) { return upload_csr(s, m, d); }

void bn_sweep_free_host(bn_sweep *w) {
  free(w->ptr_f);
  free(w->ptr_b);
  memset(w, 0, sizeof(*w));
}

/* one direction's steps of the one-launch sweep: part p holds rows
 * [p n / np, (p + 1) n / np); its piece of each dependency level is
 * contiguous in the level's ascending row list */
static int sweep_steps(struct bn_solver *s, int n, int np, int nl, const int *ptr, const int *rows, int **sp_d,
                       int **st_d) {
  long long cnt = 0;
  int *kb = (int *)bn_xmalloc((size_t)(nl > 0 ? nl : 1) * (size_t)(np + 1) * sizeof(int));
  for (int l = 0; l < nl; l++) {
    int k = 0;
    const int c = ptr[l + 1] - ptr[l];
    for (int p = 0; p <= np; p++) {
      const long long lo = (long long)p * n / np;
      while (k < c && rows[ptr[l] + k] < lo) k++;
      kb[(size_t)l * (np + 1) + p] = p == np ? c : k;
    }
    for (int p = 0; p < np; p++) cnt += kb[(size_t)l * (np + 1) + p + 1] > kb[(size_t)l * (np + 1) + p];
  }
  int *sp = (int *)bn_xmalloc((size_t)(np + 1) * sizeof(int));
  int *st = (int *)bn_xmalloc((size_t)(cnt > 0 ? cnt : 1) * 2 * sizeof(int));
  long long at = 0;
  for (int p = 0; p < np; p++) {
    sp[p] = (int)at;
    for (int l = 0; l < nl; l++) {
      const int k0 = kb[(size_t)l * (np + 1) + p], k1 = kb[(size_t)l * (np + 1) + p + 1];
      if (k1 == k0) continue;
      st[2 * at] = ptr[l] + k0;
      st[2 * at + 1] = k1 - k0;
      at++;
    }
  }
  sp[np] = (int)at;
  free(kb);
  *sp_d = (int *)dmalloc(s, (size_t)(np + 1) * sizeof(int));
  *st_d = (int *)dmalloc(s, (size_t)(cnt > 0 ? cnt : 1) * 2 * sizeof(int));
  int rc = (!*sp_d || !*st_d) ? -1 : 0;
  if (!rc && (hipMemcpy(*sp_d, sp, (size_t)(np + 1) * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
              (cnt > 0 && hipMemcpy(*st_d, st, (size_t)cnt * 2 * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)))
    rc = -1;
  free(sp);
  free(st);
  return rc;
}

/* Tables of the one-launch exact sweep (bn_sweep_parts).  Parts: as many as
 * fit on the device at once (at most one per row); BN_SWEEP_PARTS
 * overrides.  At the 1.07M-unknown network's cf_order sweeps a solve took
 * 3.8 s with 1 part, 2.5 s with 8, 1.85 s with 64, 1.66 s with 512, against
 * 1.96 s for a launch per level.  None (a launch per level) with
 * BN_SWEEP_WALK=0, when the walk is off or nothing fits; a failure here
 * only keeps the launches. */
static int sweep_walk_tables(struct bn_solver *s, const dcsr_mat *A, const int *order, const int *rf, const int *rb,
                             bn_sweep *w) {
  w->jx = NULL;
  const char *env = getenv("BN_SWEEP_WALK");
  if ((env && env[0] == '0') || !s->walk_ok || s->walk_blocks <= 0) return 0;
  const int n = A->row, nnz = A->nnz;
  if (n <= 0 || nnz <= 0) return 0;
  int fit = bn_sweep_parts_per_cu('d');
  const int fs = bn_sweep_parts_per_cu('s');
  if (fs < fit) fit = fs;
  if (fit > 2) fit = 2;
  if (fit <= 0) return 0;
  int np = fit * s->walk_blocks;
  const char *penv = getenv("BN_SWEEP_PARTS");
  if (penv && atoi(penv) > 0) np = atoi(penv);
  if (np > fit * s->walk_blocks) np = fit * s->walk_blocks;
  if (np > n) np = n;
  if (np < 1) np = 1;
  int *pos = (int *)bn_xmalloc((size_t)n * sizeof(int));
  for (int k = 0; k < n; k++) pos[order ? order[k] : k] = k;
  int *jx = (int *)bn_xmalloc((size_t)nnz * sizeof(int));
#pragma omp parallel for schedule(static) if (nnz > 65536)
  for (int i = 0; i < n; i++)
    for (int q = A->ia[i]; q < A->ia[i + 1]; q++) {
      const int j = A->ja[q];
      jx[q] = (j != i && pos[j] < pos[i]) ? (int)((unsigned)j | 0x80000000u) : j;
    }
  free(pos);
  int *d = (int *)dmalloc(s, (size_t)nnz * sizeof(int));
  void *xo = dmalloc(s, (size_t)n * sizeof(double));
  int rc = (!d || !xo || hipMemcpy(d, jx, (size_t)nnz * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) ? -1 : 0;
  free(jx);
  if (!rc) rc = sweep_steps(s, n, np, w->nf, w->ptr_f, rf, &w->sp_f, &w->st_f);
  if (!rc) rc = sweep_steps(s, n, np, w->nb, w->ptr_b, rb, &w->sp_b, &w->st_b);
  if (rc) {
    (void)hipGetLastError();
    return 0; /* the launches per level serve */
  }
  w->np = np;
  w->xold = xo;
  w->jx = d;
  return 0;
}

int bn_sweep_upload(struct bn_solver *s, const dcsr_mat *A, const int *order, bn_sweep *w) {
  const int n = A->row;
  int *rf = (int *)bn_xmalloc((size_t)(n > 0 ? n : 1) * sizeof(int));
  int *rb = (int *)bn_xmalloc((size_t)(n > 0 ? n : 1) * sizeof(int));
  int *pf = NULL, *pb = NULL;
  const int nf = bn_sweep_levels(A, order, 0, rf, &pf);
  const int nb = nf < 0 ? -1 : bn_sweep_levels(A, order, 1, rb, &pb);
  int rc = 0;
  if (nf < 0 || nb < 0) {
    bn_set_error("Gauss-Seidel order is not a permutation of the %d rows", n);
    rc = BN_ERR_ARG;
  }
  int *df = NULL, *db = NULL;
  if (!rc) {
    df = (int *)dmalloc(s, (size_t)(n > 0 ? n : 1) * sizeof(int));
    db = (int *)dmalloc(s, (size_t)(n > 0 ? n : 1) * sizeof(int));
    if (!df || !db) {
      bn_set_error("device out of memory (sweep schedule of %d rows)", n);
      rc = BN_ERR_HIP;
    }
  }
  if (!rc && (hipMemcpy(df, rf, (size_t)n * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
              hipMemcpy(db, rb, (size_t)n * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)) {
    bn_set_error("copy of a sweep schedule to the device failed");
    rc = BN_ERR_HIP;
  }
  if (rc) {
    free(rf);
    free(rb);
    free(pf);
    free(pb);
    return rc;
  }
  w->nf = nf;
  w->nb = nb;
  w->rows_f = df;
  w->rows_b = db;
  w->ptr_f = pf;
  w->ptr_b = pb;
  rc = sweep_walk_tables(s, A, order, rf, rb, w);
  free(rf);
  free(rb);
  return rc;
}

/* Chunks of the block smoother: at most max_rows rows each; at least 256
 * chunks (one per CU) while chunks keep >= 64 rows; above 256 a multiple
 * of 256 so every CU gets the same number of chunks. */
static int chunk_count(int n, int max_rows) {
  if (n <= 0) return 1;
  long long nk = ((long long)n + max_rows - 1) / max_rows;
  if (nk < 256) {
    long long small = ((long long)n + 63) / 64;
    nk = small < 256 ? small : 256;
  } else {
    nk = (nk + 255) / 256 * 256;
  }
  return (int)(nk > 0 ? nk : 1);
}

/* Upload a level's chunk schedule and its block-smoother diagonal
 * 1 / (a_ii + sum over columns outside the row's chunk of |a_ij|), the l1
 * scaling that keeps the hybrid sweep convergent for any SPD operator;
 * Ap is the level operator in chunked numbering. */
static int upload_chunks(struct bn_solver *s, bn_dlevel *d, const dcsr_mat *Ap, const int *chunk, const int *ccol,
                         const int *cbnd) {
  const int n = Ap->row, nk = d->nchunk;
  double *bd = (double *)bn_xmalloc((size_t)(n > 0 ? n : 1) * sizeof(double));
#pragma omp parallel for schedule(dynamic, 16)
  for (int k = 0; k < nk; k++) {
    const int lo = chunk[k], hi = chunk[k + 1];
    for (int i = lo; i < hi; i++) {
      double dg = 0.0, off = 0.0;
      int has = 0;
      for (int q = Ap->ia[i]; q < Ap->ia[i + 1]; q++) {
        const int j = Ap->ja[q];
        if (j == i) {
          dg = Ap->val[q];
          has = 1;
        } else if (j < lo || j >= hi) {
          off += fabs(Ap->val[q]);
        }
      }
      if (!has) dg = 1.0;
      bd[i] = 1.0 / (dg + off);
    }
  }
  int rc = upload_vec2(s, bd, n, &d->bdinv_d, &d->bdinv_s);
  free(bd);
  if (rc) return rc;
  d->chunk_d = (int *)dmalloc(s, ((size_t)nk + 1) * sizeof(int));
  d->ccol_d = (int *)dmalloc(s, ((size_t)nk + 1) * sizeof(int));
  d->cbnd_d = (int *)dmalloc(s, ((size_t)ccol[nk] + 1) * sizeof(int));
  if (!d->chunk_d || !d->ccol_d || !d->cbnd_d) {
    bn_set_error("device out of memory (chunk schedule)");
    return BN_ERR_HIP;
  }
  CK(hipMemcpy(d->chunk_d, chunk, ((size_t)nk + 1) * sizeof(int), hipMemcpyHostToDevice));
  CK(hipMemcpy(d->ccol_d, ccol, ((size_t)nk + 1) * sizeof(int), hipMemcpyHostToDevice));
  CK(hipMemcpy(d->cbnd_d, cbnd, (size_t)ccol[nk] * sizeof(int), hipMemcpyHostToDevice));
  return 0;
}

/* Tables of the one-launch sweep (bn_gs_walk) of a GS level: the column
 * words of the colour-ordered copy Ac with bit 31 set where the
 * neighbour's colour precedes the row's, and each part's rows of each
 * colour.  Part p holds level rows [p n / np, (p + 1) n / np); the rows of a
 * colour are ascending in the level numbering, so its piece of a part is a
 * contiguous range of Ac rows.  np is the most workgroups that fit on the
 * device at once for every lanes choice and both precisions (all resident),
 * at most one per 256 rows.  No tables (the colour launches serve) when
 * nothing fits, with BN_GS_WALK=0, or when the walk is off. */
static int gs_walk_tables(struct bn_solver *s, bn_dlevel *d, const dcsr_mat *Ac, const int *cperm) {
  d->gs_jx = d->gs_wsteps = NULL;
  d->gs_xold = NULL;
  d->gs_wnp = 0;
  const char *env = getenv("BN_GS_WALK");
  if ((env && env[0] == '0') || !s->walk_ok || s->walk_blocks <= 0) return 0;
  const int n = Ac->row, ncol = d->ncol;
  if (n <= 0 || ncol <= 0 || Ac->nnz <= 0) return 0;
  int fit = 2;
  const int lanes[4] = {1, 2, 4, 8};
  for (int k = 0; k < 4; k++) {
    const int fd = bn_gs_walk_per_cu('d', lanes[k]), fs = bn_gs_walk_per_cu('s', lanes[k]);
    if (fd < fit) fit = fd;
    if (fs < fit) fit = fs;
  }
  if (fit <= 0) return 0;
  int np = fit * s->walk_blocks;
  if (np > (n + 255) / 256) np = (n + 255) / 256;
  if (np < 1) np = 1;
  int *colr = (int *)bn_xmalloc((size_t)n * sizeof(int));
  for (int c = 0; c < ncol; c++)
    for (int r = d->col_ptr[c]; r < d->col_ptr[c + 1]; r++) colr[cperm[r]] = c;
  int *jx = (int *)bn_xmalloc((size_t)Ac->nnz * sizeof(int));
#pragma omp parallel for schedule(static) if (Ac->nnz > 65536)
  for (int r = 0; r < n; r++) {
    const int c = colr[cperm[r]];
    for (int q = Ac->ia[r]; q < Ac->ia[r + 1]; q++) {
      const int j = Ac->ja[q];
      jx[q] = colr[j] < c ? (int)((unsigned)j | 0x80000000u) : j;
    }
  }
  free(colr);
  int *st = (int *)bn_xmalloc((size_t)np * ncol * 2 * sizeof(int));
  for (int c = 0; c < ncol; c++) {
    const int lo = d->col_ptr[c], hi = d->col_ptr[c + 1];
    int r = lo;
    for (int p = 0; p < np; p++) {
      const long long top = (long long)(p + 1) * n / np;
      const int r0 = r;
      while (r < hi && cperm[r] < top) r++;
      st[2 * ((size_t)p * ncol + c)] = r0;
      st[2 * ((size_t)p * ncol + c) + 1] = r - r0;
    }
  }
  int rc = 0;
  d->gs_jx = (int *)dmalloc(s, (size_t)Ac->nnz * sizeof(int));
  d->gs_wsteps = (int *)dmalloc(s, (size_t)np * ncol * 2 * sizeof(int));
  d->gs_xold = dmalloc(s, (size_t)n * sizeof(double));
  if (!d->gs_jx || !d->gs_wsteps || !d->gs_xold ||
      hipMemcpy(d->gs_jx, jx, (size_t)Ac->nnz * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(d->gs_wsteps, st, (size_t)np * ncol * 2 * sizeof(int), hipMemcpyHostToDevice) != hipSuccess)
    rc = BN_ERR_HIP;
  free(jx);
  free(st);
  if (rc) return rc;
  d->gs_wnp = np;
  return 0;
}

/* Colour-ordered row copy of a GS level: row r of Ac is the off-diagonal
 * part of row cperm[r] of Ap (columns unchanged; the sweep divides by the
 * diagonal through 1/a_ii, so a row's diagonal entry is never read), plus
 * the row map, 1/a_ii in that order and the row blocks of each colour. */
static int upload_colour_rows(struct bn_solver *s, bn_dlevel *d, const dcsr_mat *Ap, const int *cperm) {
  const int n = Ap->row;
  dcsr_mat Ac;
  bn_csr_offdiag_rows(Ap, cperm, &Ac);
  int rc = upload_csr(s, &Ac, &d->Ac);
  double *dc = (double *)bn_xmalloc((size_t)(n > 0 ? n : 1) * sizeof(double));
  /* 1/a_ii of level row cperm[r]: its last diagonal entry, 1.0 when
   * absent, as diag_inverses */
#pragma omp parallel for schedule(static) if (Ap->nnz > 65536)
  for (int r = 0; r < n; r++) {
    const int i = cperm[r];
    double dg = 1.0;
    for (int q = Ap->ia[i]; q < Ap->ia[i + 1]; q++)
      if (Ap->ja[q] == i) dg = Ap->val[q];
    dc[r] = 1.0 / dg;
  }
  if (!rc) rc = upload_vec2(s, dc, n, &d->dinvc_d, &d->dinvc_s);
  free(dc);
  /* row blocks colour by colour: a block never holds rows of two colours */
  if (!rc) {
    int *bk = (int *)bn_xmalloc(((size_t)n + 2) * sizeof(int));
    int *cb = (int *)bn_xmalloc(((size_t)d->ncol + 1) * sizeof(int));
    int nb = 0;
    bk[0] = 0;
    for (int c = 0; c < d->ncol; c++) {
      const int lo = d->col_ptr[c], hi = d->col_ptr[c + 1];
      cb[c] = nb;
      if (hi > lo) {
        const int k = bn_row_blocks(Ac.ia + lo, hi - lo, BN_ROWBLOCK_NNZ, BN_ROWBLOCK_ROWS, bk + nb);
        for (int t = 1; t <= k; t++) bk[nb + t] += lo;
        bk[nb] = lo;
        nb += k;
      }
    }
    cb[d->ncol] = nb;
    d->gs_nblk = nb;
    d->gs_cblk = cb;
    bn_track_host(s, cb);
    d->gs_blk = (int *)dmalloc(s, ((size_t)nb + 1) * sizeof(int));
    d->gs_blkq = (int *)dmalloc(s, ((size_t)nb + 1) * sizeof(int));
    if (!d->gs_blk || !d->gs_blkq) {
      free(bk);
      bn_csr_free(&Ac);
      return BN_ERR_HIP;
    }
    CK(hipMemcpy(d->gs_blk, bk, ((size_t)nb + 1) * sizeof(int), hipMemcpyHostToDevice));
    for (int k = 0; k <= nb; k++) bk[k] = Ac.ia[bk[k]];
    CK(hipMemcpy(d->gs_blkq, bk, ((size_t)nb + 1) * sizeof(int), hipMemcpyHostToDevice));
    free(bk);
  }
  if (!rc) rc = gs_walk_tables(s, d, &Ac, cperm);
  bn_csr_free(&Ac);
  if (rc) return rc;
  d->cmap_d = (int *)dmalloc(s, (size_t)(n > 0 ? n : 1) * sizeof(int));
  if (!d->cmap_d) return BN_ERR_HIP;
  CK(hipMemcpy(d->cmap_d, cperm, (size_t)n * sizeof(int), hipMemcpyHostToDevice));
  /* each level row's colour, for the coalesced first colour from zero */
  d->colour_d = NULL;
  if (d->ncol <= 255) {
    unsigned char *col = (unsigned char *)bn_xmalloc((size_t)(n > 0 ? n : 1));
    for (int c = 0; c < d->ncol; c++)
      for (int r = d->col_ptr[c]; r < d->col_ptr[c + 1]; r++) col[cperm[r]] = (unsigned char)c;
    d->colour_d = (unsigned char *)dmalloc(s, (size_t)(n > 0 ? n : 1));
    if (!d->colour_d) {
      free(col);
      return BN_ERR_HIP;
    }
    const hipError_t e = hipMemcpy(d->colour_d, col, (size_t)n, hipMemcpyHostToDevice);
    free(col);
    if (e != hipSuccess) return BN_ERR_HIP;
  }
  return 0;
}

int bn_device_upload(struct bn_solver *s) {
  const double t0 = now_ms();
  if (!bn_device_ok()) {
    bn_set_error("no HIP device available: the solve runs only on the GPU");
    return BN_ERR_NO_DEVICE;
  }
  if (s->opt.device >= 0) CK(hipSetDevice(s->opt.device));
  CK(hipGetDevice(&s->device));
  if (!s->stream) CK(hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking));
  if (!s->ev0) {
    CK(hipEventCreate(&s->ev0));
    CK(hipEventCreate(&s->ev1));
  }
  /* one-launch triangular solves (bn_trsv_parts): at most one workgroup per CU, every
   * wait bounded at 0.5 s of the constant-rate clock; BN_TRI_WALK=0 keeps a
   * launch per level */
  {
    int cus = 0, khz = 0;
    const char *wenv = getenv("BN_TRI_WALK");
    s->walk_ok = !(wenv && wenv[0] == '0') &&
                 hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, s->device) == hipSuccess &&
                 hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, s->device) == hipSuccess &&
                 cus > 0 && khz > 0;
    s->walk_blocks = cus;
    s->walk_limit = (long long)khz * 500;
    /* BN_TRI_WALK_LIMIT: the wait limit in clock ticks; < 0 gives up at the
     * first wait, which exercises the re-run with a launch per level */
    const char *lenv = getenv("BN_TRI_WALK_LIMIT");
    if (lenv && lenv[0]) s->walk_limit = atoll(lenv);
  }
  {
    const char *fenv = getenv("BN_FUSE_JACOBI");
    s->jac_fuse = !(fenv && fenv[0] == '0');
    const char *denv = getenv("BN_FUSE_DOT");
    s->dot_fuse = !(denv && denv[0] == '0');
  }
  s->vt = (s->opt.precision == BN_PREC_FP32) ? 's' : 'd';
  const amg_data *amg = s->amg;
  const int nlev = amg->num_levels;
  s->nlev = nlev;
  const int split = 1; /* C-then-F colours serve both GS smoothers */
  int *inv[AMG_MAX_LEVELS];
  int *chunk_h[AMG_MAX_LEVELS], *ccol_h[AMG_MAX_LEVELS], *cbnd_h[AMG_MAX_LEVELS], *cperm_h[AMG_MAX_LEVELS];
  int rc = 0;

  double tph[7];
  tph[0] = now_ms();
  /* colour orders and inverse permutations: one task per level (each
   * level's result is its own, whatever runs first) */
#pragma omp parallel
#pragma omp single
  for (int l = 0; l < nlev; l++)
#pragma omp task firstprivate(l) shared(chunk_h, ccol_h, cbnd_h, cperm_h, inv)
  {
    const amg_level *L = &amg->levels[l];
    bn_dlevel *d = &s->lev[l];
    const int n = L->A.row;
    d->n = n;
    d->nc = (l < nlev - 1) ? L->n_coarse : 0;
    d->perm_h = (int *)bn_xmalloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    chunk_h[l] = ccol_h[l] = cbnd_h[l] = cperm_h[l] = NULL;
    if (l < nlev - 1 && s->opt.block_smoother) {
      /* chunked layout: natural order kept at the chunk scale */
      d->block = 1;
      d->use_gs = 0;
      d->ncol = 0;
      d->col_ptr = NULL;
      d->nchunk = bn_chunk_order(&L->A, L->mis, split, chunk_count(n, s->opt.chunk_rows), d->perm_h,
                                 &chunk_h[l], &ccol_h[l], &cbnd_h[l]);
      d->chunk_max = 0;
      for (int k = 0; k < d->nchunk; k++)
        if (chunk_h[l][k + 1] - chunk_h[l][k] > d->chunk_max) d->chunk_max = chunk_h[l][k + 1] - chunk_h[l][k];
    } else if (l < nlev - 1) {
      /* the level keeps its own (natural) numbering, whose neighbouring
       * rows share x cache lines; a GS level also gets its rows copied
       * in colour order for the sweep kernels */
      cperm_h[l] = (int *)bn_xmalloc((size_t)(n > 0 ? n : 1) * sizeof(int));
      d->ncol = bn_colour_order(&L->A, L->mis, split, cperm_h[l], &d->col_ptr);
      d->use_gs = d->ncol <= s->opt.max_colours;
      for (int i = 0; i < n; i++) d->perm_h[i] = i;
    } else {
      for (int i = 0; i < n; i++) d->perm_h[i] = i;
      d->ncol = 0;
      d->col_ptr = NULL;
      d->use_gs = 0;
    }
    inv[l] = (int *)bn_xmalloc((size_t)(n > 0 ? n : 1) * sizeof(int));
    for (int k = 0; k < n; k++) inv[l][d->perm_h[k]] = k;
  }
  tph[6] = now_ms();

  for (int l = 0; l < nlev && !rc; l++) {
    const amg_level *L = &amg->levels[l];
    bn_dlevel *d = &s->lev[l];
    const int n = d->n;
    dcsr_mat Ap;
    /* levels outside the block layout keep their numbering: no copy */
    const int same = is_identity(d->perm_h, n);
    if (same) Ap = L->A;
    else bn_csr_permute(&L->A, d->perm_h, inv[l], &Ap);
    rc = upload_csr(s, &Ap, &d->A);
    double *dinv = (double *)bn_xmalloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    double *l1inv = (double *)bn_xmalloc((size_t)(n > 0 ? n : 1) * sizeof(double));
    diag_inverses(&Ap, dinv, l1inv);
    if (!rc) rc = upload_vec2(s, dinv, n, &d->dinv_d, &d->dinv_s);
    if (!rc) rc = upload_vec2(s, l1inv, n, &d->l1inv_d, &d->l1inv_s);
    if (!rc && d->block) rc = upload_chunks(s, d, &Ap, chunk_h[l], ccol_h[l], cbnd_h[l]);
    if (!rc && d->use_gs) rc = upload_colour_rows(s, d, &Ap, cperm_h[l]);
    free(dinv);
    free(l1inv);
    if (!same) bn_csr_free(&Ap);
    if (!rc && l < nlev - 1) {
      dcsr_mat Pp, Rp;
      const int psame = same && is_identity(s->lev[l + 1].perm_h, s->lev[l + 1].n);
      if (psame) Pp = L->P;
      else bn_csr_permute(&L->P, d->perm_h, inv[l + 1], &Pp);
      bn_csr_transpose(&Pp, &Rp);
      rc = upload_csr(s, &Pp, &d->P);
      if (!rc) rc = upload_csr(s, &Rp, &d->R);
      if (!psame) bn_csr_free(&Pp);
      bn_csr_free(&Rp);
    }
    if (!rc) {
      d->perm_d = (int *)dmalloc(s, (size_t)(n > 0 ? n : 1) * sizeof(int));
      if (!d->perm_d) rc = BN_ERR_HIP;
      else CK(hipMemcpy(d->perm_d, d->perm_h, (size_t)n * sizeof(int), hipMemcpyHostToDevice));
    }
    if (!rc) rc = alloc_vec(s, &d->b, n);
    if (!rc) rc = alloc_vec(s, &d->x, n);
    if (!rc) rc = alloc_vec(s, &d->r, n);
    if (!rc && d->block) rc = alloc_vec(s, &d->x2, n);
  }
  for (int l = 0; l < nlev; l++) {
    free(chunk_h[l]);
    free(ccol_h[l]);
    free(cbnd_h[l]);
    free(cperm_h[l]);
  }

  /* solve matrix: shares level 0 unless set_matrix replaced it */
  if (!rc) {
    const dcsr_mat *A0 = &amg->levels[0].A;
    const dcsr_mat *As = s->A ? s->A : A0;
    int same = (As == A0) ||
               (As->row == A0->row && As->nnz == A0->nnz &&
                !memcmp(As->ia, A0->ia, ((size_t)A0->row + 1) * sizeof(int)) &&
                !memcmp(As->ja, A0->ja, (size_t)A0->nnz * sizeof(int)) &&
                !memcmp(As->val, A0->val, (size_t)A0->nnz * sizeof(double)));
    s->asolve_shared = same;
    if (same) {
      s->Asolve = s->lev[0].A;
    } else {
      dcsr_mat Ap;
      bn_csr_permute(As, s->lev[0].perm_h, inv[0], &Ap);
      rc = upload_csr(s, &Ap, &s->Asolve);
      bn_csr_free(&Ap);
    }
  }
  tph[1] = now_ms();
  /* fine IC(0) (absent for hierarchy-only device mirrors) */
  if (!rc && s->L)
    rc = upload_tri(s, s->L, is_identity(s->lev[0].perm_h, s->lev[0].n) ? NULL : s->lev[0].perm_h, inv[0],
                    &s->fine_ic);
  tph[2] = now_ms();

  /* coarsest solver */
  if (!rc) {
    const amg_level *C = &amg->levels[nlev - 1];
    const int nc = C->A.row;
    int dense = (s->opt.coarse_solver == BN_COARSE_DENSE) ||
                (s->opt.coarse_solver == BN_COARSE_AUTO && nc <= BN_DENSE_MAX);
    if (nlev == 1 && s->opt.coarse_solver == BN_COARSE_AUTO) dense = (nc <= BN_DENSE_MAX);
    s->coarse_dense = dense;
    if (dense) {
      /* the entries of bn_dense_inverse, formed on the device */
      int clamped = 0;
      s->cinv_d = (double *)dmalloc(s, (size_t)nc * (size_t)nc * sizeof(double) + 8);
      s->cinv_s = (float *)dmalloc(s, (size_t)nc * (size_t)nc * sizeof(float) + 8);
      if (!s->cinv_d || !s->cinv_s) {
        bn_set_error("device out of memory (coarsest inverse of %d rows)", nc);
        rc = BN_ERR_HIP;
      }
      if (!rc && nc <= BN_DENSE_MAX) {
        rc = bn_coarse_inverse_device(s, &C->A, s->cinv_d, s->cinv_s, &clamped);
      } else if (!rc) {
        /* a forced dense solve above the device kernel's limit: the host
         * inverse (same entries, slower) */
        double *inv_h = (double *)bn_xmalloc((size_t)nc * (size_t)nc * sizeof(double));
        float *inv_f = (float *)bn_xmalloc((size_t)nc * (size_t)nc * sizeof(float));
        clamped = bn_dense_inverse(&C->A, inv_h);
        for (size_t k = 0; k < (size_t)nc * (size_t)nc; k++) inv_f[k] = (float)inv_h[k];
        if (hipMemcpy(s->cinv_d, inv_h, (size_t)nc * (size_t)nc * sizeof(double), hipMemcpyHostToDevice) !=
                hipSuccess ||
            hipMemcpy(s->cinv_s, inv_f, (size_t)nc * (size_t)nc * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
          rc = BN_ERR_HIP;
        free(inv_h);
        free(inv_f);
      }
      if (!rc && clamped && s->opt.verbose)
        fprintf(stderr, "bnamg: coarsest operator not numerically SPD (pivot clamped)\n");
    } else {
      rc = upload_tri(s, &C->L_ichol, NULL, NULL, &s->coarse_ic);
    }
    for (int k = 0; k < 6 && !rc; k++) rc = alloc_vec(s, (void **)&s->cvec[k], nc);
  }

  tph[3] = now_ms();
  /* PCG and preconditioner work space */
  if (!rc) {
    const int n = s->lev[0].n;
    void **pv[] = {(void **)&s->x, (void **)&s->r, (void **)&s->z, (void **)&s->p, (void **)&s->q,
                   (void **)&s->bb, (void **)&s->rprev, &s->g_in, &s->u1, &s->r1, &s->icy,
                   &s->icx, &s->r2, &s->corr, &s->zout};
    for (size_t k = 0; k < sizeof(pv) / sizeof(pv[0]) && !rc; k++) rc = alloc_vec(s, pv[k], n);
    if (!rc) rc = alloc_vec(s, (void **)&s->scal, BN_S_COUNT * 2);
    if (!rc) rc = alloc_vec(s, (void **)&s->part1, BN_RED_BLOCKS);
    if (!rc) rc = alloc_vec(s, (void **)&s->part2, BN_RED_BLOCKS);
    if (!rc) rc = alloc_vec(s, (void **)&s->part3, BN_RED_BLOCKS);
    if (!rc) rc = alloc_vec(s, (void **)&s->cpart1, BN_RED_BLOCKS);
    if (!rc) rc = alloc_vec(s, (void **)&s->cpart2, BN_RED_BLOCKS);
    if (!rc && !s->host_scal) CK(hipHostMalloc((void **)&s->host_scal, BN_S_COUNT * 2 * sizeof(double), 0));
    if (!rc && !s->host_step &&
        hipHostMalloc((void **)&s->host_step, BN_S_COUNT * sizeof(double), hipHostMallocCoherent) != hipSuccess)
      s->host_step = NULL; /* the step's scalars are then copied */
  }
  for (int l = 0; l < nlev; l++) free(inv[l]);
  tph[4] = now_ms();
  if (!rc) {
    const char *env = getenv("BN_NO_AUTOTUNE");
    if (!(env && env[0] && env[0] != '0')) autotune_all(s);
    else gs_pins(s);
    if (apply_mode() == 2) pin_blocks(s);
    CK(hipDeviceSynchronize());
    s->dev_ready = 1;
  }
  tph[5] = now_ms();
  s->last.upload_ms = now_ms() - t0;
  const char *lg = getenv("BN_SETUP_LOG");
  if (lg && *lg && *lg != '0')
    fprintf(stderr,
            "[upload] %.1f ms: colour orders %.1f, level operators %.1f, fine IC(0) factor and schedules %.1f, "
            "coarsest solver %.1f, work space %.1f, autotuning %.1f\n",
            tph[5] - t0, tph[6] - tph[0], tph[1] - tph[6], tph[2] - tph[1], tph[3] - tph[2], tph[4] - tph[3],
            tph[5] - tph[4]);
  return rc;
}
