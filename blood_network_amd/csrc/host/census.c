/*
 * census.c - launch wrappers for the solve's bandwidth kernels, and the
 * kernel census built on them.
 *
 * Every sparse-operator launch of the V-cycle and of the outer PCG goes
 * through one of the bn_launch_* wrappers below.  Normally they only
 * launch.  During bn_solver_kernel_census they also record the launch: its
 * kernel symbol (spelled as rocprofv3 prints it), its algorithmic bytes
 * (each array touched once: index 4 B, value 4 B in FP32 / 8 B in FP64)
 * and its duration: a HIP event pair armed with bn_time_next_launch, which
 * the kernel's launch fills with its own start and end on the solver's
 * stream (the dispatch timestamps a profiler reads, with no marker packets
 * between the kernels).  A wrapper that launches more than one kernel
 * times its first, the one its symbol names.  After the solve
 * every distinct launch shape is launched again, each time after a read of
 * a 320 MB buffer (larger than the 256 MB last-level cache), and timed cold.
 * The census reports, per symbol: launches per solve, algorithmic bytes
 * per solve, device time per solve as the solve ran, and as it would run
 * with every launch starting cold.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bn_solver.h"

void bn_set_error(const char *fmt, ...);

enum { K_APPLY = 0, K_RESID_SCALED, K_GS, K_JACOBI, K_SPMV };

typedef struct {
  int kind, op, c, lanes;
  char vt;
  const void *m;       /* bn_dcsr (K_APPLY, K_RESID_SCALED, K_JACOBI, K_SPMV) or bn_dlevel (K_GS) */
  const void *v[7];    /* the launch's vector arguments */
  double om;           /* K_APPLY: the fused Jacobi step's damping */
  const unsigned char *colour; /* K_APPLY: the fused first colour's colour array and colour */
} ckey;

/* a K_SPMV launch that bn_launch_spmv runs with its dot fused */
static int spmv_dot_fused(const struct bn_solver *s, const ckey *k) {
  const bn_dcsr *M = (const bn_dcsr *)k->m;
  return M->adapt && M->blk && M->bpart && k->v[3] && k->v[4] && k->op == BN_OP_SET && s->dot_fuse;
}

typedef struct {
  ckey key;
  char name[80];
  double bytes;        /* algorithmic bytes of one launch */
  int count;           /* launches in the census solve */
  double us_solve;     /* their summed durations, as the solve ran */
  double us_cold;      /* one cold launch, averaged */
} cshape;

struct bn_census {
  int nshape, capshape;
  cshape *shape;
  int nev, capev;
  hipEvent_t *ev;      /* two per recorded launch */
  int *ev_shape;
  int overflow;
};

static const char *tname(char vt) { return vt == 's' ? "float" : "double"; }
static double vbytes(char vt) { return vt == 's' ? 4.0 : 8.0; }

/* the lanes kernel's load rounds (k_csr's R parameter) */
static int csr_rounds(int lanes_in) {
  const int l = lanes_in & BN_LANES_MASK;
  if (lanes_in & BN_LANES_DEEP16) return 16;
  if (lanes_in & BN_LANES_DEEP8) return 8;
  return l == 2 ? 8 : 4;
}

static void shape_describe(struct bn_solver *s, cshape *sh) {
  const ckey *k = &sh->key;
  const double v = vbytes(k->vt);
  switch (k->kind) {
    case K_APPLY:
    case K_RESID_SCALED:
    case K_SPMV: {
      const bn_dcsr *M = (const bn_dcsr *)k->m;
      const int blocks = (k->kind == K_RESID_SCALED) ||
                         (k->kind == K_SPMV ? (M->adapt && M->blk) : (k->vt == 's' ? M->adapt_s : M->adapt));
      if (blocks && k->kind == K_SPMV && spmv_dot_fused(s, k))
        snprintf(sh->name, sizeof(sh->name), "k_csr_blocks_dot<double>");
      else if (blocks)
        snprintf(sh->name, sizeof(sh->name), "k_csr_blocks<%s, %d, %s>", tname(k->vt), k->op,
                 k->kind == K_RESID_SCALED ? "true" : "false");
      else {
        const int ln = k->kind == K_SPMV ? (M->lanes | M->deep) : bn_spmv_lanes(M, k->vt);
        /* the fused dot runs while the row groups fit the partial slots */
        const long long rpb = 256 / (ln & BN_LANES_MASK);
        const int fused = k->kind == K_SPMV && k->v[3] && ((long long)M->rows + rpb - 1) / rpb <= BN_RED_BLOCKS;
        snprintf(sh->name, sizeof(sh->name), "k_csr<%s, %d, %d, %s, %d>", tname(k->vt), ln & BN_LANES_MASK, k->op,
                 fused ? "true" : "false", csr_rounds(ln));
      }
      /* entries and row pointers, x (cols), y written, b / y read, and for
       * the scaled residual dsc read and ys written */
      double b = (double)M->nnz * (4.0 + v) + ((double)M->rows + 1.0) * 4.0 + (double)M->cols * v +
                 (double)M->rows * v;
      if (k->op == BN_OP_RESID || k->op == BN_OP_ADD || k->op == BN_OP_JAC) b += (double)M->rows * v;
      if (k->kind == K_SPMV && spmv_dot_fused(s, k)) b += (double)M->rows * v; /* the dot's w */
      if (k->kind == K_RESID_SCALED) b += 2.0 * M->rows * v;
      /* a fused Jacobi step: its weights read, and for SET the step written */
      if (k->kind == K_APPLY && k->v[4]) b += (double)M->rows * v * (k->op == BN_OP_SET ? 2.0 : 1.0);
      if (k->kind == K_APPLY && k->colour) b += (double)M->rows; /* the colour bytes */
      sh->bytes = b;
      break;
    }
    case K_GS: {
      const bn_dlevel *d = (const bn_dlevel *)k->m;
      const int lo = d->col_ptr[k->c], hi = d->col_ptr[k->c + 1];
      double nz = 0.0;
      int ia[2];
      if (hipMemcpy(&ia[0], d->Ac.ia + lo, sizeof(int), hipMemcpyDeviceToHost) == hipSuccess &&
          hipMemcpy(&ia[1], d->Ac.ia + hi, sizeof(int), hipMemcpyDeviceToHost) == hipSuccess)
        nz = (double)(ia[1] - ia[0]);
      const double rows = hi - lo;
      if (k->lanes == BN_GS_BLOCKS && d->gs_cblk)
        snprintf(sh->name, sizeof(sh->name), "k_gs_blocks<%s>", tname(k->vt));
      else
        snprintf(sh->name, sizeof(sh->name), "k_rowsolve<%s, %d, 2>", tname(k->vt),
                 k->lanes > 0 && k->lanes != BN_GS_BLOCKS ? k->lanes : bn_lanes(&d->Ac, k->vt));
      /* the colour's off-diagonal entries and row pointers, row map, b,
       * 1/a_ii, its share of x read and x written (+ zbase, zsum) */
      sh->bytes = nz * (4.0 + v) + (rows + 1.0) * 4.0 + rows * 4.0 + 4.0 * rows * v + (k->v[2] ? 2.0 * rows * v : 0.0);
      (void)s;
      break;
    }
    case K_JACOBI: {
      const bn_dcsr *T = (const bn_dcsr *)k->m;
      const int scaled = k->v[6] != NULL;
      snprintf(sh->name, sizeof(sh->name), "k_jacobi_rows<%s, %d, %s>", tname(k->vt), bn_lanes(T, k->vt),
               scaled ? "true" : "false");
      const double n = T->rows;
      /* factor entries and row pointers, b, 1/l_ii, x in (when given), x out
       * or zbase + zsum, and the scaled copy */
      double b = (double)T->nnz * (4.0 + v) + (n + 1.0) * 4.0 + 2.0 * n * v + n * v;
      if (k->v[2]) b += n * v;
      if (k->v[4]) b += n * v;
      if (scaled) b += n * v;
      sh->bytes = b;
      break;
    }
  }
}

static int census_begin(struct bn_solver *s, const ckey *key) {
  struct bn_census *c = s->census;
  if (!c) return -1;
  int slot = -1;
  for (int i = 0; i < c->nshape; i++)
    if (!memcmp(&c->shape[i].key, key, sizeof(ckey))) {
      slot = i;
      break;
    }
  if (slot < 0) {
    if (c->nshape == c->capshape) {
      c->capshape = c->capshape ? 2 * c->capshape : 64;
      c->shape = (cshape *)realloc(c->shape, (size_t)c->capshape * sizeof(cshape));
    }
    slot = c->nshape++;
    memset(&c->shape[slot], 0, sizeof(cshape));
    c->shape[slot].key = *key;
    shape_describe(s, &c->shape[slot]);
  }
  c->shape[slot].count++;
  if (c->nev == c->capev) {
    c->overflow = 1;
    return -1;
  }
  bn_time_next_launch(c->ev[2 * c->nev], c->ev[2 * c->nev + 1]);
  c->ev_shape[c->nev] = slot;
  return c->nev++;
}

static void census_end(struct bn_solver *s, int ev) {
  (void)s;
  if (ev >= 0) bn_time_next_launch(NULL, NULL); /* a launch that never came */
}

static ckey mk(int kind, char vt, int op, const void *m) {
  ckey k;
  memset(&k, 0, sizeof(k));
  k.kind = kind;
  k.vt = vt;
  k.op = op;
  k.m = m;
  return k;
}

/* ---------------------------------------------------------------------
 * launch wrappers
 * ------------------------------------------------------------------- */
void bn_launch_apply_jac(struct bn_solver *s, const bn_dcsr *M, char vt, int op, const void *x, void *y,
                         const void *b, void *xz, const void *w, double om, const unsigned char *colour, int c) {
  int ev = -1;
  if (s->census) {
    ckey k = mk(K_APPLY, vt, op, M);
    k.v[0] = x, k.v[1] = y, k.v[2] = b, k.v[3] = xz, k.v[4] = w;
    k.om = om;
    k.colour = colour;
    k.c = c;
    ev = census_begin(s, &k);
  }
  bn_op_apply_jac(M, vt, op, x, y, b, xz, w, om, colour, c, s->stream);
  census_end(s, ev);
}

void bn_launch_apply(struct bn_solver *s, const bn_dcsr *M, char vt, int op, const void *x, void *y, const void *b) {
  bn_launch_apply_jac(s, M, vt, op, x, y, b, NULL, NULL, 0.0, NULL, 0);
}

void bn_launch_resid_scaled(struct bn_solver *s, const bn_dcsr *M, char vt, const void *x, void *y, const void *b,
                            const void *dsc, void *ys) {
  int ev = -1;
  if (s->census) {
    ckey k = mk(K_RESID_SCALED, vt, BN_OP_RESID, M);
    k.v[0] = x, k.v[1] = y, k.v[2] = b, k.v[3] = dsc, k.v[4] = ys;
    ev = census_begin(s, &k);
  }
  bn_csr_resid_blocks_scaled(vt, M->nblk, M->blk, M->blkq, M->ia, M->ja,
                             vt == 's' ? (const void *)M->vs : (const void *)M->vd, x, y, b, dsc, ys, s->stream);
  census_end(s, ev);
}

void bn_launch_gs(struct bn_solver *s, const bn_dlevel *d, char vt, int c, int lanes, const void *b, void *x,
                  const void *zbase, void *zsum) {
  int ev = -1;
  if (s->census) {
    ckey k = mk(K_GS, vt, 0, d);
    k.c = c, k.lanes = lanes;
    k.v[0] = b, k.v[1] = x, k.v[2] = zsum, k.v[3] = zbase;
    ev = census_begin(s, &k);
  }
  bn_gs_colour(s, d, vt, c, lanes, b, x, zbase, zsum);
  census_end(s, ev);
}

void bn_launch_jacobi(struct bn_solver *s, const bn_dcsr *T, char vt, const void *dinv, const void *b, const void *xin,
                      void *xout, const void *zbase, void *zsum, void *xs) {
  int ev = -1;
  if (s->census) {
    ckey k = mk(K_JACOBI, vt, 0, T);
    k.v[0] = dinv, k.v[1] = b, k.v[2] = xin, k.v[3] = xout, k.v[4] = zsum, k.v[5] = zbase, k.v[6] = xs;
    ev = census_begin(s, &k);
  }
  const void *val = vt == 's' ? (const void *)T->vs : (const void *)T->vd;
  if (xs)
    bn_jacobi_rows_scaled(vt, T->rows, T->ia, T->ja, val, dinv, b, xin, xout, xs, bn_lanes(T, vt), s->stream);
  else
    bn_jacobi_rows(vt, T->rows, T->ia, T->ja, val, dinv, b, xin, xout, bn_lanes(T, vt), zbase, zsum, s->stream);
  census_end(s, ev);
}

int bn_launch_spmv(struct bn_solver *s, const bn_dcsr *A, int op, const void *x, void *y, const void *b,
                   const void *dotw, double *partial) {
  int ev = -1;
  if (s->census) {
    ckey k = mk(K_SPMV, 'd', op, A);
    k.v[0] = x, k.v[1] = y, k.v[2] = b, k.v[3] = partial, k.v[4] = dotw;
    ev = census_begin(s, &k);
  }
  if (A->adapt && A->blk) {
    if (partial && dotw && op == BN_OP_SET && s->dot_fuse && A->bpart) {
      /* the product with the dot's row-block partials fused */
      bn_csr_apply_blocks_dot(A->nblk, A->blk, A->blkq, A->ia, A->ja, A->vd, (const double *)x, (double *)y,
                              (const double *)dotw, A->bpart, s->stream);
      census_end(s, ev);
      return A->nblk;
    }
    /* the row-block product, then (outside the record) the deterministic
     * dot partials it feeds */
    bn_csr_apply_blocks('d', op, A->nblk, A->blk, A->blkq, A->ia, A->ja, A->vd, x, y, b, s->stream);
    census_end(s, ev);
    if (partial)
      bn_dot_partial(A->rows, (const double *)dotw, (const double *)y, NULL, NULL, partial, NULL, s->stream);
    return 0;
  }
  bn_dcsr_apply_dot(A, 'd', op, x, y, b, dotw, partial, s->stream);
  census_end(s, ev);
  return 0;
}

/* ---------------------------------------------------------------------
 * the census
 * ------------------------------------------------------------------- */
static void relaunch(struct bn_solver *s, const ckey *k) {
  switch (k->kind) {
    case K_APPLY:
      bn_launch_apply_jac(s, (const bn_dcsr *)k->m, k->vt, k->op, k->v[0], (void *)k->v[1], k->v[2],
                          (void *)k->v[3], k->v[4], k->om, k->colour, k->c);
      break;
    case K_RESID_SCALED:
      bn_launch_resid_scaled(s, (const bn_dcsr *)k->m, k->vt, k->v[0], (void *)k->v[1], k->v[2], k->v[3],
                             (void *)k->v[4]);
      break;
    case K_GS:
      bn_launch_gs(s, (const bn_dlevel *)k->m, k->vt, k->c, k->lanes, k->v[0], (void *)k->v[1], k->v[3],
                   (void *)k->v[2]);
      break;
    case K_JACOBI:
      bn_launch_jacobi(s, (const bn_dcsr *)k->m, k->vt, k->v[0], k->v[1], k->v[2], (void *)k->v[3], k->v[5],
                       (void *)k->v[4], (void *)k->v[6]);
      break;
    case K_SPMV:
      /* the product alone (the dot partials that follow it are not part
       * of the kernel timed), or with the dot when it is fused */
      if (spmv_dot_fused(s, k))
        bn_launch_spmv(s, (const bn_dcsr *)k->m, k->op, k->v[0], (void *)k->v[1], k->v[2], k->v[4],
                       (double *)k->v[3]);
      else
        bn_launch_spmv(s, (const bn_dcsr *)k->m, k->op, k->v[0], (void *)k->v[1], k->v[2], NULL, NULL);
      break;
  }
}

int bn_solver_kernel_census(void *h, const double *d_b, double *d_x, double tol, int maxiter, int reps,
                            bn_kernel_stat *out, int max_out, int *n_out) {
  struct bn_solver *s = (struct bn_solver *)h;
  if (!s || !d_b || !d_x || !out || !n_out || max_out <= 0) {
    bn_set_error("kernel census: NULL argument");
    return BN_ERR_ARG;
  }
  int rc = bn_ensure_device(s);
  if (rc) return rc;
  if (reps < 1) reps = 1;
  struct bn_census *c = (struct bn_census *)calloc(1, sizeof(*c));
  c->capev = 16384;
  c->ev = (hipEvent_t *)calloc((size_t)2 * c->capev, sizeof(hipEvent_t));
  c->ev_shape = (int *)calloc((size_t)c->capev, sizeof(int));
  for (int i = 0; i < 2 * c->capev && !rc; i++)
    if (hipEventCreate(&c->ev[i]) != hipSuccess) rc = BN_ERR_HIP;
  /* the solve, eager (every launch recorded), graphs untouched */
  const int graph = s->opt.graph;
  if (!rc) {
    s->opt.graph = 0;
    s->census = c;
    const int it = bn_pcg_run(s, d_b, d_x, tol, maxiter, 0, NULL);
    s->census = NULL;
    s->opt.graph = graph;
    if (it == BN_ERR_HIP || it == BN_ERR_ARG || it == BN_ERR_NO_DEVICE) rc = it;
  }
  if (!rc && hipStreamSynchronize(s->stream) != hipSuccess) rc = BN_ERR_HIP;
  for (int e = 0; e < c->nev && !rc; e++) {
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, c->ev[2 * e], c->ev[2 * e + 1]) == hipSuccess)
      c->shape[c->ev_shape[e]].us_solve += 1e3 * ms;
  }
  /* each shape again, cold: a 320 MB read before every timed launch */
  void *flush = NULL;
  const size_t nflush = (size_t)40 << 20; /* doubles */
  if (!rc && hipMalloc(&flush, nflush * sizeof(double)) != hipSuccess) rc = BN_ERR_HIP;
  if (!rc) bn_fill('d', (int)nflush, flush, 1.0, s->stream);
  for (int i = 0; i < c->nshape && !rc; i++) {
    double tot = 0.0;
    for (int r = 0; r < reps; r++) {
      bn_dot_partial((int)nflush, (const double *)flush, (const double *)flush, NULL, NULL, s->part1, NULL,
                     s->stream);
      bn_time_next_launch(c->ev[2 * r], c->ev[2 * r + 1]);
      relaunch(s, &c->shape[i].key);
      bn_time_next_launch(NULL, NULL);
      hipStreamSynchronize(s->stream);
      float ms = 0.0f;
      if (hipEventElapsedTime(&ms, c->ev[2 * r], c->ev[2 * r + 1]) == hipSuccess) tot += 1e3 * ms;
    }
    c->shape[i].us_cold = tot / reps;
  }
  if (flush) hipFree(flush);
  /* per symbol */
  int nout = 0;
  for (int i = 0; i < c->nshape && !rc; i++) {
    const cshape *sh = &c->shape[i];
    int j = 0;
    while (j < nout && strcmp(out[j].name, sh->name)) j++;
    if (j == nout) {
      if (nout == max_out) continue;
      memset(&out[j], 0, sizeof(out[j]));
      snprintf(out[j].name, sizeof(out[j].name), "%s", sh->name);
      nout++;
    }
    out[j].launches += sh->count;
    out[j].shapes++;
    out[j].bytes += sh->bytes * sh->count;
    out[j].us_solve += sh->us_solve;
    out[j].us_cold += sh->us_cold * sh->count;
  }
  *n_out = nout;
  if (!rc && c->overflow) {
    bn_set_error("kernel census: more launches than event slots; the in-solve times are partial");
    rc = BN_ERR_ARG;
  }
  for (int i = 0; i < 2 * c->capev; i++)
    if (c->ev[i]) hipEventDestroy(c->ev[i]);
  free(c->ev);
  free(c->ev_shape);
  free(c->shape);
  free(c);
  return rc;
}
