/*
 * libentry.c - the library-level solve entry points of C_amg1
 * (include/amg.h:81-96,111-115) over the caller's host structures.
 *
 * A device mirror of the caller's amg_data / ichol_factor is built on
 * first use and cached by address.  The library's own routines that change
 * such a structure (amg_build_hierarchy, amg_rebuild_values, amg_free,
 * ichol_compute, ichol_free) drop the mirrors that reference it.  A caller's
 * dcsr_mat has no library routine that frees or changes it (dcsr_mat_free
 * is header-only), so every call compares it with the host copy its mirror
 * was made from: a new pattern builds the mirror again, new values are
 * copied to the device, as the reference reads A afresh on every call.
 *
 * Each cached mirror has a lock held for the whole of a call that uses it,
 * so two threads may call these entry points on the same structures (they
 * run one after the other) or on different ones (they run concurrently, one
 * HIP stream per mirror).  Vectors cross the host/device boundary on every
 * call (these entry points take host pointers); the handle API and
 * bn_solver_solve_device avoid that.
 */
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bn_solver.h"
#include "csr.h"

void bn_set_error(const char *fmt, ...);
void bn_track_device(struct bn_solver *s, void *p);
int bn_tri_upload(struct bn_solver *s, const ichol_factor *F, bn_dtri *t);
int bn_csr_upload(struct bn_solver *s, const dcsr_mat *m, bn_dcsr *d);
void bn_vc_tri_apply(struct bn_solver *s, bn_dtri *t, char vt, const void *b, void *y, void *x);
void bn_device_release(struct bn_solver *s);

enum { KIND_HIER = 0, KIND_TRI = 1, KIND_OP = 2 };

typedef struct entry {
  int kind;
  const void *k1, *k2, *k3;
  struct bn_solver *s;
  pthread_mutex_t mu;  /* held for the whole of a call on this mirror */
  dcsr_mat held;       /* host copy of the caller's dcsr_mat the mirror holds */
  unsigned gen;        /* library-option generation it was built under */
  bn_dtri tri;         /* KIND_TRI */
  double *v[6];        /* KIND_TRI/KIND_OP device vectors */
  double *hv;          /* pinned host staging */
  int n;
  struct entry *next;
} entry;

static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;
static entry *g_head = NULL;
/* options of the hidden solvers behind library-level mirrors */
static bn_options g_lib_opt;
static int g_lib_opt_set = 0;
static unsigned g_lib_gen = 0;

static void fatal(const char *what) {
  fprintf(stderr, "libbnamg: %s: %s\n", what, bn_last_error());
  abort();
}

static void destroy(entry *e) {
  if (e->hv) hipHostFree(e->hv);
  if (e->s) {
    if (e->kind == KIND_HIER) {
      amg_solver_free(e->s); /* borrowed: frees device state only */
    } else {
      bn_device_release(e->s);
      if (e->s->host_scal) hipHostFree(e->s->host_scal);
      if (e->s->stream) hipStreamDestroy(e->s->stream);
      free(e->s->tune);
      free(e->s);
    }
  }
  bn_csr_free(&e->held);
  pthread_mutex_destroy(&e->mu);
  free(e);
}

static void unlink_entry(entry *e) {
  for (entry **pp = &g_head; *pp; pp = &(*pp)->next)
    if (*pp == e) {
      *pp = e->next;
      return;
    }
}

/* drop every cached mirror that references p */
static void drop(const void *p) {
  pthread_mutex_lock(&g_mu);
  entry **pp = &g_head;
  while (*pp) {
    entry *e = *pp;
    if (e->k1 == p || e->k2 == p || e->k3 == p) {
      *pp = e->next;
      pthread_mutex_lock(&e->mu); /* wait for a call in flight on it */
      pthread_mutex_unlock(&e->mu);
      destroy(e);
    } else {
      pp = &e->next;
    }
  }
  pthread_mutex_unlock(&g_mu);
}

void bn_cache_drop(const void *p) { drop(p); }

/* Options for the solvers behind the library-level entry points (half
 * cycles, V-cycle, AMG+IC preconditioner on a caller's amg_data): NULL
 * restores bn_options_default.  The smoother, nu, omega and coarsening
 * always come from the amg_data.  Mirrors built under other options are
 * dropped. */
void bn_set_library_options(const bn_options *o) {
  pthread_mutex_lock(&g_mu);
  if (o) {
    g_lib_opt = *o;
    g_lib_opt_set = 1;
  } else {
    g_lib_opt_set = 0;
  }
  g_lib_gen++;
  entry **pp = &g_head;
  while (*pp) {
    entry *e = *pp;
    if (e->kind == KIND_HIER) {
      *pp = e->next;
      pthread_mutex_lock(&e->mu);
      pthread_mutex_unlock(&e->mu);
      destroy(e);
    } else {
      pp = &e->next;
    }
  }
  pthread_mutex_unlock(&g_mu);
}

static entry *find(int kind, const void *k1, const void *k2, const void *k3) {
  for (entry *e = g_head; e; e = e->next)
    if (e->kind == kind && e->k1 == k1 && e->k2 == k2 && e->k3 == k3) return e;
  return NULL;
}

static entry *new_entry(int kind, const void *k1, const void *k2, const void *k3) {
  entry *e = (entry *)calloc(1, sizeof(entry));
  if (!e) fatal("out of host memory");
  e->kind = kind;
  e->k1 = k1;
  e->k2 = k2;
  e->k3 = k3;
  pthread_mutex_init(&e->mu, NULL);
  return e;
}

/* 0: A is the matrix `held` copies; 1: same pattern, other values;
 * 2: another size or pattern */
static int csr_change(const dcsr_mat *held, const dcsr_mat *A) {
  if (held->row != A->row || held->col != A->col || held->nnz != A->nnz) return 2;
  if (memcmp(held->ia, A->ia, ((size_t)A->row + 1) * sizeof(int)) ||
      memcmp(held->ja, A->ja, (size_t)A->nnz * sizeof(int)))
    return 2;
  return memcmp(held->val, A->val, (size_t)A->nnz * sizeof(double)) ? 1 : 0;
}

static int new_stream_solver(struct bn_solver **out) {
  struct bn_solver *s = (struct bn_solver *)calloc(1, sizeof(struct bn_solver));
  if (!bn_device_ok()) {
    free(s);
    bn_set_error("no HIP device available: libbnamg solves run only on the GPU");
    return BN_ERR_NO_DEVICE;
  }
  if (hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking) != hipSuccess ||
      hipHostMalloc((void **)&s->host_scal, 2 * BN_S_COUNT * sizeof(double), 0) != hipSuccess) {
    free(s);
    bn_set_error("HIP stream/pinned allocation failed");
    return BN_ERR_HIP;
  }
  *out = s;
  return 0;
}

static double *dvec(struct bn_solver *s, int n) {
  void *p = NULL;
  if (hipMalloc(&p, (size_t)(n > 0 ? n : 1) * sizeof(double)) != hipSuccess) return NULL;
  bn_track_device(s, p);
  return (double *)p;
}

/* Hierarchy mirror (optionally with solve matrix A and fine factor L),
 * returned locked: the caller unlocks e->mu when its call is done. */
static entry *hier_entry(const amg_data *amg, const dcsr_mat *A, const ichol_factor *L) {
  pthread_mutex_lock(&g_mu);
  entry *e = find(KIND_HIER, amg, A, L);
  if (e && (e->gen != g_lib_gen || (A && csr_change(&e->held, A)))) {
    /* built under other options, or the caller's solve matrix changed */
    unlink_entry(e);
    pthread_mutex_lock(&e->mu);
    pthread_mutex_unlock(&e->mu);
    destroy(e);
    e = NULL;
  }
  if (e) {
    pthread_mutex_lock(&e->mu);
    pthread_mutex_unlock(&g_mu);
    return e;
  }
  struct bn_solver *s = (struct bn_solver *)calloc(1, sizeof(struct bn_solver));
  if (g_lib_opt_set) s->opt = g_lib_opt;
  else bn_options_default(&s->opt);
  s->opt.smoother = amg->smoother;
  s->opt.nu = amg->nu > 0 ? amg->nu : 1;
  s->opt.omega = amg->omega_jacobi;
  s->opt.coarsening = amg->coarsening;
  s->opt.precision = BN_PREC_FP64;
  s->amg = (amg_data *)amg;
  s->A = A ? A : &amg->levels[0].A;
  s->L = L;
  s->borrowed = 1;
  int rc = bn_device_upload(s);
  if (rc) {
    pthread_mutex_unlock(&g_mu);
    fatal("device upload of the AMG hierarchy failed");
  }
  e = new_entry(KIND_HIER, amg, A, L);
  e->s = s;
  e->gen = g_lib_gen;
  e->n = amg->levels[0].n;
  if (A) bn_csr_copy(A, &e->held);
  if (hipHostMalloc((void **)&e->hv, 2 * (size_t)e->n * sizeof(double), 0) != hipSuccess) {
    pthread_mutex_unlock(&g_mu);
    fatal("pinned allocation failed");
  }
  e->next = g_head;
  g_head = e;
  pthread_mutex_lock(&e->mu);
  pthread_mutex_unlock(&g_mu);
  return e;
}

/* the smoother settings of the amg_data, and what they need on the device */
static void use_amg_settings(struct bn_solver *s, const amg_data *amg, int nu) {
  const int sm = amg->smoother;
  const double om = amg->omega_jacobi;
  if (sm != s->opt.smoother || (sm == SMOOTHER_JACOBI && om != s->opt.omega)) bn_drop_graphs(s);
  s->opt.nu = nu > 0 ? nu : 1;
  s->opt.smoother = sm;
  s->opt.omega = om;
  s->vt = 'd';
  if (bn_vc_prepare(s)) fatal("device set-up of the smoother failed");
}

/* copy host vector into level-0 numbering on the device */
static void to_dev(struct bn_solver *s, double *stage_h, const double *h, int n, double *d_tmp, double *d_out) {
  memcpy(stage_h, h, (size_t)n * sizeof(double));
  hipMemcpyAsync(d_tmp, stage_h, (size_t)n * sizeof(double), hipMemcpyHostToDevice, s->stream);
  bn_permute(n, s->lev[0].perm_d, d_tmp, d_out, 0, s->stream);
}

static void from_dev(struct bn_solver *s, double *stage_h, const double *d_in, int n, double *d_tmp, double *h) {
  bn_permute(n, s->lev[0].perm_d, d_in, d_tmp, 1, s->stream);
  hipMemcpyAsync(stage_h, d_tmp, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s->stream);
  if (hipStreamSynchronize(s->stream) != hipSuccess) fatal("device failure");
  memcpy(h, stage_h, (size_t)n * sizeof(double));
}

/* ---------------------------------------------------------------------
 * amg.h:81-86 half cycles and V-cycle
 * ------------------------------------------------------------------- */
static void half_cycle(const amg_data *amg, INT lev, const REAL *b, REAL *x, INT nu, int fwd) {
  if (lev < 0 || lev >= amg->num_levels) {
    bn_set_error("level %d outside the hierarchy's %d levels", lev, amg->num_levels);
    fatal(fwd ? "amg_fwdslash" : "amg_backslash");
  }
  entry *e = hier_entry(amg, NULL, NULL);
  struct bn_solver *s = e->s;
  use_amg_settings(s, amg, nu);
  const int nl = s->lev[lev].n;
  double *hb = e->hv, *hx = e->hv + e->n;
  /* level `lev` vectors in that level's numbering */
  double *db = (double *)s->g_in, *dx = (double *)s->u1, *tb = (double *)s->r1, *tx = (double *)s->icy;
  memcpy(hb, b, (size_t)nl * sizeof(double));
  memcpy(hx, x, (size_t)nl * sizeof(double));
  hipMemcpyAsync(tb, hb, (size_t)nl * sizeof(double), hipMemcpyHostToDevice, s->stream);
  hipMemcpyAsync(tx, hx, (size_t)nl * sizeof(double), hipMemcpyHostToDevice, s->stream);
  bn_permute(nl, s->lev[lev].perm_d, tb, db, 0, s->stream);
  bn_permute(nl, s->lev[lev].perm_d, tx, dx, 0, s->stream);
  if (fwd) bn_vc_fwdslash(s, lev, db, dx, 0);
  else bn_vc_backslash(s, lev, db, dx, 0);
  bn_permute(nl, s->lev[lev].perm_d, dx, tx, 1, s->stream);
  hipMemcpyAsync(hx, tx, (size_t)nl * sizeof(double), hipMemcpyDeviceToHost, s->stream);
  if (hipStreamSynchronize(s->stream) != hipSuccess) fatal("device failure in a V-cycle half");
  memcpy(x, hx, (size_t)nl * sizeof(double));
  pthread_mutex_unlock(&e->mu);
}

void amg_backslash(const amg_data *amg, INT lev, const REAL *b, REAL *x, INT nu) {
  half_cycle(amg, lev, b, x, nu, 0);
}

void amg_fwdslash(const amg_data *amg, INT lev, const REAL *b, REAL *x, INT nu) {
  half_cycle(amg, lev, b, x, nu, 1);
}

void amg_vcycle_precond(const amg_data *amg, const REAL *g, REAL *x) {
  entry *e = hier_entry(amg, NULL, NULL);
  struct bn_solver *s = e->s;
  use_amg_settings(s, amg, amg->nu);
  to_dev(s, e->hv, g, e->n, s->q, s->r);
  bn_vc_vcycle_precond(s, s->r, s->z);
  from_dev(s, e->hv, s->z, e->n, s->q, x);
  pthread_mutex_unlock(&e->mu);
}

void amg_ichol_precond(const amg_data *amg, const dcsr_mat *A, const ichol_factor *L, const REAL *g,
                       REAL *x) {
  entry *e = hier_entry(amg, A, L);
  struct bn_solver *s = e->s;
  use_amg_settings(s, amg, amg->nu);
  to_dev(s, e->hv, g, e->n, s->q, s->r);
  bn_vc_ichol_precond(s, s->r, s->z);
  from_dev(s, e->hv, s->z, e->n, s->q, x);
  pthread_mutex_unlock(&e->mu);
}

/* ---------------------------------------------------------------------
 * ichol_solve (amg.h:90): level-scheduled sweeps of a cached factor
 * ------------------------------------------------------------------- */
void ichol_solve(const ichol_factor *L, const REAL *b, REAL *x) {
  pthread_mutex_lock(&g_mu);
  entry *e = find(KIND_TRI, L, NULL, NULL);
  if (!e) {
    e = new_entry(KIND_TRI, L, NULL, NULL);
    e->n = L->n;
    if (new_stream_solver(&e->s) || bn_tri_upload(e->s, L, &e->tri) || bn_tri_ell_build(e->s, &e->tri)) {
      pthread_mutex_unlock(&g_mu);
      fatal("ichol_solve device set-up failed");
    }
    for (int k = 0; k < 3; k++) e->v[k] = dvec(e->s, L->n);
    hipHostMalloc((void **)&e->hv, (size_t)(L->n > 0 ? L->n : 1) * sizeof(double), 0);
    e->next = g_head;
    g_head = e;
  }
  pthread_mutex_lock(&e->mu);
  pthread_mutex_unlock(&g_mu);
  struct bn_solver *s = e->s;
  const int n = e->n;
  memcpy(e->hv, b, (size_t)n * sizeof(double));
  hipMemcpyAsync(e->v[0], e->hv, (size_t)n * sizeof(double), hipMemcpyHostToDevice, s->stream);
  bn_vc_tri_apply(s, &e->tri, 'd', e->v[0], e->v[1], e->v[2]);
  hipMemcpyAsync(e->hv, e->v[2], (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s->stream);
  if (hipStreamSynchronize(s->stream) != hipSuccess) fatal("device failure in ichol_solve");
  memcpy(x, e->hv, (size_t)n * sizeof(double));
  pthread_mutex_unlock(&e->mu);
}

/* ---------------------------------------------------------------------
 * amg_pcg (amg.h:111): PCG with a host preconditioner callback; SpMV and
 * BLAS-1 on the device, r and z cross to the host once per step.
 * ------------------------------------------------------------------- */
/* the device copy of A for amg_pcg, locked; built again when A's size or
 * pattern changed since the last call, values refreshed when only they
 * changed */
static entry *op_entry(const dcsr_mat *A) {
  pthread_mutex_lock(&g_mu);
  entry *e = find(KIND_OP, A, NULL, NULL);
  int change = e ? csr_change(&e->held, A) : 2;
  if (e && change == 2) {
    unlink_entry(e);
    pthread_mutex_lock(&e->mu);
    pthread_mutex_unlock(&e->mu);
    destroy(e);
    e = NULL;
  }
  if (!e) {
    e = new_entry(KIND_OP, A, NULL, NULL);
    e->n = A->row;
    if (new_stream_solver(&e->s) || bn_csr_upload(e->s, A, &e->s->Asolve)) {
      pthread_mutex_unlock(&g_mu);
      fatal("amg_pcg device set-up failed");
    }
    for (int k = 0; k < 6; k++) e->v[k] = dvec(e->s, A->row);
    e->s->scal = dvec(e->s, 2 * BN_S_COUNT);
    e->s->part1 = dvec(e->s, BN_RED_BLOCKS);
    e->s->part2 = dvec(e->s, BN_RED_BLOCKS);
    hipHostMalloc((void **)&e->hv, 2 * (size_t)(A->row > 0 ? A->row : 1) * sizeof(double), 0);
    bn_csr_copy(A, &e->held);
    e->next = g_head;
    g_head = e;
    change = 0;
  }
  pthread_mutex_lock(&e->mu);
  pthread_mutex_unlock(&g_mu);
  if (change == 1) {
    /* same pattern: the new values replace the device copy's FP64 values
     * (the only ones amg_pcg multiplies by) */
    memcpy(e->held.val, A->val, (size_t)A->nnz * sizeof(double));
    if (hipMemcpy(e->s->Asolve.vd, A->val, (size_t)A->nnz * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      fatal("amg_pcg: copy of new matrix values failed");
  }
  return e;
}

INT amg_pcg(const dcsr_mat *A, const REAL *b, REAL *x,
            void (*precond)(const void *data, const REAL *g, REAL *Bg), const void *precond_data,
            REAL tol, INT maxiter, INT print_level) {
  entry *e = op_entry(A);
  struct bn_solver *s = e->s;
  hipStream_t st = s->stream;
  const int n = e->n;
  const bn_dcsr *M = &s->Asolve;
  double *db = e->v[0], *dx = e->v[1], *r = e->v[2], *z = e->v[3], *p = e->v[4], *q = e->v[5];
  double *hr = e->hv, *hz = e->hv + n, *sc = s->scal, *hs = s->host_scal;
  hipMemcpyAsync(db, b, (size_t)n * sizeof(double), hipMemcpyHostToDevice, st);
  hipMemcpyAsync(dx, x, (size_t)n * sizeof(double), hipMemcpyHostToDevice, st);
  bn_csr_apply('d', BN_OP_RESID, n, M->ia, M->ja, M->vd, dx, r, db, M->lanes | M->deep, NULL, NULL, st);
  bn_dot_partial(n, db, db, NULL, NULL, s->part1, NULL, st);
  bn_scalar_finish(3, s->part1, NULL, NULL, sc, 0, st);
#define CALLBACK()                                                                            \
  do {                                                                                        \
    hipMemcpyAsync(hr, r, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st);             \
    if (hipStreamSynchronize(st) != hipSuccess) fatal("device failure in amg_pcg");           \
    if (precond) precond(precond_data, hr, hz);                                               \
    else memcpy(hz, hr, (size_t)n * sizeof(double));                                          \
    hipMemcpyAsync(z, hz, (size_t)n * sizeof(double), hipMemcpyHostToDevice, st);             \
  } while (0)
  CALLBACK();
  bn_copy('d', n, z, p, st);
  bn_dot_partial(n, r, z, r, r, s->part1, s->part2, st);
  bn_scalar_finish(0, s->part1, s->part2, NULL, sc, 0, st);
  hipMemcpyAsync(hs, sc, BN_S_COUNT * sizeof(double), hipMemcpyDeviceToHost, st);
  if (hipStreamSynchronize(st) != hipSuccess) fatal("device failure in amg_pcg");
  double bnorm = sqrt(hs[BN_S_BB]);
  if (bnorm == 0.0) bnorm = 1.0;
  double rnorm = sqrt(hs[BN_S_RR]);
  if (print_level >= 1) {
    fprintf(stderr, "  %4s  %8s  %8s  %6s\n", "iter", "||r||_B", "rel", "rho");
    fprintf(stderr, "  %4d  %8.2e  %8.2e\n", 0, sqrt(hs[BN_S_RZ0]), 1.0);
  }
  INT it = 0;
  if (hs[BN_S_RZ0] != 0.0) {
    for (it = 1; it <= maxiter; it++) {
      bn_csr_apply('d', BN_OP_SET, n, M->ia, M->ja, M->vd, p, q, NULL, M->lanes | M->deep, p, s->part1, st);
      bn_scalar_finish(1, s->part1, NULL, NULL, sc, 0, st);
      bn_pcg_update_xr(n, dx, p, r, q, sc, s->part2, st);
      CALLBACK();
      bn_dot_partial(n, r, z, NULL, NULL, s->part1, NULL, st);
      bn_scalar_finish(2, s->part1, s->part2, NULL, sc, 0, st);
      hipMemcpyAsync(hs, sc, BN_S_COUNT * sizeof(double), hipMemcpyDeviceToHost, st);
      if (hipStreamSynchronize(st) != hipSuccess) fatal("device failure in amg_pcg");
      rnorm = sqrt(hs[BN_S_RR]);
      if (print_level >= 1)
        fprintf(stderr, "  %4d  %8.2e  %8.2e  %6.4f\n", it, sqrt(hs[BN_S_RZNEW]), hs[BN_S_REL],
                hs[BN_S_RZNEW] / hs[BN_S_RZOLD]);
      if (hs[BN_S_REL] < tol) break;
      bn_pcg_update_p(n, p, z, sc, st);
    }
  }
#undef CALLBACK
  if (it > maxiter) {
    fprintf(stderr, "  PCG: no convergence after %d iterations, relres = %.2e\n", maxiter, rnorm / bnorm);
    it = -maxiter;
  }
  hipMemcpyAsync(hr, dx, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st);
  if (hipStreamSynchronize(st) != hipSuccess) fatal("device failure in amg_pcg");
  memcpy(x, hr, (size_t)n * sizeof(double));
  pthread_mutex_unlock(&e->mu);
  return it;
}
