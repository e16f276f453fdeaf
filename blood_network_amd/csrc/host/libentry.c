/*
 * libentry.c - the library-level solve entry points of C_amg1
 * (include/amg.h:81-96,111-115) over the caller's host structures.
 *
 * A device mirror of the caller's amg_data / dcsr_mat / ichol_factor is
 * built on first use and cached by address; amg_free, amg_rebuild_values
 * and ichol_free drop the entries that reference the structure they
 * change.  Vectors cross the host/device boundary on every call (these
 * entry points take host pointers); the handle API and
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
  bn_dtri tri;       /* KIND_TRI */
  double *v[6];      /* KIND_TRI/KIND_OP device vectors */
  double *hv;        /* pinned host staging */
  int n;
  struct entry *next;
} entry;

static pthread_mutex_t g_mu = PTHREAD_MUTEX_INITIALIZER;
static entry *g_head = NULL;

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
  free(e);
}

/* drop every cached mirror that references p */
static void drop(const void *p) {
  pthread_mutex_lock(&g_mu);
  entry **pp = &g_head;
  while (*pp) {
    entry *e = *pp;
    if (e->k1 == p || e->k2 == p || e->k3 == p) {
      *pp = e->next;
      destroy(e);
    } else {
      pp = &e->next;
    }
  }
  pthread_mutex_unlock(&g_mu);
}

void bn_cache_drop(const void *p) { drop(p); }

static entry *find(int kind, const void *k1, const void *k2, const void *k3) {
  for (entry *e = g_head; e; e = e->next)
    if (e->kind == kind && e->k1 == k1 && e->k2 == k2 && e->k3 == k3) return e;
  return NULL;
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

/* hierarchy mirror (optionally with solve matrix A and fine factor L) */
static entry *hier_entry(const amg_data *amg, const dcsr_mat *A, const ichol_factor *L) {
  pthread_mutex_lock(&g_mu);
  entry *e = find(KIND_HIER, amg, A, L);
  if (e) {
    pthread_mutex_unlock(&g_mu);
    return e;
  }
  struct bn_solver *s = (struct bn_solver *)calloc(1, sizeof(struct bn_solver));
  bn_options_default(&s->opt);
  s->opt.smoother = amg->smoother;
  s->opt.nu = amg->nu > 0 ? amg->nu : 1;
  s->opt.omega = amg->omega_jacobi;
  s->opt.coarsening = amg->coarsening;
  s->amg = (amg_data *)amg;
  s->A = A ? A : &amg->levels[0].A;
  s->L = L;
  s->borrowed = 1;
  int rc = bn_device_upload(s);
  if (rc) {
    pthread_mutex_unlock(&g_mu);
    fatal("device upload of the AMG hierarchy failed");
  }
  e = (entry *)calloc(1, sizeof(entry));
  e->kind = KIND_HIER;
  e->k1 = amg;
  e->k2 = A;
  e->k3 = L;
  e->s = s;
  e->n = amg->levels[0].n;
  if (hipHostMalloc((void **)&e->hv, 2 * (size_t)e->n * sizeof(double), 0) != hipSuccess) {
    pthread_mutex_unlock(&g_mu);
    fatal("pinned allocation failed");
  }
  e->next = g_head;
  g_head = e;
  pthread_mutex_unlock(&g_mu);
  return e;
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
  entry *e = hier_entry(amg, NULL, NULL);
  struct bn_solver *s = e->s;
  s->opt.nu = nu > 0 ? nu : 1;
  s->opt.smoother = amg->smoother;
  s->opt.omega = amg->omega_jacobi;
  s->vt = 'd';
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
  s->opt.nu = amg->nu > 0 ? amg->nu : 1;
  s->opt.smoother = amg->smoother;
  s->opt.omega = amg->omega_jacobi;
  s->vt = 'd';
  to_dev(s, e->hv, g, e->n, s->q, s->r);
  bn_vc_vcycle_precond(s, s->r, s->z);
  from_dev(s, e->hv, s->z, e->n, s->q, x);
}

void amg_ichol_precond(const amg_data *amg, const dcsr_mat *A, const ichol_factor *L, const REAL *g,
                       REAL *x) {
  entry *e = hier_entry(amg, A, L);
  struct bn_solver *s = e->s;
  s->opt.nu = amg->nu > 0 ? amg->nu : 1;
  s->opt.smoother = amg->smoother;
  s->opt.omega = amg->omega_jacobi;
  s->vt = 'd';
  to_dev(s, e->hv, g, e->n, s->q, s->r);
  bn_vc_ichol_precond(s, s->r, s->z);
  from_dev(s, e->hv, s->z, e->n, s->q, x);
}

/* ---------------------------------------------------------------------
 * ichol_solve (amg.h:90): level-scheduled sweeps of a cached factor
 * ------------------------------------------------------------------- */
void ichol_solve(const ichol_factor *L, const REAL *b, REAL *x) {
  pthread_mutex_lock(&g_mu);
  entry *e = find(KIND_TRI, L, NULL, NULL);
  if (!e) {
    e = (entry *)calloc(1, sizeof(entry));
    e->kind = KIND_TRI;
    e->k1 = L;
    e->n = L->n;
    if (new_stream_solver(&e->s) || bn_tri_upload(e->s, L, &e->tri)) {
      pthread_mutex_unlock(&g_mu);
      fatal("ichol_solve device set-up failed");
    }
    for (int k = 0; k < 3; k++) e->v[k] = dvec(e->s, L->n);
    hipHostMalloc((void **)&e->hv, (size_t)(L->n > 0 ? L->n : 1) * sizeof(double), 0);
    e->next = g_head;
    g_head = e;
  }
  pthread_mutex_unlock(&g_mu);
  struct bn_solver *s = e->s;
  const int n = e->n;
  memcpy(e->hv, b, (size_t)n * sizeof(double));
  hipMemcpyAsync(e->v[0], e->hv, (size_t)n * sizeof(double), hipMemcpyHostToDevice, s->stream);
  bn_vc_tri_apply(s, &e->tri, 'd', e->v[0], e->v[1], e->v[2]);
  hipMemcpyAsync(e->hv, e->v[2], (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s->stream);
  if (hipStreamSynchronize(s->stream) != hipSuccess) fatal("device failure in ichol_solve");
  memcpy(x, e->hv, (size_t)n * sizeof(double));
}

/* ---------------------------------------------------------------------
 * amg_pcg (amg.h:111): PCG with a host preconditioner callback; SpMV and
 * BLAS-1 on the device, r and z cross to the host once per step.
 * ------------------------------------------------------------------- */
INT amg_pcg(const dcsr_mat *A, const REAL *b, REAL *x,
            void (*precond)(const void *data, const REAL *g, REAL *Bg), const void *precond_data,
            REAL tol, INT maxiter, INT print_level) {
  pthread_mutex_lock(&g_mu);
  entry *e = find(KIND_OP, A, NULL, NULL);
  if (!e) {
    e = (entry *)calloc(1, sizeof(entry));
    e->kind = KIND_OP;
    e->k1 = A;
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
    e->next = g_head;
    g_head = e;
  }
  pthread_mutex_unlock(&g_mu);
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
  return it;
}
