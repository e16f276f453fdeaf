/*
 * api.c - the exported C ABI: the opaque-handle solver of
 * C_amg1/src/amg_wrapper.c:51-268 on the GPU, the bn_* extensions, and
 * error reporting.
 */
#include <math.h>
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "bn_solver.h"
#include "csr.h"

/* ---------------------------------------------------------------------
 * errors
 * ------------------------------------------------------------------- */
static __thread char g_err[512];

void bn_set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

const char *bn_last_error(void) { return g_err; }

int bn_device_available(void) { return bn_device_ok(); }

static double now_ms(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return 1e3 * (double)t.tv_sec + 1e-6 * (double)t.tv_nsec;
}

/* ---------------------------------------------------------------------
 * device bookkeeping
 * ------------------------------------------------------------------- */
void bn_track_device(struct bn_solver *s, void *p) {
  if (s->ndptr == s->capdptr) {
    s->capdptr = s->capdptr ? 2 * s->capdptr : 64;
    s->dptr = (void **)realloc(s->dptr, (size_t)s->capdptr * sizeof(void *));
  }
  s->dptr[s->ndptr++] = p;
}

void bn_track_host(struct bn_solver *s, void *p) {
  if (s->nhptr == s->caphptr) {
    s->caphptr = s->caphptr ? 2 * s->caphptr : 64;
    s->hptr = (void **)realloc(s->hptr, (size_t)s->caphptr * sizeof(void *));
  }
  s->hptr[s->nhptr++] = p;
}

static void free_tri_host(bn_dtri *t) {
  free(t->ptr_f);
  free(t->ptr_b);
  memset(t, 0, sizeof(*t));
}

/* every captured step graph goes (they hold device pointers as kernel
 * arguments); the next solve captures again */
void bn_drop_graphs(struct bn_solver *s) {
  if (s->stream) hipStreamSynchronize(s->stream);
  bn_segments_release(s);
  for (int k = 0; k < s->ngraph; k++) {
    if (s->gexec[k]) hipGraphExecDestroy(s->gexec[k]);
    if (s->graph[k]) hipGraphDestroy(s->graph[k]);
    s->gexec[k] = NULL;
    s->graph[k] = NULL;
  }
  s->ngraph = 0;
  for (int k = 0; k < s->blk.ngraph; k++) {
    if (s->blk.gexec[k]) hipGraphExecDestroy(s->blk.gexec[k]);
    if (s->blk.graph[k]) hipGraphDestroy(s->blk.graph[k]);
    s->blk.gexec[k] = NULL;
    s->blk.graph[k] = NULL;
  }
  s->blk.ngraph = 0;
}

void bn_device_release(struct bn_solver *s) {
  bn_drop_graphs(s);
  bn_rebuild_release(s);
  for (int k = 0; k < s->ndptr; k++) hipFree(s->dptr[k]);
  free(s->dptr);
  s->dptr = NULL;
  s->ndptr = s->capdptr = 0;
  for (int k = 0; k < s->nhptr; k++) free(s->hptr[k]);
  free(s->hptr);
  s->hptr = NULL;
  s->nhptr = s->caphptr = 0;
  if (s->blk.host_scal) hipHostFree(s->blk.host_scal);
  memset(&s->blk, 0, sizeof(s->blk));
  for (int l = 0; l < AMG_MAX_LEVELS; l++) {
    free(s->lev[l].col_ptr);
    free(s->lev[l].perm_h);
    bn_sweep_free_host(&s->lev[l].xs[0]);
    bn_sweep_free_host(&s->lev[l].xs[1]);
  }
  memset(s->lev, 0, sizeof(s->lev));
  free_tri_host(&s->fine_ic);
  free_tri_host(&s->coarse_ic);
  memset(&s->Asolve, 0, sizeof(s->Asolve));
  s->asolve_shared = 0;
  s->vperm_state = 0;
  s->vperm_d = NULL;
  s->cinv_d = NULL;
  s->cinv_s = NULL;
  s->dev_ready = 0;
}

static void release_all(struct bn_solver *s) {
  bn_device_release(s);
  if (s->host_scal) hipHostFree(s->host_scal);
  s->host_scal = NULL;
  if (s->host_step) hipHostFree(s->host_step);
  s->host_step = NULL;
  if (s->ev0) hipEventDestroy(s->ev0);
  if (s->ev1) hipEventDestroy(s->ev1);
  s->ev0 = s->ev1 = NULL;
  if (s->stream) hipStreamDestroy(s->stream);
  s->stream = NULL;
}

int bn_ensure_device(struct bn_solver *s);
static int ensure_device(struct bn_solver *s) { return bn_ensure_device(s); }

int bn_ensure_device(struct bn_solver *s) {
  if (s->dev_ready) return 0;
  if (!bn_device_ok()) {
    bn_set_error("no HIP device available: libbnamg solves run only on the GPU");
    return BN_ERR_NO_DEVICE;
  }
  int rc = bn_device_upload(s);
  if (rc) bn_device_release(s);
  return rc;
}

/* ---------------------------------------------------------------------
 * options and creation
 * ------------------------------------------------------------------- */
void bn_options_default(bn_options *o) {
  memset(o, 0, sizeof(*o));
  o->coarsening = COARSEN_PMIS;  /* amg_wrapper.c:89 */
  o->threshold = 0.25;
  o->min_size = 4096;            /* test_amg.c:90 */
  o->max_levels = AMG_MAX_LEVELS;
  o->nu = 1;
  o->smoother = SMOOTHER_CF_GS;  /* amg_wrapper.c:87 */
  o->omega = 0.5;
  o->precision = BN_PREC_FP64;
  o->coarse_solver = BN_COARSE_AUTO;
  o->fine_ic = BN_IC_EXACT;
  o->ic_sweeps = 3;
  o->block_smoother = 0;
  o->chunk_rows = BN_CHUNK_MAX_ROWS;
  o->max_colours = 32;
  o->flexible = 0;
  o->stop = BN_STOP_BNORM;
  o->symmetrize = 1;
  o->device = -1;
  o->verbose = 0;
  o->graph = 1;
}

static void copy_input(int n, int nnz, const int *ia, const int *ja, const double *val, int sym,
                       dcsr_mat *out) {
  dcsr_mat in = {n, n, nnz, (int *)ia, (int *)ja, (double *)val};
  if (sym) bn_csr_symmetrize(&in, out);
  else bn_csr_copy(&in, out);
}

static int valid_csr(int n, int nnz, const int *ia, const int *ja) {
  if (n <= 0 || nnz < 0 || !ia || (nnz > 0 && !ja)) return 0;
  if (ia[0] != 0 || ia[n] != nnz) return 0;
  for (int i = 0; i < n; i++)
    if (ia[i + 1] < ia[i]) return 0;
  for (int k = 0; k < nnz; k++)
    if (ja[k] < 0 || ja[k] >= n) return 0;
  return 1;
}

void *bn_solver_create(int n, int nnz, const int *ia, const int *ja, const double *val,
                       const bn_options *opt) {
  if (!valid_csr(n, nnz, ia, ja) || !val) {
    bn_set_error("bn_solver_create: malformed CSR input (n=%d, nnz=%d)", n, nnz);
    return NULL;
  }
  struct bn_solver *s = (struct bn_solver *)calloc(1, sizeof(struct bn_solver));
  if (!s) return NULL;
  bn_options_default(&s->opt);
  if (opt) s->opt = *opt;
  if (s->opt.max_levels <= 0 || s->opt.max_levels > AMG_MAX_LEVELS) s->opt.max_levels = AMG_MAX_LEVELS;
  if (s->opt.nu <= 0) s->opt.nu = 1;
  if (s->opt.max_colours <= 0) s->opt.max_colours = 32;
  if (s->opt.chunk_rows <= 0 || s->opt.chunk_rows > BN_CHUNK_MAX_ROWS) s->opt.chunk_rows = BN_CHUNK_MAX_ROWS;
  if (s->opt.ic_sweeps <= 0) s->opt.ic_sweeps = 3;
  const double t0 = now_ms();
  copy_input(n, nnz, ia, ja, val, s->opt.symmetrize, &s->A_own);
  s->A = &s->A_own;
  s->amg = &s->amg_own;
  s->L = &s->L_own;
  amg_data *a = s->amg;
  a->threshold = s->opt.threshold;
  a->min_size = s->opt.min_size;
  a->max_levels = s->opt.max_levels;
  a->nu = s->opt.nu;
  a->smoother = s->opt.smoother;
  a->omega_jacobi = s->opt.omega;
  a->coarsening = s->opt.coarsening;
  bn_build_hierarchy(a, s->A, s->opt.strength);
  const double th = now_ms();
  ichol_compute(s->A, &s->L_own);
  s->last.setup_ms = now_ms() - t0;
  const char *lg = getenv("BN_SETUP_LOG");
  if (lg && *lg && *lg != '0')
    fprintf(stderr, "[setup] hierarchy %.1f ms, fine IC(0) %.1f ms\n", th - t0, now_ms() - th);
  s->last.num_levels = a->num_levels;
  if (s->opt.verbose) {
    fprintf(stderr, "bnamg hierarchy (%s, theta=%.2f, min_size=%d): %d levels\n",
            a->coarsening == COARSEN_PMIS ? "PMIS" : a->coarsening == COARSEN_HMIS ? "HMIS" : "RS",
            a->threshold, a->min_size, a->num_levels);
    for (int l = 0; l < a->num_levels; l++)
      fprintf(stderr, "  level %d: n = %d, nnz = %d\n", l + 1, a->levels[l].A.row, a->levels[l].A.nnz);
  }
  if (bn_device_ok()) {
    if (ensure_device(s) && s->opt.verbose) fprintf(stderr, "bnamg: upload failed: %s\n", g_err);
  }
  return s;
}

/* C_amg1/src/amg_wrapper.c:51-97: PMIS, CF-GS, symmetrised input */
void *amg_solver_create(int n, int nnz, const int *ia, const int *ja, const double *val,
                        double threshold, int min_size, int max_levels, int nu) {
  bn_options o;
  bn_options_default(&o);
  o.threshold = threshold;
  o.min_size = min_size;
  o.max_levels = (max_levels > 0 && max_levels <= AMG_MAX_LEVELS) ? max_levels : AMG_MAX_LEVELS;
  o.nu = nu > 0 ? nu : 1;
  return bn_solver_create(n, nnz, ia, ja, val, &o);
}

/* Upload after update / set_matrix.  Those reference entry points always
 * return 0 (C_amg1/src/amg_wrapper.c:148,198), so a failed upload is
 * reported on stderr and through bn_last_error(); the next solve tries the
 * upload again and returns -maxiter if it still fails. */
static void reupload(struct bn_solver *s, const char *who) {
  if (ensure_device(s)) fprintf(stderr, "%s: device upload failed: %s\n", who, g_err);
}

/* C_amg1/src/amg_wrapper.c:114-149: same pattern, values rebuilt on the
 * stored split and P pattern, new fine IC(0); the device mirror is
 * uploaded again. */
int amg_solver_update(void *h, int n, int nnz, const int *ia, const int *ja, const double *val) {
  struct bn_solver *s = (struct bn_solver *)h;
  if (!s || s->borrowed) {
    bn_set_error("amg_solver_update: not a solver handle");
    fprintf(stderr, "amg_solver_update: %s\n", g_err);
    return 0;
  }
  dcsr_mat nw;
  copy_input(n, nnz, ia, ja, val, s->opt.symmetrize, &nw);
  const double t0 = now_ms();
  ichol_free(&s->L_own);
  bn_csr_free(&s->A_own);
  s->A_own = nw;
  amg_rebuild_values(s->amg, s->A);
  ichol_compute(s->A, &s->L_own);
  s->last.setup_ms = now_ms() - t0;
  int was = s->dev_ready;
  bn_device_release(s);
  if (was || bn_device_ok()) reupload(s, "amg_solver_update");
  return 0;
}

/* C_amg1/src/amg_wrapper.c:168-199: replace the solve matrix and its
 * IC(0), keep the hierarchy (two-matrix preconditioning) */
int amg_solver_set_matrix(void *h, int n, int nnz, const int *ia, const int *ja, const double *val) {
  struct bn_solver *s = (struct bn_solver *)h;
  if (!s || s->borrowed) {
    bn_set_error("amg_solver_set_matrix: not a solver handle");
    fprintf(stderr, "amg_solver_set_matrix: %s\n", g_err);
    return 0;
  }
  dcsr_mat nw;
  copy_input(n, nnz, ia, ja, val, s->opt.symmetrize, &nw);
  ichol_free(&s->L_own);
  bn_csr_free(&s->A_own);
  s->A_own = nw;
  ichol_compute(s->A, &s->L_own);
  int was = s->dev_ready;
  bn_device_release(s);
  if (was || bn_device_ok()) reupload(s, "amg_solver_set_matrix");
  return 0;
}

/* device copies of caller vectors in caller numbering */
static int stage_in(struct bn_solver *s, const double *hb, const double *hx, double **db, double **dx) {
  const int n = s->lev[0].n;
  *db = (double *)s->lev[0].b;
  *dx = (double *)s->lev[0].x;
  if (hb && hipMemcpyAsync(*db, hb, (size_t)n * sizeof(double), hipMemcpyHostToDevice, s->stream) != hipSuccess)
    return BN_ERR_HIP;
  if (hx && hipMemcpyAsync(*dx, hx, (size_t)n * sizeof(double), hipMemcpyHostToDevice, s->stream) != hipSuccess)
    return BN_ERR_HIP;
  return 0;
}

int bn_solver_solve_info(void *h, const double *b, double *x, double tol, int maxiter, int print_level,
                         bn_solve_info *info) {
  struct bn_solver *s = (struct bn_solver *)h;
  if (!s || !b || !x) return BN_ERR_ARG;
  int rc = ensure_device(s);
  if (rc) return rc;
  double *db, *dx;
  /* the V-cycle never touches level 0's b/x vectors (level 0 works on the
   * PCG's own vectors), so they hold the caller-numbered b and x */
  if (stage_in(s, b, x, &db, &dx)) return BN_ERR_HIP;
  int it = bn_pcg_run(s, db, dx, tol, maxiter, print_level, info);
  if (it == BN_ERR_HIP) return it;
  if (hipMemcpy(x, dx, (size_t)s->lev[0].n * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    return BN_ERR_HIP;
  return it;
}

/* C_amg1/src/amg_wrapper.c:214-225.  Returns the iteration count or
 * -maxiter, as the reference does: a solve that cannot run (no device, a
 * HIP failure, a NULL argument) counts as not converged, leaves x as it
 * was, and says why on stderr and through bn_last_error().  The bn_*
 * entry points return the BN_ERR_* code instead. */
int amg_solver_solve(void *h, const double *b, double *x, double tol, int maxiter, int print_level) {
  int rc = bn_solver_solve_info(h, b, x, tol, maxiter, print_level, NULL);
  if (rc == BN_ERR_NO_DEVICE || rc == BN_ERR_HIP || rc == BN_ERR_ARG) {
    if (rc == BN_ERR_ARG && !g_err[0]) bn_set_error("NULL handle or vector");
    fprintf(stderr, "amg_solver_solve: %s\n", bn_last_error());
    return -maxiter;
  }
  return rc;
}

int bn_solver_solve_device(void *h, const double *d_b, double *d_x, double tol, int maxiter, int print_level,
                           bn_solve_info *info) {
  struct bn_solver *s = (struct bn_solver *)h;
  if (!s) return BN_ERR_ARG;
  int rc = ensure_device(s);
  if (rc) return rc;
  return bn_pcg_run(s, d_b, d_x, tol, maxiter, print_level, info);
}

int bn_solver_precond(void *h, const double *g, double *z, int which) {
  struct bn_solver *s = (struct bn_solver *)h;
  if (!s || !g || !z) return BN_ERR_ARG;
  int rc = ensure_device(s);
  if (rc) return rc;
  if (which == 0 && !s->L) return BN_ERR_ARG;
  if ((rc = bn_vc_prepare(s))) return rc;
  const int n = s->lev[0].n;
  double *db, *dx;
  if (stage_in(s, g, NULL, &db, &dx)) return BN_ERR_HIP;
  bn_permute(n, s->lev[0].perm_d, db, s->r, 0, s->stream);
  if (which == 1) bn_vc_vcycle_precond(s, s->r, s->z);
  else bn_vc_ichol_precond(s, s->r, s->z);
  bn_permute(n, s->lev[0].perm_d, s->z, s->q, 1, s->stream);
  if (hipMemcpyAsync(z, s->q, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, s->stream) != hipSuccess)
    return BN_ERR_HIP;
  return hipStreamSynchronize(s->stream) == hipSuccess ? 0 : BN_ERR_HIP;
}

int bn_solver_last_info(void *h, bn_solve_info *info) {
  struct bn_solver *s = (struct bn_solver *)h;
  if (!s || !info) return BN_ERR_ARG;
  *info = s->last;
  return 0;
}

int bn_solver_set_mode(void *h, int precision, int fine_ic, int ic_sweeps, int flexible, int stop) {
  struct bn_solver *s = (struct bn_solver *)h;
  if (!s) return BN_ERR_ARG;
  s->opt.precision = precision;
  s->vt = precision == BN_PREC_FP32 ? 's' : 'd';
  s->opt.fine_ic = fine_ic;
  if (ic_sweeps > 0) s->opt.ic_sweeps = ic_sweeps;
  s->opt.flexible = flexible;
  s->opt.stop = stop;
  return 0;
}

/* C_amg1/src/amg_wrapper.c:241-246 */
void amg_solver_set_smoother(void *h, int smoother, double omega) {
  struct bn_solver *s = (struct bn_solver *)h;
  if (!s) return;
  /* captured step graphs hold the smoother's launches and omega as kernel
   * arguments: a change drops them, the next solve captures again */
  if (smoother != s->opt.smoother || (smoother == SMOOTHER_JACOBI && omega != s->opt.omega)) bn_drop_graphs(s);
  s->opt.smoother = smoother;
  s->amg->smoother = smoother;
  if (smoother == SMOOTHER_JACOBI) {
    s->opt.omega = omega;
    s->amg->omega_jacobi = omega;
  }
}

int amg_solver_size(void *h) { return h ? ((struct bn_solver *)h)->A->row : 0; }
int amg_solver_num_levels(void *h) { return h ? ((struct bn_solver *)h)->amg->num_levels : 0; }

void amg_solver_free(void *h) {
  struct bn_solver *s = (struct bn_solver *)h;
  if (!s) return;
  release_all(s);
  free(s->tune);
  if (!s->borrowed) {
    ichol_free(&s->L_own);
    amg_free(&s->amg_own);
    bn_csr_free(&s->A_own);
  }
  free(s);
}

/* ---------------------------------------------------------------------
 * introspection
 * ------------------------------------------------------------------- */
int bn_solver_level_dims(void *h, int lev, int *dims) {
  struct bn_solver *s = (struct bn_solver *)h;
  if (!s || lev < 0 || lev >= s->amg->num_levels) return BN_ERR_ARG;
  const amg_level *L = &s->amg->levels[lev];
  const int last = (lev == s->amg->num_levels - 1);
  dims[0] = L->A.row;
  dims[1] = L->A.nnz;
  dims[2] = last ? 0 : L->n_coarse;
  dims[3] = last ? 0 : L->P.nnz;
  dims[4] = last ? 0 : L->A_filtered.nnz;
  dims[5] = last ? L->L_ichol.nnz : 0;
  dims[6] = s->L ? s->L->nnz : 0;
  dims[7] = (!last && L->cf_order) ? 1 : 0;
  return 0;
}

static void copy_csr(const dcsr_mat *m, int *ia, int *ja, double *val) {
  if (ia) memcpy(ia, m->ia, ((size_t)m->row + 1) * sizeof(int));
  if (ja) memcpy(ja, m->ja, (size_t)m->nnz * sizeof(int));
  if (val) memcpy(val, m->val, (size_t)m->nnz * sizeof(double));
}

int bn_solver_level_csr(void *h, int lev, int what, int *ia, int *ja, double *val) {
  struct bn_solver *s = (struct bn_solver *)h;
  if (!s || lev < 0 || lev >= s->amg->num_levels) return BN_ERR_ARG;
  const amg_level *L = &s->amg->levels[lev];
  if (what != 0 && lev == s->amg->num_levels - 1) return BN_ERR_ARG;
  copy_csr(what == 0 ? &L->A : what == 1 ? &L->P : &L->A_filtered, ia, ja, val);
  return 0;
}

int bn_solver_level_split(void *h, int lev, int *mis, int *cf_order, double *l1) {
  struct bn_solver *s = (struct bn_solver *)h;
  if (!s || lev < 0 || lev >= s->amg->num_levels) return BN_ERR_ARG;
  const amg_level *L = &s->amg->levels[lev];
  if (mis && L->mis) memcpy(mis, L->mis, (size_t)L->n * sizeof(int));
  if (cf_order && L->cf_order) memcpy(cf_order, L->cf_order, (size_t)L->n * sizeof(int));
  if (l1) memcpy(l1, L->l1_diag, (size_t)L->n * sizeof(double));
  return 0;
}

int bn_solver_ichol(void *h, int which, int *ia, int *ja, double *val) {
  struct bn_solver *s = (struct bn_solver *)h;
  if (!s) return BN_ERR_ARG;
  const ichol_factor *F = which ? &s->amg->levels[s->amg->num_levels - 1].L_ichol : s->L;
  if (!F) return BN_ERR_ARG;
  if (ia) memcpy(ia, F->ia, ((size_t)F->n + 1) * sizeof(int));
  if (ja) memcpy(ja, F->ja, (size_t)F->nnz * sizeof(int));
  if (val) memcpy(val, F->val, (size_t)F->nnz * sizeof(double));
  return 0;
}

int bn_solver_schedule(void *h, int lev, int *ncolours, int *use_gs, int *lanes) {
  struct bn_solver *s = (struct bn_solver *)h;
  if (!s || lev < 0 || lev >= s->amg->num_levels) return BN_ERR_ARG;
  int rc = ensure_device(s);
  if (rc) return rc;
  if (ncolours) *ncolours = s->lev[lev].ncol;
  if (use_gs) *use_gs = s->lev[lev].use_gs;
  if (lanes) *lanes = s->lev[lev].A.lanes;
  return 0;
}

int bn_solver_ic_levels(void *h, int *fwd, int *bwd) {
  struct bn_solver *s = (struct bn_solver *)h;
  if (!s) return BN_ERR_ARG;
  int rc = ensure_device(s);
  if (rc) return rc;
  if (fwd) *fwd = s->fine_ic.nlev_f;
  if (bwd) *bwd = s->fine_ic.nlev_b;
  return 0;
}
