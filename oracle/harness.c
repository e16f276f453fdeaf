/*
 * harness.c - flat C entry points over an AMG implementation with the
 * C_amg1 types (amg_data, dcsr_mat, ichol_factor), for ctypes tests.
 *
 * TEST INFRASTRUCTURE ONLY.  Compiled twice by oracle/Makefile: against the
 * oracle restatement (liboracle.so) and against the reference sources under
 * /root/reference/C_amg1 (oracle/_ref/libamgref.so, built with
 * -DHARNESS_REF), so tests can check one against the other.
 *
 * The set-up sequence mirrors C_amg1/test_amg.c:52-63,154-201: symmetrise
 * the input, build the hierarchy, IC(0) of the symmetrised fine matrix.
 */
#ifdef HARNESS_REF
#include "amg.h"
static void h_symmetrize(const dcsr_mat *a, dcsr_mat *out) {
  dcsr_mat at;
  dcsr_mat_alloc(&at, a->col, a->row, a->nnz);
  dcsr_transz(a, NULL, &at);
  dcsr_mat_add((dcsr_mat *)a, &at, out);
  for (INT k = 0; k < out->nnz; k++) out->val[k] *= 0.5;
  dcsr_mat_free(&at);
}
static void h_csr_free(dcsr_mat *a) { dcsr_mat_free(a); }
#else
#include "oracle.h"
static void h_symmetrize(const dcsr_mat *a, dcsr_mat *out) { oracle_symmetrize(a, out); }
static void h_csr_free(dcsr_mat *a) { oracle_csr_free(a); }
#endif
#include <time.h>

typedef struct {
  dcsr_mat A;
  amg_data amg;
  ichol_factor L;
} hsolver;

typedef struct { const amg_data *amg; const dcsr_mat *A; const ichol_factor *L; } hctx;

static void h_precond_cb(const void *d, const REAL *g, REAL *z) {
  const hctx *c = (const hctx *)d;
  amg_ichol_precond(c->amg, c->A, c->L, g, z);
}
static void h_vcycle_cb(const void *d, const REAL *g, REAL *z) {
  amg_vcycle_precond((const amg_data *)d, g, z);
}

/* Build: symmetrise (optional), hierarchy, fine IC(0). */
void *h_build(int n, int nnz, const int *ia, const int *ja, const double *val,
              double theta, int min_size, int max_levels, int nu,
              int coarsening, int smoother, int symmetrize) {
  hsolver *s = (hsolver *)calloc(1, sizeof(hsolver));
  dcsr_mat in;
  in.row = n; in.col = n; in.nnz = nnz;
  in.ia = (INT *)ia; in.ja = (INT *)ja; in.val = (REAL *)val;
  if (symmetrize) {
    h_symmetrize(&in, &s->A);
  } else {
    s->A.row = n; s->A.col = n; s->A.nnz = nnz;
    s->A.ia = (INT *)malloc(((size_t)n + 1) * sizeof(INT));
    s->A.ja = (INT *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(INT));
    s->A.val = (REAL *)malloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(REAL));
    memcpy(s->A.ia, ia, ((size_t)n + 1) * sizeof(INT));
    memcpy(s->A.ja, ja, (size_t)nnz * sizeof(INT));
    memcpy(s->A.val, val, (size_t)nnz * sizeof(REAL));
  }
  memset(&s->amg, 0, sizeof(amg_data));
  s->amg.threshold = theta;
  s->amg.min_size = min_size;
  s->amg.max_levels = max_levels;
  s->amg.nu = nu;
  s->amg.smoother = smoother;
  s->amg.omega_jacobi = 0.5;
  s->amg.coarsening = coarsening;
  amg_build_hierarchy(&s->amg, &s->A);
  ichol_compute(&s->A, &s->L);
  return s;
}

/* value-only rebuild on the same pattern (amg_solver_update semantics) */
void h_update(void *h, const double *val, int symmetrize) {
  hsolver *s = (hsolver *)h;
  dcsr_mat in = s->A;
  dcsr_mat nw;
  in.val = (REAL *)val;
  if (symmetrize) {
    h_symmetrize(&in, &nw);
  } else {
    nw = s->A;
    nw.ia = (INT *)malloc(((size_t)in.row + 1) * sizeof(INT));
    nw.ja = (INT *)malloc((size_t)in.nnz * sizeof(INT));
    nw.val = (REAL *)malloc((size_t)in.nnz * sizeof(REAL));
    memcpy(nw.ia, in.ia, ((size_t)in.row + 1) * sizeof(INT));
    memcpy(nw.ja, in.ja, (size_t)in.nnz * sizeof(INT));
    memcpy(nw.val, val, (size_t)in.nnz * sizeof(REAL));
  }
  ichol_free(&s->L);
  h_csr_free(&s->A);
  s->A = nw;
  amg_rebuild_values(&s->amg, &s->A);
  ichol_compute(&s->A, &s->L);
}

void h_free(void *h) {
  hsolver *s = (hsolver *)h;
  if (!s) return;
  ichol_free(&s->L);
  amg_free(&s->amg);
  h_csr_free(&s->A);
  free(s);
}

int h_num_levels(void *h) { return ((hsolver *)h)->amg.num_levels; }

/* dims[0..7] = n, nnz(A), n_coarse, nnz(P), nnz(S), nnz(L_coarsest or 0),
 *              nnz(fine L), has_split */
void h_level_dims(void *h, int lev, int *dims) {
  hsolver *s = (hsolver *)h;
  amg_level *L = &s->amg.levels[lev];
  int last = (lev == s->amg.num_levels - 1);
  dims[0] = L->n;
  dims[1] = L->A.nnz;
  dims[2] = last ? 0 : L->n_coarse;
  dims[3] = last ? 0 : L->P.nnz;
  dims[4] = last ? 0 : L->A_filtered.nnz;
  dims[5] = last ? L->L_ichol.nnz : 0;
  dims[6] = s->L.nnz;
  dims[7] = (!last && L->cf_order) ? 1 : 0;
}

static void copy_csr(const dcsr_mat *m, int *ia, int *ja, double *val) {
  if (ia) memcpy(ia, m->ia, ((size_t)m->row + 1) * sizeof(INT));
  if (ja) memcpy(ja, m->ja, (size_t)m->nnz * sizeof(INT));
  if (val) memcpy(val, m->val, (size_t)m->nnz * sizeof(REAL));
}

/* what: 0 = A, 1 = P, 2 = S */
void h_level_csr(void *h, int lev, int what, int *ia, int *ja, double *val) {
  amg_level *L = &((hsolver *)h)->amg.levels[lev];
  copy_csr(what == 0 ? &L->A : what == 1 ? &L->P : &L->A_filtered, ia, ja, val);
}

void h_level_split(void *h, int lev, int *mis, int *cf_order, double *l1) {
  amg_level *L = &((hsolver *)h)->amg.levels[lev];
  if (mis && L->mis) memcpy(mis, L->mis, (size_t)L->n * sizeof(INT));
  if (cf_order && L->cf_order) memcpy(cf_order, L->cf_order, (size_t)L->n * sizeof(INT));
  if (l1) memcpy(l1, L->l1_diag, (size_t)L->n * sizeof(REAL));
}

/* which: 0 = fine-level IC factor, 1 = coarsest-level IC factor */
void h_ichol(void *h, int which, int *ia, int *ja, double *val) {
  hsolver *s = (hsolver *)h;
  ichol_factor *F = which ? &s->amg.levels[s->amg.num_levels - 1].L_ichol : &s->L;
  if (ia) memcpy(ia, F->ia, ((size_t)F->n + 1) * sizeof(INT));
  if (ja) memcpy(ja, F->ja, (size_t)F->nnz * sizeof(INT));
  if (val) memcpy(val, F->val, (size_t)F->nnz * sizeof(REAL));
}

void h_set_smoother(void *h, int smoother, double omega) {
  hsolver *s = (hsolver *)h;
  s->amg.smoother = smoother;
  if (smoother == SMOOTHER_JACOBI) s->amg.omega_jacobi = omega;
}

/* PCG with the AMG+IC preconditioner (precond = 0) or the plain V-cycle
 * (precond = 1); x is the initial guess on entry.  seconds gets the wall
 * time of the amg_pcg call alone. */
int h_solve(void *h, const double *b, double *x, double tol, int maxiter,
            int print_level, int precond, double *seconds) {
  hsolver *s = (hsolver *)h;
  hctx c = {&s->amg, &s->A, &s->L};
  struct timespec t0, t1;
  clock_gettime(CLOCK_MONOTONIC, &t0);
  int it = precond == 1
               ? amg_pcg(&s->A, b, x, h_vcycle_cb, &s->amg, tol, maxiter, print_level)
               : amg_pcg(&s->A, b, x, h_precond_cb, &c, tol, maxiter, print_level);
  clock_gettime(CLOCK_MONOTONIC, &t1);
  if (seconds) *seconds = (double)(t1.tv_sec - t0.tv_sec) + 1e-9 * (double)(t1.tv_nsec - t0.tv_nsec);
  return it;
}

void h_precond(void *h, const double *g, double *z) {
  hsolver *s = (hsolver *)h;
  amg_ichol_precond(&s->amg, &s->A, &s->L, g, z);
}
void h_vcycle(void *h, const double *g, double *z) {
  amg_vcycle_precond(&((hsolver *)h)->amg, g, z);
}
void h_backslash(void *h, int lev, const double *b, double *x) {
  hsolver *s = (hsolver *)h;
  amg_backslash(&s->amg, lev, b, x, s->amg.nu);
}
void h_fwdslash(void *h, int lev, const double *b, double *x) {
  hsolver *s = (hsolver *)h;
  amg_fwdslash(&s->amg, lev, b, x, s->amg.nu);
}
void h_ichol_solve(void *h, int which, const double *b, double *x) {
  hsolver *s = (hsolver *)h;
  ichol_solve(which ? &s->amg.levels[s->amg.num_levels - 1].L_ichol : &s->L, b, x);
}
/* symmetrised fine matrix as used by the solve */
void h_fine_csr(void *h, int *ia, int *ja, double *val) {
  copy_csr(&((hsolver *)h)->A, ia, ja, val);
}
int h_fine_nnz(void *h) { return ((hsolver *)h)->A.nnz; }
