/*
 * block.c - block PCG: k right-hand sides solved together with one
 * hierarchy.  Every vector is a block of K interleaved columns (K = k
 * rounded up to a power of two, the extra columns zero), so each sparse
 * kernel streams a matrix once for all K columns (CSR SpMM).  The
 * V-cycle is the single-column one (vcycle.c) with block kernels; the
 * outer PCG keeps one scalar set per column and freezes a column once it
 * meets its stop test (alpha = beta = 0), so every column follows the
 * single-column iteration C_amg1/src/amg_pcg.c:12-101 until it stops.
 * One PCG step is captured as a hipGraph per (K, mode) and replayed.
 * The fine IC(0) is applied by Jacobi sweeps (flexible outer iteration);
 * the coarsest solve is the dense inverse.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "bn_solver.h"

void bn_set_error(const char *fmt, ...);
void bn_track_device(struct bn_solver *s, void *p);
int bn_ensure_device(struct bn_solver *s);

static double now_ms(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return 1e3 * (double)t.tv_sec + 1e-6 * (double)t.tv_nsec;
}

static inline const void *opv(const bn_dcsr *m, char vt) {
  return vt == 's' ? (const void *)m->vs : (const void *)m->vd;
}
static inline const void *vecv(const double *vd, const float *vs, char vt) {
  return vt == 's' ? (const void *)vs : (const void *)vd;
}

/* rows over this many entries go to the segmented long-row kernels */
#define BLOCK_LONG_LANES 4

static void apply_k(struct bn_solver *s, const bn_dcsr *M, char vt, int op, const void *X, void *Y, const void *B) {
  const int K = s->blk.K;
  const int nl = bn_nlong(M, BLOCK_LONG_LANES);
  const int nseg = nl > 0 ? M->seg_ptr_h[nl] : 0;
  const int nmulti = nl < M->nmulti_all ? nl : M->nmulti_all;
  bn_spmm(vt, K, op, M->rows, M->ia, M->ja, opv(M, vt), X, Y, B, bn_long_len(BLOCK_LONG_LANES), nl, M->long_rows,
          M->seg_ptr, M->seg_row, nseg, nmulti, s->blk.segpart, s->stream);
}

/* ---------------------------------------------------------------------
 * work space
 * ------------------------------------------------------------------- */
static void *bmalloc(struct bn_solver *s, size_t bytes) {
  void *p = NULL;
  if (hipMalloc(&p, bytes ? bytes : 8) != hipSuccess) return NULL;
  bn_track_device(s, p);
  return p;
}

static int block_alloc(struct bn_solver *s, int K) {
  bn_block *b = &s->blk;
  if (b->Kcap >= K) {
    b->K = K;
    return 0;
  }
  /* buffers of a smaller K stay allocated until the mirror is released */
  for (int k = 0; k < b->ngraph; k++) {
    if (b->gexec[k]) hipGraphExecDestroy(b->gexec[k]);
    if (b->graph[k]) hipGraphDestroy(b->graph[k]);
    b->gexec[k] = NULL;
    b->graph[k] = NULL;
  }
  b->ngraph = 0;
  const size_t n0 = (size_t)s->lev[0].n;
  int ok = 1;
#define BA(ptr, count) ok = ok && ((ptr) = bmalloc(s, (size_t)(count) * sizeof(double))) != NULL
  for (int l = 0; l < s->nlev; l++) {
    const size_t nl = (size_t)s->lev[l].n * K;
    BA(b->b[l], nl);
    BA(b->x[l], nl);
    BA(b->r[l], nl);
  }
  void **pv[] = {(void **)&b->x0, (void **)&b->r0, (void **)&b->z0, (void **)&b->p0, (void **)&b->q0,
                 (void **)&b->bb0, &b->g_in, &b->u1, &b->r1, &b->icy, &b->icx, &b->r2, &b->corr, &b->zout};
  for (size_t k = 0; k < sizeof(pv) / sizeof(pv[0]); k++) BA(*pv[k], n0 * K);
  const size_t nc = (size_t)s->lev[s->nlev - 1].n;
  BA(b->cin, nc * K);
  BA(b->cout, nc * K);
  BA(b->scal, (size_t)BN_S_COUNT * K);
  BA(b->part1, (size_t)BN_BLOCK_RED * K);
  BA(b->part2, (size_t)BN_BLOCK_RED * K);
  BA(b->part3, (size_t)BN_BLOCK_RED * K);
  /* segment partials: the largest multi-segment prefix of any operator */
  size_t segmax = 1;
  for (int l = 0; l < s->nlev; l++) {
    const bn_dcsr *ops[3] = {&s->lev[l].A, &s->lev[l].P, &s->lev[l].R};
    for (int q = 0; q < 3; q++)
      if (ops[q]->seg_ptr_h && (size_t)ops[q]->seg_ptr_h[ops[q]->nmulti_all] > segmax)
        segmax = (size_t)ops[q]->seg_ptr_h[ops[q]->nmulti_all];
  }
  if (s->Asolve.seg_ptr_h && (size_t)s->Asolve.seg_ptr_h[s->Asolve.nmulti_all] > segmax)
    segmax = (size_t)s->Asolve.seg_ptr_h[s->Asolve.nmulti_all];
  BA(b->segpart, segmax * K);
#undef BA
  if (!ok) {
    bn_set_error("device out of memory allocating the block solve work space (K = %d)", K);
    return BN_ERR_HIP;
  }
  if (b->host_scal) hipHostFree(b->host_scal);
  b->host_scal = NULL;
  if (hipHostMalloc((void **)&b->host_scal, (size_t)BN_S_COUNT * K * sizeof(double), 0) != hipSuccess) {
    bn_set_error("pinned host allocation failed (block scalars)");
    return BN_ERR_HIP;
  }
  b->Kcap = K;
  b->K = K;
  return 0;
}

/* ---------------------------------------------------------------------
 * block V-cycle (same structure as vcycle.c)
 * ------------------------------------------------------------------- */
static void smooth_k(struct bn_solver *s, int lev, const void *B, void *X, int backward, int x_zero,
                     const void *zbase, void *zsum) {
  bn_dlevel *d = &s->lev[lev];
  const char vt = s->vt;
  const int K = s->blk.K;
  const int nu = s->opt.nu > 0 ? s->opt.nu : 1;
  const int n = d->n;
  const long long ne = (long long)n * K;
  const int gs = d->use_gs && s->opt.smoother == SMOOTHER_CF_GS;
  if (gs) {
    const int fused_zero = x_zero && !(nu == 1 && zsum) && d->ncol > 0;
    if (x_zero && !fused_zero) bn_fill(vt, (int)ne, X, 0.0, s->stream);
    const void *dinvc = vecv(d->dinvc_d, d->dinvc_s, vt);
    for (int sw = 0; sw < nu; sw++) {
      const int last = (sw == nu - 1);
      for (int t = 0; t < d->ncol; t++) {
        const int c = backward ? d->ncol - 1 - t : t;
        const int lo = d->col_ptr[c], hi = d->col_ptr[c + 1];
        if (sw == 0 && t == 0 && fused_zero) {
          bn_gs_zero_map_k(vt, K, lo, hi, n, d->cmap_d, dinvc, B, X, s->stream);
          continue;
        }
        bn_gs_color_map_k(vt, K, lo, hi, d->cmap_d, d->Ac.ia, d->Ac.ja, opv(&d->Ac, vt), dinvc, B, X,
                          last ? zbase : NULL, last ? zsum : NULL, s->stream);
      }
    }
    return;
  }
  const int jac = (s->opt.smoother == SMOOTHER_JACOBI);
  const void *w = jac ? vecv(d->dinv_d, d->dinv_s, vt) : vecv(d->l1inv_d, d->l1inv_s, vt);
  const double om = jac ? s->opt.omega : 1.0;
  for (int sw = 0; sw < nu; sw++) {
    if (sw == 0 && x_zero) {
      bn_row_scale_k(vt, K, n, w, om, B, X, 0, s->stream);
      continue;
    }
    apply_k(s, &d->A, vt, BN_OP_RESID, X, s->blk.r[lev], B);
    bn_row_scale_k(vt, K, n, w, om, s->blk.r[lev], X, 1, s->stream);
  }
  if (zsum) {
    bn_copy(vt, (int)ne, zbase, zsum, s->stream);
    bn_axpy1(vt, (int)ne, X, zsum, s->stream);
  }
}

static void coarse_k(struct bn_solver *s, const void *B, void *X) {
  const int n = s->lev[s->nlev - 1].n, K = s->blk.K;
  if (s->vt == 'd') {
    bn_gemm_k(K, n, s->cinv_d, (const double *)B, (double *)X, s->stream);
    return;
  }
  bn_convert('s', 'd', n * K, B, s->blk.cin, s->stream);
  bn_gemm_k(K, n, s->cinv_d, s->blk.cin, s->blk.cout, s->stream);
  bn_convert('d', 's', n * K, s->blk.cout, X, s->stream);
}

static void backslash_k(struct bn_solver *s, int lev, const void *B, void *X, int x_zero) {
  if (lev == s->nlev - 1) {
    coarse_k(s, B, X);
    return;
  }
  bn_dlevel *d = &s->lev[lev];
  smooth_k(s, lev, B, X, 0, x_zero, NULL, NULL);
  apply_k(s, &d->A, s->vt, BN_OP_RESID, X, s->blk.r[lev], B);
  apply_k(s, &d->R, s->vt, BN_OP_SET, s->blk.r[lev], s->blk.b[lev + 1], NULL);
  backslash_k(s, lev + 1, s->blk.b[lev + 1], s->blk.x[lev + 1], 1);
  apply_k(s, &d->P, s->vt, BN_OP_ADD, s->blk.x[lev + 1], X, NULL);
}

/* fwdslash from x = 0 at every level (the only form the preconditioners use) */
static void fwdslash_k(struct bn_solver *s, int lev, const void *B, void *X, const void *zbase, void *zsum) {
  const int K = s->blk.K;
  if (lev == s->nlev - 1) {
    coarse_k(s, B, X);
    if (zsum) {
      const int ne = s->lev[lev].n * K;
      bn_copy(s->vt, ne, zbase, zsum, s->stream);
      bn_axpy1(s->vt, ne, X, zsum, s->stream);
    }
    return;
  }
  bn_dlevel *d = &s->lev[lev];
  apply_k(s, &d->R, s->vt, BN_OP_SET, B, s->blk.b[lev + 1], NULL);
  fwdslash_k(s, lev + 1, s->blk.b[lev + 1], s->blk.x[lev + 1], NULL, NULL);
  apply_k(s, &d->P, s->vt, BN_OP_SET, s->blk.x[lev + 1], X, NULL);
  smooth_k(s, lev, B, X, 1, 0, zbase, zsum);
}

/* ic_sweeps Jacobi sweeps per triangle of the fine IC(0); writes only
 * zsum = zbase + result */
static void ic_jacobi_k(struct bn_solver *s, const void *B, const void *zbase, void *zsum) {
  const char vt = s->vt;
  const int K = s->blk.K;
  bn_dtri *t = &s->fine_ic;
  const int n = t->n;
  const void *dinv = vt == 's' ? (const void *)t->dinv_s : (const void *)t->dinv_d;
  const int sweeps = s->opt.ic_sweeps > 0 ? s->opt.ic_sweeps : 1;
  void *bufs[2] = {s->blk.icy, s->blk.icx};
  int cur = 0;
  bn_row_scale_k(vt, K, n, dinv, 1.0, B, bufs[cur], 0, s->stream);
  for (int k = 1; k < sweeps; k++) {
    bn_jacobi_rows_k(vt, K, n, t->lo.ia, t->lo.ja, opv(&t->lo, vt), dinv, B, bufs[cur], bufs[1 - cur], NULL, NULL,
                     s->stream);
    cur = 1 - cur;
  }
  const void *yy = bufs[cur];
  void *xb[2] = {bufs[1 - cur], s->blk.r[0]};
  int cx = 0;
  if (sweeps == 1) {
    /* one sweep per triangle: x = D^{-1} y, then zsum = zbase + x */
    bn_row_scale_k(vt, K, n, dinv, 1.0, yy, xb[0], 0, s->stream);
    bn_copy(vt, n * K, zbase, zsum, s->stream);
    bn_axpy1(vt, n * K, xb[0], zsum, s->stream);
    return;
  }
  bn_row_scale_k(vt, K, n, dinv, 1.0, yy, xb[cx], 0, s->stream);
  for (int k = 1; k < sweeps; k++) {
    const int last = (k == sweeps - 1);
    bn_jacobi_rows_k(vt, K, n, t->up.ia, t->up.ja, opv(&t->up, vt), dinv, yy, xb[cx], xb[1 - cx],
                     last ? zbase : NULL, last ? zsum : NULL, s->stream);
    cx = 1 - cx;
  }
}

/* z = B g in the V-cycle's precision (FP32 blocks b->g_in / b->zout when
 * vt = 's') */
static void precond_core_k(struct bn_solver *s, const void *gv, void *zv) {
  const char vt = s->vt;
  bn_block *b = &s->blk;
  const bn_dcsr *A = &s->Asolve;
  backslash_k(s, 0, gv, b->u1, 1);
  apply_k(s, A, vt, BN_OP_RESID, b->u1, b->r1, gv);
  if (s->opt.fine_ic == BN_IC_NONE) {
    fwdslash_k(s, 0, b->r1, b->corr, b->u1, zv);
  } else {
    ic_jacobi_k(s, b->r1, b->u1, b->r2);
    apply_k(s, A, vt, BN_OP_RESID, b->r2, b->r1, gv);
    fwdslash_k(s, 0, b->r1, b->corr, b->r2, zv);
  }
}

/* z = B g for the K columns of g (FP64 block vectors) */
static void precond_k(struct bn_solver *s, const double *g, double *z) {
  const char vt = s->vt;
  const int n = s->lev[0].n, K = s->blk.K;
  bn_block *b = &s->blk;
  const void *gv = g;
  if (vt == 's') {
    bn_convert('d', 's', n * K, g, b->g_in, s->stream);
    gv = b->g_in;
  }
  precond_core_k(s, gv, (vt == 's') ? b->zout : (void *)z);
  if (vt == 's') bn_convert('s', 'd', n * K, b->zout, z, s->stream);
}

/* ---------------------------------------------------------------------
 * block PCG
 * ------------------------------------------------------------------- */
static void enqueue_step_k(struct bn_solver *s) {
  bn_block *b = &s->blk;
  const int n = s->lev[0].n, K = b->K;
  hipStream_t st = s->stream;
  /* FP32 V-cycle: r's FP32 copy comes out of the x/r update and z stays
   * FP32 (b->zout, widened where it is read); the first step's z is the
   * one the start-up preconditioning left there */
  const int f32 = (s->vt == 's');
  if (f32) bn_update_p_k_s(K, n, b->p0, (const float *)b->zout, b->scal, st);
  else bn_update_p_k(K, n, b->p0, b->z0, b->scal, st);
  apply_k(s, &s->Asolve, 'd', BN_OP_SET, b->p0, b->q0, NULL);
  bn_dot_k(K, n, b->p0, b->q0, NULL, NULL, b->part1, NULL, st);
  bn_finish_k(K, 1, n, b->part1, NULL, NULL, b->scal, st);
  if (f32) {
    bn_update_xr_k_s(K, n, b->x0, b->p0, b->r0, b->q0, b->scal, b->part2, (float *)b->g_in, st);
    precond_core_k(s, b->g_in, b->zout);
    bn_dot_rz_k_s(K, n, b->r0, (const float *)b->zout, b->q0, b->part1, b->part3, st);
  } else {
    bn_update_xr_k(K, n, b->x0, b->p0, b->r0, b->q0, b->scal, b->part2, st);
    precond_k(s, b->r0, b->z0);
    bn_dot_k(K, n, b->r0, b->z0, b->z0, b->q0, b->part1, b->part3, st);
  }
  bn_finish_k(K, 4, n, b->part1, b->part2, b->part3, b->scal, st);
}

static hipGraphExec_t step_graph_k(struct bn_solver *s) {
  bn_block *b = &s->blk;
  if (!s->opt.graph || !s->coarse_dense) return NULL;
  const char *env = getenv("BN_NO_GRAPH");
  if (env && env[0] && env[0] != '0') return NULL;
  const int key = (b->K << 8) | (s->vt == 's') | (s->opt.fine_ic << 1) | (s->opt.smoother << 3) |
                  (s->opt.ic_sweeps << 16) | (s->opt.nu << 24);
  for (int k = 0; k < b->ngraph; k++)
    if (b->graph_key[k] == key) return b->gexec[k];
  if (b->ngraph == (int)(sizeof(b->gexec) / sizeof(b->gexec[0]))) return NULL;
  hipGraph_t g = NULL;
  hipGraphExec_t e = NULL;
  if (hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return NULL;
  enqueue_step_k(s);
  if (hipStreamEndCapture(s->stream, &g) != hipSuccess || !g) return NULL;
  if (hipGraphInstantiate(&e, g, NULL, NULL, 0) != hipSuccess) {
    hipGraphDestroy(g);
    return NULL;
  }
  b->graph[b->ngraph] = g;
  b->gexec[b->ngraph] = e;
  b->graph_key[b->ngraph] = key;
  b->ngraph++;
  return e;
}

static int pow2_at_least(int k) {
  int K = 1;
  while (K < k) K *= 2;
  return K;
}

/* d_B, d_X: k right-hand sides / initial guesses stored one after another
 * (k x n, caller numbering = level-0 numbering), device pointers */
static int pcg_block(struct bn_solver *s, int k, const double *d_B, double *d_X, double tol, int maxiter, int *its,
                     bn_solve_info *info) {
  const double w0 = now_ms();
  const int K = pow2_at_least(k);
  int rc = block_alloc(s, K);
  if (rc) return rc;
  bn_block *b = &s->blk;
  const int n = s->lev[0].n;
  hipStream_t st = s->stream;
  hipGraphExec_t gexec = step_graph_k(s);
  /* stop parameters per column; the start kernels fill the rest */
  double *hs = b->host_scal;
  memset(hs, 0, (size_t)BN_S_COUNT * K * sizeof(double));
  for (int c = 0; c < K; c++) {
    hs[c * BN_S_COUNT + BN_S_TOL] = tol;
    hs[c * BN_S_COUNT + BN_S_STOPREL] = (s->opt.stop == BN_STOP_RELRES) ? 1.0 : 0.0;
  }
  hipEventRecord(s->ev0, st);
  hipMemcpyAsync(b->scal, hs, (size_t)BN_S_COUNT * K * sizeof(double), hipMemcpyHostToDevice, st);
  bn_interleave(n, k, K, d_B, b->bb0, 0, st);
  bn_interleave(n, k, K, d_X, b->x0, 0, st);
  apply_k(s, &s->Asolve, 'd', BN_OP_RESID, b->x0, b->r0, b->bb0);
  bn_dot_k(K, n, b->bb0, b->bb0, NULL, NULL, b->part1, NULL, st);
  bn_finish_k(K, 3, n, b->part1, NULL, NULL, b->scal, st);
  precond_k(s, b->r0, b->z0);
  bn_fill('d', n * K, b->p0, 0.0, st);
  bn_dot_k(K, n, b->r0, b->z0, b->r0, b->r0, b->part1, b->part2, st);
  bn_finish_k(K, 0, n, b->part1, b->part2, NULL, b->scal, st);
  int it = 0, all_done = 0;
  for (it = 1; it <= maxiter; it++) {
    hipMemcpyAsync(hs, b->scal, (size_t)BN_S_COUNT * K * sizeof(double), hipMemcpyDeviceToHost, st);
    if (hipStreamSynchronize(st) != hipSuccess) {
      bn_set_error("device failure in block PCG step %d", it);
      return BN_ERR_HIP;
    }
    all_done = 1;
    for (int c = 0; c < k; c++) all_done &= hs[c * BN_S_COUNT + BN_S_DONE] != 0.0;
    if (all_done) break;
    if (gexec) hipGraphLaunch(gexec, st);
    else enqueue_step_k(s);
  }
  if (!all_done) {
    hipMemcpyAsync(hs, b->scal, (size_t)BN_S_COUNT * K * sizeof(double), hipMemcpyDeviceToHost, st);
    if (hipStreamSynchronize(st) != hipSuccess) return BN_ERR_HIP;
  }
  int worst = 0, failed = 0;
  double rel_max = 0.0, relres_max = 0.0;
  for (int c = 0; c < k; c++) {
    const double *sc = hs + (size_t)c * BN_S_COUNT;
    const int done = sc[BN_S_DONE] != 0.0;
    const int ic = done ? (int)sc[BN_S_ITS] : -maxiter;
    if (its) its[c] = ic;
    if (!done) failed = 1;
    if ((ic < 0 ? -ic : ic) > worst) worst = ic < 0 ? -ic : ic;
    double bn = sqrt(sc[BN_S_BB]);
    if (bn == 0.0) bn = 1.0;
    if (sc[BN_S_REL] > rel_max) rel_max = sc[BN_S_REL];
    if (sqrt(sc[BN_S_RR]) / bn > relres_max) relres_max = sqrt(sc[BN_S_RR]) / bn;
  }
  if (failed) fprintf(stderr, "  block PCG: %d-column solve did not converge in %d iterations\n", k, maxiter);
  bn_interleave(n, k, K, b->x0, d_X, 1, st);
  hipEventRecord(s->ev1, st);
  if (hipStreamSynchronize(st) != hipSuccess) {
    bn_set_error("device failure finishing block PCG");
    return BN_ERR_HIP;
  }
  float ms = 0.0f;
  hipEventElapsedTime(&ms, s->ev0, s->ev1);
  s->last.iterations = failed ? -worst : worst;
  s->last.rel_bnorm = rel_max;
  s->last.relres = relres_max;
  s->last.solve_ms = ms;
  s->last.wall_ms = now_ms() - w0;
  s->last.num_levels = s->nlev;
  if (info) *info = s->last;
  return failed ? -worst : worst;
}

/* ---------------------------------------------------------------------
 * entry points
 * ------------------------------------------------------------------- */
static int block_check(struct bn_solver *s, int k) {
  if (!s || k < 1 || k > BN_BLOCK_KMAX) {
    bn_set_error("block solve: k must be in 1..%d", BN_BLOCK_KMAX);
    return BN_ERR_ARG;
  }
  int rc = bn_ensure_device(s);
  if (rc) return rc;
  if (!s->coarse_dense) {
    bn_set_error("block solve: needs the dense coarsest solve (BN_COARSE_DENSE / AUTO with n_c <= %d)",
                 BN_DENSE_MAX);
    return BN_ERR_ARG;
  }
  for (int l = 0; l < s->nlev - 1; l++)
    if (s->lev[l].block) {
      bn_set_error("block solve: the block Gauss-Seidel smoother option is not supported here");
      return BN_ERR_ARG;
    }
  if (s->opt.fine_ic == BN_IC_EXACT) {
    bn_set_error("block solve: the fine IC(0) must be BN_IC_JACOBI or BN_IC_NONE");
    return BN_ERR_ARG;
  }
  if (bn_vc_exact_kind(s) >= 0) {
    bn_set_error("block solve: exact-order Gauss-Seidel sweeps (SMOOTHER_GS, gs_exact) are not supported; "
                 "use the multicolour SMOOTHER_CF_GS, SMOOTHER_L1 or SMOOTHER_JACOBI");
    return BN_ERR_ARG;
  }
  return 0;
}

int bn_solver_solve_block_device(void *h, int k, const double *d_b, double *d_x, double tol, int maxiter, int *its,
                                 bn_solve_info *info) {
  struct bn_solver *s = (struct bn_solver *)h;
  int rc = block_check(s, k);
  if (rc) return rc;
  if (!d_b || !d_x) return BN_ERR_ARG;
  return pcg_block(s, k, d_b, d_x, tol, maxiter, its, info);
}

int bn_solver_solve_block(void *h, int k, const double *b, double *x, double tol, int maxiter, int *its,
                          bn_solve_info *info) {
  struct bn_solver *s = (struct bn_solver *)h;
  int rc = block_check(s, k);
  if (rc) return rc;
  if (!b || !x) return BN_ERR_ARG;
  const size_t n = (size_t)s->lev[0].n, bytes = n * (size_t)k * sizeof(double);
  double *db = NULL, *dx = NULL;
  if (hipMalloc((void **)&db, bytes) != hipSuccess || hipMalloc((void **)&dx, bytes) != hipSuccess) {
    hipFree(db);
    bn_set_error("device out of memory (block right-hand sides)");
    return BN_ERR_HIP;
  }
  hipMemcpy(db, b, bytes, hipMemcpyHostToDevice);
  hipMemcpy(dx, x, bytes, hipMemcpyHostToDevice);
  rc = pcg_block(s, k, db, dx, tol, maxiter, its, info);
  if (rc != BN_ERR_HIP) hipMemcpy(x, dx, bytes, hipMemcpyDeviceToHost);
  hipFree(db);
  hipFree(dx);
  return rc;
}
