/*
 * segments.c - many independent systems in one solve.  The caller stacks
 * nsys systems into one block-diagonal matrix (rows of system g are
 * seg_ptr[g] .. seg_ptr[g+1]-1, no coupling between systems) and builds
 * one solver on it: the AMG set-up never couples the blocks (strength,
 * C/F splits, interpolation and Galerkin products stay block-diagonal),
 * so one V-cycle preconditions every system at once.  The outer PCG keeps
 * one scalar set per system through segmented reductions and freezes a
 * system when it meets its stop test, so each system follows its own PCG
 * iteration (C_amg1/src/amg_pcg.c:12-101) while every kernel launch
 * serves all of them.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "bn_solver.h"
#include "csr.h"

void bn_set_error(const char *fmt, ...);
void bn_track_device(struct bn_solver *s, void *p);
int bn_ensure_device(struct bn_solver *s);

#define ITEM_ROWS 4096

static double now_ms(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return 1e3 * (double)t.tv_sec + 1e-6 * (double)t.tv_nsec;
}

/* device state for one segmentation (cached per solver) */
typedef struct bn_segs {
  int nsys, nitem;
  int *seg_ptr_h;        /* host copy of the segmentation */
  int *item_lo, *item_seg, *item_ptr;
  double *scal, *part1, *part2, *part3;
  double *host_scal;
  hipGraph_t graph;
  hipGraphExec_t gexec;
  int graph_key;
  struct bn_segs *next;
} bn_segs;

static void *smalloc(struct bn_solver *s, size_t bytes) {
  void *p = NULL;
  if (hipMalloc(&p, bytes ? bytes : 8) != hipSuccess) return NULL;
  bn_track_device(s, p);
  return p;
}

/* drop the cached segmentations of a solver whose device state goes */
void bn_segments_release(struct bn_solver *s) {
  bn_segs *g = (bn_segs *)s->segs;
  while (g) {
    bn_segs *nx = g->next;
    if (g->gexec) hipGraphExecDestroy(g->gexec);
    if (g->graph) hipGraphDestroy(g->graph);
    if (g->host_scal) hipHostFree(g->host_scal);
    free(g->seg_ptr_h);
    free(g); /* device arrays are freed with the solver's allocations */
    g = nx;
  }
  s->segs = NULL;
}

static bn_segs *segs_for(struct bn_solver *s, int nsys, const int *seg_ptr) {
  bn_segs **head = (bn_segs **)&s->segs;
  for (bn_segs *g = *head; g; g = g->next)
    if (g->nsys == nsys && !memcmp(g->seg_ptr_h, seg_ptr, ((size_t)nsys + 1) * sizeof(int))) return g;
  bn_segs *g = (bn_segs *)calloc(1, sizeof(bn_segs));
  g->nsys = nsys;
  g->seg_ptr_h = (int *)bn_xmalloc(((size_t)nsys + 1) * sizeof(int));
  memcpy(g->seg_ptr_h, seg_ptr, ((size_t)nsys + 1) * sizeof(int));
  /* items: <= ITEM_ROWS rows, never crossing a system */
  int nitem = 0;
  for (int q = 0; q < nsys; q++) {
    const int len = seg_ptr[q + 1] - seg_ptr[q];
    nitem += len > 0 ? (len + ITEM_ROWS - 1) / ITEM_ROWS : 0;
  }
  int *lo = (int *)bn_xmalloc(((size_t)nitem + 1) * sizeof(int));
  int *sg = (int *)bn_xmalloc((size_t)(nitem > 0 ? nitem : 1) * sizeof(int));
  int *ip = (int *)bn_xmalloc(((size_t)nsys + 1) * sizeof(int));
  int w = 0;
  for (int q = 0; q < nsys; q++) {
    ip[q] = w;
    for (int r = seg_ptr[q]; r < seg_ptr[q + 1]; r += ITEM_ROWS) {
      lo[w] = r;
      sg[w] = q;
      w++;
    }
  }
  ip[nsys] = w;
  lo[w] = seg_ptr[nsys];
  g->nitem = nitem;
  g->item_lo = (int *)smalloc(s, ((size_t)nitem + 1) * sizeof(int));
  g->item_seg = (int *)smalloc(s, (size_t)(nitem > 0 ? nitem : 1) * sizeof(int));
  g->item_ptr = (int *)smalloc(s, ((size_t)nsys + 1) * sizeof(int));
  g->scal = (double *)smalloc(s, (size_t)BN_S_COUNT * nsys * sizeof(double));
  g->part1 = (double *)smalloc(s, (size_t)(nitem > 0 ? nitem : 1) * sizeof(double));
  g->part2 = (double *)smalloc(s, (size_t)(nitem > 0 ? nitem : 1) * sizeof(double));
  g->part3 = (double *)smalloc(s, (size_t)(nitem > 0 ? nitem : 1) * sizeof(double));
  int ok = g->item_lo && g->item_seg && g->item_ptr && g->scal && g->part1 && g->part2 && g->part3 &&
           hipHostMalloc((void **)&g->host_scal, (size_t)BN_S_COUNT * nsys * sizeof(double), 0) == hipSuccess;
  if (ok) {
    ok = hipMemcpy(g->item_lo, lo, ((size_t)nitem + 1) * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(g->item_seg, sg, (size_t)nitem * sizeof(int), hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(g->item_ptr, ip, ((size_t)nsys + 1) * sizeof(int), hipMemcpyHostToDevice) == hipSuccess;
  }
  free(lo);
  free(sg);
  free(ip);
  if (!ok) {
    bn_set_error("device allocation failed (segmented solve state)");
    if (g->host_scal) hipHostFree(g->host_scal);
    free(g->seg_ptr_h);
    free(g);
    return NULL;
  }
  g->next = *head;
  *head = g;
  return g;
}

static void enqueue_step_seg(struct bn_solver *s, bn_segs *g) {
  hipStream_t st = s->stream;
  const bn_dcsr *A = &s->Asolve;
  bn_update_p_seg(g->nitem, g->item_lo, g->item_seg, s->p, s->z, g->scal, st);
  bn_dcsr_apply_dot(A, 'd', BN_OP_SET, s->p, s->q, NULL, NULL, NULL, st);
  bn_dot_seg(g->nitem, g->item_lo, s->p, s->q, NULL, NULL, g->part1, NULL, st);
  bn_finish_seg(g->nsys, 1, g->item_ptr, g->part1, NULL, NULL, g->scal, st);
  bn_update_xr_seg(g->nitem, g->item_lo, g->item_seg, s->x, s->p, s->r, s->q, g->scal, g->part2, st);
  bn_vc_precond(s, s->r, s->z);
  bn_dot_seg(g->nitem, g->item_lo, s->r, s->z, s->z, s->q, g->part1, g->part3, st);
  bn_finish_seg(g->nsys, 4, g->item_ptr, g->part1, g->part2, g->part3, g->scal, st);
}

static hipGraphExec_t step_graph_seg(struct bn_solver *s, bn_segs *g) {
  if (!s->opt.graph || !s->coarse_dense) return NULL;
  const char *env = getenv("BN_NO_GRAPH");
  if (env && env[0] && env[0] != '0') return NULL;
  const int key = (s->vt == 's') | (s->opt.fine_ic << 1) | (s->opt.smoother << 3) | (s->opt.nu << 8) |
                  (s->opt.ic_sweeps << 16) | ((s->opt.gs_exact != 0) << 28);
  if (g->gexec && g->graph_key == key) return g->gexec;
  if (g->gexec) return NULL; /* one captured mode per segmentation; other modes run eagerly */
  hipGraph_t gr = NULL;
  hipGraphExec_t e = NULL;
  if (hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return NULL;
  enqueue_step_seg(s, g);
  if (hipStreamEndCapture(s->stream, &gr) != hipSuccess || !gr) return NULL;
  if (hipGraphInstantiate(&e, gr, NULL, NULL, 0) != hipSuccess) {
    hipGraphDestroy(gr);
    return NULL;
  }
  g->graph = gr;
  g->gexec = e;
  g->graph_key = key;
  return e;
}

static int seg_check(struct bn_solver *s, int nsys, const int *seg_ptr) {
  if (!s || nsys < 1 || !seg_ptr) {
    bn_set_error("segmented solve: need nsys >= 1 and seg_ptr");
    return BN_ERR_ARG;
  }
  const int n = s->A->row;
  if (seg_ptr[0] != 0 || seg_ptr[nsys] != n) {
    bn_set_error("segmented solve: seg_ptr must run from 0 to n = %d", n);
    return BN_ERR_ARG;
  }
  for (int q = 0; q < nsys; q++)
    if (seg_ptr[q + 1] < seg_ptr[q]) {
      bn_set_error("segmented solve: seg_ptr must be non-decreasing");
      return BN_ERR_ARG;
    }
  int rc = bn_ensure_device(s);
  if (rc) return rc;
  for (int l = 0; l < s->nlev - 1; l++)
    if (s->lev[l].block) {
      bn_set_error("segmented solve: the block Gauss-Seidel smoother option is not supported here");
      return BN_ERR_ARG;
    }
  return bn_vc_prepare(s);
}

/* d_b, d_x: n-vectors in caller numbering (= level-0 numbering) */
static int pcg_segments(struct bn_solver *s, int nsys, const int *seg_ptr, const double *d_b, double *d_x,
                        double tol, int maxiter, int *its, bn_solve_info *info) {
  const double w0 = now_ms();
  bn_segs *g = segs_for(s, nsys, seg_ptr);
  if (!g) return BN_ERR_HIP;
  hipStream_t st = s->stream;
  const int n = s->lev[0].n;
  hipGraphExec_t gexec = step_graph_seg(s, g);
  double *hs = g->host_scal;
  memset(hs, 0, (size_t)BN_S_COUNT * nsys * sizeof(double));
  for (int q = 0; q < nsys; q++) {
    hs[q * BN_S_COUNT + BN_S_TOL] = tol;
    hs[q * BN_S_COUNT + BN_S_STOPREL] = (s->opt.stop == BN_STOP_RELRES) ? 1.0 : 0.0;
  }
  const bn_dcsr *A = &s->Asolve;
  hipEventRecord(s->ev0, st);
  hipMemcpyAsync(g->scal, hs, (size_t)BN_S_COUNT * nsys * sizeof(double), hipMemcpyHostToDevice, st);
  bn_copy('d', n, d_b, s->bb, st);
  bn_copy('d', n, d_x, s->x, st);
  bn_dcsr_apply_dot(A, 'd', BN_OP_RESID, s->x, s->r, s->bb, NULL, NULL, st);
  bn_dot_seg(g->nitem, g->item_lo, s->bb, s->bb, NULL, NULL, g->part1, NULL, st);
  bn_finish_seg(nsys, 3, g->item_ptr, g->part1, NULL, NULL, g->scal, st);
  bn_vc_precond(s, s->r, s->z);
  bn_fill('d', n, s->p, 0.0, st);
  bn_dot_seg(g->nitem, g->item_lo, s->r, s->z, s->r, s->r, g->part1, g->part2, st);
  bn_finish_seg(nsys, 0, g->item_ptr, g->part1, g->part2, NULL, g->scal, st);
  int all_done = 0;
  for (int it = 1; it <= maxiter + 1; it++) {
    hipMemcpyAsync(hs, g->scal, (size_t)BN_S_COUNT * nsys * sizeof(double), hipMemcpyDeviceToHost, st);
    if (hipStreamSynchronize(st) != hipSuccess) {
      bn_set_error("device failure in segmented PCG step %d", it);
      return BN_ERR_HIP;
    }
    all_done = 1;
    for (int q = 0; q < nsys; q++) all_done &= hs[q * BN_S_COUNT + BN_S_DONE] != 0.0;
    if (all_done || it > maxiter) break;
    if (gexec) hipGraphLaunch(gexec, st);
    else enqueue_step_seg(s, g);
  }
  int worst = 0, failed = 0;
  double rel_max = 0.0, relres_max = 0.0;
  for (int q = 0; q < nsys; q++) {
    const double *sc = hs + (size_t)q * BN_S_COUNT;
    const int done = sc[BN_S_DONE] != 0.0;
    const int ic = done ? (int)sc[BN_S_ITS] : -maxiter;
    if (its) its[q] = ic;
    if (!done) failed = 1;
    if ((ic < 0 ? -ic : ic) > worst) worst = ic < 0 ? -ic : ic;
    double bn = sqrt(sc[BN_S_BB]);
    if (bn == 0.0) bn = 1.0;
    if (sc[BN_S_REL] > rel_max) rel_max = sc[BN_S_REL];
    if (sqrt(sc[BN_S_RR]) / bn > relres_max) relres_max = sqrt(sc[BN_S_RR]) / bn;
  }
  if (failed) fprintf(stderr, "  segmented PCG: not every system converged in %d iterations\n", maxiter);
  bn_copy('d', n, s->x, d_x, st);
  hipEventRecord(s->ev1, st);
  if (hipStreamSynchronize(st) != hipSuccess) {
    bn_set_error("device failure finishing segmented PCG");
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

int bn_solver_solve_segments_device(void *h, int nsys, const int *seg_ptr, const double *d_b, double *d_x,
                                    double tol, int maxiter, int *its, bn_solve_info *info) {
  struct bn_solver *s = (struct bn_solver *)h;
  int rc = seg_check(s, nsys, seg_ptr);
  if (rc) return rc;
  if (!d_b || !d_x) return BN_ERR_ARG;
  return pcg_segments(s, nsys, seg_ptr, d_b, d_x, tol, maxiter, its, info);
}

int bn_solver_solve_segments(void *h, int nsys, const int *seg_ptr, const double *b, double *x, double tol,
                             int maxiter, int *its, bn_solve_info *info) {
  struct bn_solver *s = (struct bn_solver *)h;
  int rc = seg_check(s, nsys, seg_ptr);
  if (rc) return rc;
  if (!b || !x) return BN_ERR_ARG;
  const size_t bytes = (size_t)s->lev[0].n * sizeof(double);
  double *db = NULL, *dx = NULL;
  if (hipMalloc((void **)&db, bytes) != hipSuccess || hipMalloc((void **)&dx, bytes) != hipSuccess) {
    hipFree(db);
    bn_set_error("device out of memory (segmented right-hand side)");
    return BN_ERR_HIP;
  }
  hipMemcpy(db, b, bytes, hipMemcpyHostToDevice);
  hipMemcpy(dx, x, bytes, hipMemcpyHostToDevice);
  rc = pcg_segments(s, nsys, seg_ptr, db, dx, tol, maxiter, its, info);
  if (rc != BN_ERR_HIP) hipMemcpy(x, dx, bytes, hipMemcpyDeviceToHost);
  hipFree(db);
  hipFree(dx);
  return rc;
}
