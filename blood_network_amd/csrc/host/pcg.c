/*
 * pcg.c - outer preconditioned conjugate gradients on the device.
 *
 * Follows C_amg1/src/amg_pcg.c:12-101: r = b - A x, z = B r, p = z;
 * each step q = A p, alpha = r'z / p'q, x += alpha p, r -= alpha q,
 * z = B r, stop when sqrt(r'z / r0'z0) < tol, beta = r'z_new / r'z_old,
 * p = z + beta p; returns the step count or -maxiter.
 * All vectors and scalars live on the device.  A step is enqueued as one
 * unit that starts with the p update (beta = 0 and p = 0 before the first
 * step, so p = z then), which makes every step the same kernel sequence:
 * it is captured once as a hipGraph and replayed.  The only device-to-host
 * traffic per step is the scalar block that decides whether to stop.
 * Options: the ||r||/||b|| stop test and the flexible (Polak-Ribiere)
 * beta = z_new'(r_new - r_old) / r_old'z_old, needed when the
 * preconditioner is not a fixed symmetric operator (FP32 V-cycle,
 * Jacobi-approximated IC).
 */
#include <math.h>
#include <stdio.h>
#include <string.h>
#include <stdlib.h>
#include <time.h>

#include "bn_solver.h"

void bn_set_error(const char *fmt, ...);

static double now_ms(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return 1e3 * (double)t.tv_sec + 1e-6 * (double)t.tv_nsec;
}

/* one PCG step, enqueued on the solver stream */
static void enqueue_step(struct bn_solver *s) {
  hipStream_t st = s->stream;
  const int n = s->lev[0].n;
  const bn_dcsr *A = &s->Asolve;
  double *sc = s->scal;
  const int flex = s->opt.flexible;
  /* FP32 V-cycle: r's FP32 copy comes out of the x/r update and z stays
   * FP32 (s->zout, widened where it is read), so no conversion pass runs;
   * the first step's z is the one the start-up preconditioning left there */
  const int f32 = (s->vt == 's');
  if (f32) bn_pcg_update_p_s(n, s->p, (const float *)s->zout, sc, st);
  else bn_pcg_update_p(n, s->p, s->z, sc, st);
  /* q = A p, p'q -> alpha */
  const int nslot = bn_launch_spmv(s, A, BN_OP_SET, s->p, s->q, NULL, s->p, s->part1);
  if (nslot > 0) bn_scalar_finish_pap(A->bpart, nslot, sc, st);
  else bn_scalar_finish(1, s->part1, NULL, NULL, sc, 0, st);
  if (f32) {
    const bn_dlevel *d0 = &s->lev[0];
    if (bn_vc_first_colour_fused(s)) {
      /* the preconditioner's first colour rides on the update */
      bn_pcg_update_xr_s0(n, s->x, s->p, s->r, s->q, sc, s->part2, (float *)s->g_in, d0->colour_d, 0, d0->dinv_s,
                          (float *)s->u1, st);
      s->u1_first = 1;
    } else {
      bn_pcg_update_xr_s(n, s->x, s->p, s->r, s->q, sc, s->part2, (float *)s->g_in, st);
    }
    bn_vc_precond_s(s);
    bn_dot_rz_s(n, s->r, (const float *)s->zout, flex ? s->q : NULL, s->part1, s->part3, st);
  } else {
    bn_pcg_update_xr(n, s->x, s->p, s->r, s->q, sc, s->part2, st);
    bn_vc_precond(s, s->r, s->z);
    bn_dot_partial(n, s->r, s->z, flex ? s->z : NULL, flex ? s->q : NULL, s->part1, s->part3, st);
  }
  /* the scalars also go straight to the host (no copy after the step) */
  if (s->host_step)
    bn_scalar_finish_host(flex ? 4 : 2, s->part1, s->part2, flex ? s->part3 : NULL, sc, s->host_step, st);
  else
    bn_scalar_finish(flex ? 4 : 2, s->part1, s->part2, flex ? s->part3 : NULL, sc, 0, st);
}

static int mode_key(const struct bn_solver *s) {
  return (s->vt == 's') | (s->opt.fine_ic << 1) | (s->opt.flexible << 3) | (s->opt.smoother << 4) |
         (s->opt.nu << 8) | (s->opt.ic_sweeps << 16) | ((s->opt.gs_exact != 0) << 28) |
         (s->walk_on << 29);
}

/* The captured step for the current mode, capturing it on first use.
 * Graphs are kept per mode until the device state is released.  Returns
 * NULL when graphs are off (option, BN_NO_GRAPH in the environment, or a
 * host-driven coarsest solve). */
static hipGraphExec_t step_graph(struct bn_solver *s) {
  if (!s->opt.graph || !s->coarse_dense) return NULL;
  const char *env = getenv("BN_NO_GRAPH");
  if (env && env[0] && env[0] != '0') return NULL;
  const int key = mode_key(s);
  for (int k = 0; k < s->ngraph; k++)
    if (s->graph_key[k] == key) return s->gexec[k];
  if (s->noalloc) return NULL; /* capturing allocates: run the step eagerly */
  if (s->ngraph == (int)(sizeof(s->gexec) / sizeof(s->gexec[0]))) return NULL;
  hipGraph_t g = NULL;
  hipGraphExec_t e = NULL;
  if (hipStreamBeginCapture(s->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return NULL;
  enqueue_step(s);
  if (hipStreamEndCapture(s->stream, &g) != hipSuccess || !g) return NULL;
  if (hipGraphInstantiate(&e, g, NULL, NULL, 0) != hipSuccess) {
    hipGraphDestroy(g);
    return NULL;
  }
  s->graph[s->ngraph] = g;
  s->gexec[s->ngraph] = e;
  s->graph_key[s->ngraph] = key;
  s->ngraph++;
  return e;
}

/* A one-launch triangular walk gave up waiting (its grid was not resident
 * at once): nothing of the caller's has been written yet (d_x is only
 * written at the end), so the solve starts again with a launch per level,
 * which the solver keeps from now on. */
static int walk_gave_up(struct bn_solver *s, const double *d_b, double *d_x, double tol, int maxiter,
                        int print_level, bn_solve_info *info) {
  fprintf(stderr, "  PCG: a one-launch triangular solve timed out; solving again with a launch per level\n");
  s->walk_on = 0;
  s->walk_ok = 0;
  bn_drop_graphs(s);
  return bn_pcg_run(s, d_b, d_x, tol, maxiter, print_level, info);
}

static int pcg_body(struct bn_solver *s, const double *d_b, double *d_x, double tol, int maxiter,
                    int print_level, bn_solve_info *info);

/* d_b, d_x in caller (original) numbering; x is the initial guess.  Exact
 * triangular solves walk their levels in one launch per run of ELL levels
 * inside this solve only; the launches flag a wait that timed out in the
 * scalar block, which is read back with every step. */
int bn_pcg_run(struct bn_solver *s, const double *d_b, double *d_x, double tol, int maxiter,
               int print_level, bn_solve_info *info) {
  s->walk_on = s->walk_ok;
  const int it = pcg_body(s, d_b, d_x, tol, maxiter, print_level, info);
  s->walk_on = 0;
  return it;
}

static int pcg_body(struct bn_solver *s, const double *d_b, double *d_x, double tol, int maxiter,
                    int print_level, bn_solve_info *info) {
  const double w0 = now_ms();
  hipStream_t st = s->stream;
  const int n = s->lev[0].n;
  const bn_dcsr *A = &s->Asolve;
  double *sc = s->scal;
  double *hs = s->host_scal;
  const int relres_stop = (s->opt.stop == BN_STOP_RELRES);
  /* schedules the smoother needs are built before anything is captured */
  const int prc = bn_vc_prepare(s);
  if (prc) return prc;
  hipGraphExec_t gexec = step_graph(s);

  hipEventRecord(s->ev0, st);
  bn_fill('d', 1, sc + BN_S_WALKERR, 0.0, st);
  bn_permute(n, s->lev[0].perm_d, d_b, s->bb, 0, st);
  bn_permute(n, s->lev[0].perm_d, d_x, s->x, 0, st);
  /* r = b - A x ; ||b||^2 */
  bn_launch_spmv(s, A, BN_OP_RESID, s->x, s->r, s->bb, NULL, NULL);
  bn_dot_partial(n, s->bb, s->bb, NULL, NULL, s->part1, NULL, st);
  bn_scalar_finish(3, s->part1, NULL, NULL, sc, 0, st);
  bn_vc_precond(s, s->r, s->z);
  bn_fill('d', n, s->p, 0.0, st);
  bn_dot_partial(n, s->r, s->z, s->r, s->r, s->part1, s->part2, st);
  bn_scalar_finish(0, s->part1, s->part2, NULL, sc, 0, st);
  hipMemcpyAsync(hs, sc, BN_S_COUNT * sizeof(double), hipMemcpyDeviceToHost, st);
  if (hipStreamSynchronize(st) != hipSuccess) {
    bn_set_error("device failure during PCG set-up");
    return BN_ERR_HIP;
  }
  if (hs[BN_S_WALKERR] != 0.0) return walk_gave_up(s, d_b, d_x, tol, maxiter, print_level, info);
  double bnorm = sqrt(hs[BN_S_BB]);
  if (bnorm == 0.0) bnorm = 1.0; /* amg_pcg.c:29 */
  double rnorm = sqrt(hs[BN_S_RR]);
  double rel = 1.0;
  if (print_level >= 1) {
    fprintf(stderr, "  %4s  %8s  %8s  %6s\n", "iter", "||r||_B", "rel", "rho");
    fprintf(stderr, "  %4d  %8.2e  %8.2e\n", 0, sqrt(hs[BN_S_RZ0]), 1.0);
  }
  int it = 0;
  if (hs[BN_S_RZ0] == 0.0) {
    /* r0 = 0 (e.g. b = 0 with x0 = 0): x0 already solves the system; the
     * reference would divide 0/0 here (amg_pcg.c:59) */
    rel = 0.0;
  } else {
    for (it = 1; it <= maxiter; it++) {
      if (gexec) hipGraphLaunch(gexec, st);
      else enqueue_step(s);
      if (!s->host_step) hipMemcpyAsync(hs, sc, BN_S_COUNT * sizeof(double), hipMemcpyDeviceToHost, st);
      if (hipStreamSynchronize(st) != hipSuccess) {
        bn_set_error("device failure in PCG step %d", it);
        return BN_ERR_HIP;
      }
      if (s->host_step) memcpy(hs, s->host_step, BN_S_COUNT * sizeof(double));
      if (hs[BN_S_WALKERR] != 0.0) return walk_gave_up(s, d_b, d_x, tol, maxiter, print_level, info);
      rel = hs[BN_S_REL];
      rnorm = sqrt(hs[BN_S_RR]);
      if (print_level >= 1)
        fprintf(stderr, "  %4d  %8.2e  %8.2e  %6.4f\n", it, sqrt(hs[BN_S_RZNEW]), rel,
                hs[BN_S_RZNEW] / hs[BN_S_RZOLD]);
      const double test = relres_stop ? rnorm / bnorm : rel;
      if (test < tol) break;
    }
  }
  if (it > maxiter) {
    fprintf(stderr, "  PCG: no convergence after %d iterations, relres = %.2e\n", maxiter, rnorm / bnorm);
    it = -maxiter;
  }
  bn_permute(n, s->lev[0].perm_d, s->x, d_x, 1, st);
  hipEventRecord(s->ev1, st);
  if (hipStreamSynchronize(st) != hipSuccess) {
    bn_set_error("device failure finishing PCG");
    return BN_ERR_HIP;
  }
  float ms = 0.0f;
  hipEventElapsedTime(&ms, s->ev0, s->ev1);
  s->last.iterations = it;
  s->last.rel_bnorm = rel;
  s->last.relres = rnorm / bnorm;
  s->last.solve_ms = ms;
  s->last.wall_ms = now_ms() - w0;
  s->last.num_levels = s->nlev;
  if (info) *info = s->last;
  return it;
}
