/*
 * pcg.c - outer preconditioned conjugate gradients on the device.
 *
 * Follows C_amg1/src/amg_pcg.c:12-101: r = b - A x, z = B r, p = z;
 * each step q = A p, alpha = r'z / p'q, x += alpha p, r -= alpha q,
 * z = B r, stop when sqrt(r'z / r0'z0) < tol, beta = r'z_new / r'z_old,
 * p = z + beta p; returns the step count or -maxiter.
 * All vectors and scalars live on the device; the only device-to-host
 * traffic per step is the scalar block (to decide whether to stop).
 * Options: the ||r||/||b|| stop test and the flexible (Polak-Ribiere)
 * beta = z_new'(r_new - r_old) / r_old'z_old, needed when the
 * preconditioner is not a fixed symmetric operator (FP32 V-cycle,
 * Jacobi-approximated IC).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include "bn_solver.h"

void bn_set_error(const char *fmt, ...);

static double now_ms(void) {
  struct timespec t;
  clock_gettime(CLOCK_MONOTONIC, &t);
  return 1e3 * (double)t.tv_sec + 1e-6 * (double)t.tv_nsec;
}

static void precond(struct bn_solver *s, const double *g, double *z, int which) {
  if (which == 1) bn_vc_vcycle_precond(s, g, z);
  else bn_vc_ichol_precond(s, g, z);
}

/* d_b, d_x in caller (original) numbering; x is the initial guess. */
int bn_pcg_run(struct bn_solver *s, const double *d_b, double *d_x, double tol, int maxiter,
               int print_level, bn_solve_info *info) {
  const double w0 = now_ms();
  hipStream_t st = s->stream;
  const int n = s->lev[0].n;
  const bn_dcsr *A = &s->Asolve;
  const int which = 0;
  double *sc = s->scal;
  double *hs = s->host_scal;
  const int flex = s->opt.flexible;
  const int relres_stop = (s->opt.stop == BN_STOP_RELRES);

  hipEventRecord(s->ev0, st);
  bn_permute(n, s->lev[0].perm_d, d_b, s->bb, 0, st);
  bn_permute(n, s->lev[0].perm_d, d_x, s->x, 0, st);
  /* r = b - A x ; ||b||^2 */
  bn_csr_apply('d', BN_OP_RESID, n, A->ia, A->ja, A->vd, s->x, s->r, s->bb, A->lanes, NULL, NULL, st);
  bn_dot_partial(n, s->bb, s->bb, NULL, NULL, s->part1, NULL, st);
  bn_scalar_finish(3, s->part1, NULL, NULL, sc, 0, st);
  precond(s, s->r, s->z, which);
  bn_copy('d', n, s->z, s->p, st);
  bn_dot_partial(n, s->r, s->z, s->r, s->r, s->part1, s->part2, st);
  bn_scalar_finish(0, s->part1, s->part2, NULL, sc, 0, st);
  hipMemcpyAsync(hs, sc, BN_S_COUNT * sizeof(double), hipMemcpyDeviceToHost, st);
  if (hipStreamSynchronize(st) != hipSuccess) {
    bn_set_error("device failure during PCG set-up");
    return BN_ERR_HIP;
  }
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
      /* q = A p, p'q */
      bn_csr_apply('d', BN_OP_SET, n, A->ia, A->ja, A->vd, s->p, s->q, NULL, A->lanes, s->p, s->part1, st);
      bn_scalar_finish(1, s->part1, NULL, NULL, sc, 0, st);
      bn_pcg_update_xr(n, s->x, s->p, s->r, s->q, sc, s->part2, st);
      precond(s, s->r, s->z, which);
      bn_dot_partial(n, s->r, s->z, flex ? s->z : NULL, flex ? s->q : NULL, s->part1, s->part3, st);
      bn_scalar_finish(flex ? 4 : 2, s->part1, s->part2, flex ? s->part3 : NULL, sc, 0, st);
      hipMemcpyAsync(hs, sc, BN_S_COUNT * sizeof(double), hipMemcpyDeviceToHost, st);
      if (hipStreamSynchronize(st) != hipSuccess) {
        bn_set_error("device failure in PCG step %d", it);
        return BN_ERR_HIP;
      }
      rel = hs[BN_S_REL];
      rnorm = sqrt(hs[BN_S_RR]);
      if (print_level >= 1)
        fprintf(stderr, "  %4d  %8.2e  %8.2e  %6.4f\n", it, sqrt(hs[BN_S_RZNEW]), rel,
                hs[BN_S_RZNEW] / hs[BN_S_RZOLD]);
      const double test = relres_stop ? rnorm / bnorm : rel;
      if (test < tol) break;
      bn_pcg_update_p(n, s->p, s->z, sc, st);
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
