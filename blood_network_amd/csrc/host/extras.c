/*
 * extras.c - entry points that reuse one hierarchy for more than the
 * plain solve: the adjoint gradient and multi-right-hand-side solves.
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bn_solver.h"
#include "csr.h"
#include "schedule.h"

void bn_set_error(const char *fmt, ...);
void bn_track_device(struct bn_solver *s, void *p);
int bn_ensure_device(struct bn_solver *s);

/* A is symmetric, so the adjoint system A lam = dL/dp is solved with the
 * same hierarchy and PCG; the edge-scatter kernel then forms
 *   dL/dg_e = -(lam_u - lam_v)(p_u - p_v)   (Dirichlet ends: lam = 0,
 *                                            p = pdir)
 *   dL/dlog r_e = 4 g_e dL/dg_e             (g = pi r^4 / (8 mu len)). */
int bn_solver_adjoint_grad(void *h, const double *p, const double *dLdp, int m, const int *eu, const int *ev,
                           const double *g, const double *pdir, double tol, int maxiter, double *lam,
                           double *dLdg, double *dLdlogr) {
  struct bn_solver *s = (struct bn_solver *)h;
  if (!s || !p || !dLdp || !lam || m < 0 || (m > 0 && (!eu || !ev || !g))) return BN_ERR_ARG;
  const int n = s->A->row;
  memset(lam, 0, (size_t)n * sizeof(double));
  int it = bn_solver_solve_info(h, dLdp, lam, tol, maxiter, 0, NULL);
  if (it == BN_ERR_HIP || it == BN_ERR_NO_DEVICE || it == BN_ERR_ARG) return it;
  if (m == 0 || (!dLdg && !dLdlogr)) return it;
  int nd = 0;
  for (int e = 0; e < m; e++) {
    if (eu[e] < -nd) nd = -eu[e];
    if (ev[e] < -nd) nd = -ev[e];
  }
  if (nd > 0 && !pdir) return BN_ERR_ARG;
  void *buf[8] = {0};
  size_t sz[8] = {(size_t)m * 4, (size_t)m * 4, (size_t)m * 8, (size_t)n * 8, (size_t)n * 8,
                  (size_t)(nd > 0 ? nd : 1) * 8, (size_t)m * 8, (size_t)m * 8};
  for (int k = 0; k < 8; k++)
    if (hipMalloc(&buf[k], sz[k] ? sz[k] : 8) != hipSuccess) {
      for (int q = 0; q < k; q++) hipFree(buf[q]);
      bn_set_error("device out of memory in the adjoint gradient");
      return BN_ERR_HIP;
    }
  hipStream_t st = s->stream;
  hipMemcpyAsync(buf[0], eu, sz[0], hipMemcpyHostToDevice, st);
  hipMemcpyAsync(buf[1], ev, sz[1], hipMemcpyHostToDevice, st);
  hipMemcpyAsync(buf[2], g, sz[2], hipMemcpyHostToDevice, st);
  hipMemcpyAsync(buf[3], p, sz[3], hipMemcpyHostToDevice, st);
  hipMemcpyAsync(buf[4], lam, sz[4], hipMemcpyHostToDevice, st);
  if (nd > 0) hipMemcpyAsync(buf[5], pdir, (size_t)nd * 8, hipMemcpyHostToDevice, st);
  bn_edge_grad(m, (const int *)buf[0], (const int *)buf[1], (const double *)buf[2], (const double *)buf[3],
               (const double *)buf[4], (const double *)buf[5], (double *)buf[6], (double *)buf[7], st);
  if (dLdg) hipMemcpyAsync(dLdg, buf[6], sz[6], hipMemcpyDeviceToHost, st);
  if (dLdlogr) hipMemcpyAsync(dLdlogr, buf[7], sz[7], hipMemcpyDeviceToHost, st);
  hipError_t e = hipStreamSynchronize(st);
  for (int k = 0; k < 8; k++) hipFree(buf[k]);
  return e == hipSuccess ? it : BN_ERR_HIP;
}

/* The same gradient with every array in device memory: no allocation and
 * no host copies, for descent loops whose pressures already live in HBM.
 * d_lam holds the initial guess on entry (zeros for a cold start; the
 * previous step's multipliers warm-start the next).  nd is the number of
 * Dirichlet pressures in d_pdir (endpoints -1-d with d < nd). */
int bn_solver_adjoint_grad_device(void *h, const double *d_p, const double *d_dLdp, int m, const int *d_eu,
                                  const int *d_ev, const double *d_g, const double *d_pdir, int nd, double tol,
                                  int maxiter, double *d_lam, double *d_dLdg, double *d_dLdlogr) {
  struct bn_solver *s = (struct bn_solver *)h;
  if (!s || !d_p || !d_dLdp || !d_lam || m < 0 || nd < 0 || (nd > 0 && !d_pdir) ||
      (m > 0 && (!d_eu || !d_ev || (d_dLdlogr && !d_g))))
    return BN_ERR_ARG;
  int it = bn_solver_solve_device(h, d_dLdp, d_lam, tol, maxiter, 0, NULL);
  if (it == BN_ERR_HIP || it == BN_ERR_NO_DEVICE || it == BN_ERR_ARG) return it;
  if (m == 0 || (!d_dLdg && !d_dLdlogr)) return it;
  bn_edge_grad(m, d_eu, d_ev, d_g, d_p, d_lam, d_pdir, d_dLdg, d_dLdlogr, s->stream);
  return hipStreamSynchronize(s->stream) == hipSuccess ? it : BN_ERR_HIP;
}

/* Caller slot of every stored solve-matrix entry: the stored matrix is
 * the caller's pattern relabelled by the level-0 row order, so relabelling
 * a matrix whose values are the slot numbers (exact in FP64) gives the
 * map.  The same order needs no gather. */
int bn_value_map(struct bn_solver *s) {
  const dcsr_mat *A = s->A;
  const int n = A->row, nnz = A->nnz;
  int *inv = (int *)bn_xmalloc((size_t)(n > 0 ? n : 1) * sizeof(int));
  for (int k = 0; k < n; k++) inv[s->lev[0].perm_h[k]] = k;
  dcsr_mat I = *A, Ip;
  I.val = (double *)bn_xmalloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(double));
  for (int k = 0; k < nnz; k++) I.val[k] = (double)k;
  bn_csr_permute(&I, s->lev[0].perm_h, inv, &Ip);
  free(inv);
  free(I.val);
  int *map = (int *)bn_xmalloc((size_t)(nnz > 0 ? nnz : 1) * sizeof(int));
  int same = 1;
  for (int k = 0; k < nnz; k++) {
    map[k] = (int)Ip.val[k];
    same &= (map[k] == k);
  }
  bn_csr_free(&Ip);
  int rc = 0;
  if (same) {
    s->vperm_state = 1;
  } else {
    void *d = NULL;
    if (hipMalloc(&d, (size_t)nnz * sizeof(int)) != hipSuccess) {
      bn_set_error("device out of memory for the value map");
      rc = BN_ERR_HIP;
    } else {
      bn_track_device(s, d);
      if (hipMemcpy(d, map, (size_t)nnz * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) rc = BN_ERR_HIP;
      s->vperm_d = (int *)d;
      s->vperm_state = 2;
    }
  }
  free(map);
  return rc;
}

/* bn_solver_adjoint_grad without any device allocation: the solve runs on
 * the solver's own vectors (step graphs already captured are replayed,
 * none is captured), p sits in one of the solver's work vectors, and the
 * segment arrays pass through the static device scratch (noalloc.hip) in
 * chunks of BN_SCRATCH_EDGES, so m is unbounded; at most
 * BN_SCRATCH_DIRICHLET Dirichlet pressures.  Every output element is
 * written by the kernel, so nothing in the scratch needs clearing.  Calls
 * that overlap on one GPU share its scratch and race. */
int bn_solver_adjoint_grad_noalloc(void *h, const double *p, const double *dLdp, int m, const int *eu,
                                   const int *ev, const double *g, const double *pdir, double tol, int maxiter,
                                   double *lam, double *dLdg, double *dLdlogr) {
  struct bn_solver *s = (struct bn_solver *)h;
  if (!s || !p || !dLdp || !lam || m < 0 || (m > 0 && (!eu || !ev || !g))) {
    bn_set_error("adjoint_grad_noalloc: NULL argument or m < 0");
    return BN_ERR_ARG;
  }
  if (!s->dev_ready) {
    bn_set_error("adjoint_grad_noalloc: the solver has no device mirror (it is uploaded at creation when a GPU "
                 "is present; this entry point does not upload)");
    return BN_ERR_ARG;
  }
  int nd = 0;
  for (int e = 0; e < m; e++) {
    if (eu[e] < -nd) nd = -eu[e];
    if (ev[e] < -nd) nd = -ev[e];
  }
  if (nd > 0 && !pdir) {
    bn_set_error("adjoint_grad_noalloc: Dirichlet ends given but no pdir");
    return BN_ERR_ARG;
  }
  if (nd > BN_SCRATCH_DIRICHLET) {
    bn_set_error("adjoint_grad_noalloc: %d Dirichlet pressures, more than the %d the static scratch holds", nd,
                 BN_SCRATCH_DIRICHLET);
    return BN_ERR_ARG;
  }
  if (hipSetDevice(s->device) != hipSuccess) return BN_ERR_HIP;
  unsigned char *scr = NULL;
  size_t bytes = 0;
  int rc = bn_scratch_ptr((void **)&scr, &bytes);
  if (rc) {
    bn_set_error("adjoint_grad_noalloc: device scratch unavailable");
    return rc;
  }
  const int n = s->lev[0].n;
  hipStream_t st = s->stream;
  /* lambda = A^{-1} dL/dp from zero, on the solver's level-0 vectors */
  double *db = (double *)s->lev[0].b, *dx = (double *)s->lev[0].x;
  if (hipMemcpyAsync(db, dLdp, (size_t)n * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess)
    return BN_ERR_HIP;
  bn_fill('d', n, dx, 0.0, st);
  s->noalloc = 1;
  const int it = bn_pcg_run(s, db, dx, tol, maxiter, 0, NULL);
  s->noalloc = 0;
  if (it == BN_ERR_HIP || it == BN_ERR_NO_DEVICE || it == BN_ERR_ARG) return it;
  if (hipMemcpyAsync(lam, dx, (size_t)n * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess)
    return BN_ERR_HIP;
  if (m > 0 && (dLdg || dLdlogr)) {
    /* scratch: pdir, then per pass eu | ev | g | dg | dlogr */
    double *pdir_d = (double *)scr;
    unsigned char *chunk = scr + (size_t)BN_SCRATCH_DIRICHLET * sizeof(double);
    int *eu_d = (int *)chunk;
    int *ev_d = eu_d + BN_SCRATCH_EDGES;
    double *g_d = (double *)(ev_d + BN_SCRATCH_EDGES);
    double *dg_d = g_d + BN_SCRATCH_EDGES;
    double *dr_d = dg_d + BN_SCRATCH_EDGES;
    double *p_d = s->q; /* a free FP64 n-vector of the solver after the solve */
    if (hipMemcpyAsync(p_d, p, (size_t)n * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess ||
        (nd > 0 && hipMemcpyAsync(pdir_d, pdir, (size_t)nd * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess))
      return BN_ERR_HIP;
    /* stream order keeps each pass's copies behind the previous pass's reads */
    for (int e0 = 0; e0 < m; e0 += BN_SCRATCH_EDGES) {
      const int mc = (m - e0 < BN_SCRATCH_EDGES) ? m - e0 : BN_SCRATCH_EDGES;
      if (hipMemcpyAsync(eu_d, eu + e0, (size_t)mc * sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess ||
          hipMemcpyAsync(ev_d, ev + e0, (size_t)mc * sizeof(int), hipMemcpyHostToDevice, st) != hipSuccess ||
          hipMemcpyAsync(g_d, g + e0, (size_t)mc * sizeof(double), hipMemcpyHostToDevice, st) != hipSuccess)
        return BN_ERR_HIP;
      bn_edge_grad(mc, eu_d, ev_d, g_d, p_d, dx, pdir_d, dg_d, dr_d, st);
      if ((dLdg && hipMemcpyAsync(dLdg + e0, dg_d, (size_t)mc * sizeof(double), hipMemcpyDeviceToHost, st) !=
                       hipSuccess) ||
          (dLdlogr && hipMemcpyAsync(dLdlogr + e0, dr_d, (size_t)mc * sizeof(double), hipMemcpyDeviceToHost, st) !=
                          hipSuccess))
        return BN_ERR_HIP;
    }
  }
  return hipStreamSynchronize(st) == hipSuccess ? it : BN_ERR_HIP;
}

/* New FP64 values for the solve matrix, in device memory, on the pattern
 * and in the entry order of the stored (symmetrised) matrix: the creation
 * order whenever the input pattern was symmetric without duplicates, as
 * for every assembled network; nnz must equal the stored count.  Only the matrix
 * PCG multiplies by changes: the hierarchy and the fine IC(0) stay those
 * of the matrix they were built from (two-matrix preconditioning, as
 * C_amg1/src/amg_wrapper.c:168-199 set_matrix, here without the host
 * factorisation and the upload).  The host copy of the matrix is not
 * updated; amg_solver_update / amg_solver_set_matrix replace both. */
int bn_solver_set_values_device(void *h, int nnz_in, const double *d_val) {
  struct bn_solver *s = (struct bn_solver *)h;
  if (!s || !d_val || s->borrowed) return BN_ERR_ARG;
  if (nnz_in != s->A->nnz) {
    bn_set_error("set_values: %d values given, the stored solve matrix has %d", nnz_in, s->A->nnz);
    return BN_ERR_ARG;
  }
  int rc = bn_ensure_device(s);
  if (rc) return rc;
  if (s->Asolve.nnz != s->A->nnz) return BN_ERR_ARG;
  if (!s->vperm_state && (rc = bn_value_map(s))) return rc;
  const size_t nnz = (size_t)s->A->nnz;
  if (s->asolve_shared) {
    /* the level-0 V-cycle operator keeps its values: the solve matrix
       gets FP64 and FP32 values of its own (pattern arrays stay shared;
       the FP32 copy serves the residuals inside an FP32 preconditioner,
       which use the solve matrix as amg_ichol_precond.c:37,44 do) */
    void *v = NULL, *v32 = NULL;
    if (hipMalloc(&v, (nnz ? nnz : 1) * sizeof(double)) != hipSuccess ||
        hipMalloc(&v32, (nnz ? nnz : 1) * sizeof(float)) != hipSuccess) {
      if (v) hipFree(v);
      bn_set_error("device out of memory for the solve-matrix values");
      return BN_ERR_HIP;
    }
    bn_track_device(s, v);
    bn_track_device(s, v32);
    bn_drop_graphs(s);
    s->Asolve.vd = (double *)v;
    s->Asolve.vs = (float *)v32;
    s->asolve_shared = 0;
  }
  hipStream_t st = s->stream;
  if (s->vperm_state == 1) {
    if (hipMemcpyAsync(s->Asolve.vd, d_val, nnz * sizeof(double), hipMemcpyDeviceToDevice, st) != hipSuccess)
      return BN_ERR_HIP;
  } else {
    bn_permute((int)nnz, s->vperm_d, d_val, s->Asolve.vd, 0, st);
  }
  /* the FP32 copy follows (same rounding as at upload) */
  bn_rb_to_float((int)nnz, s->Asolve.vd, s->Asolve.vs, st);
  return hipStreamSynchronize(st) == hipSuccess ? 0 : BN_ERR_HIP;
}

/* k right-hand sides, column-major; every column is solved to tol with
 * the shared hierarchy.  its[j] gets column j's iteration count; the
 * return value is the largest count (negative if any column failed). */
int bn_solver_coarse_inverse(void *h, int *n, double *out) {
  struct bn_solver *s = (struct bn_solver *)h;
  if (!s || !n) return BN_ERR_ARG;
  int rc = bn_ensure_device(s);
  if (rc) return rc;
  if (!s->coarse_dense) {
    bn_set_error("the coarsest solve is not the dense inverse");
    return BN_ERR_ARG;
  }
  const int nc = s->lev[s->nlev - 1].n;
  *n = nc;
  if (out && hipMemcpy(out, s->cinv_d, (size_t)nc * (size_t)nc * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    return BN_ERR_HIP;
  return 0;
}

int bn_solver_solve_batch(void *h, int k, const double *b, double *x, double tol, int maxiter, int *its,
                          bn_solve_info *info) {
  struct bn_solver *s = (struct bn_solver *)h;
  if (!s || k < 0 || (k > 0 && (!b || !x))) return BN_ERR_ARG;
  const size_t n = (size_t)s->A->row;
  int worst = 0, failed = 0;
  double ms = 0.0, wall = 0.0;
  for (int j = 0; j < k; j++) {
    bn_solve_info inf;
    int it = bn_solver_solve_info(h, b + j * n, x + j * n, tol, maxiter, 0, &inf);
    if (it == BN_ERR_HIP || it == BN_ERR_NO_DEVICE || it == BN_ERR_ARG) return it;
    if (its) its[j] = it;
    if (it < 0) failed = 1;
    if ((it < 0 ? -it : it) > worst) worst = it < 0 ? -it : it;
    ms += inf.solve_ms;
    wall += inf.wall_ms;
  }
  if (info) {
    *info = s->last;
    info->solve_ms = ms;
    info->wall_ms = wall;
  }
  return failed ? -worst : worst;
}

/* ---------------------------------------------------------------------
 * live kernel timing for the roofline report
 * ------------------------------------------------------------------- */
int bn_solver_time_kernel(void *h, int which, int reps, double *ms_avg, double *bytes) {
  struct bn_solver *s = (struct bn_solver *)h;
  if (!s || reps <= 0) return BN_ERR_ARG;
  int rc = bn_ensure_device(s);
  if (rc) return rc;
  const bn_dlevel *d = &s->lev[0];
  const int n = d->n;
  const char vt = s->vt;
  const double v = (vt == 's') ? 4.0 : 8.0;
  const bn_dcsr *A = &s->Asolve;
  hipStream_t st = s->stream;
  double b = 0.0;
  for (int r = 0; r <= reps; r++) {
    if (r == 1) hipEventRecord(s->ev0, st);
    switch (which) {
      case 0:
        bn_dcsr_apply_dot(A, 'd', BN_OP_SET, s->p, s->q, NULL, NULL, NULL, st);
        b = (double)A->nnz * 12.0 + (n + 1.0) * 4.0 + 2.0 * n * 8.0;
        break;
      case 1: {
        const void *val = vt == 's' ? (const void *)d->A.vs : (const void *)d->A.vd;
        if (!d->block && !d->use_gs) {
          /* l1-Jacobi level: residual + update */
          bn_dcsr_apply(&d->A, vt, BN_OP_RESID, bn_lanes(&d->A, vt), s->u1, d->r, s->r1, st);
          bn_jacobi_update(vt, n, d->r, vt == 's' ? (const void *)d->l1inv_s : (const void *)d->l1inv_d, 1.0,
                           s->u1, st);
        } else if (d->block)
          bn_block_gs(vt, d->nchunk, d->chunk_max, d->chunk_d, d->ccol_d, d->cbnd_d, d->A.ia, d->A.ja, val,
                      vt == 's' ? (const void *)d->bdinv_s : (const void *)d->bdinv_d, s->r1, s->u1, d->x2, 0,
                      bn_lanes(&d->A, vt), NULL, NULL, st);
        else
          for (int c = 0; c < d->ncol; c++)
            bn_gs_colour(s, d, vt, c, vt == 's' ? d->gs_lanes_s : d->gs_lanes, s->r1, s->u1, NULL, NULL);
        /* the colour-ordered copy's off-diagonal entries and row pointers,
         * the row map, b, x read and written, 1/a_ii */
        if (d->use_gs && !d->block)
          b = (double)d->Ac.nnz * (4.0 + v) + (n + 1.0) * 4.0 + 4.0 * n + 3.0 * n * v + n * v;
        else
          b = (double)d->A.nnz * (4.0 + v) + (n + 1.0) * 4.0 + 3.0 * n * v + n * v;
        break;
      }
      case 2: {
        bn_op_apply(&d->A, vt, BN_OP_RESID, s->u1, s->r1, s->g_in, st);
        b = (double)d->A.nnz * (4.0 + v) + (n + 1.0) * 4.0 + 3.0 * n * v;
        break;
      }
      case 3:
        bn_vc_fine_ic(s, s->r1, s->icx);
        b = 2.0 * ((double)s->fine_ic.lo.nnz * (4.0 + v) + (n + 1.0) * 4.0 + 3.0 * n * v + 4.0 * n);
        break;
      case 4:
        bn_vc_ichol_precond(s, s->r, s->z);
        b = 0.0;
        break;
      case 5: /* plain FP64 copy of n values: the byte-counter calibration */
        bn_copy('d', n, s->p, s->q, st);
        b = 16.0 * n;
        break;
      case 6: { /* one Jacobi sweep over the fine IC(0) lower triangle */
        const bn_dtri *t = &s->fine_ic;
        if (!t->lo.ia) return BN_ERR_ARG;
        bn_jacobi_rows(vt, n, t->lo.ia, t->lo.ja, vt == 's' ? (const void *)t->lo.vs : (const void *)t->lo.vd,
                       vt == 's' ? (const void *)t->dinv_s : (const void *)t->dinv_d, s->r1, s->icy, s->icx,
                       bn_lanes(&t->lo, vt), NULL, NULL, st);
        /* L entries (index + value), row pointers, b, dinv, x gathered once, x written */
        b = (double)t->lo.nnz * (4.0 + v) + (n + 1.0) * 4.0 + 4.0 * n * v;
        break;
      }
      default:
        return BN_ERR_ARG;
    }
  }
  hipEventRecord(s->ev1, st);
  if (hipEventSynchronize(s->ev1) != hipSuccess) return BN_ERR_HIP;
  float ms = 0.0f;
  hipEventElapsedTime(&ms, s->ev0, s->ev1);
  if (ms_avg) *ms_avg = ms / reps;
  if (bytes) *bytes = b;
  return 0;
}

/* ---------------------------------------------------------------------
 * device memory helpers (one HIP runtime per process: the library's)
 * ------------------------------------------------------------------- */
int bn_dev_alloc(void **ptr, size_t bytes) {
  if (!ptr) return BN_ERR_ARG;
  return hipMalloc(ptr, bytes ? bytes : 8) == hipSuccess ? 0 : BN_ERR_HIP;
}

int bn_dev_free(void *ptr) { return hipFree(ptr) == hipSuccess ? 0 : BN_ERR_HIP; }

int bn_dev_copy(void *dst, const void *src, size_t bytes, int kind) {
  hipMemcpyKind k = kind == 1 ? hipMemcpyHostToDevice : kind == 2 ? hipMemcpyDeviceToHost
                                                                  : hipMemcpyDeviceToDevice;
  return hipMemcpy(dst, src, bytes, k) == hipSuccess ? 0 : BN_ERR_HIP;
}

int bn_dev_sync(void) { return hipDeviceSynchronize() == hipSuccess ? 0 : BN_ERR_HIP; }
