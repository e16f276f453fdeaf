/*
 * vcycle.c - device V-cycle halves and the AMG + IC(0) preconditioner.
 *
 * Host C that only enqueues kernels on the solver's stream; all vectors
 * stay on the device.  Structure follows the reference:
 *   backslash  C_amg1/src/amg_cycle.c:180-217 (pre-smooth, residual,
 *              restrict, recurse from zero, prolong)
 *   fwdslash   C_amg1/src/amg_cycle.c:222-258 (residual, restrict,
 *              recurse, prolong, post-smooth in reverse order)
 *   AMG+IC     C_amg1/src/amg_ichol_precond.c:22-59
 *   V-cycle    C_amg1/src/amg_cycle.c:263-284
 * Differences by design: the CF-ordered Gauss-Seidel sweep runs as a
 * C-then-F multicolour sweep (levels needing more than max_colours
 * colours fall back to l1-Jacobi) unless the exact order is asked for
 * (gs_exact; SMOOTHER_GS always sweeps in exact natural order): then the
 * sequential sweep runs level by level over its dependency DAG and reads
 * exactly the values the reference's sweep reads (amg_cycle.c:21-80),
 * restriction gathers over a stored
 * R = P^T instead of scattering, the coarsest solve is a dense inverse
 * (or the reference IC-PCG), and zero initial guesses skip the work they
 * make redundant (a residual equal to b, the first colour's row sums).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "bn_solver.h"

void bn_set_error(const char *fmt, ...);

static inline const void *opv(const bn_dcsr *m, char vt) {
  return vt == 's' ? (const void *)m->vs : (const void *)m->vd;
}
static inline const void *vecv(const double *vd, const float *vs, char vt) {
  return vt == 's' ? (const void *)vs : (const void *)vd;
}

/* y = op(M, x [, b]) in precision vt */
static void apply(struct bn_solver *s, const bn_dcsr *M, int op, const void *x, void *y, const void *b) {
  bn_launch_apply(s, M, s->vt, op, x, y, b);
}

/* Block smoother: nu sweeps that start from xin (NULL: zero) and end in
 * x.  Sweep k writes x when nu-1-k is even, else the level's x2, so xin
 * must be block_input(): x2 for odd nu, x for even nu (never the buffer
 * the first sweep writes). */
static void *block_input(const bn_dlevel *d, int nu, void *x) { return (nu & 1) ? d->x2 : x; }

static void smooth_block(struct bn_solver *s, int lev, const void *b, const void *xin, void *x, int backward,
                         const void *zbase, void *zsum) {
  bn_dlevel *d = &s->lev[lev];
  const char vt = s->vt;
  const int nu = s->opt.nu > 0 ? s->opt.nu : 1;
  const void *src = xin;
  for (int k = 0; k < nu; k++) {
    void *dst = ((nu - 1 - k) % 2 == 0) ? x : d->x2;
    const int last = (k == nu - 1);
    bn_block_gs(vt, d->nchunk, d->chunk_max, d->chunk_d, d->ccol_d, d->cbnd_d, d->A.ia, d->A.ja, opv(&d->A, vt),
                vecv(d->bdinv_d, d->bdinv_s, vt), b, src, dst, backward, bn_lanes(&d->A, vt),
                last ? zbase : NULL, last ? zsum : NULL, s->stream);
    src = dst;
  }
}

int bn_vc_exact_kind(const struct bn_solver *s) {
  if (s->opt.smoother == SMOOTHER_GS) return 0;
  if (s->opt.smoother == SMOOTHER_CF_GS && s->opt.gs_exact) return 1;
  return -1;
}

int bn_vc_prepare(struct bn_solver *s) {
  /* ELL copies of the triangles the exact IC(0) solves walk (the CSR rows
   * serve until they exist, with the same results) */
  if (!s->noalloc) {
    int rc = 0;
    if (s->opt.fine_ic == BN_IC_EXACT && s->fine_ic.lo.ia && !s->fine_ic.ell) rc = bn_tri_ell_build(s, &s->fine_ic);
    if (!rc && !s->coarse_dense && s->coarse_ic.lo.ia && !s->coarse_ic.ell) rc = bn_tri_ell_build(s, &s->coarse_ic);
    if (rc) return rc;
  }
  const int kind = bn_vc_exact_kind(s);
  if (kind < 0) return 0;
  for (int l = 0; l < s->nlev - 1; l++) {
    bn_dlevel *d = &s->lev[l];
    if (d->block || d->xs[kind].ptr_f) continue;
    if (s->noalloc) {
      bn_set_error("the exact Gauss-Seidel schedules are not built yet (an allocation-free call does not "
                   "build them): run one ordinary solve with this smoother first");
      return BN_ERR_ARG;
    }
    const amg_level *L = &s->amg->levels[l];
    if (kind == 1 && !L->cf_order) {
      bn_set_error("level %d has no cf_order for the exact C/F Gauss-Seidel sweep", l);
      return BN_ERR_ARG;
    }
    /* level rows keep the host numbering outside the block layout */
    const int rc = bn_sweep_upload(s, &L->A, kind == 1 ? L->cf_order : NULL, &d->xs[kind]);
    if (rc) return rc;
  }
  return 0;
}

/* nu exact-order sweeps (the sequential sweep, level by level over its
 * dependency DAG): forward visits the order, backward its reverse
 * (amg_cycle.c:21-80) */
static void smooth_exact(struct bn_solver *s, int lev, int kind, const void *b, void *x, int backward, int x_zero,
                         const void *zbase, void *zsum) {
  bn_dlevel *d = &s->lev[lev];
  const char vt = s->vt;
  const int nu = s->opt.nu > 0 ? s->opt.nu : 1;
  bn_sweep *w = &d->xs[kind];
  if (!w->ptr_f && bn_vc_prepare(s)) {
    fprintf(stderr, "libbnamg: exact Gauss-Seidel schedule unavailable: %s\n", bn_last_error());
    abort();
  }
  if (x_zero) bn_fill(vt, d->n, x, 0.0, s->stream);
  const int nl = backward ? w->nb : w->nf;
  const int *rows = backward ? w->rows_b : w->rows_f;
  const int *ptr = backward ? w->ptr_b : w->ptr_f;
  const void *dinv = vecv(d->dinv_d, d->dinv_s, vt);
  if (s->walk_on && w->jx) {
    /* each sweep in one launch (bn_sweep_parts): the input copied aside, x
     * filled with the sentinel */
    for (int sw = 0; sw < nu; sw++) {
      bn_gs_prep(vt, d->n, x, w->xold, s->stream);
      bn_sweep_parts(vt, bn_lanes(&d->A, vt), w->np, backward, backward ? w->sp_b : w->sp_f,
                     backward ? w->st_b : w->st_f, rows, d->A.ia, w->jx, opv(&d->A, vt), dinv, b, x, w->xold,
                     s->scal + BN_S_WALKERR, s->walk_limit, s->stream);
    }
  } else {
    for (int sw = 0; sw < nu; sw++)
      for (int l = 0; l < nl; l++)
        bn_trsv_level(vt, ptr[l + 1] - ptr[l], rows + ptr[l], d->A.ia, d->A.ja, opv(&d->A, vt), dinv, b, x,
                      bn_lanes(&d->A, vt), NULL, NULL, s->stream);
  }
  if (zsum) {
    bn_copy(vt, d->n, zbase, zsum, s->stream);
    bn_axpy1(vt, d->n, x, zsum, s->stream);
  }
}

void bn_gs_colour(struct bn_solver *s, const bn_dlevel *d, char vt, int c, int lanes, const void *b, void *x,
                  const void *zbase, void *zsum) {
  const void *dinvc = vecv(d->dinvc_d, d->dinvc_s, vt);
  if (lanes == BN_GS_BLOCKS && d->gs_cblk) {
    const int b0 = d->gs_cblk[c];
    bn_gs_color_blocks(vt, b0, d->gs_cblk[c + 1] - b0, d->gs_blk, d->gs_blkq, d->Ac.ia, d->Ac.ja, opv(&d->Ac, vt),
                       d->cmap_d, dinvc, b, x, zbase, zsum, s->stream);
    return;
  }
  if (lanes <= 0 || lanes == BN_GS_BLOCKS) lanes = bn_lanes(&d->Ac, vt);
  bn_gs_color_map(vt, d->col_ptr[c], d->col_ptr[c + 1], d->cmap_d, d->Ac.ia, d->Ac.ja, opv(&d->Ac, vt), dinvc, b, x,
                  lanes, zbase, zsum, s->stream);
}

/* nu smoothing sweeps on level `lev`; backward runs colours in reverse.
 * x_zero: x is zero on entry (first colour then needs no row sums); 2:
 * and the first colour's values are already in x (gs_first_fused).
 * zsum != NULL: the final sweep also writes zsum = zbase + x. */
static void smooth(struct bn_solver *s, int lev, const void *b, void *x, int backward, int x_zero,
                   const void *zbase, void *zsum) {
  bn_dlevel *d = &s->lev[lev];
  const char vt = s->vt;
  const int nu = s->opt.nu > 0 ? s->opt.nu : 1;
  const int n = d->n;
  if (d->block) {
    if (x_zero) {
      smooth_block(s, lev, b, NULL, x, backward, zbase, zsum);
    } else {
      void *in = block_input(d, nu, x);
      if (in != x) bn_copy(vt, n, x, in, s->stream);
      smooth_block(s, lev, b, in, x, backward, zbase, zsum);
    }
    return;
  }
  const int xk = bn_vc_exact_kind(s);
  if (xk >= 0) {
    smooth_exact(s, lev, xk, b, x, backward, x_zero, zbase, zsum);
    return;
  }
  const int gs = d->use_gs && s->opt.smoother == SMOOTHER_CF_GS;
  const int walk = vt == 's' ? d->gs_walk_s : d->gs_walk;
  if (gs && s->walk_on && walk > 0 && d->gs_jx) {
    /* each sweep in one launch (bn_gs_walk): the input copied aside, x
     * filled with the sentinel */
    const void *dinvc = vecv(d->dinvc_d, d->dinvc_s, vt);
    for (int sw = 0; sw < nu; sw++) {
      const int last = (sw == nu - 1), zero = x_zero && sw == 0;
      bn_gs_prep(vt, n, x, zero ? NULL : d->gs_xold, s->stream);
      bn_gs_walk(vt, walk, d->gs_wnp, d->ncol, backward, d->gs_wsteps, d->Ac.ia, d->gs_jx, opv(&d->Ac, vt),
                 d->cmap_d, dinvc, b, x, d->gs_xold, zero, last ? zbase : NULL, last ? zsum : NULL,
                 s->scal + BN_S_WALKERR, s->walk_limit, s->stream);
    }
    return;
  }
  if (gs) {
    /* from zero, the first colour's kernel also zeroes the other rows
     * (unless the first sweep is also the last and must write zsum) */
    const int fused_zero = x_zero && !(nu == 1 && zsum) && d->ncol > 0;
    if (x_zero && !fused_zero) bn_fill(vt, n, x, 0.0, s->stream);
    for (int sw = 0; sw < nu; sw++) {
      const int last = (sw == nu - 1);
      for (int t = 0; t < d->ncol; t++) {
        const int c = backward ? d->ncol - 1 - t : t;
        const int lo = d->col_ptr[c], hi = d->col_ptr[c + 1];
        const void *dinvc = vecv(d->dinvc_d, d->dinvc_s, vt);
        if (sw == 0 && t == 0 && fused_zero) {
          if (x_zero == 2) continue; /* written by the restriction that formed b */
          if (d->colour_d)
            bn_gs_color_zero_colour(vt, n, c, d->colour_d, vecv(d->dinv_d, d->dinv_s, vt), b, x, s->stream);
          else
            bn_gs_color_zero_map(vt, lo, hi, n, d->cmap_d, dinvc, b, x, s->stream);
          continue;
        }
        bn_launch_gs(s, d, vt, c, vt == 's' ? d->gs_lanes_s : d->gs_lanes, b, x, last ? zbase : NULL,
                     last ? zsum : NULL);
      }
    }
  } else {
    const int jac = (s->opt.smoother == SMOOTHER_JACOBI);
    const void *w = jac ? vecv(d->dinv_d, d->dinv_s, vt) : vecv(d->l1inv_d, d->l1inv_s, vt);
    const double om = jac ? s->opt.omega : 1.0;
    for (int sw = 0; sw < nu; sw++) {
      if (sw == 0 && x_zero) {
        bn_jacobi_zero(vt, n, b, w, om, x, s->stream); /* x = om * w * b */
        continue;
      }
      apply(s, &d->A, BN_OP_RESID, x, d->r, b);
      bn_jacobi_update(vt, n, d->r, w, om, x, s->stream);
    }
    if (zsum) {
      bn_copy(vt, n, zbase, zsum, s->stream);
      bn_axpy1(vt, n, x, zsum, s->stream);
    }
  }
}

/* x = A_c^{-1} b on the coarsest level */
static void coarse_solve(struct bn_solver *s, const void *b, void *x);

/* Level l is smoothed by smooth()'s Jacobi steps, and those steps are
 * fused into the launches around them (jac_fuse): the first step from zero
 * into the restriction that forms the level's b (its SET writes x = om w b
 * beside b), each further step into one launch (BN_OP_JAC: the residual and
 * the update that RESID + bn_jacobi_update form, into the other of x and
 * the level's r, so the steps alternate and the last lands in x).  The
 * values are those of the separate launches, to the last bit. */
static int jacobi_fused(const struct bn_solver *s, int l) {
  if (!s->jac_fuse || l <= 0 || l >= s->nlev - 1 || bn_vc_exact_kind(s) >= 0) return 0;
  const bn_dlevel *d = &s->lev[l];
  return !d->block && !(d->use_gs && s->opt.smoother == SMOOTHER_CF_GS);
}

/* Level l's forward multicolour sweep from zero starts with the kernel
 * that writes b_i / a_ii on colour 0 and zero elsewhere (smooth()'s
 * fused_zero), and that too rides on the restriction forming b. */
static int gs_first_fused(const struct bn_solver *s, int l) {
  if (!s->jac_fuse || l < 0 || l >= s->nlev - 1 || bn_vc_exact_kind(s) >= 0) return 0;
  const bn_dlevel *d = &s->lev[l];
  if (d->block || !(d->use_gs && s->opt.smoother == SMOOTHER_CF_GS) || d->ncol <= 0 || !d->colour_d) return 0;
  const int walk = s->vt == 's' ? d->gs_walk_s : d->gs_walk;
  return !(s->walk_on && walk > 0 && d->gs_jx); /* the one-launch sweep starts from the sentinel */
}

int bn_vc_first_colour_fused(const struct bn_solver *s) { return gs_first_fused(s, 0); }

static void jacobi_weights(const struct bn_solver *s, const bn_dlevel *d, const void **w, double *om) {
  const int jac = (s->opt.smoother == SMOOTHER_JACOBI);
  const char vt = s->vt;
  if (jac) *w = vt == 's' ? (const void *)d->dinv_s : (const void *)d->dinv_d;
  else *w = vt == 's' ? (const void *)d->l1inv_s : (const void *)d->l1inv_d;
  *om = jac ? s->opt.omega : 1.0;
}

/* k fused Jacobi steps on level l starting from cur (x or the level's r),
 * alternating between the two; returns where the last step wrote */
static void *jacobi_steps(struct bn_solver *s, int l, const void *b, void *cur, void *x, int k) {
  bn_dlevel *d = &s->lev[l];
  const void *w;
  double om;
  jacobi_weights(s, d, &w, &om);
  for (int i = 0; i < k; i++) {
    void *out = cur == x ? d->r : x;
    bn_launch_apply_jac(s, &d->A, s->vt, BN_OP_JAC, cur, out, b, NULL, w, om, NULL, 0);
    cur = out;
  }
  return cur;
}

/* x_zero: 1 x is zero on entry; 2 the smoother's first step from zero
 * was written by the restriction that formed b: on a fused Jacobi level
 * into x, or into the level's r when the number of further steps is odd,
 * on a Gauss-Seidel level the first colour into x (gs_first_fused) */
void bn_vc_backslash(struct bn_solver *s, int lev, const void *b, void *x, int x_zero) {
  bn_dlevel *d = &s->lev[lev];
  if (lev == s->nlev - 1) {
    coarse_solve(s, b, x);
    return;
  }
  bn_dlevel *c = &s->lev[lev + 1];
  const int nu = s->opt.nu > 0 ? s->opt.nu : 1;
  if (x_zero && jacobi_fused(s, lev)) {
    void *first = ((nu - 1) & 1) ? d->r : x;
    if (x_zero != 2) {
      const void *w;
      double om;
      jacobi_weights(s, d, &w, &om);
      bn_jacobi_zero(s->vt, d->n, b, w, om, first, s->stream);
    }
    jacobi_steps(s, lev, b, first, x, nu - 1);
  } else {
    smooth(s, lev, b, x, 0, x_zero, NULL, NULL);
  }
  apply(s, &d->A, BN_OP_RESID, x, d->r, b);
  if (jacobi_fused(s, lev + 1)) {
    /* the child's first Jacobi step from zero rides on its restriction */
    const void *w;
    double om;
    jacobi_weights(s, c, &w, &om);
    bn_launch_apply_jac(s, &d->R, s->vt, BN_OP_SET, d->r, c->b, NULL, ((nu - 1) & 1) ? c->r : c->x, w, om, NULL,
                        0);
    bn_vc_backslash(s, lev + 1, c->b, c->x, 2);
  } else if (gs_first_fused(s, lev + 1)) {
    /* and so does the first colour of a Gauss-Seidel child's sweep from
     * zero (b_i / a_ii on colour 0, zero elsewhere) */
    bn_launch_apply_jac(s, &d->R, s->vt, BN_OP_SET, d->r, c->b, NULL, c->x, vecv(c->dinv_d, c->dinv_s, s->vt), 1.0,
                        c->colour_d, 0);
    bn_vc_backslash(s, lev + 1, c->b, c->x, 2);
  } else {
    apply(s, &d->R, BN_OP_SET, d->r, c->b, NULL);
    bn_vc_backslash(s, lev + 1, c->b, c->x, 1);
  }
  apply(s, &d->P, BN_OP_ADD, c->x, x, NULL);
}

static void fwdslash_rec(struct bn_solver *s, int lev, const void *b, void *x, int x_zero,
                         const void *zbase, void *zsum) {
  bn_dlevel *d = &s->lev[lev];
  if (lev == s->nlev - 1) {
    coarse_solve(s, b, x);
    if (zsum) {
      bn_copy(s->vt, d->n, zbase, zsum, s->stream);
      bn_axpy1(s->vt, d->n, x, zsum, s->stream);
    }
    return;
  }
  bn_dlevel *c = &s->lev[lev + 1];
  if (x_zero) {
    apply(s, &d->R, BN_OP_SET, b, c->b, NULL);
  } else {
    apply(s, &d->A, BN_OP_RESID, x, d->r, b);
    apply(s, &d->R, BN_OP_SET, d->r, c->b, NULL);
  }
  fwdslash_rec(s, lev + 1, c->b, c->x, 1, NULL, NULL);
  if (x_zero && !zsum && jacobi_fused(s, lev)) {
    /* interpolate into the buffer the nu fused steps start from */
    const int nu = s->opt.nu > 0 ? s->opt.nu : 1;
    void *first = (nu & 1) ? d->r : x;
    apply(s, &d->P, BN_OP_SET, c->x, first, NULL);
    jacobi_steps(s, lev, b, first, x, nu);
    return;
  }
  if (d->block && x_zero) {
    /* interpolate straight into the buffer the first sweep reads */
    void *in = block_input(d, s->opt.nu > 0 ? s->opt.nu : 1, x);
    apply(s, &d->P, BN_OP_SET, c->x, in, NULL);
    smooth_block(s, lev, b, in, x, 1, zbase, zsum);
    return;
  }
  apply(s, &d->P, x_zero ? BN_OP_SET : BN_OP_ADD, c->x, x, NULL);
  smooth(s, lev, b, x, 1, 0, zbase, zsum);
}

void bn_vc_fwdslash(struct bn_solver *s, int lev, const void *b, void *x, int x_zero) {
  fwdslash_rec(s, lev, b, x, x_zero, NULL, NULL);
}

/* ---------------------------------------------------------------------
 * coarsest level
 * ------------------------------------------------------------------- */
static int coarse_icpcg(struct bn_solver *s, const double *b, double *x);

/* x = A_c^{-1} b.  The dense path ignores x; the IC-PCG path starts from
 * x, which every internal caller has zeroed (the reference recursion
 * calloc's it, C_amg1/src/amg_cycle.c:205,243). */
static void coarse_solve(struct bn_solver *s, const void *b, void *x) {
  const int n = s->lev[s->nlev - 1].n;
  if (s->vt == 'd') {
    if (s->coarse_dense) {
      bn_dense_gemv('d', n, s->cinv_d, b, x, s->stream);
    } else {
      bn_fill('d', n, x, 0.0, s->stream);
      coarse_icpcg(s, (const double *)b, (double *)x);
    }
    return;
  }
  /* FP32 V-cycle: the coarsest solve stays FP64 */
  if (s->coarse_dense) {
    bn_dense_gemv_ds(n, s->cinv_d, (const float *)b, (float *)x, s->stream);
    return;
  }
  double *bd = s->cvec[4], *xd = s->cvec[5];
  bn_convert('s', 'd', n, b, bd, s->stream);
  bn_fill('d', n, xd, 0.0, s->stream);
  coarse_icpcg(s, bd, xd);
  bn_convert('d', 's', n, xd, x, s->stream);
}

/* one triangle, level by level: from the ELL copy where it exists,
 * otherwise from the CSR rows.  The ELL kernels form the one-lane CSR
 * kernel's sums, whatever lanes the CSR rows would run with (the tuner
 * picks one lane per row for large factors).  (Runs of thin levels in one
 * workgroup, a barrier between levels, measured no faster than a launch
 * per level: 116.7 against 115.3 ms for the plain FP64 solve of the
 * 4.36M-unknown network.) */
static void tri_dir(struct bn_solver *s, const bn_dtri *t, const bn_dcsr *m, const bn_tell *e, int nl,
                    const int *ptr, const int *rows, char vt, const void *dinv, const void *b, void *y,
                    const void *zbase, void *zsum) {
  const int lanes = bn_lanes(m, vt);
  const size_t vb = vt == 's' ? sizeof(float) : sizeof(double);
  const void *ev = vt == 's' ? (const void *)e->vs : (const void *)e->vd;
  const void *dl = vt == 's' ? (const void *)e->ds : (const void *)e->dd;
  if (s->walk_on && t->ell && e->steps) {
    /* each run of ELL levels in one launch (bn_trsv_parts); levels with CSR
     * rows one launch each */
    bn_walk_fill(vt, t->n, y, s->stream);
    int r = 0;
    for (int l = 0; l < nl;) {
      if (r < e->nrun && e->run_lo[r] == l) {
        bn_trsv_parts(vt, e->np, e->sp + (size_t)r * (e->np + 1), e->steps, rows, e->ja, ev, dl, b, y, zbase, zsum,
                      s->scal + BN_S_WALKERR, s->walk_limit, s->stream);
        l = e->run_hi[r++];
        continue;
      }
      bn_trsv_level(vt, ptr[l + 1] - ptr[l], rows + ptr[l], m->ia, m->ja, opv(m, vt), dinv, b, y, lanes, zbase,
                    zsum, s->stream);
      l++;
    }
    return;
  }
  for (int l = 0; l < nl; l++) {
    const int cnt = ptr[l + 1] - ptr[l];
    if (t->ell && e->w[l] > 0)
      bn_trsv_ell(vt, e->w[l], cnt, rows + ptr[l], e->ja + e->q[l], (const char *)ev + (size_t)e->q[l] * vb,
                  (const char *)dl + (size_t)ptr[l] * vb, b, y, zbase, zsum, s->stream);
    else
      bn_trsv_level(vt, cnt, rows + ptr[l], m->ia, m->ja, opv(m, vt), dinv, b, y, lanes, zbase, zsum, s->stream);
  }
}

/* IC(0) application with level-scheduled forward and backward sweeps:
 * x = (L L^T)^{-1} b  (C_amg1/src/ichol.c:121-165).  If zsum, also
 * zsum = zbase + x. */
static void tri_solve(struct bn_solver *s, bn_dtri *t, char vt, const void *b, void *y, void *x,
                      const void *zbase, void *zsum) {
  const void *dinv = vt == 's' ? (const void *)t->dinv_s : (const void *)t->dinv_d;
  tri_dir(s, t, &t->lo, &t->ef, t->nlev_f, t->ptr_f, t->rows_f, vt, dinv, b, y, NULL, NULL);
  tri_dir(s, t, &t->up, &t->eb, t->nlev_b, t->ptr_b, t->rows_b, vt, dinv, y, x, zbase, zsum);
}

/* k Jacobi sweeps per triangle (approximate IC application; only used in
 * the flexible outer iteration).  The result goes to x, or, when zsum is
 * given, only zsum = zbase + result is written (fused into the last
 * sweep). */
/* first_done: y already holds the lower triangle's first step b_i dinv_i
 * (written by the residual kernel that formed b) */
static void tri_jacobi(struct bn_solver *s, bn_dtri *t, char vt, const void *b, void *y, void *x,
                       void *tmp, int sweeps, const void *zbase, void *zsum, int first_done) {
  const void *dinv = vt == 's' ? (const void *)t->dinv_s : (const void *)t->dinv_d;
  const int n = t->n;
  void *bufs[2] = {y, tmp};
  int cur = 0;
  if (!first_done) bn_launch_jacobi(s, &t->lo, vt, dinv, b, NULL, bufs[cur], NULL, NULL, NULL);
  /* the last lower sweep also writes the upper triangle's first step
   * (yy_i dinv_i) into x, which the upper sweeps start from */
  for (int k = 1; k < sweeps; k++) {
    if (k == sweeps - 1) bn_launch_jacobi(s, &t->lo, vt, dinv, b, bufs[cur], bufs[1 - cur], NULL, NULL, x);
    else bn_launch_jacobi(s, &t->lo, vt, dinv, b, bufs[cur], bufs[1 - cur], NULL, NULL, NULL);
    cur = 1 - cur;
  }
  const void *yy = bufs[cur];
  void *xb[2] = {x, bufs[1 - cur]};
  int cx = 0;
  const int fuse = zsum && sweeps > 1;
  if (sweeps < 2) bn_launch_jacobi(s, &t->up, vt, dinv, yy, NULL, xb[cx], NULL, NULL, NULL);
  for (int k = 1; k < sweeps; k++) {
    const int last = (k == sweeps - 1);
    bn_launch_jacobi(s, &t->up, vt, dinv, yy, xb[cx], xb[1 - cx], fuse && last ? zbase : NULL,
                     fuse && last ? zsum : NULL, NULL);
    cx = 1 - cx;
  }
  if (fuse) return;
  if (xb[cx] != x) bn_copy(vt, n, xb[cx], x, s->stream);
  if (zsum) {
    bn_copy(vt, n, zbase, zsum, s->stream);
    bn_axpy1(vt, n, x, zsum, s->stream);
  }
}

void bn_vc_tri_apply(struct bn_solver *s, bn_dtri *t, char vt, const void *b, void *y, void *x) {
  tri_solve(s, t, vt, b, y, x, NULL, NULL);
}

void bn_vc_fine_ic(struct bn_solver *s, const void *b, void *x) {
  if (s->opt.fine_ic == BN_IC_JACOBI)
    tri_jacobi(s, &s->fine_ic, s->vt, b, s->icy, x, s->lev[0].r, s->opt.ic_sweeps, NULL, NULL, 0);
  else
    tri_solve(s, &s->fine_ic, s->vt, b, s->icy, x, NULL, NULL);
}

/* IC-PCG on the coarsest level, FP64, to 1e-12 with at most n_c steps
 * (C_amg1/src/amg_cycle.c:186-190, amg_pcg.c:12-101).  Host-driven loop
 * with its own scalars and partial buffers (it runs inside the outer
 * PCG's preconditioner). */
static int coarse_icpcg(struct bn_solver *s, const double *b, double *x) {
  const bn_dlevel *C = &s->lev[s->nlev - 1];
  const int n = C->n;
  double *r = s->cvec[0], *z = s->cvec[1], *p = s->cvec[2], *q = s->cvec[3];
  double *y = (double *)C->r; /* forward-sweep buffer */
  double *sc = s->scal + BN_S_COUNT;
  hipStream_t st = s->stream;
  const double *Av = C->A.vd;
  /* x is the initial guess: r = b - A x */
  bn_csr_apply('d', BN_OP_RESID, n, C->A.ia, C->A.ja, Av, x, r, b, C->A.lanes, NULL, NULL, st);
  tri_solve(s, &s->coarse_ic, 'd', r, y, z, NULL, NULL);
  bn_copy('d', n, z, p, st);
  double *pa = s->cpart1, *pb = s->cpart2;
  bn_dot_partial(n, r, z, r, r, pa, pb, st);
  bn_scalar_finish(0, pa, pb, NULL, sc, 0, st);
  int it;
  for (it = 1; it <= n; it++) {
    bn_csr_apply('d', BN_OP_SET, n, C->A.ia, C->A.ja, Av, p, q, NULL, C->A.lanes, p, pa, st);
    bn_scalar_finish(1, pa, NULL, NULL, sc, 0, st);
    bn_pcg_update_xr(n, x, p, r, q, sc, pb, st);
    tri_solve(s, &s->coarse_ic, 'd', r, y, z, NULL, NULL);
    bn_dot_partial(n, r, z, NULL, NULL, pa, NULL, st);
    bn_scalar_finish(2, pa, pb, NULL, sc, 0, st);
    hipMemcpyAsync(s->host_scal + BN_S_COUNT, sc, BN_S_COUNT * sizeof(double), hipMemcpyDeviceToHost, st);
    hipStreamSynchronize(st);
    const double rel = s->host_scal[BN_S_COUNT + BN_S_REL];
    if (rel < 1e-12 || !(rel == rel)) break;
    bn_pcg_update_p(n, p, z, sc, st);
  }
  return it;
}

/* ---------------------------------------------------------------------
 * preconditioners (input g and output z are FP64, level-0 numbering)
 * ------------------------------------------------------------------- */
/* gv, zv in the V-cycle's precision (FP32 copies when vt = 's') */
/* x_zero of the level-0 half from zero: 2 when the caller's x/r update
 * already wrote the sweep's first colour into u1 (u1_first) */
static int u1_start(struct bn_solver *s) {
  const int z = s->u1_first ? 2 : 1;
  s->u1_first = 0;
  return z;
}

static void ichol_core(struct bn_solver *s, const void *gv, void *zv) {
  const char vt = s->vt;
  const bn_dcsr *A = &s->Asolve;
  /* u1 = backslash(g) from zero */
  bn_vc_backslash(s, 0, gv, s->u1, u1_start(s));
  /* r1 = g - A u1; u2 = u1 + (L L^T)^{-1} r1 (u2 kept in icx's partner).
   * With Jacobi-swept IC and the row-block kernel for A, the residual
   * kernel also writes the lower triangle's first step r1_i / l_ii */
  const int jac = (s->opt.fine_ic == BN_IC_JACOBI);
  const int first = jac && (vt == 's' ? A->adapt_s : A->adapt) && A->blk && s->fine_ic.n == s->lev[0].n;
  if (first)
    bn_launch_resid_scaled(s, A, vt, s->u1, s->r1, gv,
                           vt == 's' ? (const void *)s->fine_ic.dinv_s : (const void *)s->fine_ic.dinv_d, s->icy);
  else
    apply(s, A, BN_OP_RESID, s->u1, s->r1, gv);
  if (jac)
    tri_jacobi(s, &s->fine_ic, vt, s->r1, s->icy, s->icx, s->lev[0].r, s->opt.ic_sweeps, s->u1, s->r2, first);
  else
    tri_solve(s, &s->fine_ic, vt, s->r1, s->icy, s->icx, s->u1, s->r2);
  /* s->r2 now holds u2; r1 <- g - A u2 */
  apply(s, A, BN_OP_RESID, s->r2, s->r1, gv);
  /* z = u2 + fwdslash(g - A u2) from zero, final update fused */
  fwdslash_rec(s, 0, s->r1, s->corr, 1, s->r2, zv);
}

static void vcycle_core(struct bn_solver *s, const void *gv, void *zv) {
  bn_vc_backslash(s, 0, gv, s->u1, u1_start(s));
  apply(s, &s->lev[0].A, BN_OP_RESID, s->u1, s->r1, gv);
  fwdslash_rec(s, 0, s->r1, s->corr, 1, s->u1, zv);
}

static void precond_core(struct bn_solver *s, const void *gv, void *zv) {
  if (s->opt.fine_ic == BN_IC_NONE) vcycle_core(s, gv, zv);
  else ichol_core(s, gv, zv);
}

/* FP64 g, z; an FP32 V-cycle works on FP32 copies */
static void precond64(struct bn_solver *s, const double *g, double *z, int ic) {
  const char vt = s->vt;
  const int n = s->lev[0].n;
  const void *gv = g;
  if (vt == 's') {
    bn_convert('d', 's', n, g, s->g_in, s->stream);
    gv = s->g_in;
  }
  void *zv = (vt == 's') ? s->zout : (void *)z;
  if (ic) ichol_core(s, gv, zv);
  else vcycle_core(s, gv, zv);
  if (vt == 's') bn_convert('s', 'd', n, s->zout, z, s->stream);
}

void bn_vc_ichol_precond(struct bn_solver *s, const double *g, double *z) { precond64(s, g, z, 1); }

void bn_vc_vcycle_precond(struct bn_solver *s, const double *g, double *z) { precond64(s, g, z, 0); }

/* FP32 V-cycle only: input already in s->g_in (FP32), result left in
 * s->zout (FP32) for consumers that widen it on the fly */
void bn_vc_precond_s(struct bn_solver *s) { precond_core(s, s->g_in, s->zout); }

void bn_vc_precond(struct bn_solver *s, const double *g, double *z) {
  if (s->opt.fine_ic == BN_IC_NONE) bn_vc_vcycle_precond(s, g, z);
  else bn_vc_ichol_precond(s, g, z);
}
