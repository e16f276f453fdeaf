/*
 * rebuild.c - value-only rebuild of the hierarchy on the device.
 *
 * amg_solver_update (C_amg1/src/amg_wrapper.c:114-149) rebuilds the
 * values of a hierarchy on its stored C/F split and patterns on the host
 * (C_amg1/src/build_hierarchy.c:170-242) and uploads it again.  Here the
 * new fine values come from device memory and the level operators and
 * interpolation weights are recomputed by the kernels of rebuild.hip, in
 * the host's order and without contraction, so they are bit-identical to
 * the host rebuild.  Every derived device array (FP32 copies, diagonal and
 * l1 inverses, the colour-ordered GS rows, R = P^T) is refreshed in place
 * through index maps built once from the host layout, so captured graphs
 * stay valid.  What stays on the host: the fine IC(0) and the coarsest
 * solve (dense inverse or IC(0)), whose inputs are downloaded; the host
 * hierarchy is brought up to date as well, so a later re-upload gives the
 * same state.
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
int bn_dense_inverse(const dcsr_mat *A, double *inv);

typedef struct {
  int *s_ia, *s_col, *s_slot, *cpt; /* strength rows sorted, P slot of each strong C point */
  int *rmap;                        /* R entry -> P entry */
  int *c_col, *c_pos, *tpos;        /* sorted row copy and transpose positions (levels >= 1) */
  int *acmap;                       /* colour-ordered GS row entry -> level entry */
} rb_level;

typedef struct {
  rb_level lev[AMG_MAX_LEVELS];
  int *tpos0;    /* fine level: transpose positions (input symmetrisation) */
  double *work;  /* product values / gathered input, max nnz over levels */
  int maxnnz;
} rb_plan;

void bn_rebuild_release(struct bn_solver *s) {
  free(s->rb); /* the device arrays go with the solver's allocations */
  s->rb = NULL;
}

static void *rb_dev(struct bn_solver *s, size_t bytes) {
  void *p = NULL;
  if (hipMalloc(&p, bytes ? bytes : 8) != hipSuccess) return NULL;
  bn_track_device(s, p);
  return p;
}

static int *rb_upload_int(struct bn_solver *s, const int *h, size_t n) {
  int *d = (int *)rb_dev(s, n * sizeof(int));
  if (d && n && hipMemcpy(d, h, n * sizeof(int), hipMemcpyHostToDevice) != hipSuccess) return NULL;
  return d;
}

typedef struct {
  int col, val;
} pair_t;

static int cmp_pair(const void *a, const void *b) {
  const pair_t *x = (const pair_t *)a, *y = (const pair_t *)b;
  return (x->col > y->col) - (x->col < y->col);
}

static void sort_row(pair_t *p, int len) {
  if (len > 24) {
    qsort(p, (size_t)len, sizeof(pair_t), cmp_pair);
    return;
  }
  for (int i = 1; i < len; i++) {
    pair_t t = p[i];
    int j = i - 1;
    while (j >= 0 && p[j].col > t.col) p[j + 1] = p[j], j--;
    p[j + 1] = t;
  }
}

/* per-row sorted copy of a CSR pattern: col[] sorted within each row, pos[]
 * the original position of each entry */
static void sorted_rows(const dcsr_mat *A, int *col, int *pos) {
#pragma omp parallel for schedule(dynamic, 1024)
  for (int i = 0; i < A->row; i++) {
    const int lo = A->ia[i], len = A->ia[i + 1] - lo;
    pair_t *p = (pair_t *)malloc((size_t)(len > 0 ? len : 1) * sizeof(pair_t));
    for (int k = 0; k < len; k++) p[k].col = A->ja[lo + k], p[k].val = lo + k;
    sort_row(p, len);
    for (int k = 0; k < len; k++) col[lo + k] = p[k].col, pos[lo + k] = p[k].val;
    free(p);
  }
}

static int find_sorted(const int *col, int lo, int hi, int c) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (col[mid] == c) return mid;
    if (col[mid] < c) lo = mid + 1;
    else hi = mid;
  }
  return -1;
}

/* position of entry (j, i) for every entry (i, j); -1 if a matrix is not
 * structurally symmetric */
static int transpose_positions(const dcsr_mat *A, const int *col, const int *pos, int *tp) {
  int bad = 0;
#pragma omp parallel for schedule(static) reduction(| : bad)
  for (int i = 0; i < A->row; i++)
    for (int q = A->ia[i]; q < A->ia[i + 1]; q++) {
      const int j = A->ja[q];
      const int at = find_sorted(col, A->ia[j], A->ia[j + 1], i);
      if (at < 0) bad = 1;
      else tp[q] = pos[at];
    }
  return bad ? -1 : 0;
}

static int plan_build(struct bn_solver *s) {
  rb_plan *pl = (rb_plan *)calloc(1, sizeof(rb_plan));
  const amg_data *amg = s->amg;
  const int nlev = s->nlev;
  int rc = 0;
  for (int l = 0; l < nlev && !rc; l++) {
    const amg_level *L = &amg->levels[l];
    const dcsr_mat *A = &L->A;
    rb_level *r = &pl->lev[l];
    if (A->nnz > pl->maxnnz) pl->maxnnz = A->nnz;
    int *col = (int *)bn_xmalloc((size_t)(A->nnz > 0 ? A->nnz : 1) * sizeof(int));
    int *pos = (int *)bn_xmalloc((size_t)(A->nnz > 0 ? A->nnz : 1) * sizeof(int));
    int *tp = (int *)bn_xmalloc((size_t)(A->nnz > 0 ? A->nnz : 1) * sizeof(int));
    sorted_rows(A, col, pos);
    if (transpose_positions(A, col, pos, tp)) {
      bn_set_error("level %d operator is not structurally symmetric", l);
      rc = BN_ERR_ARG;
    }
    if (!rc && l == 0) pl->tpos0 = rb_upload_int(s, tp, (size_t)A->nnz);
    if (!rc && l > 0) {
      r->c_col = rb_upload_int(s, col, (size_t)A->nnz);
      r->c_pos = rb_upload_int(s, pos, (size_t)A->nnz);
      r->tpos = rb_upload_int(s, tp, (size_t)A->nnz);
    }
    free(col);
    free(pos);
    free(tp);
    /* colour-ordered GS rows: relabel a matrix whose values are the slot
     * numbers exactly as the upload relabelled the values */
    const bn_dlevel *d = &s->lev[l];
    if (!rc && d->use_gs) {
      int *cperm = (int *)bn_xmalloc((size_t)(d->n > 0 ? d->n : 1) * sizeof(int));
      if (hipMemcpy(cperm, d->cmap_d, (size_t)d->n * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) rc = BN_ERR_HIP;
      dcsr_mat I = *A, Ic;
      I.val = (double *)bn_xmalloc((size_t)(A->nnz > 0 ? A->nnz : 1) * sizeof(double));
      for (int k = 0; k < A->nnz; k++) I.val[k] = (double)k;
      bn_csr_permute(&I, cperm, NULL, &Ic);
      int *m = (int *)bn_xmalloc((size_t)(A->nnz > 0 ? A->nnz : 1) * sizeof(int));
      for (int k = 0; k < Ic.nnz; k++) m[k] = (int)Ic.val[k];
      if (!rc) r->acmap = rb_upload_int(s, m, (size_t)Ic.nnz);
      free(m);
      bn_csr_free(&Ic);
      free(I.val);
      free(cperm);
    }
    if (rc || l == nlev - 1) continue;
    /* strength rows sorted, with the P slot of each strong C point (slots
     * in the order the interpolation's sparsity pass assigns them) */
    const dcsr_mat *S = &L->A_filtered;
    const dcsr_mat *P = &L->P;
    int *scol = (int *)bn_xmalloc((size_t)(S->nnz > 0 ? S->nnz : 1) * sizeof(int));
    int *sslot = (int *)bn_xmalloc((size_t)(S->nnz > 0 ? S->nnz : 1) * sizeof(int));
    int *cpt = (int *)bn_xmalloc((size_t)(L->n > 0 ? L->n : 1) * sizeof(int));
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 1024) reduction(| : bad)
    for (int i = 0; i < S->row; i++) {
      cpt[i] = L->mis[i] ? 1 : 0;
      const int lo = S->ia[i], len = S->ia[i + 1] - lo;
      pair_t *pp = (pair_t *)malloc((size_t)(len > 0 ? len : 1) * sizeof(pair_t));
      int slot = P->ia[i];
      for (int k = 0; k < len; k++) {
        const int j = S->ja[lo + k];
        pp[k].col = j;
        pp[k].val = L->mis[j] ? slot++ : -1;
      }
      if (!L->mis[i] && slot != P->ia[i + 1]) bad = 1;
      sort_row(pp, len);
      for (int k = 0; k < len; k++) scol[lo + k] = pp[k].col, sslot[lo + k] = pp[k].val;
      free(pp);
    }
    if (bad) {
      bn_set_error("level %d: interpolation pattern does not match the strong C points", l);
      rc = BN_ERR_ARG;
    }
    if (!rc) {
      r->s_ia = rb_upload_int(s, S->ia, (size_t)S->row + 1);
      r->s_col = rb_upload_int(s, scol, (size_t)S->nnz);
      r->s_slot = rb_upload_int(s, sslot, (size_t)S->nnz);
      r->cpt = rb_upload_int(s, cpt, (size_t)L->n);
    }
    free(scol);
    free(sslot);
    free(cpt);
    /* R = P^T as the upload built it (stable transpose): entry -> P entry */
    if (!rc) {
      dcsr_mat Pi = *P, Ri;
      Pi.val = (double *)bn_xmalloc((size_t)(P->nnz > 0 ? P->nnz : 1) * sizeof(double));
      for (int k = 0; k < P->nnz; k++) Pi.val[k] = (double)k;
      bn_csr_transpose(&Pi, &Ri);
      int *m = (int *)bn_xmalloc((size_t)(Ri.nnz > 0 ? Ri.nnz : 1) * sizeof(int));
      for (int k = 0; k < Ri.nnz; k++) m[k] = (int)Ri.val[k];
      r->rmap = rb_upload_int(s, m, (size_t)Ri.nnz);
      free(m);
      bn_csr_free(&Ri);
      free(Pi.val);
    }
  }
  if (!rc) pl->work = (double *)rb_dev(s, (size_t)(pl->maxnnz > 0 ? pl->maxnnz : 1) * sizeof(double));
  for (int l = 0; l < nlev && !rc; l++) {
    const rb_level *r = &pl->lev[l];
    const int last = (l == nlev - 1);
    if ((l == 0 && !pl->tpos0) || (l > 0 && (!r->c_col || !r->c_pos || !r->tpos)) ||
        (s->lev[l].use_gs && !r->acmap) ||
        (!last && (!r->s_ia || !r->s_col || !r->s_slot || !r->cpt || !r->rmap)))
      rc = BN_ERR_HIP;
  }
  if (!rc && !pl->work) rc = BN_ERR_HIP;
  if (rc) {
    free(pl);
    if (rc == BN_ERR_HIP) bn_set_error("device memory for the rebuild plan");
    return rc;
  }
  s->rb = pl;
  return 0;
}

/* host factor values of an IC(0) factor into its device copy (lower rows,
 * their transpose and the inverse diagonal, as the upload forms them) */
static int refresh_tri(struct bn_solver *s, const ichol_factor *F, bn_dtri *t) {
  const int n = F->n;
  dcsr_mat Lh = {n, n, F->nnz, F->ia, F->ja, F->val}, Uh;
  bn_csr_transpose(&Lh, &Uh);
  double *dinv = (double *)bn_xmalloc((size_t)(n > 0 ? n : 1) * sizeof(double));
  for (int i = 0; i < n; i++) {
    double d = 1.0;
    for (int k = Lh.ia[i]; k < Lh.ia[i + 1]; k++)
      if (Lh.ja[k] == i) {
        d = Lh.val[k];
        break;
      }
    dinv[i] = 1.0 / d;
  }
  const size_t nz = (size_t)F->nnz;
  float *fs = (float *)bn_xmalloc((nz > (size_t)n ? nz : (size_t)n) * sizeof(float) + 4);
  hipError_t e = hipMemcpy(t->lo.vd, Lh.val, nz * sizeof(double), hipMemcpyHostToDevice);
  for (size_t k = 0; k < nz; k++) fs[k] = (float)Lh.val[k];
  if (e == hipSuccess) e = hipMemcpy(t->lo.vs, fs, nz * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(t->up.vd, Uh.val, nz * sizeof(double), hipMemcpyHostToDevice);
  for (size_t k = 0; k < nz; k++) fs[k] = (float)Uh.val[k];
  if (e == hipSuccess) e = hipMemcpy(t->up.vs, fs, nz * sizeof(float), hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(t->dinv_d, dinv, (size_t)n * sizeof(double), hipMemcpyHostToDevice);
  for (int i = 0; i < n; i++) fs[i] = (float)dinv[i];
  if (e == hipSuccess) e = hipMemcpy(t->dinv_s, fs, (size_t)n * sizeof(float), hipMemcpyHostToDevice);
  free(fs);
  free(dinv);
  bn_csr_free(&Uh);
  return e == hipSuccess ? 0 : BN_ERR_HIP;
}

/* New values of the solve matrix (device memory, nnz entries in the
 * stored order, as bn_solver_set_values_device takes them), and a
 * value-only rebuild of the whole hierarchy from them on the device: the
 * same split and patterns, interpolation weights and Galerkin operators
 * recomputed bit for bit as amg_rebuild_values would, fine IC(0) and
 * coarsest solve refreshed. */
int bn_solver_rebuild_values_device(void *h, int nnz_in, const double *d_val) {
  struct bn_solver *s = (struct bn_solver *)h;
  if (!s || !d_val || s->borrowed) return BN_ERR_ARG;
  if (nnz_in != s->A->nnz) {
    bn_set_error("rebuild: %d values given, the stored solve matrix has %d", nnz_in, s->A->nnz);
    return BN_ERR_ARG;
  }
  int rc = bn_ensure_device(s);
  if (rc) return rc;
  const int nlev = s->nlev;
  amg_data *amg = s->amg;
  for (int l = 0; l < nlev; l++)
    if (s->lev[l].block) {
      bn_set_error("rebuild on the device does not support the block smoother layout");
      return BN_ERR_ARG;
    }
  if (amg->levels[0].A.nnz != s->A->nnz ||
      (!s->asolve_shared && (s->Asolve.ia != s->lev[0].A.ia || s->Asolve.ja != s->lev[0].A.ja))) {
    bn_set_error("rebuild: the solver holds a solve matrix apart from its hierarchy");
    return BN_ERR_ARG;
  }
  if (!s->vperm_state && (rc = bn_value_map(s))) return rc;
  if (!s->rb && (rc = plan_build(s))) return rc;
  rb_plan *pl = (rb_plan *)s->rb;
  hipStream_t st = s->stream;
  const int n0 = s->lev[0].n, nz0 = s->lev[0].A.nnz;

  /* fine values in stored order, symmetrised as the handle stores them */
  const double *in = d_val;
  if (s->vperm_state == 2) {
    bn_permute(nz0, s->vperm_d, d_val, pl->work, 0, st);
    in = pl->work;
  }
  if (s->opt.symmetrize) {
    if (in == pl->work) {
      /* symmetrise from a copy: the product scratch doubles as input */
      if (hipMemcpyAsync(s->lev[0].A.vd, in, (size_t)nz0 * sizeof(double), hipMemcpyDeviceToDevice, st) !=
          hipSuccess)
        return BN_ERR_HIP;
      bn_rb_symmetrize(nz0, pl->tpos0, s->lev[0].A.vd, pl->work, st);
      if (hipMemcpyAsync(s->lev[0].A.vd, pl->work, (size_t)nz0 * sizeof(double), hipMemcpyDeviceToDevice, st) !=
          hipSuccess)
        return BN_ERR_HIP;
    } else {
      bn_rb_symmetrize(nz0, pl->tpos0, in, s->lev[0].A.vd, st);
    }
  } else if (hipMemcpyAsync(s->lev[0].A.vd, in, (size_t)nz0 * sizeof(double), hipMemcpyDeviceToDevice, st) !=
             hipSuccess) {
    return BN_ERR_HIP;
  }

  /* level by level: interpolation weights, R = P^T, Galerkin product,
   * symmetrisation */
  for (int l = 0; l < nlev - 1; l++) {
    bn_dlevel *d = &s->lev[l], *c = &s->lev[l + 1];
    const rb_level *r = &pl->lev[l], *rc1 = &pl->lev[l + 1];
    bn_rb_interp(d->n, d->A.ia, d->A.ja, d->A.vd, r->s_ia, r->s_col, r->s_slot, r->cpt, d->P.ia, d->P.vd, st);
    bn_rb_to_float(d->P.nnz, d->P.vd, d->P.vs, st);
    bn_rb_gather(d->R.nnz, r->rmap, d->P.vd, d->R.vd, d->R.vs, st);
    bn_rb_rap(c->n, d->R.ia, d->R.ja, d->R.vd, d->A.ia, d->A.ja, d->A.vd, d->P.ia, d->P.ja, d->P.vd, c->A.ia,
              rc1->c_col, rc1->c_pos, pl->work, st);
    bn_rb_symmetrize(c->A.nnz, rc1->tpos, pl->work, c->A.vd, st);
  }
  /* derived per-level arrays */
  for (int l = 0; l < nlev; l++) {
    bn_dlevel *d = &s->lev[l];
    bn_rb_to_float(d->A.nnz, d->A.vd, d->A.vs, st);
    bn_rb_row_scalars(d->n, d->A.ia, d->A.ja, d->A.vd, d->dinv_d, d->dinv_s, d->l1inv_d, d->l1inv_s, NULL, st);
    if (d->use_gs) {
      bn_rb_gather(d->Ac.nnz, pl->lev[l].acmap, d->A.vd, d->Ac.vd, d->Ac.vs, st);
      bn_rb_dinv_map(d->n, d->Ac.ia, d->Ac.ja, d->Ac.vd, d->cmap_d, d->dinvc_d, d->dinvc_s, st);
    }
  }
  if (!s->asolve_shared &&
      hipMemcpyAsync(s->Asolve.vd, s->lev[0].A.vd, (size_t)nz0 * sizeof(double), hipMemcpyDeviceToDevice, st) !=
          hipSuccess)
    return BN_ERR_HIP;
  if (hipStreamSynchronize(st) != hipSuccess) {
    bn_set_error("device failure in the value-only rebuild");
    return BN_ERR_HIP;
  }

  /* the host hierarchy follows (level and interpolation values, l1 norms) */
  hipError_t e = hipSuccess;
  for (int l = 0; l < nlev && e == hipSuccess; l++) {
    amg_level *L = &amg->levels[l];
    e = hipMemcpy(L->A.val, s->lev[l].A.vd, (size_t)L->A.nnz * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess && l < nlev - 1)
      e = hipMemcpy(L->P.val, s->lev[l].P.vd, (size_t)L->P.nnz * sizeof(double), hipMemcpyDeviceToHost);
    if (L->l1_diag) bn_csr_l1(&L->A, L->l1_diag);
  }
  if (e != hipSuccess) return BN_ERR_HIP;
  memcpy(s->A_own.val, amg->levels[0].A.val, (size_t)nz0 * sizeof(double));
  (void)n0;

  /* fine IC(0) and the coarsest solve, as amg_solver_update forms them */
  ichol_free(&s->L_own);
  ichol_compute(s->A, &s->L_own);
  if (s->fine_ic.lo.ia && (rc = refresh_tri(s, &s->L_own, &s->fine_ic))) return rc;
  amg_level *C = &amg->levels[nlev - 1];
  ichol_free(&C->L_ichol);
  ichol_compute(&C->A, &C->L_ichol);
  if (s->coarse_dense) {
    const int nc = C->A.row;
    double *inv = (double *)bn_xmalloc((size_t)nc * (size_t)nc * sizeof(double));
    float *invs = (float *)bn_xmalloc((size_t)nc * (size_t)nc * sizeof(float));
    bn_dense_inverse(&C->A, inv);
    for (size_t k = 0; k < (size_t)nc * (size_t)nc; k++) invs[k] = (float)inv[k];
    e = hipMemcpy(s->cinv_d, inv, (size_t)nc * (size_t)nc * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(s->cinv_s, invs, (size_t)nc * (size_t)nc * sizeof(float), hipMemcpyHostToDevice);
    free(inv);
    free(invs);
    if (e != hipSuccess) return BN_ERR_HIP;
  } else if ((rc = refresh_tri(s, &C->L_ichol, &s->coarse_ic))) {
    return rc;
  }
  return 0;
}
