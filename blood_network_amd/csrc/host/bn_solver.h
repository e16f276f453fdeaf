/*
 * bn_solver.h - internal solver state: host hierarchy plus its device
 * mirror.
 */
#ifndef BN_SOLVER_H
#define BN_SOLVER_H

#include <hip/hip_runtime_api.h>

#include "amg.h"
#include "bn_amg_ext.h"
#include "bn_kernels.h"

#ifdef __cplusplus
extern "C" {
#endif

/* a CSR operator on the device, FP64 and/or FP32 values */
typedef struct {
  int rows, cols, nnz;
  int lanes;       /* FP64 lanes per row (autotuned at upload)          */
  int lanes_s;     /* FP32 lanes per row                                */
  int *ia, *ja;
  double *vd;
  float *vs;
  int nlong_for[7];/* rows longer than 16 * 2^k entries (lanes 2^k go */
                   /* to the block-per-row bin with that many rows)     */
  int *long_rows;  /* device: rows longer than 16 entries, longest     */
                   /* first, so every bin is a prefix                  */
  int *seg_ptr;    /* device: long row t owns segments seg_ptr[t]..    */
  int *seg_row;    /* device: segment -> long row index                */
  int *seg_ptr_h;  /* host copy of seg_ptr                             */
  int nmulti_all;  /* long rows with more than one segment (a prefix)  */
  double *seg_part;/* device: partial sums of multi-segment rows (the  */
                   /* sentinel between launches)                       */
  int *seg_cnt;    /* device: finished segments per multi-segment row  */
  int nblk;        /* row blocks for the adaptive kernel               */
  int *blk;        /* device: nblk + 1 row boundaries                  */
  int *blkq;       /* device: their entry offsets, blkq[k] = ia[blk[k]] */
  double *bpart;   /* device: nblk partials of a dot fused into the     */
                   /* row-block product                                */
  int adapt;       /* FP64 apply uses the row-block kernel (autotuned) */
  int adapt_s;     /* FP32 likewise                                    */
  int deep;        /* FP64 SpMV load-round variant (BN_LANES_DEEP* bit) */
  int deep_s;      /* FP32 likewise                                    */
} bn_dcsr;

/* Level-ordered ELL copy of one triangle for its exact solve.  Level l
 * (rows rows[ptr[l] .. ptr[l+1]) of the level schedule) holds w[l] slots
 * per row; slot s of the level's k-th row sits at q[l] + s * cnt + k (cnt:
 * the level's rows), so a wavefront's slot loads are contiguous and do
 * not wait for the row index.  Padding slots name the row itself with a
 * zero coefficient.  w[l] = 0: the level keeps the CSR rows (wider than
 * BN_ELL_MAX_W). */
typedef struct {
  int *w, *q;          /* host, per level */
  int *src, *ja;       /* device, per slot: CSR entry (-1: padding), column */
  /* one-launch solve (bn_trsv_parts) of each maximal run of levels with
   * ELL slots: run r covers levels run_lo[r] .. run_hi[r]-1 (host); the np
   * parts' steps of run r start at sp[r * (np + 1) + p] (device) in steps
   * (device, int4 per step) */
  int np, nrun;
  int *run_lo, *run_hi;
  int *sp, *steps;
  double *vd;          /* device, per slot: values gathered from the CSR copy */
  float *vs;
  double *dd;          /* device: 1/l_ii in level order */
  float *ds;
  long long slots;
} bn_tell;

/* level-scheduled triangular factor L (forward) and L^T (backward) */
typedef struct {
  int n, nlev_f, nlev_b;
  bn_dcsr lo, up;  /* L rows (cols <= i) and L^T rows (cols >= i) */
  double *dinv_d;
  float *dinv_s;
  int *rows_f, *rows_b;   /* device: rows grouped by level */
  int *ptr_f, *ptr_b;     /* host: level boundaries */
  int ell;                /* ELL copies of both triangles built */
  int ell_tried;          /* a build was attempted (not retried) */
  bn_tell ef, eb;         /* forward (L) and backward (L^T) copies */
} bn_dtri;

/* exact-order Gauss-Seidel sweep of a level: the rows of the sequential
 * sweep grouped by dependency level, forward and backward (built on first
 * use by bn_vc_prepare) */
typedef struct {
  int nf, nb;           /* levels of the forward / backward sweep */
  int *rows_f, *rows_b; /* device: rows grouped by level */
  int *ptr_f, *ptr_b;   /* host: level boundaries (NULL: not built) */
  /* one-launch sweep (bn_sweep_parts; NULL jx: a launch per level): the
   * operator's column words with bit 31 where the forward sweep visits the
   * neighbour earlier, parts, each part's steps (device: offsets np + 1 and
   * int2 steps, per direction), the sweep input's copy */
  int *jx;
  int np;
  int *sp_f, *sp_b, *st_f, *st_b;
  void *xold;
} bn_sweep;

typedef struct {
  int n, nc;
  bn_dcsr A;            /* level operator (relabelled) */
  bn_dcsr P;            /* n x nc prolongation (relabelled both sides) */
  bn_dcsr R;            /* nc x n restriction = P^T */
  double *dinv_d;       /* 1/a_ii (1.0 when the diagonal is absent) */
  float *dinv_s;
  double *l1inv_d;      /* 1/sum_j |a_ij| */
  float *l1inv_s;
  int ncol;             /* colours of the multicolour sweep */
  int *col_ptr;         /* host: colour boundaries (rows of Ac) */
  int use_gs;           /* colour count allows the multicolour sweep */
  bn_dcsr Ac;           /* GS levels: A's rows copied in colour order,
                           columns in the level numbering */
  int *cmap_d;          /* device: row r of Ac is level row cmap_d[r] */
  unsigned char *colour_d; /* device: colour of each level row (NULL above 255 colours) */
  double *dinvc_d;      /* 1/a_ii in Ac row order */
  float *dinvc_s;
  int gs_lanes, gs_lanes_s; /* lanes per row of the colour kernels, timed on
                               whole sweeps at upload (0: Ac's SpMV choice;
                               BN_GS_BLOCKS: the row-block colour kernel) */
  int gs_nblk;          /* row blocks of Ac, colour by colour */
  int *gs_blk, *gs_blkq;/* device: their row bounds and entry offsets */
  int *gs_cblk;         /* host: first block of each colour (ncol + 1) */
  /* one-launch sweep (bn_gs_walk): Ac's column words with bit 31 where the
   * neighbour's colour precedes the row's, each part's rows of each colour
   * (int2 per part and colour), parts, the sweep input's copy, and the
   * lanes per row timed fastest (FP64, FP32; 0: a launch per colour) */
  int *gs_jx, *gs_wsteps;
  int gs_wnp;
  void *gs_xold;
  int gs_walk, gs_walk_s;
  int *perm_h;          /* host perm[new] = old */
  int *perm_d;
  void *b, *x, *r;      /* level work vectors in V-cycle precision */
  /* block smoother (chunked layout) */
  int block;            /* rows chunked; smoothing by bn_block_gs */
  int nchunk;           /* chunks */
  int chunk_max;        /* rows of the largest chunk */
  int *chunk_d;         /* device: nchunk + 1 row starts */
  int *ccol_d, *cbnd_d; /* device: per-chunk local colour boundaries */
  double *bdinv_d;      /* 1 / (a_ii + sum_{j outside the chunk} |a_ij|) */
  float *bdinv_s;
  void *x2;             /* second buffer for sweeps from a nonzero start */
  bn_sweep xs[2];       /* exact sweeps: [0] natural order (SMOOTHER_GS),
                           [1] cf_order (SMOOTHER_CF_GS with gs_exact) */
} bn_dlevel;

/* device work space of the block (multi-right-hand-side) solve for K
 * interleaved columns; allocated on the first block solve with that K */
typedef struct {
  int Kcap;                        /* columns the buffers hold (0: none) */
  int K;                           /* columns of the solve in progress */
  void *b[AMG_MAX_LEVELS];         /* per level n_l x K, V-cycle precision */
  void *x[AMG_MAX_LEVELS];
  void *r[AMG_MAX_LEVELS];
  double *x0, *r0, *z0, *p0, *q0, *bb0; /* outer PCG, FP64 n x K */
  void *g_in, *u1, *r1, *icy, *icx, *r2, *corr, *zout;
  double *cin, *cout;              /* coarsest FP64 n_c x K */
  double *scal;                    /* BN_S_COUNT per column */
  double *part1, *part2, *part3;   /* BN_BLOCK_RED x K */
  double *segpart;                 /* long-row segment partials x K */
  double *host_scal;               /* pinned mirror of scal */
  hipGraph_t graph[4];
  hipGraphExec_t gexec[4];
  int graph_key[4];
  int ngraph;
} bn_block;

struct bn_solver {
  /* host side */
  dcsr_mat A_own;       /* symmetrised solve matrix (handle API)     */
  amg_data amg_own;     /* hierarchy (levels[0].A is the AMG matrix) */
  ichol_factor L_own;   /* IC(0) of the solve matrix                 */
  const dcsr_mat *A;    /* -> A_own, or a caller's matrix (cache)    */
  amg_data *amg;        /* -> amg_own, or a caller's hierarchy       */
  const ichol_factor *L;/* -> L_own, caller's factor, or NULL        */
  int borrowed;         /* host structures belong to the caller      */
  bn_options opt;
  int device;           /* HIP device ordinal */
  int dev_ready;        /* device mirror uploaded */
  int noalloc;          /* inside an allocation-free call: step graphs are
                           replayed if captured but none is captured, and
                           no smoother schedule is built */
  /* device side */
  hipStream_t stream;
  char vt;              /* V-cycle precision: 'd' or 's' */
  int nlev;
  bn_dlevel lev[AMG_MAX_LEVELS];
  bn_dcsr Asolve;       /* FP64 solve matrix in level-0 numbering */
  int asolve_shared;    /* Asolve aliases lev[0].A's FP64 arrays */
  int *tune;            /* tuned kernel choices of the last upload (6 ints
                           per operator), reused while the shapes match */
  int ntune;
  int tune_gs[AMG_MAX_LEVELS][4]; /* the levels' sweep lanes: colour launches
                                     (FP64, FP32), one-launch sweep (FP64, FP32) */
  int vperm_state;      /* caller value order -> Asolve slots: 0 not yet
                           known, 1 the same order, 2 gathered by vperm_d */
  int *vperm_d;         /* device: Asolve slot k takes caller slot vperm_d[k] */
  bn_dtri fine_ic;      /* IC(0) of the solve matrix, level-0 numbering */
  /* exact triangular solves inside bn_pcg_run as one launch per run of ELL
   * levels (bn_trsv_parts): usable, in use by the running solve, compute
   * units (the most parts), wait limit in wall_clock64() ticks */
  int walk_ok, walk_on, walk_blocks;
  long long walk_limit;
  int u1_first;         /* the next preconditioner's level-0 sweep from zero
                           finds its first colour in u1 (written by the
                           outer x/r update) */
  int jac_fuse;         /* Jacobi-smoothed levels fuse their steps into the
                           launches around them (vcycle.c; BN_FUSE_JACOBI=0
                           keeps separate launches) */
  int dot_fuse;         /* the PCG step's p.Ap rides on the row-block
                           product q = A p (BN_FUSE_DOT=0: k_dot2 after it) */
  int coarse_dense;     /* coarsest solve: 1 dense inverse, 0 IC-PCG */
  double *cinv_d;       /* dense inverse of the coarsest operator */
  float *cinv_s;
  bn_dtri coarse_ic;    /* IC(0) of the coarsest operator (IC-PCG mode) */
  double *cvec[6];      /* FP64 coarsest work vectors (IC-PCG, FP32 mode) */
  /* PCG state (FP64, level-0 numbering) */
  double *x, *r, *z, *p, *q, *bb, *rprev;
  /* preconditioner work (V-cycle precision) */
  void *g_in, *u1, *r1, *icy, *icx, *r2, *corr, *zout;
  double *scal;         /* BN_S_COUNT scalars */
  double *part1, *part2, *part3;
  double *cpart1, *cpart2;  /* partials of the coarsest IC-PCG */
  double *host_scal;    /* pinned mirror of scal */
  double *host_step;    /* coherent host memory the step's last kernel
                           copies the scalars into (NULL: a copy per step) */
  double *hbuf;         /* host staging (n doubles, pinned) */
  hipEvent_t ev0, ev1;
  /* captured PCG steps, one per mode (never re-captured in place) */
  hipGraph_t graph[8];
  hipGraphExec_t gexec[8];
  int graph_key[8];
  int ngraph;
  bn_solve_info last;
  /* device allocations owned by this solver */
  void **dptr;
  int ndptr, capdptr;
  void **hptr;          /* host tables that live as long as the mirror */
  int nhptr, caphptr;
  bn_block blk;         /* block solve work space */
  void *segs;           /* cached segmentations of segmented solves (segments.c) */
  void *rb;             /* index maps of the device value-only rebuild (rebuild.c) */
  struct bn_census *census; /* set while bn_solver_kernel_census records a solve */
};

/* gs_lanes value selecting the row-block colour kernel */
#define BN_GS_BLOCKS 1024

/* lanes per row for precision vt */
static inline int bn_lanes(const bn_dcsr *m, char vt) { return vt == 's' ? m->lanes_s : m->lanes; }

/* long-row bin for a lanes-per-row kernel: rows over 16 strides */
/* Rows longer than 16 x this width go to the long-row bin.  The default
 * kernel covers 4 load rounds of `lanes` entries per pass; the 8- and
 * 16-round variants cover 2x and 4x as many, so their bin starts 2x and
 * 4x further out (the precomputed bins reach 16 x 64 entries). */
static inline int bn_long_lanes(int lanes) {
  const int l = lanes & BN_LANES_MASK;
  const int w = (lanes & BN_LANES_DEEP16) ? 4 * l : (lanes & BN_LANES_DEEP8) ? 2 * l : l;
  return w > 64 ? 64 : w;
}
static inline int bn_long_len(int lanes) { return 16 * bn_long_lanes(lanes); }
static inline int bn_nlong(const bn_dcsr *m, int lanes) {
  lanes = bn_long_lanes(lanes);
  int k = 0;
  while ((1 << k) < lanes && k < 6) k++;
  return m->nlong_for[k];
}
/* lanes for the SpMV of M, with its tuned load-round variant */
static inline int bn_spmv_lanes(const bn_dcsr *m, char vt) {
  return vt == 's' ? (m->lanes_s | m->deep_s) : (m->lanes | m->deep);
}
/* y = op(M, x [, b]), optionally with the fused w.y partials, by the
 * tuned lanes kernel (no long-row bin: for the fine solve matrix, whose
 * rows the lanes kernel covers) */
static inline void bn_dcsr_apply_dot(const bn_dcsr *m, char vt, int op, const void *x, void *y, const void *b,
                                     const void *dotw, double *partial, hipStream_t st) {
  if (vt == 'd' && m->adapt && m->blk) {
    /* the row-block kernel was timed fastest: run it, then the dot (the
     * same deterministic block partials the fused form writes) */
    bn_csr_apply_blocks('d', op, m->nblk, m->blk, m->blkq, m->ia, m->ja, m->vd, x, y, b, st);
    if (partial) bn_dot_partial(m->rows, (const double *)dotw, (const double *)y, NULL, NULL, partial, NULL, st);
    return;
  }
  bn_csr_apply(vt, op, m->rows, m->ia, m->ja, vt == 's' ? (const void *)m->vs : (const void *)m->vd, x, y, b,
               bn_spmv_lanes(m, vt), dotw, partial, st);
}
/* y = op(M, x [, b]) with a lanes-per-row kernel and the segmented bin
 * (lanes may carry a BN_LANES_DEEP* bit) */
static inline void bn_dcsr_apply(const bn_dcsr *m, char vt, int op, int lanes, const void *x, void *y,
                                 const void *b, hipStream_t st) {
  const int nl = bn_nlong(m, lanes);
  const int nseg = nl > 0 ? m->seg_ptr_h[nl] : 0;
  bn_csr_apply_seg(vt, op, m->rows, m->ia, m->ja, vt == 's' ? (const void *)m->vs : (const void *)m->vd, x, y, b,
                   lanes, bn_long_len(lanes), nl, m->long_rows, m->seg_ptr, m->seg_row, nseg, m->seg_part,
                   m->seg_cnt, st);
}
/* y = op(M, x [, b]) with M's tuned kernel: row blocks, or lanes per row
 * with the long-row bin; with a Jacobi step fused in (xz, w, om as
 * bn_csr_apply_seg_jac; op BN_OP_JAC) */
static inline void bn_op_apply_jac(const bn_dcsr *m, char vt, int op, const void *x, void *y, const void *b,
                                   void *xz, const void *w, double om, const unsigned char *colour, int c,
                                   hipStream_t st) {
  const void *val = vt == 's' ? (const void *)m->vs : (const void *)m->vd;
  if (vt == 's' ? m->adapt_s : m->adapt) {
    bn_csr_apply_blocks_jac(vt, op, m->nblk, m->blk, m->blkq, m->ia, m->ja, val, x, y, b, xz, w, om, colour, c,
                            st);
    return;
  }
  const int lanes = bn_spmv_lanes(m, vt);
  const int nl = bn_nlong(m, lanes);
  const int nseg = nl > 0 ? m->seg_ptr_h[nl] : 0;
  bn_csr_apply_seg_jac(vt, op, m->rows, m->ia, m->ja, val, x, y, b, lanes, bn_long_len(lanes), nl, m->long_rows,
                       m->seg_ptr, m->seg_row, nseg, m->seg_part, m->seg_cnt, xz, w, om, colour, c, st);
}
static inline void bn_op_apply(const bn_dcsr *m, char vt, int op, const void *x, void *y, const void *b,
                               hipStream_t st) {
  bn_op_apply_jac(m, vt, op, x, y, b, NULL, NULL, 0.0, NULL, 0, st);
}

/* census.c: the V-cycle's and the outer PCG's bandwidth kernels, launched
 * on the solver's stream (recorded while a kernel census runs) */
void bn_launch_apply(struct bn_solver *s, const bn_dcsr *M, char vt, int op, const void *x, void *y, const void *b);
/* the same with a smoother's first step or a Jacobi step fused in
 * (bn_op_apply_jac) */
void bn_launch_apply_jac(struct bn_solver *s, const bn_dcsr *M, char vt, int op, const void *x, void *y,
                         const void *b, void *xz, const void *w, double om, const unsigned char *colour, int c);
void bn_launch_resid_scaled(struct bn_solver *s, const bn_dcsr *M, char vt, const void *x, void *y, const void *b,
                            const void *dsc, void *ys);
void bn_launch_gs(struct bn_solver *s, const bn_dlevel *d, char vt, int c, int lanes, const void *b, void *x,
                  const void *zbase, void *zsum);
/* one Jacobi sweep of triangle T: xout = (b - T_offdiag xin) dinv, or
 * zsum = zbase + that, or with xs also xs = xout dinv */
void bn_launch_jacobi(struct bn_solver *s, const bn_dcsr *T, char vt, const void *dinv, const void *b, const void *xin,
                      void *xout, const void *zbase, void *zsum, void *xs);
/* FP64 y = op(A, x [, b]) of the outer PCG, with the dotw.y partials */
/* FP64 y = op(A) x; with partial, also the dotw.y partials: returns 0 when
 * they are in partial (BN_RED_BLOCKS slots, bn_scalar_finish), or the
 * number of slots of A->bpart that hold them (the row-block product with
 * the dot fused: bn_scalar_finish_pap) */
int bn_launch_spmv(struct bn_solver *s, const bn_dcsr *A, int op, const void *x, void *y, const void *b,
                   const void *dotw, double *partial);
int bn_ensure_device(struct bn_solver *s);

/* segments.c */
void bn_segments_release(struct bn_solver *s);
void bn_drop_graphs(struct bn_solver *s);
/* extras.c: caller value order -> stored solve-matrix slots (vperm_state) */
int bn_value_map(struct bn_solver *s);
/* rebuild.c: frees the device rebuild plan's host side */
void bn_rebuild_release(struct bn_solver *s);
int bn_coarse_inverse_device(struct bn_solver *s, const dcsr_mat *A, double *cinv_d, float *cinv_s, int *clamped);

/* device.c */
int bn_device_upload(struct bn_solver *s);
/* host level sets of the sequential sweep of A in `order` (NULL: natural)
 * and their device row lists; 0 or a BN_ERR_* code */
int bn_sweep_upload(struct bn_solver *s, const dcsr_mat *A, const int *order, bn_sweep *w);
void bn_sweep_free_host(bn_sweep *w);
/* default lanes per row for an operator of that size */
int bn_lanes_for(long long nnz, long long rows);
void bn_device_release(struct bn_solver *s);
int bn_device_ok(void);

/* vcycle.c */
/* Builds what the current smoother needs before any V-cycle is enqueued
 * or captured (the exact sweeps' schedules); 0 or a BN_ERR_* code. */
int bn_vc_prepare(struct bn_solver *s);
/* the ELL copies of a triangle (device.c): built from the device CSR
 * copies, values refreshed from them after a value change */
int bn_tri_ell_build(struct bn_solver *s, bn_dtri *t);
void bn_tri_ell_refresh(struct bn_solver *s, bn_dtri *t);
/* -1: the smoother runs no exact-order sweep, else the bn_sweep slot */
int bn_vc_exact_kind(const struct bn_solver *s);
/* one colour c of level d's multicolour sweep with the given colour-kernel
 * choice (lanes per row, or BN_GS_BLOCKS) */
void bn_gs_colour(struct bn_solver *s, const bn_dlevel *d, char vt, int c, int lanes, const void *b, void *x,
                  const void *zbase, void *zsum);
void bn_vc_backslash(struct bn_solver *s, int lev, const void *b, void *x, int x_zero);
void bn_vc_fwdslash(struct bn_solver *s, int lev, const void *b, void *x, int x_zero);
void bn_vc_ichol_precond(struct bn_solver *s, const double *g, double *z);
/* the outer PCG's preconditioner: AMG + fine IC, or the plain V-cycle when
 * fine_ic == BN_IC_NONE */
void bn_vc_precond(struct bn_solver *s, const double *g, double *z);
void bn_vc_vcycle_precond(struct bn_solver *s, const double *g, double *z);
/* FP32 V-cycle: FP32 input in s->g_in, FP32 result left in s->zout */
void bn_vc_precond_s(struct bn_solver *s);
/* level 0's sweep from zero starts with the first-colour kernel that the
 * outer x/r update can write (bn_pcg_update_xr_s0, then u1_first = 1) */
int bn_vc_first_colour_fused(const struct bn_solver *s);
void bn_vc_fine_ic(struct bn_solver *s, const void *b, void *x);

/* pcg.c */
int bn_pcg_run(struct bn_solver *s, const double *d_b, double *d_x, double tol, int maxiter,
               int print_level, bn_solve_info *info);

#ifdef __cplusplus
}
#endif
#endif
