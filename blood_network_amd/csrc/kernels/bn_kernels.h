/*
 * bn_kernels.h - C launchers for the gfx950 HIP kernels (internal ABI
 * between the C host code and kernels.hip).  All pointers are device
 * pointers; every launcher enqueues on `stream` and returns at once.
 * Value type codes: 'd' = FP64 storage and arithmetic, 's' = FP32.
 */
#ifndef BN_KERNELS_H
#define BN_KERNELS_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

/* scalar slots in the per-solver device scalar block */
enum {
  BN_S_RZ = 0,     /* r^T z of the current iteration            */
  BN_S_RZ0 = 1,    /* r0^T z0                                   */
  BN_S_PAP = 2,    /* p^T A p                                   */
  BN_S_ALPHA = 3,
  BN_S_BETA = 4,
  BN_S_RR = 5,     /* r^T r                                     */
  BN_S_RZNEW = 6,
  BN_S_REL = 7,    /* sqrt(rz_new / rz0)                        */
  BN_S_BB = 8,     /* b^T b                                     */
  BN_S_RZOLD = 9,  /* previous r^T z (flexible beta)            */
  BN_S_ZRPREV = 10,/* z^T r_prev (flexible beta)                */
  BN_S_DONE = 11,  /* block solve: column converged (frozen)    */
  BN_S_WALKERR = 11,/* single-vector solve: 1 when a one-launch
                       triangular solve gave up (bn_trsv_parts)   */
  BN_S_ITS = 12,   /* block solve: iteration it converged at    */
  BN_S_ITER = 13,  /* block solve: iterations done              */
  BN_S_TOL = 14,   /* block solve: stop tolerance               */
  BN_S_STOPREL = 15,/* block solve: 1 = ||r||/||b|| stop test   */
  BN_S_COUNT = 16
};

/* reduction partial buffer holds at most this many blocks per dot */
#define BN_RED_BLOCKS 2048

/* Static device scratch of the allocation-free entry points (noalloc.hip):
 * BN_SCRATCH_EDGES segments staged per pass (4 + 4 + 8 bytes in, 8 + 8
 * out) plus BN_SCRATCH_DIRICHLET Dirichlet pressures; one copy per GPU. */
#define BN_SCRATCH_EDGES (1 << 20)
#define BN_SCRATCH_DIRICHLET (1 << 16)
#define BN_SCRATCH_BYTES ((size_t)BN_SCRATCH_EDGES * 32 + (size_t)BN_SCRATCH_DIRICHLET * 8)
#define BN_SCRATCH_MAX_DEVICES 64
/* this device's copy of the scratch (resolved once per device, the
 * device current); 0 or a BN_ERR_* code */
int bn_scratch_ptr(void **ptr, size_t *bytes);

/* Set-up helpers (devsort.hip), synchronous on `st`, 0 or -1: exclusive
 * prefix sums of n int32 / int64 values, and a stable radix sort of n
 * (key, value) pairs by the keys' low end_bit bits (higher bits ignored);
 * inputs and outputs must not overlap. */
int bn_dev_scan_i32(const int *in, int *out, long long n, hipStream_t st);
int bn_dev_scan_i64(const long long *in, long long *out, long long n, hipStream_t st);
int bn_dev_sort_pairs(const unsigned long long *kin, unsigned long long *kout, const int *vin, int *vout, int n,
                      int end_bit, hipStream_t st);

/* Arms a kernel timer: the calling thread's next launch from kernels.hip
 * records its start and end into the two events (the dispatch's own
 * timestamps); NULL, NULL disarms it. */
void bn_time_next_launch(hipEvent_t start, hipEvent_t stop);

/* Lanes-per-row SpMV variants: the lanes argument of bn_csr_apply and
 * bn_csr_apply_seg may carry one of these bits.  Each lane then issues 8
 * (DEEP8) or 16 (DEEP16) index and value loads before its x gathers
 * instead of 4 (8 for 2 lanes): fewer dependent memory round trips per
 * row. */
#define BN_LANES_DEEP8 0x100
#define BN_LANES_DEEP16 0x200
#define BN_LANES_MASK 0xff

/* CSR operator kinds for bn_csr_apply */
enum {
  BN_OP_SET = 0,    /* y = A x                 */
  BN_OP_RESID = 1,  /* y = b - A x             */
  BN_OP_ADD = 2,    /* y = y + A x             */
  BN_OP_JAC = 3     /* y = x + om w (b - A x): one Jacobi step (the *_jac entry points) */
};

/* y = op(A, x [, b]); lanes = threads per row (1..64, power of two).
 * If partial != NULL also writes per-block partial sums of w_i * y_i
 * (w = dotw) into partial[0..BN_RED_BLOCKS). */
void bn_csr_apply(char vt, int op, int n, const int *ia, const int *ja,
                  const void *val, const void *x, void *y, const void *b,
                  int lanes, const void *dotw, double *partial,
                  hipStream_t stream);

/* Same, split by row length: rows longer than long_len (listed in
 * long_rows, nlong of them) are computed by a block-per-row kernel, the
 * rest by the lanes-per-row kernel.  No fused dot. */
void bn_csr_apply_binned(char vt, int op, int n, const int *ia, const int *ja,
                         const void *val, const void *x, void *y, const void *b,
                         int lanes, int long_len, int nlong, const int *long_rows,
                         hipStream_t stream);

/* Same split, long rows cut into segments of BN_SEG_NNZ entries, one
 * block per segment: long row t = long_rows[t] owns segments
 * seg_ptr[t] .. seg_ptr[t+1]-1 (seg_row[s] = t); nseg = seg_ptr[nlong].
 * The rows with more than one segment (a prefix of long_rows) pass their
 * partial sums through part[seg_ptr[t] ..] and count their segments in
 * cnt[t]; the last segment of a row to finish adds the partials in
 * segment order.  part's slots must hold the walk's sentinel and cnt
 * zeros (bn_seg_reset) before the first call; every call leaves them so.
 * Calls that share part and cnt must not overlap. */
#define BN_SEG_NNZ 1024
void bn_csr_apply_seg(char vt, int op, int n, const int *ia, const int *ja,
                      const void *val, const void *x, void *y, const void *b,
                      int lanes, int long_len, int nlong, const int *long_rows,
                      const int *seg_ptr, const int *seg_row, int nseg,
                      double *part, int *cnt, hipStream_t stream);
/* The same with a smoother's first step fused in (op BN_OP_SET with xz !=
 * NULL: also xz_r = om y_r w_r, bn_jacobi_zero's x for right-hand side y,
 * and with colour != NULL that only where colour_r == c and zero elsewhere,
 * bn_gs_color_zero_colour's x when om = 1 and w = 1/a_ii), or a Jacobi step
 * (op BN_OP_JAC: y_r = x_r + om (b_r - (A x)_r) w_r, what BN_OP_RESID into
 * r followed by bn_jacobi_update(r, w, om, x) leave in x; y must not be x),
 * to the last bit. */
void bn_csr_apply_seg_jac(char vt, int op, int n, const int *ia, const int *ja,
                          const void *val, const void *x, void *y, const void *b,
                          int lanes, int long_len, int nlong, const int *long_rows,
                          const int *seg_ptr, const int *seg_row, int nseg,
                          double *part, int *cnt, void *xz, const void *w, double om,
                          const unsigned char *colour, int c, hipStream_t stream);
/* part[0..nslot) = the sentinel, cnt[0..nmulti) = 0 */
void bn_seg_reset(int nslot, double *part, int nmulti, int *cnt, hipStream_t stream);

/* Row-block split for bn_csr_apply_blocks: consecutive rows grouped so a
 * block holds at most BN_ROWBLOCK_NNZ entries and BN_ROWBLOCK_ROWS rows; a
 * row longer than BN_ROWBLOCK_NNZ is a block of its own. */
#define BN_ROWBLOCK_NNZ 2048
#define BN_ROWBLOCK_ROWS 256

/* Same operator, one 256-thread block per row block (blocks[0..nblk],
 * row indices): the block's entries are streamed with coalesced loads,
 * their products staged in LDS and summed per row by a power-of-two
 * group of threads; a single over-long row is summed by the whole block.
 * Load balance is independent of the row-length distribution. */
/* blkq (may be NULL) holds each block boundary's entry offset,
 * blkq[k] = ia[blocks[k]], which saves a block one dependent load. */
void bn_csr_apply_blocks(char vt, int op, int nblk, const int *blocks, const int *blkq,
                         const int *ia, const int *ja, const void *val,
                         const void *x, void *y, const void *b,
                         hipStream_t stream);
/* FP64 y = A x with the w.y dot fused: row block k also writes
 * partial[k] = sum of w_r y_r over its rows, so partial holds nblk
 * slots; bn_scalar_finish_pap sums them. */
void bn_csr_apply_blocks_dot(int nblk, const int *blocks, const int *blkq, const int *ia, const int *ja,
                             const double *val, const double *x, double *y, const double *w, double *partial,
                             hipStream_t stream);
/* The same with a Jacobi step fused in, as bn_csr_apply_seg_jac. */
void bn_csr_apply_blocks_jac(char vt, int op, int nblk, const int *blocks, const int *blkq,
                             const int *ia, const int *ja, const void *val,
                             const void *x, void *y, const void *b, void *xz, const void *w, double om,
                             const unsigned char *colour, int c, hipStream_t stream);
/* y = b - A x by the row-block kernel, also writing ys_r = y_r dsc_r */
void bn_csr_resid_blocks_scaled(char vt, int nblk, const int *blocks, const int *blkq, const int *ia,
                                const int *ja, const void *val, const void *x, void *y, const void *b,
                                const void *dsc, void *ys, hipStream_t stream);

/* One multicolour Gauss-Seidel colour: rows lo..hi-1 of a level whose
 * rows are stored in colour order:
 *   x_i <- (b_i - sum_{j != i} a_ij x_j) * dinv_i.
 * If zsum != NULL, also writes z_i = zbase_i + x_i (fused final update). */
void bn_gs_color(char vt, int lo, int hi, const int *ia, const int *ja,
                 const void *val, const void *dinv, const void *b, void *x,
                 int lanes, const void *zbase, void *zsum,
                 hipStream_t stream);
/* One block Gauss-Seidel sweep (l1 hybrid) over a chunked level: chunk
 * k = rows [chunk[k], chunk[k+1]), its local colours delimited by
 * cbnd[ccol[k] .. ccol[k+1]-1]; xold = NULL means a zero start.  xnew
 * must not alias xold.  max_rows (largest chunk) sizes the LDS copy of
 * the chunk's iterate and must be <= 8192.  zsum: as bn_gs_color. */
void bn_block_gs(char vt, int nchunk, int max_rows, const int *chunk, const int *ccol,
                 const int *cbnd, const int *ia, const int *ja, const void *val,
                 const void *dinv, const void *b, const void *xold, void *xnew,
                 int backward, int lanes, const void *zbase, void *zsum,
                 hipStream_t stream);
#define BN_CHUNK_MAX_ROWS 8192

/* The same colour with the level's rows copied in colour order (row r of
 * ia/ja/val, columns in natural numbering) updating natural entry
 * i = map[r]: x_i <- (b_i - sum_{j != i} a_rj x_j) * dinv_r, dinv in
 * colour order.  Vectors stay in natural order. */
/* One multicolour Gauss-Seidel colour by row blocks: blocks b0 .. b0+nb-1
 * of a row-block split of the colour-ordered copy (blocks never straddle
 * two colours; blkq their entry offsets) hold that colour's rows; the copy
 * holds off-diagonal entries only, row r updating x[map[r]]:
 * x_i = (b_i - sum_j a_rj x_j) dinv_r.  zsum != NULL: also zsum_i = zbase_i
 * + x_i. */
void bn_gs_color_blocks(char vt, int b0, int nb, const int *blocks, const int *blkq, const int *ia, const int *ja,
                        const void *val, const int *map, const void *dinv, const void *b, void *x, const void *zbase,
                        void *zsum, hipStream_t stream);
void bn_gs_color_map(char vt, int lo, int hi, const int *map, const int *ia, const int *ja,
                     const void *val, const void *dinv, const void *b, void *x,
                     int lanes, const void *zbase, void *zsum, hipStream_t stream);
/* first colour from x = 0 in that layout: x_map[r] = b_map[r] * dinv_r
 * for lo <= r < hi, and x_map[r] = 0 for hi <= r < all_n (the rows of
 * the later colours), in one pass */
void bn_gs_color_zero_map(char vt, int lo, int hi, int all_n, const int *map, const void *dinv,
                          const void *b, void *x, hipStream_t stream);
/* the same from level-row order: x_i = b_i dinv_i on rows whose colour[i]
 * is c, 0 elsewhere (dinv in level-row order) */
void bn_gs_color_zero_colour(char vt, int n, int c, const unsigned char *colour, const void *dinv, const void *b,
                             void *x, hipStream_t stream);

/* first colour of a sweep that starts from x = 0: x_i = b_i * dinv_i */
void bn_gs_color_zero(char vt, int lo, int hi, const void *dinv,
                      const void *b, void *x, hipStream_t stream);

/* Jacobi-type sweep pieces: x_i += w_i * (b_i - (A x)_i), w_i = scale/d_i;
 * the residual must come from a separate bn_csr_apply into r. */
void bn_jacobi_update(char vt, int n, const void *r, const void *dscale,
                      double omega, void *x, hipStream_t stream);
/* first sweep from x = 0: x_i = omega * dscale_i * b_i */
void bn_jacobi_zero(char vt, int n, const void *b, const void *dscale,
                    double omega, void *x, hipStream_t stream);

/* One level of a level-scheduled triangular solve: for the `cnt` rows in
 * rows[], y_i = (b_i - sum_{j != i} T_ij y_j) * dinv_i.  If zsum != NULL,
 * also z_i = zbase_i + y_i. */
void bn_trsv_level(char vt, int cnt, const int *rows, const int *ia,
                   const int *ja, const void *val, const void *dinv,
                   const void *b, void *y, int lanes, const void *zbase,
                   void *zsum, hipStream_t stream);

#define BN_ELL_MAX_W 8    /* widest level of the ELL triangular kernels */

/* The same level from a level-ordered ELL copy: rows[0..cnt) are the
 * level's rows, slot s of its k-th row at ja/val[s * cnt + k] (w slots,
 * w <= 8, padding naming the row itself), dl[k] its 1/t_ii.  Row sums as
 * bn_trsv_level's with one lane per row, to the last bit. */
void bn_trsv_ell(char vt, int w, int cnt, const int *rows, const int *ja, const void *val, const void *dl,
                 const void *b, void *y, const void *zbase, void *zsum, hipStream_t stream);
/* A run of levels of the same ELL copy in one launch of np workgroups of
 * BN_PART_BLOCK threads (np at most one per CU, so all resident).  The
 * rows are cut into np contiguous ranges; workgroup p computes its range's
 * rows level by level, with a barrier between levels: its steps are
 * steps[sp[p] .. sp[p+1]) (int4 {first entry of the level's row list,
 * rows, level rows * 16 + slots per row (1..8), first slot}); rows, dl and
 * ja/val are the copy's base arrays.  y must hold the sentinel
 * (bn_walk_fill) on every row the steps compute and the final values on
 * every row they read from earlier levels.  A row reads its y_j with plain
 * loads and, where one still holds the sentinel, waits for it with
 * agent-scope atomic loads; a wait past `limit` ticks of wall_clock64()
 * (limit < 0: any wait) sets *ctl = 1.0 and the launch gives up (y then
 * holds sentinels); a launch finding *ctl != 0 does nothing.  Values as
 * bn_trsv_ell's, to the last bit. */
#define BN_PART_BLOCK 256
/* workgroups of bn_trsv_parts that fit on one CU at once (0: unknown) */
int bn_trsv_parts_per_cu(char vt);
void bn_trsv_parts(char vt, int np, const int *sp, const int *steps, const int *rows, const int *ja,
                   const void *val, const void *dl, const void *b, void *y, const void *zbase, void *zsum,
                   double *ctl, long long limit, hipStream_t stream);
/* One multicolour Gauss-Seidel sweep of a level in one launch of np
 * workgroups of BN_GSW_BLOCK threads (np no more than fit on the device at
 * once).  Part p's rows of colour c are rows steps[2*(p*ncol+c)] .. + its
 * count (int2) of the colour-ordered copy (ia, jx, val; row r updates
 * level row map[r] with 1/a_ii dinvc[r]); jx is the copy's column array
 * with bit 31 set where the neighbour's colour precedes the row's.
 * Colours run in reverse when backward.  x must hold the sentinel
 * (bn_gs_prep) and xold the sweep's input (zero: the input is zero, xold
 * unused); the sweep writes x (and zsum = zbase + x when zsum != NULL).
 * Rows of `lanes` lanes (1, 2, 4 or 8) form the colour launches' sums with
 * the same lanes, to the last bit.  Waits, limit and *ctl as
 * bn_trsv_parts. */
#define BN_GSW_BLOCK 1024
/* workgroups of bn_gs_walk that fit on one CU at once (0: unknown) */
int bn_gs_walk_per_cu(char vt, int lanes);
void bn_gs_walk(char vt, int lanes, int np, int ncol, int backward, const int *steps, const int *ia, const int *jx,
                const void *val, const int *map, const void *dinvc, const void *b, void *x, void *xold, int zero,
                const void *zbase, void *zsum, double *ctl, long long limit, hipStream_t stream);
/* An exact-order Gauss-Seidel sweep in one launch of np workgroups of
 * BN_PART_BLOCK threads (np no more than fit at once): part p's steps are
 * steps[sp[p] .. sp[p+1]) (int2 {first entry of the dependency level's row
 * list `rows`, rows}); row i of (ia, jx, val) updates x_i with dinv[i]; jx
 * is the column array with bit 31 set where the forward sweep visits the
 * neighbour earlier (the backward sweep: where it visits it later).  x must
 * hold the sentinel and xold the sweep's input (bn_gs_prep).  Rows of
 * `lanes` lanes (1..64; other values as 64, like bn_trsv_level) form
 * bn_trsv_level's sums, to the last bit.  Waits, limit and *ctl as
 * bn_trsv_parts. */
int bn_sweep_parts_per_cu(char vt);
void bn_sweep_parts(char vt, int lanes, int np, int backward, const int *sp, const int *steps, const int *rows,
                    const int *ia, const int *jx, const void *val, const void *dinv, const void *b, void *x,
                    const void *xold, double *ctl, long long limit, hipStream_t stream);
/* xold = x (xold may be NULL), then x = the sentinel */
void bn_gs_prep(char vt, int n, void *x, void *xold, hipStream_t stream);
/* y[0..n) = the walk's sentinel (a signalling-NaN bit pattern) */
void bn_walk_fill(char vt, int n, void *y, hipStream_t stream);
/* out[q] = from[src[q]] (0 where src[q] < 0), and its FP32 rounding */
void bn_ell_gather(long long n, const int *src, const double *from, double *out, float *outs, hipStream_t stream);

/* One Jacobi sweep of a triangular (or any) system over all n rows:
 * xout_i = (b_i - sum_{j != i} T_ij xin_j) * dinv_i.  xin == NULL means
 * xin = 0 (xout = b * dinv).  With zsum != NULL (and xin != NULL) the
 * sweep writes zsum_i = zbase_i + xout_i instead of xout. */
void bn_jacobi_rows(char vt, int n, const int *ia, const int *ja, const void *val,
                    const void *dinv, const void *b, const void *xin, void *xout,
                    int lanes, const void *zbase, void *zsum, hipStream_t stream);
/* bn_jacobi_rows (xin != NULL, no zsum) that also writes xs_i = xout_i
 * dinv_i: the first step from zero of a triangle whose right-hand side
 * is xout and whose diagonal inverse is the same dinv */
void bn_jacobi_rows_scaled(char vt, int n, const int *ia, const int *ja, const void *val,
                           const void *dinv, const void *b, const void *xin, void *xout,
                           void *xs, int lanes, hipStream_t stream);

/* dense y = M x (row-major n x n), M in vt, x,y in vt */
void bn_dense_gemv(char vt, int n, const void *M, const void *x, void *y,
                   hipStream_t stream);
/* y = (float)(M (double)x): the FP64 dense product on FP32 vectors */
void bn_dense_gemv_ds(int n, const double *M, const float *x, float *y, hipStream_t stream);

/* ---- BLAS-1 (FP64 unless noted) ---- */
void bn_fill(char vt, int n, void *x, double v, hipStream_t stream);
void bn_copy(char vt, int n, const void *x, void *y, hipStream_t stream);
/* y(vt_out) = x(vt_in) elementwise conversion */
void bn_convert(char vt_in, char vt_out, int n, const void *x, void *y,
                hipStream_t stream);
/* partial sums of x.y (and of u.v if u != NULL into partial2) */
void bn_dot_partial(int n, const double *x, const double *y,
                    const double *u, const double *v, double *partial,
                    double *partial2, hipStream_t stream);
/* x += alpha p; r -= alpha q; partial of r.r.  alpha = s[BN_S_ALPHA] */
void bn_pcg_update_xr(int n, double *x, const double *p, double *r,
                      const double *q, const double *s, double *partial,
                      hipStream_t stream);
/* p = z + beta p, beta = s[BN_S_BETA] */
void bn_pcg_update_p(int n, double *p, const double *z, const double *s,
                     hipStream_t stream);
/* FP32-preconditioner forms: the x/r update also stores r rounded to
 * FP32 in rs; the r.z (and z.q if q != NULL) partials and the p update
 * read z in FP32 (exactly widened, so results equal the FP64-z forms') */
void bn_pcg_update_xr_s(int n, double *x, const double *p, double *r,
                        const double *q, const double *s, double *partial,
                        float *rs, hipStream_t stream);
/* bn_pcg_update_xr_s that also writes the first colour of a multicolour
 * sweep from zero with the FP32 r as right-hand side: x0_i = rs_i dinv_i
 * where colour_i == c, zero elsewhere (bn_gs_color_zero_colour's x, to the
 * last bit) */
void bn_pcg_update_xr_s0(int n, double *x, const double *p, double *r, const double *q, const double *s,
                         double *partial, float *rs, const unsigned char *colour, int c, const float *dinv,
                         float *x0, hipStream_t stream);
void bn_dot_rz_s(int n, const double *r, const float *zs, const double *q,
                 double *partial, double *partial2, hipStream_t stream);
void bn_pcg_update_p_s(int n, double *p, const float *zs, const double *s,
                       hipStream_t stream);
/* y = y + x (vt) */
void bn_axpy1(char vt, int n, const void *x, void *y, hipStream_t stream);

/* Scalar-finish kernels (single block): reduce partials and update s[].
 * which:
 *   0 start:   s[RZ]=s[RZ0]=sum(p1), s[RR]=sum(p2), s[BETA]=0, s[BB] kept
 *   1 alpha:   s[PAP]=sum(p1), s[ALPHA]=s[RZ]/s[PAP]
 *   2 iterate: s[RZNEW]=sum(p1), s[RR]=sum(p2), s[REL]=sqrt(RZNEW/RZ0),
 *              s[BETA]=RZNEW/RZ, s[RZ]=RZNEW
 *   3 bnorm:   s[BB]=sum(p1)
 *   4 iterate-flexible: like 2 but with the Polak-Ribiere beta
 *              -ALPHA * sum(p3) / RZ, where p3 holds z_new . (A p), i.e.
 *              z_new . (r_new - r_old) / (r_old . z_old)
 *   5 plain:   s[slot]=sum(p1) with slot passed in `slot` */
void bn_scalar_finish(int which, const double *p1, const double *p2,
                      const double *p3, double *s, int slot,
                      hipStream_t stream);
/* the same (slot 0), also copying the BN_S_COUNT scalars to hs, host
 * memory the device writes coherently */
void bn_scalar_finish_host(int which, const double *p1, const double *p2, const double *p3, double *s, double *hs,
                           hipStream_t stream);
/* which = 1 (PAP and ALPHA) over nslot partials of p.Ap, as written by
 * bn_csr_apply_blocks_dot (slots summed in a fixed order) */
void bn_scalar_finish_pap(const double *part, int nslot, double *s, hipStream_t stream);

/* gather/scatter for permuted level numbering: y[k] = x[perm[k]] (gather)
 * or y[perm[k]] = x[k] (scatter), FP64 */
void bn_permute(int n, const int *perm, const double *x, double *y,
                int scatter, hipStream_t stream);

/* edge-scatter adjoint gradient: for each segment e with endpoints
 * (u,v) (index or -1-d for Dirichlet node d):
 *   dg_e = -(lam_u - lam_v)(p_u - p_v),  grad_e = 4 g_e dg_e */
void bn_edge_grad(int m, const int *eu, const int *ev, const double *g,
                  const double *p, const double *lam, const double *pdir,
                  double *dg, double *grad_logr, hipStream_t stream);

/* ---------------------------------------------------------------------
 * Block (multi-right-hand-side) kernels, batch.hip.  A block vector holds
 * K columns interleaved by row (element (i, c) at i*K + c); K is a power
 * of two <= 32 and every call of one solve uses the same K.  Reductions
 * use bn_block_nb(n) blocks; partial buffers hold BN_BLOCK_RED * K
 * doubles; scalar blocks hold BN_S_COUNT doubles per column.
 * ------------------------------------------------------------------- */
#define BN_BLOCK_RED 1024
#define BN_BLOCK_KMAX 32

/* Y = op(A, X [, B]); rows longer than long_len go through the segmented
 * long-row bin (same arrays as bn_csr_apply_seg; part holds
 * seg_ptr[nmulti] * K doubles). */
void bn_spmm(char vt, int K, int op, int n, const int *ia, const int *ja, const void *val,
             const void *X, void *Y, const void *B, int long_len, int nlong, const int *rows,
             const int *seg_ptr, const int *seg_row, int nseg, int nmulti, double *part,
             hipStream_t stream);
/* bn_gs_color_map / bn_gs_color_zero_map on block vectors */
void bn_gs_color_map_k(char vt, int K, int lo, int hi, const int *map, const int *ia, const int *ja,
                       const void *val, const void *dinv, const void *B, void *X, const void *zb,
                       void *zs, hipStream_t stream);
void bn_gs_zero_map_k(char vt, int K, int lo, int hi, int all_n, const int *map, const void *dinv,
                      const void *B, void *X, hipStream_t stream);
/* bn_jacobi_rows on block vectors (Xin must not be NULL) */
void bn_jacobi_rows_k(char vt, int K, int n, const int *ia, const int *ja, const void *val,
                      const void *dinv, const void *B, const void *Xin, void *Xout, const void *zb,
                      void *zs, hipStream_t stream);
/* X = om * d_i * B (add = 0) or X += om * d_i * B (add = 1), d per row */
void bn_row_scale_k(char vt, int K, int n, const void *d, double om, const void *B, void *X, int add,
                    hipStream_t stream);
/* dense Y = M X, M n x n row-major FP64 */
void bn_gemm_k(int K, int n, const double *M, const double *X, double *Y, hipStream_t stream);
int bn_block_nb(int n);
/* per-column partials of x.y (and u.v) */
void bn_dot_k(int K, int n, const double *x, const double *y, const double *u, const double *v,
              double *p1, double *p2, hipStream_t stream);
/* x += alpha_c p, r -= alpha_c q, partials of r.r */
void bn_update_xr_k(int K, int n, double *x, const double *p, double *r, const double *q,
                    const double *s, double *part, hipStream_t stream);
/* p = z + beta_c p */
void bn_update_p_k(int K, int n, double *p, const double *z, const double *s, hipStream_t stream);
/* FP32-V-cycle forms (as bn_pcg_update_xr_s, bn_dot_rz_s, bn_pcg_update_p_s) */
void bn_update_xr_k_s(int K, int n, double *x, const double *p, double *r, const double *q,
                      const double *s, double *part, float *rs, hipStream_t stream);
void bn_dot_rz_k_s(int K, int n, const double *r, const float *zs, const double *q, double *p1,
                   double *p2, hipStream_t stream);
void bn_update_p_k_s(int K, int n, double *p, const float *zs, const double *s, hipStream_t stream);
/* per-column scalar logic (modes as bn_scalar_finish) with the stop test
 * and freezing of converged columns in modes 2 and 4 */
void bn_finish_k(int K, int which, int n, const double *p1, const double *p2, const double *p3,
                 double *s, hipStream_t stream);
/* Segmented BLAS-1 and scalar logic for a block-diagonal union of nsys
 * independent systems: work item w = rows [item_lo[w], item_lo[w+1]) of
 * system item_seg[w]; system g owns items item_ptr[g] .. item_ptr[g+1]-1;
 * scalars BN_S_COUNT per system; partials one per item. */
void bn_dot_seg(int nitem, const int *item_lo, const double *x, const double *y, const double *u,
                const double *v, double *p1, double *p2, hipStream_t stream);
void bn_update_xr_seg(int nitem, const int *item_lo, const int *item_seg, double *x, const double *p,
                      double *r, const double *q, const double *s, double *part, hipStream_t stream);
void bn_update_p_seg(int nitem, const int *item_lo, const int *item_seg, double *p, const double *z,
                     const double *s, hipStream_t stream);
void bn_finish_seg(int nsys, int which, const int *item_ptr, const double *p1, const double *p2,
                   const double *p3, double *s, hipStream_t stream);

/* k right-hand sides stored one after another (k x n) <-> interleaved
 * n x K (columns k..K-1 zero); back = 1 copies the first k columns out */
void bn_interleave(int n, int k, int K, const double *src, double *dst, int back, hipStream_t stream);

/* ---- value-only rebuild of a hierarchy (rebuild.hip) ---- */
/* (v + v^T) / 2 through transpose positions; diag0 = 1 for Galerkin sums
 * formed from -0.0 (a zero diagonal is taken as +0.0) */
void bn_rb_symmetrize(int nnz, const int *tpos, const double *v, double *out, int diag0, hipStream_t st);
void bn_rb_gather(int n, const int *map, const double *src, double *out, float *outs, hipStream_t st);
void bn_rb_to_float(int n, const double *src, float *out, hipStream_t st);
void bn_rb_row_scalars(int n, const int *ia, const int *ja, const double *a, double *dinv, float *dinvs,
                       double *l1inv, float *l1invs, double *l1, hipStream_t st);
void bn_rb_dinv_map(int n, const int *ia, const int *ja, const double *a, const int *map, double *dinv, float *dinvs,
                    hipStream_t st);
/* standard interpolation from contribution lists: den per (F row, strong F
 * neighbour) pair, dhat per F row, then every P entry */
void bn_rb_den(int npairs, const int *doff, const int *dq, const double *a, double *den, hipStream_t st);
void bn_rb_dhat(int n, const int *aia, const double *a, const unsigned char *cls, const int *cpt, const int *rp,
                const int *pair_k, const double *den, double *dhat, hipStream_t st);
void bn_rb_pval(int nnzp, const int *prow, const int *cpt, const int *pia, const int *poff, const int *pa,
                const int *pq, const int *pair_k, const double *a, const double *den, const double *dhat, double *pv,
                hipStream_t st);
/* lists longer than this take a wavefront each */
#define BN_RB_LONG 64
/* Galerkin entries from per-entry triple lists (toff, (t1, t2, t3) = R, A,
 * P entry); the nlong entries in elist, whose lists exceed 64 terms, take
 * one wavefront each */
void bn_rb_rap_pull(int nnz, const int *toff, const int *t1, const int *t2, const int *t3, int nlong,
                    const int *elist, const double *rv, const double *av, const double *pv, double *out,
                    hipStream_t st);
/* IC(0) of the rows rows[0..cnt) of a lower factor (strictly ascending
 * columns, diagonal last), in place on lv holding the operator's lower
 * entries; every row a row depends on must be final (one schedule level
 * per call) */
void bn_ic0_rows(int cnt, const int *rows, const int *lia, const int *lja, double *lv, hipStream_t st);
/* Dense inverse of an SPD CSR matrix of n <= 8192 rows on the device,
 * entry for entry as dense.c's bn_dense_inverse forms it: Lc and Lr are
 * n x n scratch (the factor column-major and its row-major copy), inv the
 * row-major inverse and invs (may be NULL) its FP32 copy; *clamped (may
 * be NULL) is set to 1 if a pivot had to be clamped.  Returns -1 if n is
 * out of range. */
int bn_dn_inverse(int n, const int *ia, const int *ja, const double *a, double *Lc, double *Lr, double *inv,
                  float *invs, int *clamped, hipStream_t st);
/* Galerkin entries, one thread per coarse row */
void bn_rb_rap(int nc, const int *ria, const int *rja, const double *rv, const int *aia, const int *aja,
               const double *av, const int *pia, const int *pja, const double *pv, const int *cia, const int *ccol,
               const int *cpos, double *out, hipStream_t st);

#ifdef __cplusplus
}
#endif
#endif
