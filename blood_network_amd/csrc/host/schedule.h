/* schedule.h - colourings and level sets for device sweeps (internal) */
#ifndef BN_SCHEDULE_H
#define BN_SCHEDULE_H

#include "amg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Colour order of a level: perm[new] = old lists colour 0 rows ascending,
 * then colour 1, ...  With split != 0 the C points (mis[i] != 0) are
 * coloured among themselves first and the F points after them.  Returns
 * the colour count; *col_ptr (malloc'ed, ncol+1 entries) delimits colours
 * in the new numbering. */
int bn_colour_order(const dcsr_mat *A, const int *mis, int split, int *perm, int **col_ptr);

/* Level sets of a triangular sweep over CSR rows: lower != 0 means row i
 * depends on its columns j < i (forward), else on j > i (backward).
 * rows[] receives the rows grouped by level (ascending within a level);
 * *lev_ptr (malloc'ed) delimits levels.  Returns the level count. */
int bn_level_sets(const int *ia, const int *ja, int n, int lower, int *rows, int **lev_ptr);

/* Level sets of a sequential Gauss-Seidel sweep over the rows of A in the
 * visiting order order[0], order[1], ... (order NULL: 0, 1, ...;
 * backward != 0 visits the order from its end).  Row i depends on every
 * row j coupled to it (A_ij or A_ji stored, j != i) that the sweep visits
 * before i, so a level's rows are mutually uncoupled: every row they read
 * was either visited earlier (an earlier level, already updated) or is
 * visited later (a later level, still old) - the values the sequential
 * sweep reads.  rows[] receives the rows grouped by level (ascending within
 * a level); *lev_ptr (malloc'ed) delimits the levels.  Returns the level
 * count, or -1 when order is not a permutation of 0..n-1. */
int bn_sweep_levels(const dcsr_mat *A, const int *order, int backward, int *rows, int **lev_ptr);

/* Chunked colour order for the block Gauss-Seidel smoother.  The rows are
 * cut into nchunks contiguous ranges of the current numbering (chunk k =
 * rows [k*S, min(n, (k+1)*S)), S = ceil(n / nchunks)); inside each chunk
 * the rows are coloured greedily over the chunk's own couplings (C points
 * first, then F points, when split != 0) and listed colour by colour,
 * ascending within a colour.  perm[new] = old never leaves a chunk.
 * Outputs (malloc'ed): chunk_ptr (nchunks+1 row starts), ccol_ptr
 * (nchunks+1 offsets into cbnd), cbnd (chunk k's colour boundaries are
 * cbnd[ccol_ptr[k] .. ccol_ptr[k+1]-1], first = chunk start, last = chunk
 * end).  Returns the number of chunks actually used (<= nchunks). */
int bn_chunk_order(const dcsr_mat *A, const int *mis, int split, int nchunks, int *perm, int **chunk_ptr,
                   int **ccol_ptr, int **cbnd);

/* Row blocks for the adaptive CSR kernel: blocks[0] = 0 < blocks[1] < ...
 * < blocks[nblk] = n.  Consecutive rows are packed while the block holds
 * at most max_nnz entries and max_rows rows; a row with more than max_nnz
 * entries forms a block alone.  blocks must hold n + 1 ints.  Returns
 * nblk. */
int bn_row_blocks(const int *ia, int n, int max_nnz, int max_rows, int *blocks);

/* B = A's off-diagonal part with rows taken in perm_r order (NULL:
 * identity): row k of B is row perm_r[k] of A without its entries in
 * column perm_r[k], in A's entry order; columns keep A's numbering. */
void bn_csr_offdiag_rows(const dcsr_mat *A, const int *perm_r, dcsr_mat *B);

/* B = A with rows taken in perm_r order and columns renamed by inv_c
 * (either may be NULL for identity); entry order within a row is kept. */
void bn_csr_permute(const dcsr_mat *A, const int *perm_r, const int *inv_c, dcsr_mat *B);

#ifdef __cplusplus
}
#endif
#endif
