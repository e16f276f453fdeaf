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

/* B = A with rows taken in perm_r order and columns renamed by inv_c
 * (either may be NULL for identity); entry order within a row is kept. */
void bn_csr_permute(const dcsr_mat *A, const int *perm_r, const int *inv_c, dcsr_mat *B);

#ifdef __cplusplus
}
#endif
#endif
