/*
 * csr.h - host CSR helpers shared by the set-up code (internal).
 */
#ifndef BN_CSR_H
#define BN_CSR_H

#include <stddef.h>
#include "amg.h"

#ifdef __cplusplus
extern "C" {
#endif

/* allocation helpers that never return NULL for zero sizes */
void *bn_xmalloc(size_t bytes);
void *bn_xcalloc(size_t count, size_t size);

void bn_csr_alloc(dcsr_mat *a, INT rows, INT cols, INT nnz);
void bn_csr_free(dcsr_mat *a);
void bn_csr_copy(const dcsr_mat *src, dcsr_mat *dst);
/* stable transpose: rows of A^T list source rows ascending */
void bn_csr_transpose(const dcsr_mat *a, dcsr_mat *at);
/* C = A + B, A's entries first in A's order, then B's new columns */
void bn_csr_add(const dcsr_mat *a, const dcsr_mat *b, dcsr_mat *c);
/* out = (A + A^T) / 2 */
void bn_csr_symmetrize(const dcsr_mat *a, dcsr_mat *out);
/* l1 row norms */
void bn_csr_l1(const dcsr_mat *a, REAL *l1);
/* exclusive prefix sum over counts[0..n) into ptr[0..n] */
void bn_prefix(const INT *counts, INT n, INT *ptr);

int bn_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
