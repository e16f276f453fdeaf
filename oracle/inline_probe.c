/*
 * inline_probe.c - test infrastructure: runs the header-only CSR helpers
 * of whichever amg.h it is compiled against (the reference's
 * C_amg1/include/agg_lib.h:97-242 or libbnamg's include/amg.h) on a
 * matrix read from a COO file, and prints the results, so the two
 * headers' helpers can be compared bit for bit.
 *
 *   inline_probe <coo file> <perm file or ->
 *
 * Prints A^T (dcsr_transz, p = NULL), the transpose of the permuted
 * matrix (p from the file: n integers), and A + A^T (dcsr_mat_add), each
 * as "ia..." / "ja..." / "val..." lines with %.17g values.
 */
#include "amg.h"

static void dump(const char *tag, const dcsr_mat *m) {
  printf("%s %d %d %d\nia", tag, m->row, m->col, m->nnz);
  for (INT i = 0; i <= m->row; i++) printf(" %d", m->ia[i]);
  printf("\nja");
  for (INT k = 0; k < m->nnz; k++) printf(" %d", m->ja[k]);
  printf("\nval");
  for (INT k = 0; k < m->nnz; k++) printf(" %.17g", m->val[k]);
  printf("\n");
}

int main(int argc, char **argv) {
  if (argc < 3) return 2;
  FILE *fp = fopen(argv[1], "r");
  if (!fp) return 2;
  dcsr_mat A;
  if (dcsr_mat_dcoo(fp, &A)) return 3;
  fclose(fp);
  dcsr_mat At;
  dcsr_mat_alloc(&At, A.col, A.row, A.nnz);
  dcsr_transz(&A, NULL, &At);
  dump("transpose", &At);
  if (argv[2][0] != '-') {
    FILE *fq = fopen(argv[2], "r");
    if (!fq) return 2;
    INT *p = (INT *)malloc((size_t)A.row * sizeof(INT));
    for (INT i = 0; i < A.row; i++)
      if (fscanf(fq, "%d", &p[i]) != 1) return 4;
    fclose(fq);
    dcsr_mat Pt;
    dcsr_mat_alloc(&Pt, A.col, A.row, A.nnz);
    dcsr_transz(&A, p, &Pt);
    dump("permuted_transpose", &Pt);
    dcsr_mat_free(&Pt);
    free(p);
  }
  if (A.row == A.col) {
    dcsr_mat S;
    if (dcsr_mat_add(&A, &At, &S) != 0) return 5;
    dump("sum", &S);
    dcsr_mat_free(&S);
  }
  dcsr_mat_free(&At);
  dcsr_mat_free(&A);
  return 0;
}
