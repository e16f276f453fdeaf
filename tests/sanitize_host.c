/*
 * sanitize_host.c - drives every host-side routine of libbnamg that needs
 * no GPU (network generator, hierarchy set-up for RS/PMIS/HMIS, value-only
 * rebuild, IC(0), dense coarsest inverse, colour/level/chunk/row-block
 * schedules, COO reader, network case reader, CSR helpers) on small
 * inputs.  tests/test_sanitize_host.py builds it with the host sources
 * under -fsanitize=address,undefined and runs it: any out-of-bounds
 * access, use after free, leak or undefined behaviour fails the test.
 *
 *   sanitize_host <scratch directory>
 */
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "amg.h"
#include "bn_netgen.h"
#include "schedule.h"

/* the two symbols the host sources take from the GPU-side files */
void bn_set_error(const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  va_end(ap);
}
void bn_cache_drop(const void *p) { (void)p; }

int bn_dense_inverse(const dcsr_mat *A, double *inv);

static int fails = 0;
#define CHECK(c)                                                  \
  do {                                                            \
    if (!(c)) {                                                   \
      fprintf(stderr, "check failed: %s (line %d)\n", #c, __LINE__); \
      fails++;                                                    \
    }                                                             \
  } while (0)

static void schedules(const dcsr_mat *A, const int *mis) {
  const int n = A->row;
  int *perm = (int *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
  int *cp = NULL;
  const int nc = bn_colour_order(A, mis, 1, perm, &cp);
  CHECK(nc >= 1 && cp[nc] == n);
  free(cp);
  int *chunk = NULL, *ccol = NULL, *cbnd = NULL;
  const int nch = bn_chunk_order(A, mis, 1, n > 512 ? 8 : 1, perm, &chunk, &ccol, &cbnd);
  CHECK(nch >= 1 && chunk[nch] == n);
  free(chunk);
  free(ccol);
  free(cbnd);
  int *lv = NULL;
  const int nl = bn_level_sets(A->ia, A->ja, n, 1, perm, &lv);
  CHECK(nl >= 1 && lv[nl] == n);
  free(lv);
  int *blocks = (int *)malloc(((size_t)n + 1) * sizeof(int));
  const int nb = bn_row_blocks(A->ia, n, 2048, 256, blocks);
  CHECK(nb >= 1 && blocks[nb] == n);
  free(blocks);
  free(perm);
}

static void hierarchy(const bn_network *net, int coarsening) {
  dcsr_mat A = {net->n, net->n, net->nnz, net->ia, net->ja, net->val};
  amg_data amg;
  memset(&amg, 0, sizeof(amg));
  amg.threshold = 0.25;
  amg.min_size = 300;
  amg.max_levels = AMG_MAX_LEVELS;
  amg.nu = 1;
  amg.smoother = SMOOTHER_CF_GS;
  amg.omega_jacobi = 0.5;
  amg.coarsening = coarsening;
  amg_build_hierarchy(&amg, &A);
  CHECK(amg.num_levels >= 2);
  for (int l = 0; l < amg.num_levels - 1; l++) schedules(&amg.levels[l].A, amg.levels[l].mis);
  const dcsr_mat *C = &amg.levels[amg.num_levels - 1].A;
  double *inv = (double *)malloc((size_t)C->row * (size_t)C->row * sizeof(double));
  bn_dense_inverse(C, inv);
  free(inv);
  ichol_factor L;
  ichol_compute(&A, &L);
  CHECK(L.n == A.row && L.nnz > 0);
  ichol_free(&L);
  /* value-only rebuild on other conductances */
  double *g = (double *)malloc((size_t)net->n_edges * sizeof(double));
  for (int e = 0; e < net->n_edges; e++) g[e] = net->g[e] * (1.0 + 0.3 * ((e * 7919) % 13) / 13.0);
  dcsr_mat A2;
  dcsr_mat_alloc(&A2, net->n, net->n, net->nnz);
  memcpy(A2.ia, net->ia, ((size_t)net->n + 1) * sizeof(int));
  memcpy(A2.ja, net->ja, (size_t)net->nnz * sizeof(int));
  double *b2 = (double *)malloc((size_t)net->n * sizeof(double));
  bn_netgen_assemble(net, g, NULL, A2.val, b2);
  amg_rebuild_values(&amg, &A2);
  amg_free(&amg);
  dcsr_mat_free(&A2);
  free(b2);
  free(g);
}

static void coo_and_helpers(const char *dir, const bn_network *net) {
  char path[4096];
  snprintf(path, sizeof(path), "%s/a.txt", dir);
  FILE *fp = fopen(path, "w");
  fprintf(fp, "%d %d %d\n", net->n, net->n, net->nnz);
  for (int i = 0; i < net->n; i++)
    for (int q = net->ia[i]; q < net->ia[i + 1]; q++) fprintf(fp, "%d %d %.17g\n", i + 1, net->ja[q] + 1, net->val[q]);
  fclose(fp);
  fp = fopen(path, "r");
  dcsr_mat A;
  CHECK(dcsr_mat_dcoo(fp, &A) == 0);
  fclose(fp);
  CHECK(A.nnz == net->nnz && !memcmp(A.ja, net->ja, (size_t)A.nnz * sizeof(int)));
  dcsr_mat At, S;
  dcsr_mat_alloc(&At, A.col, A.row, A.nnz);
  dcsr_transz(&A, NULL, &At);
  CHECK(dcsr_mat_add(&A, &At, &S) == 0);
  double *x = (double *)calloc((size_t)A.row, sizeof(double)), *y = (double *)calloc((size_t)A.row, sizeof(double));
  randv(x, A.row);
  abyvg(A.ia, A.ja, A.val, x, A.row, y);
  abyvam(x, A.ia, A.ja, A.val, y, A.row, y);
  abyvcs(A.ia, A.ja, A.val, x, A.row, y);
  vbya(A.ia, A.ja, A.val, x, A.row, A.col, y);
  double s;
  l2norm(y, &s, A.row);
  free(x);
  free(y);
  dcsr_mat_free(&S);
  dcsr_mat_free(&At);
  dcsr_mat_free(&A);
  /* malformed files */
  snprintf(path, sizeof(path), "%s/bad.txt", dir);
  fp = fopen(path, "w");
  fprintf(fp, "3 3 4\n1 1 1.0\n9 9 1.0\n");
  fclose(fp);
  fp = fopen(path, "r");
  dcsr_mat B;
  CHECK(dcsr_mat_dcoo(fp, &B) == 1);
  fclose(fp);
}

static void network_case(const char *dir) {
  /* 4 x 3 x 3 Cartesian grid, x = 0 and x = 3 planes Dirichlet */
  char path[4096];
  const int nx = 4, ny = 3, nz = 3, np = nx * ny * nz;
  snprintf(path, sizeof(path), "%s/c.pMx", dir);
  FILE *fp = fopen(path, "w");
  for (int z = 0; z < nz; z++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) fprintf(fp, "%g %g %g\n", 1e-5 * x, 2e-5 * y, 3e-5 * z);
  fclose(fp);
  snprintf(path, sizeof(path), "%s/c.fMx", dir);
  fp = fopen(path, "w");
  int f = 0;
  for (int z = 0; z < nz; z++)
    for (int y = 0; y < ny; y++)
      for (int x = 0; x < nx; x++) {
        const int id = x + nx * (y + ny * z) + 1;
        if (x + 1 < nx) fprintf(fp, "%d %d %d\n", ++f, id, id + 1);
        if (y + 1 < ny) fprintf(fp, "%d %d %d\n", ++f, id, id + nx);
        if (z + 1 < nz) fprintf(fp, "%d %d %d\n", ++f, id, id + nx * ny);
      }
  fclose(fp);
  snprintf(path, sizeof(path), "%s/c.dia", dir);
  fclose(fopen(path, "w"));
  snprintf(path, sizeof(path), "%s/c.BC", dir);
  fp = fopen(path, "w");
  for (int z = 0; z < nz; z++)
    for (int y = 0; y < ny; y++) {
      fprintf(fp, "%d 1 100\n", nx * (y + ny * z) + 1);
      fprintf(fp, "%d 1 10\n", nx * (y + ny * z) + nx);
    }
  fclose(fp);
  snprintf(path, sizeof(path), "%s/c", dir);
  bn_nwk w;
  bn_nwk_sys s;
  CHECK(bn_nwk_read(path, &w) == 0);
  CHECK(w.np == np && w.dia == NULL);
  CHECK(bn_nwk_system(&w, &s) == 0);
  CHECK(s.n == np - 2 * ny * nz && s.nnz > 0);
  snprintf(path, sizeof(path), "%s/p.txt", dir);
  CHECK(bn_write_pressures(path, s.b, s.n) == 0);
  snprintf(path, sizeof(path), "%s/pl.txt", dir);
  CHECK(bn_write_labelled(path, s.interior, s.b, s.n) == 0);
  bn_nwk_sys_free(&s);
  bn_nwk_free(&w);
  snprintf(path, sizeof(path), "%s/absent", dir);
  CHECK(bn_nwk_read(path, &w) != 0);
}

int main(int argc, char **argv) {
  if (argc < 2) return 2;
  bn_netgen_params p;
  bn_netgen_default_params(&p);
  p.L = 10;
  p.depth = 6;
  p.n_trees = 1;
  p.seed = 3;
  bn_network net;
  if (bn_netgen_build(&p, &net)) return 3;
  hierarchy(&net, COARSEN_RS);
  hierarchy(&net, COARSEN_PMIS);
  hierarchy(&net, COARSEN_HMIS);
  coo_and_helpers(argv[1], &net);
  network_case(argv[1]);
  bn_netgen_free(&net);
  if (fails) return 1;
  printf("ok\n");
  return 0;
}
