/*
 * nwk.c - vessel-network case files and solution files.
 *
 * A case <base> is four numeric text files, read as MATLAB's load() reads
 * them (Matlab_amg/caseReaderMJ.m:1-33): rows of whitespace-separated
 * numbers, the same count on every row, '%' starting a comment.
 *   <base>.fMx  faces: column 2 and 3 are the 1-based end points
 *   <base>.pMx  point coordinates x y z
 *   <base>.dia  face diameters (may be empty; the system does not use them)
 *   <base>.BC   boundary conditions: column 1 node (1-based), column 3
 *               pressure
 * The pressure system is Matlab_amg/read_nwk.m:15-37: with C the face-by-
 * node incidence (+1 at the first end point, -1 at the second) and W the
 * diffusive face conductances of makeDiffusiveConductance.m (Cartesian
 * grids: an x face conducts hy hz / hx, where hx, hy, hz are the lengths of
 * the first face along each axis), A = C_I' W C_I and b = -C_I' W C_B p_B
 * over the interior nodes, numbered in ascending node order.  Duplicate
 * entries are summed and exact zeros dropped, as MATLAB's sparse products
 * do; rows list their columns ascending.
 * Solutions are written as Matlab_amg/test_amg.m:68-72 (one %.16e per
 * line, interior order) and label_solution.m:42-47 ("%d %.16e", 1-based
 * node index).
 */
#include <ctype.h>
#include <errno.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "bn_netgen.h"

void bn_set_error(const char *fmt, ...);

/* numeric text matrix; *rows = 0 for an empty file */
static int read_matrix(const char *path, int *rows, int *cols, double **data) {
  *rows = *cols = 0;
  *data = NULL;
  FILE *fp = fopen(path, "r");
  if (!fp) {
    bn_set_error("cannot open %s: %s", path, strerror(errno));
    return -1;
  }
  size_t cap = 1024, n = 0;
  double *v = (double *)malloc(cap * sizeof(double));
  char *line = NULL;
  size_t lcap = 0;
  int r = 0, c = -1, rc = 0;
  while (getline(&line, &lcap, fp) >= 0) {
    char *pct = strchr(line, '%');
    if (pct) *pct = '\0';
    int k = 0;
    char *p = line;
    for (;;) {
      while (*p && isspace((unsigned char)*p)) p++;
      if (!*p) break;
      char *end = NULL;
      const double x = strtod(p, &end);
      if (end == p) {
        bn_set_error("%s: not a number at row %d", path, r + 1);
        rc = -1;
        break;
      }
      if (n == cap) {
        cap *= 2;
        v = (double *)realloc(v, cap * sizeof(double));
      }
      v[n++] = x;
      k++;
      p = end;
    }
    if (rc) break;
    if (k == 0) continue;
    if (c >= 0 && k != c) {
      bn_set_error("%s: row %d has %d numbers, row 1 has %d", path, r + 1, k, c);
      rc = -1;
      break;
    }
    c = k;
    r++;
  }
  free(line);
  fclose(fp);
  if (rc) {
    free(v);
    return -1;
  }
  *rows = r;
  *cols = c < 0 ? 0 : c;
  *data = v;
  return 0;
}

static int as_index(double x, int limit, int *out) {
  if (!(x >= 1.0) || x > (double)limit || x != floor(x)) return -1;
  *out = (int)x - 1;
  return 0;
}

void bn_nwk_free(bn_nwk *w) {
  if (!w) return;
  free(w->pts);
  free(w->face);
  free(w->dia);
  free(w->bc_node);
  free(w->bc_val);
  memset(w, 0, sizeof(*w));
}

int bn_nwk_read(const char *base, bn_nwk *w) {
  memset(w, 0, sizeof(*w));
  const size_t bl = strlen(base);
  char *path = (char *)malloc(bl + 8);
  double *F = NULL, *P = NULL, *D = NULL, *B = NULL;
  int fr, fc, pr, pc, dr, dc, br, bc;
  int rc = -1;
  sprintf(path, "%s.fMx", base);
  if (read_matrix(path, &fr, &fc, &F)) goto done;
  sprintf(path, "%s.pMx", base);
  if (read_matrix(path, &pr, &pc, &P)) goto done;
  sprintf(path, "%s.dia", base);
  if (read_matrix(path, &dr, &dc, &D)) goto done;
  sprintf(path, "%s.BC", base);
  if (read_matrix(path, &br, &bc, &B)) goto done;
  if (fr == 0 || fc < 3) {
    bn_set_error("%s.fMx: need at least one face with 3 columns", base);
    goto done;
  }
  if (pr == 0 || pc < 3) {
    bn_set_error("%s.pMx: need x y z columns", base);
    goto done;
  }
  if (br > 0 && bc < 3) {
    bn_set_error("%s.BC: need node, type and pressure columns", base);
    goto done;
  }
  if (dr > 0 && (long long)dr * dc != fr) {
    bn_set_error("%s.dia: %d values for %d faces", base, dr * dc, fr);
    goto done;
  }
  w->np = pr;
  w->nf = fr;
  w->nbc = br;
  w->pts = (double *)malloc((size_t)pr * 3 * sizeof(double));
  for (int i = 0; i < pr; i++)
    for (int d = 0; d < 3; d++) w->pts[3 * (size_t)i + d] = P[(size_t)i * pc + d];
  w->face = (int *)malloc((size_t)fr * 2 * sizeof(int));
  for (int f = 0; f < fr; f++)
    for (int e = 0; e < 2; e++)
      if (as_index(F[(size_t)f * fc + 1 + e], pr, &w->face[2 * (size_t)f + e])) {
        bn_set_error("%s.fMx: face %d names point %g of %d", base, f + 1, F[(size_t)f * fc + 1 + e], pr);
        goto done;
      }
  if (dr > 0) {
    w->dia = (double *)malloc((size_t)fr * sizeof(double));
    memcpy(w->dia, D, (size_t)fr * sizeof(double));
  }
  w->bc_node = (int *)malloc((size_t)(br > 0 ? br : 1) * sizeof(int));
  w->bc_val = (double *)malloc((size_t)(br > 0 ? br : 1) * sizeof(double));
  for (int k = 0; k < br; k++) {
    if (as_index(B[(size_t)k * bc], pr, &w->bc_node[k])) {
      bn_set_error("%s.BC: row %d names node %g of %d", base, k + 1, B[(size_t)k * bc], pr);
      goto done;
    }
    w->bc_val[k] = B[(size_t)k * bc + 2];
  }
  rc = 0;
done:
  free(path);
  free(F);
  free(P);
  free(D);
  free(B);
  if (rc) bn_nwk_free(w);
  return rc;
}

typedef struct {
  int r, c;
  size_t k; /* insertion order: duplicates are summed in face order */
  double v;
} coo_t;

static int cmp_coo(const void *a, const void *b) {
  const coo_t *x = (const coo_t *)a, *y = (const coo_t *)b;
  if (x->r != y->r) return (x->r > y->r) - (x->r < y->r);
  if (x->c != y->c) return (x->c > y->c) - (x->c < y->c);
  return (x->k > y->k) - (x->k < y->k);
}

int bn_nwk_system(const bn_nwk *w, bn_nwk_sys *s) {
  memset(s, 0, sizeof(*s));
  const int np = w->np, nf = w->nf;
  /* boundary pressures; a node listed twice would make p_B ambiguous */
  int *role = (int *)malloc((size_t)(np > 0 ? np : 1) * sizeof(int)); /* -1 interior, else BC row */
  for (int i = 0; i < np; i++) role[i] = -1;
  for (int k = 0; k < w->nbc; k++) {
    if (role[w->bc_node[k]] >= 0) {
      bn_set_error("boundary node %d is listed twice", w->bc_node[k] + 1);
      free(role);
      return -1;
    }
    role[w->bc_node[k]] = k;
  }
  /* diffusive conductances: lengths of the first face along +x, +y, +z */
  double h[3] = {NAN, NAN, NAN};
  for (int d = 0; d < 3; d++)
    for (int f = 0; f < nf; f++) {
      const double *a = w->pts + 3 * (size_t)w->face[2 * f], *b = w->pts + 3 * (size_t)w->face[2 * f + 1];
      if (b[d] - a[d] > 0) {
        const double u = b[0] - a[0], v = b[1] - a[1], z = b[2] - a[2];
        h[d] = sqrt(u * u + v * v + z * z);
        break;
      }
    }
  double *g = (double *)malloc((size_t)nf * sizeof(double));
  for (int f = 0; f < nf; f++) {
    const double *a = w->pts + 3 * (size_t)w->face[2 * f], *b = w->pts + 3 * (size_t)w->face[2 * f + 1];
    double c = NAN;
    if (b[0] - a[0] > 0) c = h[1] * h[2] / h[0];
    if (b[1] - a[1] > 0) c = h[0] * h[2] / h[1];
    if (b[2] - a[2] > 0) c = h[0] * h[1] / h[2];
    if (!isfinite(c) && w->face[2 * f] != w->face[2 * f + 1]) {
      bn_set_error("face %d has no positive x, y or z extent (its conductance would be NaN)", f + 1);
      free(g);
      free(role);
      return -1;
    }
    g[f] = c;
  }
  /* interior numbering, ascending node order */
  int *uid = (int *)malloc((size_t)(np > 0 ? np : 1) * sizeof(int));
  int n = 0;
  for (int i = 0; i < np; i++) uid[i] = role[i] < 0 ? n++ : -1;
  s->n = n;
  s->interior = (int *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int));
  for (int i = 0; i < np; i++)
    if (uid[i] >= 0) s->interior[uid[i]] = i + 1;
  s->b = (double *)calloc((size_t)(n > 0 ? n : 1), sizeof(double));
  coo_t *e = (coo_t *)malloc((size_t)(4 * (size_t)nf + 1) * sizeof(coo_t));
  size_t ne = 0;
  for (int f = 0; f < nf; f++) {
    const int p1 = w->face[2 * f], p2 = w->face[2 * f + 1];
    if (p1 == p2) continue; /* +1 and -1 in one column: no coupling */
    const int u1 = uid[p1], u2 = uid[p2];
    const double c = g[f];
    if (u1 >= 0) e[ne] = (coo_t){u1, u1, ne, c}, ne++;
    if (u2 >= 0) e[ne] = (coo_t){u2, u2, ne, c}, ne++;
    if (u1 >= 0 && u2 >= 0) {
      e[ne] = (coo_t){u1, u2, ne, -c}, ne++;
      e[ne] = (coo_t){u2, u1, ne, -c}, ne++;
    } else if (u1 >= 0) {
      s->b[u1] += c * w->bc_val[role[p2]];
    } else if (u2 >= 0) {
      s->b[u2] += c * w->bc_val[role[p1]];
    }
  }
  qsort(e, ne, sizeof(coo_t), cmp_coo);
  s->ia = (int *)calloc((size_t)n + 1, sizeof(int));
  s->ja = (int *)malloc((ne > 0 ? ne : 1) * sizeof(int));
  s->val = (double *)malloc((ne > 0 ? ne : 1) * sizeof(double));
  size_t nnz = 0;
  for (size_t k = 0; k < ne;) {
    size_t q = k;
    double v = 0.0;
    while (q < ne && e[q].r == e[k].r && e[q].c == e[k].c) v += e[q++].v;
    if (v != 0.0) {
      s->ja[nnz] = e[k].c;
      s->val[nnz] = v;
      s->ia[e[k].r + 1]++;
      nnz++;
    }
    k = q;
  }
  for (int i = 0; i < n; i++) s->ia[i + 1] += s->ia[i];
  s->nnz = (int)nnz;
  free(e);
  free(uid);
  free(g);
  free(role);
  return 0;
}

void bn_nwk_sys_free(bn_nwk_sys *s) {
  if (!s) return;
  free(s->ia);
  free(s->ja);
  free(s->val);
  free(s->b);
  free(s->interior);
  memset(s, 0, sizeof(*s));
}

int bn_write_pressures(const char *path, const double *p, int n) {
  FILE *fp = fopen(path, "w");
  if (!fp) {
    bn_set_error("cannot write %s: %s", path, strerror(errno));
    return -1;
  }
  for (int i = 0; i < n; i++) fprintf(fp, "%.16e\n", p[i]);
  return fclose(fp) ? -1 : 0;
}

int bn_write_labelled(const char *path, const int *node, const double *p, int n) {
  FILE *fp = fopen(path, "w");
  if (!fp) {
    bn_set_error("cannot write %s: %s", path, strerror(errno));
    return -1;
  }
  for (int i = 0; i < n; i++) fprintf(fp, "%d %.16e\n", node[i], p[i]);
  return fclose(fp) ? -1 : 0;
}
