/*
 * netgen.c - deterministic synthetic vessel-network generator.
 *
 * Random numbers: SplitMix64 seeds one xoshiro256** state per stream
 * (stream = one random quantity), uniform doubles take the top 53 bits,
 * normals use the cosine branch of Box-Muller.  The streams are
 *   1          capillary radii, one normal per lattice segment
 *   16 + t     tree t radius factors U(0.9,1.1), heap order
 *   8          leaf-attachment targets (partial Fisher-Yates over L^3)
 *   9          boundary_roots: each tree's boundary-face root node
 * Segment order: lattice segments (node ascending; +x, +y, +z), then for
 * each tree (arterial trees, then venous) its parent->child segments in
 * heap order of the child, then its leaf attachments in heap order.
 */
#include "bn_netgen.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef struct { uint64_t s[4]; } rng_t;

static uint64_t splitmix64(uint64_t *x) {
  uint64_t z = (*x += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

static void rng_seed(rng_t *r, uint64_t seed, uint64_t stream) {
  uint64_t x = seed ^ (stream * 0xD1B54A32D192ED03ULL);
  for (int k = 0; k < 4; k++) r->s[k] = splitmix64(&x);
}

static inline uint64_t rotl(uint64_t x, int k) { return (x << k) | (x >> (64 - k)); }

static uint64_t rng_next(rng_t *r) {
  uint64_t *s = r->s;
  uint64_t res = rotl(s[1] * 5, 7) * 9;
  uint64_t t = s[1] << 17;
  s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3];
  s[2] ^= t; s[3] = rotl(s[3], 45);
  return res;
}

static double rng_uniform(rng_t *r) {  /* [0,1) */
  return (double)(rng_next(r) >> 11) * (1.0 / 9007199254740992.0);
}

static double rng_normal(rng_t *r) {
  double u1 = 1.0 - rng_uniform(r);  /* (0,1] */
  double u2 = rng_uniform(r);
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

static uint64_t rng_below(rng_t *r, uint64_t m) {  /* unbiased in [0,m) */
  uint64_t lim = UINT64_MAX - (UINT64_MAX % m);
  uint64_t x;
  do { x = rng_next(r); } while (x >= lim);
  return x % m;
}

void bn_netgen_default_params(bn_netgen_params *p) {
  p->L = 16;
  p->depth = 9;
  p->n_trees = 1;
  p->seed = 0;
  p->p_art = 13.3e3;
  p->p_ven = 1.33e3;
  p->mu = 3.5e-3;
  p->spacing = 50e-6;
  p->cap_r = 4e-6;
  p->cap_sigma = 0.2;
  p->root_r = 500e-6;
  p->boundary_roots = 0;
}

/* boundary_roots: tree t is rooted at a distinct random node of the
 * lattice's boundary faces; its `leaves` targets are the first lattice nodes
 * not taken by an earlier tree in breadth-first order from that node
 * (neighbours visited -x, +x, -y, +y, -z, +z).  target[] receives them tree
 * by tree.  Returns 0, or -1 on allocation failure. */
static int boundary_rooted_targets(long long L, int ntrees, long long leaves, uint64_t seed, int *target) {
  const long long NL = L * L * L;
  unsigned char *taken = (unsigned char *)calloc((size_t)NL, 1);
  int *stamp = (int *)malloc((size_t)NL * sizeof(int));
  int *queue = (int *)malloc((size_t)NL * sizeof(int));
  long long nb = 0;
  for (long long id = 0; id < NL; id++) {
    const long long x = id % L, y = (id / L) % L, z = id / (L * L);
    nb += (x == 0 || x == L - 1 || y == 0 || y == L - 1 || z == 0 || z == L - 1);
  }
  int *face = (int *)malloc((size_t)(nb > 0 ? nb : 1) * sizeof(int));
  unsigned char *root_used = (unsigned char *)calloc((size_t)(nb > 0 ? nb : 1), 1);
  if (!taken || !stamp || !queue || !face || !root_used || nb < ntrees) {
    free(taken); free(stamp); free(queue); free(face); free(root_used);
    return -1;
  }
  nb = 0;
  for (long long id = 0; id < NL; id++) {
    const long long x = id % L, y = (id / L) % L, z = id / (L * L);
    if (x == 0 || x == L - 1 || y == 0 || y == L - 1 || z == 0 || z == L - 1) face[nb++] = (int)id;
  }
  for (long long i = 0; i < NL; i++) stamp[i] = -1;
  rng_t r;
  rng_seed(&r, seed, 9);
  const long long stride[3] = {1, L, L * L};
  long long k = 0;
  for (int t = 0; t < ntrees; t++) {
    long long f;
    do { f = (long long)rng_below(&r, (uint64_t)nb); } while (root_used[f]);
    root_used[f] = 1;
    long long head = 0, tail = 0, got = 0;
    queue[tail++] = face[f];
    stamp[face[f]] = t;
    while (head < tail && got < leaves) {
      const int id = queue[head++];
      if (!taken[id]) {
        taken[id] = 1;
        target[k++] = id;
        got++;
      }
      const long long c[3] = {id % L, (id / L) % L, id / (L * L)};
      for (int dir = 0; dir < 3; dir++)
        for (int sgn = -1; sgn <= 1; sgn += 2) {
          const long long cc = c[dir] + sgn;
          if (cc < 0 || cc >= L) continue;
          const long long nid = id + sgn * stride[dir];
          if (stamp[nid] == t) continue;
          stamp[nid] = t;
          queue[tail++] = (int)nid;
        }
    }
  }
  free(taken); free(stamp); free(queue); free(face); free(root_used);
  return 0;
}

void bn_netgen_conductance(const bn_network *net, const double *radius,
                           double mu, double *g) {
  const double c = 3.14159265358979323846 / (8.0 * mu);
  for (int e = 0; e < net->n_edges; e++) {
    double r2 = radius[e] * radius[e];
    g[e] = c * r2 * r2 / net->length[e];
  }
}

void bn_netgen_assemble(const bn_network *net, const double *g,
                        const double *p_dir, double *val, double *b) {
  const double *pd = p_dir ? p_dir : net->p_dirichlet;
  memset(val, 0, (size_t)net->nnz * sizeof(double));
  memset(b, 0, (size_t)net->n * sizeof(double));
  for (int e = 0; e < net->n_edges; e++) {
    int u = net->eu[e], v = net->ev[e];
    double ge = g[e];
    if (u >= 0) val[net->diag_pos[u]] += ge;
    if (v >= 0) val[net->diag_pos[v]] += ge;
    if (u >= 0 && v >= 0) {
      val[net->pos_uv[e]] -= ge;
      val[net->pos_vu[e]] -= ge;
    } else if (u >= 0) {
      b[u] += ge * pd[-1 - v];
    } else if (v >= 0) {
      b[v] += ge * pd[-1 - u];
    }
  }
}

void bn_netgen_free(bn_network *net) {
  free(net->ia); free(net->ja); free(net->val); free(net->b);
  free(net->eu); free(net->ev); free(net->radius); free(net->length);
  free(net->g); free(net->pos_uv); free(net->pos_vu); free(net->diag_pos);
  free(net->p_dirichlet); free(net->dirichlet_is_art); free(net->node_unknown);
  memset(net, 0, sizeof(*net));
}

int bn_netgen_build(const bn_netgen_params *p, bn_network *net) {
  memset(net, 0, sizeof(*net));
  const int L = p->L, d = p->depth, T = p->n_trees;
  if (L < 2 || d < 2 || T < 1 || d > 24) return -1;
  const long long NL = (long long)L * L * L;
  const long long tree_nodes = (1LL << d) - 1;
  const long long leaves = 1LL << (d - 1);
  if (2LL * T * leaves > NL) return -1;
  const long long n_nodes = NL + 2LL * T * tree_nodes;
  const long long lat_edges = 3LL * L * L * (L - 1);
  const long long n_edges = lat_edges + 2LL * T * (tree_nodes - 1 + leaves);
  if (n_nodes > 2000000000LL || n_edges > 1000000000LL) return -1;

  net->n_nodes = (int)n_nodes;
  net->n_edges = (int)n_edges;
  net->n_dirichlet = 2 * T;
  net->n = (int)(n_nodes - 2LL * T);

  net->eu = (int *)malloc((size_t)n_edges * sizeof(int));
  net->ev = (int *)malloc((size_t)n_edges * sizeof(int));
  net->radius = (double *)malloc((size_t)n_edges * sizeof(double));
  net->length = (double *)malloc((size_t)n_edges * sizeof(double));
  net->g = (double *)malloc((size_t)n_edges * sizeof(double));
  net->pos_uv = (int *)malloc((size_t)n_edges * sizeof(int));
  net->pos_vu = (int *)malloc((size_t)n_edges * sizeof(int));
  net->p_dirichlet = (double *)malloc((size_t)2 * T * sizeof(double));
  net->dirichlet_is_art = (int *)malloc((size_t)2 * T * sizeof(int));
  net->node_unknown = (int *)malloc((size_t)n_nodes * sizeof(int));
  if (!net->eu || !net->ev || !net->radius || !net->length || !net->g ||
      !net->pos_uv || !net->pos_vu || !net->p_dirichlet ||
      !net->dirichlet_is_art || !net->node_unknown) {
    bn_netgen_free(net);
    return -1;
  }

  /* original id -> unknown id (roots are Dirichlet) */
  {
    int u = 0, dct = 0;
    for (long long id = 0; id < n_nodes; id++) {
      int is_root = 0;
      if (id >= NL) is_root = ((id - NL) % tree_nodes) == 0;
      if (is_root) {
        int t = (int)((id - NL) / tree_nodes);
        net->node_unknown[id] = -1 - dct;
        net->dirichlet_is_art[dct] = (t < T);
        net->p_dirichlet[dct] = (t < T) ? p->p_art : p->p_ven;
        dct++;
      } else {
        net->node_unknown[id] = u++;
      }
    }
  }
  const int *nu = net->node_unknown;

  /* lattice segments */
  long long e = 0;
  {
    rng_t r;
    rng_seed(&r, p->seed, 1);
    const long long stride[3] = {1, L, (long long)L * L};
    for (long long id = 0; id < NL; id++) {
      long long x = id % L, y = (id / L) % L, z = id / ((long long)L * L);
      long long c[3] = {x, y, z};
      for (int dir = 0; dir < 3; dir++) {
        if (c[dir] + 1 >= L) continue;
        net->eu[e] = nu[id];
        net->ev[e] = nu[id + stride[dir]];
        net->radius[e] = p->cap_r * exp(p->cap_sigma * rng_normal(&r));
        net->length[e] = p->spacing;
        e++;
      }
    }
  }

  /* leaf-attachment targets: distinct lattice nodes (partial Fisher-Yates),
   * or the nodes nearest to each tree's boundary-face root */
  long long n_attach = 2LL * T * leaves;
  int *target = (int *)malloc((size_t)n_attach * sizeof(int));
  if (p->boundary_roots) {
    if (!target || boundary_rooted_targets(L, 2 * T, leaves, p->seed, target)) {
      free(target);
      bn_netgen_free(net);
      return -1;
    }
  } else {
    int *perm = (int *)malloc((size_t)NL * sizeof(int));
    if (!perm || !target) { free(perm); free(target); bn_netgen_free(net); return -1; }
    for (long long i = 0; i < NL; i++) perm[i] = (int)i;
    rng_t r;
    rng_seed(&r, p->seed, 8);
    for (long long k = 0; k < n_attach; k++) {
      long long j = k + (long long)rng_below(&r, (uint64_t)(NL - k));
      int t = perm[k]; perm[k] = perm[j]; perm[j] = t;
      target[k] = perm[k];
    }
    free(perm);
  }

  /* trees */
  {
    double *tr = (double *)malloc((size_t)tree_nodes * sizeof(double));
    const double shrink = pow(2.0, -1.0 / 3.0);
    long long attach_k = 0;
    for (int t = 0; t < 2 * T; t++) {
      rng_t r;
      rng_seed(&r, p->seed, 16 + (uint64_t)t);
      const long long base = NL + (long long)t * tree_nodes;
      tr[0] = p->root_r;
      for (long long k = 1; k < tree_nodes; k++) {
        double f = 0.9 + 0.2 * rng_uniform(&r);
        tr[k] = tr[(k - 1) / 2] * shrink * f;
      }
      for (long long k = 1; k < tree_nodes; k++) {
        net->eu[e] = nu[base + (k - 1) / 2];
        net->ev[e] = nu[base + k];
        net->radius[e] = tr[k];
        net->length[e] = 20.0 * tr[k];
        e++;
      }
      for (long long k = tree_nodes - leaves; k < tree_nodes; k++) {
        net->eu[e] = nu[base + k];
        net->ev[e] = nu[target[attach_k++]];
        net->radius[e] = tr[k];
        net->length[e] = 20.0 * tr[k];
        e++;
      }
    }
    free(tr);
  }
  free(target);

  bn_netgen_conductance(net, net->radius, p->mu, net->g);

  /* CSR pattern: count, fill, sort rows */
  const int n = net->n;
  net->ia = (int *)calloc((size_t)n + 1, sizeof(int));
  net->diag_pos = (int *)malloc((size_t)n * sizeof(int));
  net->b = (double *)malloc((size_t)n * sizeof(double));
  for (int i = 0; i < n; i++) net->ia[i + 1] = 1;
  for (long long k = 0; k < n_edges; k++) {
    int u = net->eu[k], v = net->ev[k];
    if (u >= 0 && v >= 0) { net->ia[u + 1]++; net->ia[v + 1]++; }
  }
  for (int i = 0; i < n; i++) net->ia[i + 1] += net->ia[i];
  net->nnz = net->ia[n];
  net->ja = (int *)malloc((size_t)net->nnz * sizeof(int));
  net->val = (double *)malloc((size_t)net->nnz * sizeof(double));
  int *fill = (int *)malloc((size_t)n * sizeof(int));
  for (int i = 0; i < n; i++) {
    fill[i] = net->ia[i];
    net->ja[fill[i]++] = i;
  }
  for (long long k = 0; k < n_edges; k++) {
    int u = net->eu[k], v = net->ev[k];
    if (u >= 0 && v >= 0) {
      net->ja[fill[u]++] = v;
      net->ja[fill[v]++] = u;
    }
  }
  free(fill);
  for (int i = 0; i < n; i++) {  /* rows are short: insertion sort */
    int lo = net->ia[i], hi = net->ia[i + 1];
    for (int a = lo + 1; a < hi; a++) {
      int key = net->ja[a], q = a - 1;
      while (q >= lo && net->ja[q] > key) { net->ja[q + 1] = net->ja[q]; q--; }
      net->ja[q + 1] = key;
    }
    for (int a = lo; a < hi; a++)
      if (net->ja[a] == i) net->diag_pos[i] = a;
  }
  /* value positions of each segment (binary search in sorted rows) */
  for (long long k = 0; k < n_edges; k++) {
    int u = net->eu[k], v = net->ev[k];
    net->pos_uv[k] = net->pos_vu[k] = -1;
    if (u < 0 || v < 0) continue;
    for (int side = 0; side < 2; side++) {
      int row = side ? v : u, col = side ? u : v;
      int lo = net->ia[row], hi = net->ia[row + 1] - 1;
      while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (net->ja[mid] < col) lo = mid + 1; else hi = mid;
      }
      if (side) net->pos_vu[k] = lo; else net->pos_uv[k] = lo;
    }
  }
  bn_netgen_assemble(net, net->g, NULL, net->val, net->b);
  return 0;
}
