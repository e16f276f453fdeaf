/*
 * netgen.h - deterministic synthetic vessel-network generator.
 *
 * A vessel network is a 3-D capillary lattice (L^3 nodes, 6-neighbour
 * segments) plus binary arterial and venous trees whose leaves attach to
 * distinct random lattice nodes.  Every segment carries a Poiseuille
 * conductance g = pi r^4 / (8 mu len).  The tree roots are Dirichlet nodes;
 * eliminating them leaves the SPD weighted graph Laplacian
 *     A = C_I^T W C_I,   b = -C_I^T W C_B p_B
 * which is the same algebra as Matlab_amg/read_nwk.m:23-37 (there C is the
 * face-node incidence, W = diag(g)).  Unknowns are the non-Dirichlet nodes
 * in ascending original id; CSR rows are sorted by column.
 */
#ifndef BN_NETGEN_H
#define BN_NETGEN_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
  int      L;          /* lattice edge (nodes per axis)                    */
  int      depth;      /* binary-tree depth: 2^depth-1 nodes per tree      */
  int      n_trees;    /* trees per kind (arterial and venous each)        */
  uint64_t seed;       /* master seed; every random quantity has a stream  */
  double   p_art;      /* arterial root pressure [Pa]                      */
  double   p_ven;      /* venous root pressure [Pa]                        */
  double   mu;         /* viscosity [Pa s]                                 */
  double   spacing;    /* lattice segment length [m]                       */
  double   cap_r;      /* capillary median radius [m]                      */
  double   cap_sigma;  /* lognormal sigma of capillary radius              */
  double   root_r;     /* tree root radius [m]                             */
  int      boundary_roots; /* 0: every leaf attaches to a distinct uniformly
                          random lattice node; 1: each tree is rooted at a
                          distinct random boundary-face lattice node and its
                          leaves attach to the free lattice nodes nearest to
                          it (breadth-first from that node), so every tree
                          perfuses the region it enters                    */
} bn_netgen_params;

typedef struct {
  int n_nodes;      /* all nodes, Dirichlet included                      */
  int n;            /* unknowns                                           */
  int n_edges;      /* segments                                           */
  int nnz;          /* nnz of A                                           */
  int n_dirichlet;  /* Dirichlet (root) nodes                             */
  /* CSR of A (rows sorted), and b for the stored Dirichlet pressures */
  int    *ia, *ja;
  double *val, *b;
  /* per segment: endpoints as unknown index, or -1-d for Dirichlet node d */
  int    *eu, *ev;
  double *radius, *length, *g;
  /* position in val[] of the (eu,ev) and (ev,eu) entries, -1 if an
   * endpoint is Dirichlet; diag position of each unknown */
  int    *pos_uv, *pos_vu, *diag_pos;
  double *p_dirichlet;   /* pressure of each Dirichlet node               */
  int    *dirichlet_is_art; /* 1 arterial root, 0 venous root            */
  int    *node_unknown;  /* original node id -> unknown index or -1       */
} bn_network;

void bn_netgen_default_params(bn_netgen_params *p);
/* Builds the network.  Returns 0, or -1 on bad parameters (more leaves than
 * lattice nodes, L < 2, depth < 2) or allocation failure. */
int  bn_netgen_build(const bn_netgen_params *p, bn_network *net);
/* Re-assembles val[] and b[] from per-segment conductances g[] on the same
 * pattern (g may alias net->g).  p_dir overrides the Dirichlet pressures if
 * non-NULL. */
void bn_netgen_assemble(const bn_network *net, const double *g,
                        const double *p_dir, double *val, double *b);
/* g = pi r^4 / (8 mu len) for every segment */
void bn_netgen_conductance(const bn_network *net, const double *radius,
                           double mu, double *g);
void bn_netgen_free(bn_network *net);

/* ---- vessel-network case files (Matlab_amg/caseReaderMJ.m:1-33) ----
 * <base>.fMx (faces: columns 2, 3 = 1-based end points), <base>.pMx (x y z
 * per point), <base>.dia (face diameters, may be empty), <base>.BC (node,
 * type, pressure), numeric text as MATLAB load() reads it.  Indices here
 * are 0-based. */
typedef struct {
  int np, nf, nbc;
  double *pts;     /* np x 3 */
  int *face;       /* nf x 2 end points */
  double *dia;     /* nf, or NULL when the file is empty */
  int *bc_node;    /* nbc boundary nodes */
  double *bc_val;  /* nbc pressures */
} bn_nwk;

/* the interior pressure system of Matlab_amg/read_nwk.m:15-37 */
typedef struct {
  int n, nnz;
  int *ia, *ja;    /* CSR, columns ascending per row */
  double *val;
  double *b;       /* n right-hand side */
  int *interior;   /* n: 1-based node of each unknown, ascending */
} bn_nwk_sys;

/* Returns 0, or -1 (bn_last_error() says why). */
int bn_nwk_read(const char *base, bn_nwk *w);
void bn_nwk_free(bn_nwk *w);
/* A = C_I' W C_I, b = -C_I' W C_B p_B with diffusive Cartesian face
 * conductances (makeDiffusiveConductance.m); duplicate entries summed,
 * exact zeros dropped.  Rejects a boundary node listed twice and faces
 * with no positive axis extent (MATLAB would give NaN).  Returns 0 or -1. */
int bn_nwk_system(const bn_nwk *w, bn_nwk_sys *s);
void bn_nwk_sys_free(bn_nwk_sys *s);
/* solution files: one %.16e per line (Matlab_amg/test_amg.m:68-72), and
 * "%d %.16e" with 1-based node indices (label_solution.m:42-47) */
int bn_write_pressures(const char *path, const double *p, int n);
int bn_write_labelled(const char *path, const int *node, const double *p, int n);

#ifdef __cplusplus
}
#endif
#endif
