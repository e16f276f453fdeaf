// rebuild.hip - value-only rebuild of a hierarchy on the device (same C/F
// split, strength and interpolation patterns; new operator values), for
// gfx950.
//
// Every value is formed by one thread in the order the host set-up uses
// (C_amg1/src/rs_classical.c standard interpolation, the row-by-row
// Galerkin product and (A + A^T)/2 of C_amg1/src/build_hierarchy.c), and
// this file is compiled without floating-point contraction, so the values
// are bit-identical to the host's amg_rebuild_values.  Lookups that the
// host does through scratch arrays of length n (owner/slot, column
// stamps) are binary searches in per-row sorted copies of the patterns.
#include <hip/hip_runtime.h>
#include <math.h>

#include "bn_kernels.h"

#pragma clang fp contract(off)

#define BLOCK 256

namespace {

inline dim3 g1(long long n) { return dim3((unsigned)((n + BLOCK - 1) / BLOCK)); }

// position of column c in a row's sorted column list cols[lo, hi), or -1
__device__ __forceinline__ int find_col(const int *__restrict__ cols, int lo, int hi, int c) {
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    const int v = cols[mid];
    if (v == c) return mid;
    if (v < c) lo = mid + 1;
    else hi = mid;
  }
  return -1;
}

// out[q] = (v[q] + v[tpos[q]]) * 0.5  (symmetrisation of a structurally
// symmetric matrix: tpos[q] holds the position of the transposed entry)
__global__ void k_rb_symmetrize(int nnz, const int *__restrict__ tpos, const double *__restrict__ v,
                                double *__restrict__ out) {
  const int q = blockIdx.x * BLOCK + threadIdx.x;
  if (q >= nnz) return;
  out[q] = (v[q] + v[tpos[q]]) * 0.5;
}

// out[q] = src[map[q]]
__global__ void k_rb_gather(int n, const int *__restrict__ map, const double *__restrict__ src,
                            double *__restrict__ out, float *__restrict__ outs) {
  const int q = blockIdx.x * BLOCK + threadIdx.x;
  if (q >= n) return;
  const double v = src[map[q]];
  if (out) out[q] = v;
  if (outs) outs[q] = (float)v;
}

__global__ void k_rb_to_float(int n, const double *__restrict__ src, float *__restrict__ out) {
  const int q = blockIdx.x * BLOCK + threadIdx.x;
  if (q < n) out[q] = (float)src[q];
}

// per row: 1/a_ii (1.0 if absent; the last diagonal entry wins), the l1
// inverse 1/sum|a_ij| (1.0 for an empty row), and optionally the host
// set-up's l1 row norms
__global__ void k_rb_row_scalars(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                 const double *__restrict__ a, double *__restrict__ dinv, float *__restrict__ dinvs,
                                 double *__restrict__ l1inv, float *__restrict__ l1invs, double *__restrict__ l1) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  double d = 1.0, s = 0.0;
  for (int k = ia[i]; k < ia[i + 1]; k++) {
    if (ja[k] == i) d = a[k];
    s += fabs(a[k]);
  }
  const double di = 1.0 / d, li = s > 0.0 ? 1.0 / s : 1.0;
  if (dinv) dinv[i] = di;
  if (dinvs) dinvs[i] = (float)di;
  if (l1inv) l1inv[i] = li;
  if (l1invs) l1invs[i] = (float)li;
  if (l1) l1[i] = s;
}

// GS row copy: row r updates level row map[r]; its diagonal is the entry in
// column map[r] (the last one wins), 1.0 if absent
__global__ void k_rb_dinv_map(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                              const double *__restrict__ a, const int *__restrict__ map, double *__restrict__ dinv,
                              float *__restrict__ dinvs) {
  const int r = blockIdx.x * BLOCK + threadIdx.x;
  if (r >= n) return;
  const int self = map[r];
  double d = 1.0;
  for (int k = ia[r]; k < ia[r + 1]; k++)
    if (ja[k] == self) d = a[k];
  dinv[r] = 1.0 / d;
  dinvs[r] = (float)(1.0 / d);
}

// Standard interpolation values on a fixed pattern, one thread per fine
// row, in the order of rs_standard_interpolation_values
// (C_amg1/src/rs_classical.c).  S rows are given sorted (scol) with the P
// slot of each strong neighbour (sslot: position in P.val for a C point,
// -1 for an F point).
__global__ void k_rb_interp(int n, const int *__restrict__ aia, const int *__restrict__ aja,
                            const double *__restrict__ a, const int *__restrict__ sia, const int *__restrict__ scol,
                            const int *__restrict__ sslot, const int *__restrict__ cpt, const int *__restrict__ pia,
                            double *__restrict__ pv) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  const int p0 = pia[i], p1 = pia[i + 1];
  if (cpt[i]) {
    pv[p0] = 1.0;
    return;
  }
  const int s0 = sia[i], s1 = sia[i + 1];
  for (int p = p0; p < p1; p++) pv[p] = 0.0;
  double diag = 0.0, weak = 0.0;
  for (int k = aia[i]; k < aia[i + 1]; k++) {
    const int j = aja[k];
    const double v = a[k];
    if (j == i) {
      diag = v;
      continue;
    }
    const int at = find_col(scol, s0, s1, j);
    if (at >= 0) {
      const int slot = sslot[at];
      if (slot >= 0) pv[slot] += v;
      continue;  // a strong F neighbour is handled below
    }
    weak += v;
  }
  double dhat = diag + weak;
  for (int k = aia[i]; k < aia[i + 1]; k++) {
    const int f = aja[k];
    if (f == i) continue;
    const int at = find_col(scol, s0, s1, f);
    if (at < 0 || sslot[at] >= 0) continue;  // not a strong F neighbour
    const double aik = a[k];
    double den = 0.0;
    for (int q = aia[f]; q < aia[f + 1]; q++) {
      const int m = find_col(scol, s0, s1, aja[q]);
      if (m >= 0 && sslot[m] >= 0) den += a[q];
    }
    if (fabs(den) < 1e-15) {
      dhat += aik;
    } else {
      for (int q = aia[f]; q < aia[f + 1]; q++) {
        const int m = find_col(scol, s0, s1, aja[q]);
        if (m >= 0 && sslot[m] >= 0) pv[sslot[m]] += aik * a[q] / den;
      }
    }
  }
  if (fabs(dhat) > 1e-15)
    for (int p = p0; p < p1; p++) pv[p] /= -dhat;
}

// Galerkin product values on the fixed pattern of the coarse operator, one
// thread per coarse row, in the order of the host's row-by-row product
// (R row, then A row, then P row).  ccol/cpos: the coarse row's columns
// sorted, with their positions in the stored row.
__global__ void k_rb_rap(int nc, const int *__restrict__ ria, const int *__restrict__ rja,
                         const double *__restrict__ rv, const int *__restrict__ aia, const int *__restrict__ aja,
                         const double *__restrict__ av, const int *__restrict__ pia, const int *__restrict__ pja,
                         const double *__restrict__ pv, const int *__restrict__ cia, const int *__restrict__ ccol,
                         const int *__restrict__ cpos, double *__restrict__ out) {
  const int ic = blockIdx.x * BLOCK + threadIdx.x;
  if (ic >= nc) return;
  const int c0 = cia[ic], c1 = cia[ic + 1];
  for (int q = c0; q < c1; q++) out[q] = 0.0;
  for (int q1 = ria[ic]; q1 < ria[ic + 1]; q1++) {
    const int i1 = rja[q1];
    const double r = rv[q1];
    for (int q2 = aia[i1]; q2 < aia[i1 + 1]; q2++) {
      const int i2 = aja[q2];
      const double ra = r * av[q2];
      for (int q3 = pia[i2]; q3 < pia[i2 + 1]; q3++) {
        const int at = find_col(ccol, c0, c1, pja[q3]);
        if (at >= 0) out[cpos[at]] += ra * pv[q3];  // always found on the host's pattern
      }
    }
  }
}

}  // namespace

extern "C" {

void bn_rb_symmetrize(int nnz, const int *tpos, const double *v, double *out, hipStream_t st) {
  if (nnz > 0) hipLaunchKernelGGL(k_rb_symmetrize, g1(nnz), dim3(BLOCK), 0, st, nnz, tpos, v, out);
}

void bn_rb_gather(int n, const int *map, const double *src, double *out, float *outs, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_rb_gather, g1(n), dim3(BLOCK), 0, st, n, map, src, out, outs);
}

void bn_rb_to_float(int n, const double *src, float *out, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_rb_to_float, g1(n), dim3(BLOCK), 0, st, n, src, out);
}

void bn_rb_row_scalars(int n, const int *ia, const int *ja, const double *a, double *dinv, float *dinvs,
                       double *l1inv, float *l1invs, double *l1, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_rb_row_scalars, g1(n), dim3(BLOCK), 0, st, n, ia, ja, a, dinv, dinvs, l1inv, l1invs, l1);
}

void bn_rb_dinv_map(int n, const int *ia, const int *ja, const double *a, const int *map, double *dinv, float *dinvs,
                    hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_rb_dinv_map, g1(n), dim3(BLOCK), 0, st, n, ia, ja, a, map, dinv, dinvs);
}

void bn_rb_interp(int n, const int *aia, const int *aja, const double *a, const int *sia, const int *scol,
                  const int *sslot, const int *cpt, const int *pia, double *pv, hipStream_t st) {
  if (n > 0)
    hipLaunchKernelGGL(k_rb_interp, g1(n), dim3(BLOCK), 0, st, n, aia, aja, a, sia, scol, sslot, cpt, pia, pv);
}

void bn_rb_rap(int nc, const int *ria, const int *rja, const double *rv, const int *aia, const int *aja,
               const double *av, const int *pia, const int *pja, const double *pv, const int *cia, const int *ccol,
               const int *cpos, double *out, hipStream_t st) {
  if (nc > 0)
    hipLaunchKernelGGL(k_rb_rap, g1(nc), dim3(BLOCK), 0, st, nc, ria, rja, rv, aia, aja, av, pia, pja, pv, cia,
                       ccol, cpos, out);
}

}  // extern "C"
