// rebuild.hip - value-only rebuild of a hierarchy on the device (same C/F
// split, strength and interpolation patterns; new operator values), for
// gfx950.
//
// Every value is a sum formed in the order the host set-up uses
// (C_amg1/src/rs_classical.c standard interpolation, the row-by-row
// Galerkin product and (A + A^T)/2 of C_amg1/src/build_hierarchy.c), and
// this file is compiled without floating-point contraction, so the values
// are bit-identical to the host's amg_rebuild_values.  The terms of each
// interpolation weight, and of each Galerkin entry on the coarse levels,
// come from lists built once from the patterns, so a value is one sum
// over its own list however long the rows of hub nodes grow; the light
// levels' product runs one thread per coarse row, with binary searches in
// per-row sorted copies of the coarse pattern where the host uses column
// stamps of length n.
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
// symmetric matrix: tpos[q] holds the position of the transposed entry).
// diag0: v holds Galerkin sums that start from -0.0, so that an
// off-diagonal entry is its first term exactly, as the host product stores
// it; the host starts the diagonal from +0.0 instead, which differs only
// in the sign of a zero sum: such a diagonal is taken as +0.0.
__global__ void k_rb_symmetrize(int nnz, const int *__restrict__ tpos, const double *__restrict__ v,
                                double *__restrict__ out, int diag0) {
  const int q = blockIdx.x * BLOCK + threadIdx.x;
  if (q >= nnz) return;
  const int t = tpos[q];
  if (diag0 && t == q) {
    const double d = v[q] == 0.0 ? 0.0 : v[q];
    out[q] = (d + d) * 0.5;
    return;
  }
  out[q] = (v[q] + v[t]) * 0.5;
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

// Standard interpolation values on a fixed pattern
// (C_amg1/src/rs_classical.c, rs_standard_interpolation_values), from
// contribution lists built once from the host hierarchy, so that each
// value is one thread's sum in the host's order:
//  - den of every pair (F row i, strong F neighbour f = column of A entry
//    pair_k): the A entries of row f whose column is a strong C point of
//    i, in row order;
//  - dhat of every F row: the diagonal (the last one wins) plus the weak
//    entries in row order, then a_if for each strong F neighbour whose den
//    is below 1e-15, in row order;
//  - every P entry: first its direct entries of row i (q < 0: a[k]), then
//    a_if a_fj / den for the strong F neighbours in row order (a = pair).
// cls per A entry: 0 strong, 1 weak, 2 diagonal.
__global__ void k_rb_den(int npairs, const int *__restrict__ doff, const int *__restrict__ dq,
                         const double *__restrict__ a, double *__restrict__ den) {
  const int p = blockIdx.x * BLOCK + threadIdx.x;
  if (p >= npairs) return;
  double d = 0.0;
  for (int t = doff[p]; t < doff[p + 1]; t++) d += a[dq[t]];
  den[p] = d;
}

__global__ void k_rb_dhat(int n, const int *__restrict__ aia, const double *__restrict__ a,
                          const unsigned char *__restrict__ cls, const int *__restrict__ cpt,
                          const int *__restrict__ rp, const int *__restrict__ pair_k, const double *__restrict__ den,
                          double *__restrict__ dhat) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n || cpt[i]) return;
  double diag = 0.0, weak = 0.0;
  for (int k = aia[i]; k < aia[i + 1]; k++) {
    const int c = cls[k];
    if (c == 2) diag = a[k];
    else if (c == 1) weak += a[k];
  }
  double h = diag + weak;
  for (int p = rp[i]; p < rp[i + 1]; p++)
    if (fabs(den[p]) < 1e-15) h += a[pair_k[p]];
  dhat[i] = h;
}

__global__ void k_rb_pval(int nnzp, const int *__restrict__ prow, const int *__restrict__ cpt,
                          const int *__restrict__ pia, const int *__restrict__ poff, const int *__restrict__ pa,
                          const int *__restrict__ pq, const int *__restrict__ pair_k, const double *__restrict__ a,
                          const double *__restrict__ den, const double *__restrict__ dhat, double *__restrict__ pv) {
  const int e = blockIdx.x * BLOCK + threadIdx.x;
  if (e >= nnzp) return;
  const int i = prow[e];
  if (cpt[i]) {
    if (e == pia[i]) pv[e] = 1.0;
    return;
  }
  double v = 0.0;
  for (int t = poff[e]; t < poff[e + 1]; t++) {
    const int q = pq[t];
    if (q < 0) {
      v += a[pa[t]];
    } else {
      const int pr = pa[t];
      const double d = den[pr];
      if (!(fabs(d) < 1e-15)) v += a[pair_k[pr]] * a[q] / d;
    }
  }
  const double h = dhat[i];
  if (fabs(h) > 1e-15) v /= -h;
  pv[e] = v;
}

// Galerkin product values on the fixed pattern of the coarse operator, one
// thread per coarse row, in the order of the host's row-by-row product
// (R row, then A row, then P row), from -0.0 (see k_rb_symmetrize).
// ccol/cpos: the coarse row's columns sorted, with their positions in the
// stored row.
__global__ void k_rb_rap(int nc, const int *__restrict__ ria, const int *__restrict__ rja,
                         const double *__restrict__ rv, const int *__restrict__ aia, const int *__restrict__ aja,
                         const double *__restrict__ av, const int *__restrict__ pia, const int *__restrict__ pja,
                         const double *__restrict__ pv, const int *__restrict__ cia, const int *__restrict__ ccol,
                         const int *__restrict__ cpos, double *__restrict__ out) {
  const int ic = blockIdx.x * BLOCK + threadIdx.x;
  if (ic >= nc) return;
  const int c0 = cia[ic], c1 = cia[ic + 1];
  for (int q = c0; q < c1; q++) out[q] = -0.0;
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


// Galerkin product values from per-entry lists: entry e of the coarse
// operator (stored order) is the sum over its triples t in toff[e] ..
// toff[e+1], in the host product's (R row, A row, P row) order, of
// (r a) p, from -0.0 (see k_rb_symmetrize).  Lists longer than BN_RB_LONG
// go to k_rb_rap_pull_wave.
__global__ void k_rb_rap_pull(int nnz, const int *__restrict__ toff, const int *__restrict__ t1,
                              const int *__restrict__ t2, const int *__restrict__ t3, const double *__restrict__ rv,
                              const double *__restrict__ av, const double *__restrict__ pv, double *__restrict__ out) {
  const int e = blockIdx.x * BLOCK + threadIdx.x;
  if (e >= nnz) return;
  const int lo = toff[e], hi = toff[e + 1];
  if (hi - lo > BN_RB_LONG) return;
  double acc = -0.0;
  for (int t = lo; t < hi; t++) {
    const double ra = rv[t1[t]] * av[t2[t]];
    acc += ra * pv[t3[t]];
  }
  out[e] = acc;
}

__device__ __forceinline__ double lane_value(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// one wavefront per long entry: the lanes form 64 terms at a time and the
// running sum takes them in lane order, so the sum is the sequential one
__global__ void __launch_bounds__(BLOCK) k_rb_rap_pull_wave(int nlong, const int *__restrict__ elist,
                                                            const int *__restrict__ toff, const int *__restrict__ t1,
                                                            const int *__restrict__ t2, const int *__restrict__ t3,
                                                            const double *__restrict__ rv, const double *__restrict__ av,
                                                            const double *__restrict__ pv, double *__restrict__ out) {
  const int w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (BLOCK / 64) + threadIdx.x / 64));
  const int lane = threadIdx.x & 63;
  if (w >= nlong) return;
  const int e = __builtin_amdgcn_readfirstlane(elist[w]);
  const int lo = __builtin_amdgcn_readfirstlane(toff[e]), hi = __builtin_amdgcn_readfirstlane(toff[e + 1]);
  double acc = -0.0;
  for (int base = lo; base < hi; base += 64) {
    const int t = base + lane;
    double c = 0.0;
    if (t < hi) {
      const double ra = rv[t1[t]] * av[t2[t]];
      c = ra * pv[t3[t]];
    }
    const int cnt = hi - base < 64 ? hi - base : 64;
    for (int l = 0; l < cnt; l++) acc += lane_value(c, l);
  }
  if (lane == 0) out[e] = acc;
}


// ---- IC(0) of the fine operator (ichol_compute, setup.c, entry for
// entry): rows of one level of the lower factor's schedule, one thread per
// row.  Rows are strictly ascending with the diagonal last (checked when
// the plan is built), so "column j is in row i" of the host's column
// stamps is a merge of the two sorted rows, and the row's own values
// serve as the host's work row. ----
__global__ void k_ic0_rows(int cnt, const int *__restrict__ rows, const int *__restrict__ lia,
                           const int *__restrict__ lja, double *__restrict__ lv) {
  const int x = blockIdx.x * BLOCK + threadIdx.x;
  if (x >= cnt) return;
  const int i = rows[x], lo = lia[i], hi = lia[i + 1];
  for (int q = lo; q < hi; q++) {
    const int k = lja[q];
    if (k >= i) continue;
    double v = lv[q];
    int at = lo;
    const int k1 = lia[k + 1];
    for (int t = lia[k]; t < k1; t++) {
      const int j = lja[t];
      if (j >= k) break;
      while (at < hi && lja[at] < j) at++;
      if (at < hi && lja[at] == j) v -= lv[at] * lv[t];
    }
    lv[q] = v / lv[k1 - 1];
  }
  double d = lv[hi - 1];
  for (int q = lo; q < hi - 1; q++) d -= lv[q] * lv[q];
  if (d <= 0.0) d = 1e-14;
  lv[hi - 1] = sqrt(d);
}

// ---- dense inverse of the coarsest operator (dense.c's bn_dense_inverse,
// entry for entry) ----
// The factor is kept column-major (Lc: L(i,p) at Lc[p n + i]) for the
// factorisation and the forward sweep, and copied row-major (Lr) for the
// backward sweep, so that every load below is coalesced or uniform.

__global__ void k_dn_zero(long long m, double *__restrict__ x) {
  for (long long q = blockIdx.x * (long long)BLOCK + threadIdx.x; q < m; q += (long long)gridDim.x * BLOCK) x[q] = 0.0;
}

// dense copy of a CSR matrix, duplicates added in row order
__global__ void k_dn_scatter(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                             const double *__restrict__ a, double *__restrict__ Lc) {
  const int i = blockIdx.x * BLOCK + threadIdx.x;
  if (i >= n) return;
  for (int k = ia[i]; k < ia[i + 1]; k++) Lc[(size_t)ja[k] * n + i] += a[k];
}

// Cholesky factor in panels of DN_B columns.  Every entry takes the terms
// of dense.c's left-looking loop in the same order, p ascending:
//   L(i,k) = (A(i,k) - sum_{p<k} L(i,p) L(k,p)) / L(k,k)   (i > k)
//   L(k,k) = sqrt(A(k,k) - sum_{p<k} L(k,p)^2)  (pivot <= 0 -> 1e-300)
// the terms of earlier panels through the trailing updates (one panel's
// terms per update, in order), those of the entry's own panel in
// k_dn_diag (the DN_B x DN_B diagonal block) or k_dn_panel (the rows
// below it).
#define DN_B 32

__global__ void __launch_bounds__(64) k_dn_diag(int n, int q0, double *__restrict__ Lc, int *__restrict__ clamped) {
  __shared__ double D[DN_B][DN_B + 1];  // D[i][k] = L(q0 + i, q0 + k), i >= k
  const int t = threadIdx.x;
  const int bw = n - q0 < DN_B ? n - q0 : DN_B;
  for (int k = 0; k < bw; k++)
    if (t < bw && t >= k) D[t][k] = Lc[(size_t)(q0 + k) * n + q0 + t];
  __syncthreads();
  for (int k = 0; k < bw; k++) {
    if (t == k) {
      double d = D[k][k];
      for (int p = 0; p < k; p++) d -= D[k][p] * D[k][p];
      if (!(d > 0.0)) {
        d = 1e-300;
        if (clamped) *clamped = 1;
      }
      D[k][k] = sqrt(d);
    }
    __syncthreads();
    if (t > k && t < bw) {
      double v = D[t][k];
      for (int p = 0; p < k; p++) v -= D[t][p] * D[k][p];
      D[t][k] = v / D[k][k];
    }
    __syncthreads();
  }
  for (int k = 0; k < bw; k++)
    if (t < bw && t >= k) Lc[(size_t)(q0 + k) * n + q0 + t] = D[t][k];
}

// rows i >= q0 + DN_B of the panel's DN_B columns, one thread per row
__global__ void __launch_bounds__(256) k_dn_panel(int n, int q0, double *__restrict__ Lc) {
  __shared__ double D[DN_B][DN_B + 1];
  for (int e = threadIdx.x; e < DN_B * DN_B; e += 256) {
    const int i = e % DN_B, k = e / DN_B;
    if (i >= k) D[i][k] = Lc[(size_t)(q0 + k) * n + q0 + i];
  }
  __syncthreads();
  const int i = q0 + DN_B + blockIdx.x * 256 + threadIdx.x;
  if (i >= n) return;
  double r[DN_B];
#pragma unroll
  for (int k = 0; k < DN_B; k++) r[k] = Lc[(size_t)(q0 + k) * n + i];
#pragma unroll
  for (int k = 0; k < DN_B; k++) {
    double v = r[k];
#pragma unroll
    for (int p = 0; p < k; p++) v -= r[p] * D[k][p];
    r[k] = v / D[k][k];
  }
#pragma unroll
  for (int k = 0; k < DN_B; k++) Lc[(size_t)(q0 + k) * n + i] = r[k];
}

// trailing update after panel q0: A(i,k) -= L(i,q) L(k,q), q over the
// panel in order, for i >= k >= t0 = q0 + DN_B; 64 x 64 tiles of the lower
// triangle, tile index -> (bi, bk) with bk <= bi
__global__ void __launch_bounds__(256) k_dn_trail(int n, int q0, double *__restrict__ Lc) {
  __shared__ double Pi[64][DN_B + 1], Pk[64][DN_B + 1];
  const int t0 = q0 + DN_B;
  int bi = (int)((sqrt(8.0 * blockIdx.x + 1.0) - 1.0) * 0.5);
  while ((bi + 1) * (bi + 2) / 2 <= (int)blockIdx.x) bi++;
  while (bi * (bi + 1) / 2 > (int)blockIdx.x) bi--;
  const int bk = (int)blockIdx.x - bi * (bi + 1) / 2;
  const int i0 = t0 + bi * 64, k0 = t0 + bk * 64;
  for (int e = threadIdx.x; e < 64 * DN_B; e += 256) {
    const int r = e & 63, q = e >> 6;
    Pi[r][q] = (i0 + r < n) ? Lc[(size_t)(q0 + q) * n + i0 + r] : 0.0;
    Pk[r][q] = (k0 + r < n) ? Lc[(size_t)(q0 + q) * n + k0 + r] : 0.0;
  }
  __syncthreads();
  const int il = threadIdx.x & 63, i = i0 + il;
  if (i >= n) return;
  for (int j = 0; j < 16; j++) {
    const int kl = (threadIdx.x >> 6) + 4 * j, k = k0 + kl;
    if (k > i) continue;  // k >= n implies k > i
    double a = Lc[(size_t)k * n + i];
#pragma unroll
    for (int q = 0; q < DN_B; q++) a -= Pi[il][q] * Pk[kl][q];
    Lc[(size_t)k * n + i] = a;
  }
}

// Lr[i n + p] = Lc[p n + i] for p <= i (the lower triangle, row-major)
__global__ void k_dn_transpose(int n, const double *__restrict__ Lc, double *__restrict__ Lr) {
  __shared__ double tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;  // bx: column block p, by: row block i
  if (bx > by + 31) return;                              // strictly upper tile
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8 threads
  for (int r = ty; r < 32; r += 8) {
    const int p = bx + r, i = by + tx;
    if (p < n && i < n) tile[r][tx] = Lc[(size_t)p * n + i];
  }
  __syncthreads();
  for (int r = ty; r < 32; r += 8) {
    const int i = by + r, p = bx + tx;
    if (i < n && p < n && p <= i) Lr[(size_t)i * n + p] = tile[tx][r];
  }
}

// G columns of the inverse per block of DN_T threads; thread t keeps rows
// t, t + DN_T, ... (R of them) of the G right-hand sides in registers.  Step p: the owner of row p finishes y_p, the block reads it
// from LDS and takes y_p times column p (forward: L(i,p), i > p) or row p
// (backward: L(p,i), i < p) out of its rows, the order of dense.c.
#define DN_T 1024
template <int R, int DN_G>
__global__ void __launch_bounds__(DN_T) k_dn_solve(int n, const double *__restrict__ Lc, const double *__restrict__ Lr,
                                                 double *__restrict__ inv) {
  __shared__ double yb[2][DN_G];
  const int t = threadIdx.x, c0 = blockIdx.x * DN_G;
  double s[R][DN_G];
#pragma unroll
  for (int r = 0; r < R; r++)
#pragma unroll
    for (int g = 0; g < DN_G; g++) s[r][g] = (t + r * DN_T == c0 + g) ? 1.0 : 0.0;
  // forward: L y = e_c, columns c0 + g start at p = c0 + g
  for (int p = c0; p < n; p++) {
    const int buf = p & 1;
    if (t == (p & (DN_T - 1))) {
      const double d = Lc[(size_t)p * n + p];
#pragma unroll
      for (int r = 0; r < R; r++)
        if (r == p / DN_T) {
#pragma unroll
          for (int g = 0; g < DN_G; g++) {
            double y = 0.0;
            if (p >= c0 + g) {
              y = s[r][g] / d;
              s[r][g] = y;
            }
            yb[buf][g] = y;
          }
        }
    }
    __syncthreads();
    double y[DN_G];
#pragma unroll
    for (int g = 0; g < DN_G; g++) y[g] = yb[buf][g];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int i = t + r * DN_T;
      if (i > p && i < n) {
        const double l = Lc[(size_t)p * n + i];
#pragma unroll
        for (int g = 0; g < DN_G; g++)
          if (p >= c0 + g) s[r][g] -= l * y[g];
      }
    }
  }
  // backward: L^T x = y
  for (int p = n - 1; p >= 0; p--) {
    const int buf = p & 1;
    if (t == (p & (DN_T - 1))) {
      const double d = Lr[(size_t)p * n + p];
#pragma unroll
      for (int r = 0; r < R; r++)
        if (r == p / DN_T) {
#pragma unroll
          for (int g = 0; g < DN_G; g++) {
            const double x = s[r][g] / d;
            s[r][g] = x;
            yb[buf][g] = x;
          }
        }
    }
    __syncthreads();
    double y[DN_G];
#pragma unroll
    for (int g = 0; g < DN_G; g++) y[g] = yb[buf][g];
#pragma unroll
    for (int r = 0; r < R; r++) {
      const int i = t + r * DN_T;
      if (i < p) {
        const double l = Lr[(size_t)p * n + i];
#pragma unroll
        for (int g = 0; g < DN_G; g++) s[r][g] -= l * y[g];
      }
    }
  }
#pragma unroll
  for (int r = 0; r < R; r++) {
    const int i = t + r * DN_T;
    if (i < n)
#pragma unroll
      for (int g = 0; g < DN_G; g++)
        if (c0 + g < n) inv[(size_t)i * n + c0 + g] = s[r][g];
  }
}

// 8 columns per block up to R = 4 rows per thread, 4 beyond (registers)
template <int R>
void launch_dn_solve(int n, const double *Lc, const double *Lr, double *inv, hipStream_t st) {
  constexpr int G = R <= 4 ? 8 : 4;
  hipLaunchKernelGGL((k_dn_solve<R, G>), dim3((unsigned)((n + G - 1) / G)), dim3(DN_T), 0, st, n, Lc, Lr, inv);
}

__global__ void k_dn_to_float(long long m, const double *__restrict__ src, float *__restrict__ out) {
  for (long long q = blockIdx.x * (long long)BLOCK + threadIdx.x; q < m; q += (long long)gridDim.x * BLOCK)
    out[q] = (float)src[q];
}

}  // namespace

extern "C" {

void bn_rb_symmetrize(int nnz, const int *tpos, const double *v, double *out, int diag0, hipStream_t st) {
  if (nnz > 0) hipLaunchKernelGGL(k_rb_symmetrize, g1(nnz), dim3(BLOCK), 0, st, nnz, tpos, v, out, diag0);
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

void bn_rb_den(int npairs, const int *doff, const int *dq, const double *a, double *den, hipStream_t st) {
  if (npairs > 0) hipLaunchKernelGGL(k_rb_den, g1(npairs), dim3(BLOCK), 0, st, npairs, doff, dq, a, den);
}

void bn_rb_dhat(int n, const int *aia, const double *a, const unsigned char *cls, const int *cpt, const int *rp,
                const int *pair_k, const double *den, double *dhat, hipStream_t st) {
  if (n > 0) hipLaunchKernelGGL(k_rb_dhat, g1(n), dim3(BLOCK), 0, st, n, aia, a, cls, cpt, rp, pair_k, den, dhat);
}

void bn_rb_pval(int nnzp, const int *prow, const int *cpt, const int *pia, const int *poff, const int *pa,
                const int *pq, const int *pair_k, const double *a, const double *den, const double *dhat, double *pv,
                hipStream_t st) {
  if (nnzp > 0)
    hipLaunchKernelGGL(k_rb_pval, g1(nnzp), dim3(BLOCK), 0, st, nnzp, prow, cpt, pia, poff, pa, pq, pair_k, a, den,
                       dhat, pv);
}

void bn_rb_rap_pull(int nnz, const int *toff, const int *t1, const int *t2, const int *t3, int nlong,
                    const int *elist, const double *rv, const double *av, const double *pv, double *out,
                    hipStream_t st) {
  if (nnz > 0) hipLaunchKernelGGL(k_rb_rap_pull, g1(nnz), dim3(BLOCK), 0, st, nnz, toff, t1, t2, t3, rv, av, pv, out);
  if (nlong > 0)
    hipLaunchKernelGGL(k_rb_rap_pull_wave, dim3((unsigned)((nlong + BLOCK / 64 - 1) / (BLOCK / 64))), dim3(BLOCK), 0,
                       st, nlong, elist, toff, t1, t2, t3, rv, av, pv, out);
}

void bn_rb_rap(int nc, const int *ria, const int *rja, const double *rv, const int *aia, const int *aja,
               const double *av, const int *pia, const int *pja, const double *pv, const int *cia, const int *ccol,
               const int *cpos, double *out, hipStream_t st) {
  if (nc > 0)
    hipLaunchKernelGGL(k_rb_rap, g1(nc), dim3(BLOCK), 0, st, nc, ria, rja, rv, aia, aja, av, pia, pja, pv, cia,
                       ccol, cpos, out);
}

void bn_ic0_rows(int cnt, const int *rows, const int *lia, const int *lja, double *lv, hipStream_t st) {
  if (cnt > 0) hipLaunchKernelGGL(k_ic0_rows, g1(cnt), dim3(BLOCK), 0, st, cnt, rows, lia, lja, lv);
}

int bn_dn_inverse(int n, const int *ia, const int *ja, const double *a, double *Lc, double *Lr, double *inv,
                  float *invs, int *clamped, hipStream_t st) {
  if (n <= 0) return 0;
  if (n > DN_T * 8) return -1;
  const long long m = (long long)n * n;
  const unsigned gm = (unsigned)((m + BLOCK - 1) / BLOCK < 4096 ? (m + BLOCK - 1) / BLOCK : 4096);
  hipLaunchKernelGGL(k_dn_zero, dim3(gm), dim3(BLOCK), 0, st, m, Lc);
  hipLaunchKernelGGL(k_dn_scatter, g1(n), dim3(BLOCK), 0, st, n, ia, ja, a, Lc);
  for (int q0 = 0; q0 < n; q0 += DN_B) {
    hipLaunchKernelGGL(k_dn_diag, dim3(1), dim3(64), 0, st, n, q0, Lc, clamped);
    const int t0 = q0 + DN_B;
    if (t0 >= n) break;
    hipLaunchKernelGGL(k_dn_panel, dim3((unsigned)((n - t0 + 255) / 256)), dim3(256), 0, st, n, q0, Lc);
    const long long nt = (n - t0 + 63) / 64;
    hipLaunchKernelGGL(k_dn_trail, dim3((unsigned)(nt * (nt + 1) / 2)), dim3(256), 0, st, n, q0, Lc);
  }
  const unsigned nb = (unsigned)((n + 31) / 32);
  hipLaunchKernelGGL(k_dn_transpose, dim3(nb, nb), dim3(256), 0, st, n, Lc, Lr);
  switch ((n + DN_T - 1) / DN_T) {
    case 1: launch_dn_solve<1>(n, Lc, Lr, inv, st); break;
    case 2: launch_dn_solve<2>(n, Lc, Lr, inv, st); break;
    case 3: launch_dn_solve<3>(n, Lc, Lr, inv, st); break;
    case 4: launch_dn_solve<4>(n, Lc, Lr, inv, st); break;
    case 5: launch_dn_solve<5>(n, Lc, Lr, inv, st); break;
    case 6: launch_dn_solve<6>(n, Lc, Lr, inv, st); break;
    case 7: launch_dn_solve<7>(n, Lc, Lr, inv, st); break;
    default: launch_dn_solve<8>(n, Lc, Lr, inv, st); break;
  }
  if (invs) hipLaunchKernelGGL(k_dn_to_float, dim3(gm), dim3(BLOCK), 0, st, m, inv, invs);
  return 0;
}

}  // extern "C"
