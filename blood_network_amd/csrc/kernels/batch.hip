// batch.hip - gfx950 kernels for the block (multi-right-hand-side) solve.
//
// A block vector holds K columns interleaved by row: element (i, c) is at
// i*K + c, so the K values a row contributes or receives are one
// contiguous K-wide segment.  Every sparse kernel maps K/V consecutive
// lanes to one row, each lane owning V consecutive columns (V = 4 FP32 or
// 2 FP64 values, one vector load): each matrix entry is read once for all
// K columns and the x-gathers are K-wide contiguous reads.  K is a power
// of two <= 32.  Reductions (per column) are deterministic: per-block
// partials in fixed order, finished by one block per column.
#include <hip/hip_runtime.h>
#include <math.h>

#include "bn_kernels.h"

#define BLOCK 256

namespace {

__device__ __forceinline__ int logical_block() {
  const int g = gridDim.x, b = blockIdx.x;
  if ((g & 7) != 0) return b;
  return (b & 7) * (g >> 3) + (b >> 3);
}

// blocks for `rows` rows at lr lanes per row
inline dim3 grid_for_rows(long long rows, int lr) {
  const long long rpb = BLOCK / lr;
  long long g = (rows + rpb - 1) / rpb;
  if (g >= 64) g = (g + 7) & ~7LL;
  return dim3((unsigned)(g > 0 ? g : 1));
}

inline dim3 g1(long long n) { return dim3((unsigned)((n + BLOCK - 1) / BLOCK)); }

// Each lane owns V consecutive columns (V = 4 for FP32, 2 for FP64,
// fewer when K is smaller) and moves them with one vector load/store, so
// a row's K-wide segment takes K/V lanes and one gather instruction
// covers several rows.
template <typename T, int K>
struct VW {
  static constexpr int V = (sizeof(T) == 4) ? (K < 4 ? K : 4) : (K < 2 ? K : 2);
  static constexpr int LR = K / V; /* lanes per row */
};

template <typename T, int V>
struct vec {
  T v[V];
};

template <typename T, int V>
__device__ __forceinline__ vec<T, V> vload(const T *p) {
  vec<T, V> o;
  if constexpr (sizeof(T) == 4 && V == 4) {
    const float4 t = *reinterpret_cast<const float4 *>(p);
    o.v[0] = t.x; o.v[1] = t.y; o.v[2] = t.z; o.v[3] = t.w;
  } else if constexpr (V == 2) {
    if constexpr (sizeof(T) == 4) {
      const float2 t = *reinterpret_cast<const float2 *>(p);
      o.v[0] = t.x; o.v[1] = t.y;
    } else {
      const double2 t = *reinterpret_cast<const double2 *>(p);
      o.v[0] = t.x; o.v[1] = t.y;
    }
  } else {
#pragma unroll
    for (int u = 0; u < V; u++) o.v[u] = p[u];
  }
  return o;
}

template <typename T, int V>
__device__ __forceinline__ void vstore(T *p, const vec<T, V> &o) {
  if constexpr (sizeof(T) == 4 && V == 4) {
    *reinterpret_cast<float4 *>(p) = make_float4(o.v[0], o.v[1], o.v[2], o.v[3]);
  } else if constexpr (V == 2) {
    if constexpr (sizeof(T) == 4) *reinterpret_cast<float2 *>(p) = make_float2(o.v[0], o.v[1]);
    else *reinterpret_cast<double2 *>(p) = make_double2(o.v[0], o.v[1]);
  } else {
#pragma unroll
    for (int u = 0; u < V; u++) p[u] = o.v[u];
  }
}

template <typename T, int V>
__device__ __forceinline__ void vfma(vec<T, V> &acc, T a, const vec<T, V> &x) {
#pragma unroll
  for (int u = 0; u < V; u++) acc.v[u] += a * x.v[u];
}

// sum over one CSR row of val[q] * X[ja[q]*K + c0 .. c0+V), SKIP_DIAG
// drops column == self.  Rounds of four entries issue all their loads
// before use; the last partial round is predicated, so a short row (7
// entries on the fine level, 4 in an IC triangle) costs at most two
// dependent rounds.
template <typename T, int K, bool SKIP_DIAG>
__device__ __forceinline__ vec<T, VW<T, K>::V> row_dot_k(int q, const int e, const int *__restrict__ ja,
                                                          const T *__restrict__ val, const T *__restrict__ X,
                                                          int c0, int self) {
  constexpr int V = VW<T, K>::V;
  vec<T, V> s0 = {}, s1 = {};
  for (; q + 3 < e; q += 4) {
    const int j0 = ja[q], j1 = ja[q + 1], j2 = ja[q + 2], j3 = ja[q + 3];
    const T a0 = (SKIP_DIAG && j0 == self) ? T(0) : val[q];
    const T a1 = (SKIP_DIAG && j1 == self) ? T(0) : val[q + 1];
    const T a2 = (SKIP_DIAG && j2 == self) ? T(0) : val[q + 2];
    const T a3 = (SKIP_DIAG && j3 == self) ? T(0) : val[q + 3];
    const vec<T, V> x0 = vload<T, V>(X + (size_t)j0 * K + c0);
    const vec<T, V> x1 = vload<T, V>(X + (size_t)j1 * K + c0);
    const vec<T, V> x2 = vload<T, V>(X + (size_t)j2 * K + c0);
    const vec<T, V> x3 = vload<T, V>(X + (size_t)j3 * K + c0);
    vfma(s0, a0, x0);
    vfma(s1, a1, x1);
    vfma(s0, a2, x2);
    vfma(s1, a3, x3);
  }
  if (q < e) {
    const bool h1 = q + 1 < e, h2 = q + 2 < e;
    const int j0 = ja[q];
    const int j1 = h1 ? ja[q + 1] : j0;
    const int j2 = h2 ? ja[q + 2] : j0;
    const T a0 = (SKIP_DIAG && j0 == self) ? T(0) : val[q];
    const T a1 = (!h1 || (SKIP_DIAG && j1 == self)) ? T(0) : val[q + 1];
    const T a2 = (!h2 || (SKIP_DIAG && j2 == self)) ? T(0) : val[q + 2];
    const vec<T, V> x0 = vload<T, V>(X + (size_t)j0 * K + c0);
    const vec<T, V> x1 = vload<T, V>(X + (size_t)j1 * K + c0);
    const vec<T, V> x2 = vload<T, V>(X + (size_t)j2 * K + c0);
    vfma(s0, a0, x0);
    vfma(s1, a1, x1);
    vfma(s0, a2, x2);
  }
#pragma unroll
  for (int u = 0; u < V; u++) s0.v[u] += s1.v[u];
  return s0;
}

// Y[row, c0..c0+V) = op(sum, B, Y)
template <typename T, int V, int OP>
__device__ __forceinline__ void put_row(T *Y, const T *B, size_t e, const vec<T, V> &sm) {
  vec<T, V> out;
  if (OP == BN_OP_SET) {
    out = sm;
  } else if (OP == BN_OP_RESID) {
    out = vload<T, V>(B + e);
#pragma unroll
    for (int u = 0; u < V; u++) out.v[u] -= sm.v[u];
  } else {
    out = vload<T, V>(Y + e);
#pragma unroll
    for (int u = 0; u < V; u++) out.v[u] += sm.v[u];
  }
  vstore<T, V>(Y + e, out);
}

// ---------------------------------------------------------------------
// SpMM: Y = op(A, X [, B]); rows longer than long_len are skipped (the
// segmented kernel below does them)
// ---------------------------------------------------------------------
template <typename T, int K, int OP>
__global__ __launch_bounds__(BLOCK) void k_spmm(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                                                const T *__restrict__ val, const T *__restrict__ X,
                                                T *__restrict__ Y, const T *__restrict__ B, int long_len) {
  constexpr int V = VW<T, K>::V, LR = VW<T, K>::LR;
  const int row = logical_block() * (BLOCK / LR) + threadIdx.x / LR;
  const int c0 = (threadIdx.x % LR) * V;
  if (row >= n) return;
  const int q0 = ia[row], q1 = ia[row + 1];
  if (q1 - q0 > long_len) return;
  const vec<T, V> sm = row_dot_k<T, K, false>(q0, q1, ja, val, X, c0, row);
  put_row<T, V, OP>(Y, B, (size_t)row * K + c0, sm);
}

// long row t = rows[t] in segments of BN_SEG_NNZ entries (one block
// each): BLOCK/K slices of the segment times K columns, reduced over the
// slices in LDS; one-segment rows finish here, longer ones via partials
template <typename T, int K, int OP>
__global__ __launch_bounds__(BLOCK) void k_spmm_seg(const int *__restrict__ seg_row, const int *__restrict__ seg_ptr,
                                                    const int *__restrict__ rows, const int *__restrict__ ia,
                                                    const int *__restrict__ ja, const T *__restrict__ val,
                                                    const T *__restrict__ X, T *__restrict__ Y,
                                                    const T *__restrict__ B, double *__restrict__ part) {
  constexpr int SL = BLOCK / K;
  __shared__ double red[BLOCK];
  const int sg = blockIdx.x;
  const int t = seg_row[sg];
  const int row = rows[t];
  const int s0 = seg_ptr[t], s1 = seg_ptr[t + 1];
  const int beg = ia[row] + (sg - s0) * BN_SEG_NNZ;
  const int end = min(beg + BN_SEG_NNZ, ia[row + 1]);
  const int sl = threadIdx.x / K, c = threadIdx.x % K;
  double acc = 0.0;
  for (int q = beg + sl; q < end; q += SL) acc += (double)(val[q] * X[(size_t)ja[q] * K + c]);
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = SL / 2; w > 0; w >>= 1) {
    if (sl < w) red[threadIdx.x] += red[threadIdx.x + w * K];
    __syncthreads();
  }
  if (sl != 0) return;
  const double tot = red[c];
  if (s1 - s0 > 1) {
    part[(size_t)sg * K + c] = tot;
    return;
  }
  const size_t e = (size_t)row * K + c;
  if (OP == BN_OP_SET) Y[e] = (T)tot;
  else if (OP == BN_OP_RESID) Y[e] = B[e] - (T)tot;
  else Y[e] = Y[e] + (T)tot;
}

template <typename T, int K, int OP>
__global__ void k_spmm_seg_finish(int nmulti, const int *__restrict__ seg_ptr, const int *__restrict__ rows,
                                  const double *__restrict__ part, T *__restrict__ Y, const T *__restrict__ B) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  const int t = g / K, c = g % K;
  if (t >= nmulti) return;
  double acc = 0.0;
  for (int k = seg_ptr[t]; k < seg_ptr[t + 1]; k++) acc += part[(size_t)k * K + c];
  const size_t e = (size_t)rows[t] * K + c;
  if (OP == BN_OP_SET) Y[e] = (T)acc;
  else if (OP == BN_OP_RESID) Y[e] = B[e] - (T)acc;
  else Y[e] = Y[e] + (T)acc;
}

// ---------------------------------------------------------------------
// multicolour GS colour over a colour-ordered row copy (row r updates
// level row map[r]) and its zero start
// ---------------------------------------------------------------------
template <typename T, int K>
__global__ __launch_bounds__(BLOCK) void k_gs_map_k(int lo, int cnt, const int *__restrict__ map,
                                                    const int *__restrict__ ia, const int *__restrict__ ja,
                                                    const T *__restrict__ val, const T *__restrict__ dinv,
                                                    const T *__restrict__ B, T *__restrict__ X,
                                                    const T *__restrict__ zb, T *__restrict__ zs) {
  constexpr int V = VW<T, K>::V, LR = VW<T, K>::LR;
  const int k = logical_block() * (BLOCK / LR) + threadIdx.x / LR;
  const int c0 = (threadIdx.x % LR) * V;
  if (k >= cnt) return;
  const int r = lo + k, i = map[r];
  const vec<T, V> sm = row_dot_k<T, K, true>(ia[r], ia[r + 1], ja, val, X, c0, i);
  const size_t e = (size_t)i * K + c0;
  const T dv = dinv[r];
  vec<T, V> v = vload<T, V>(B + e);
#pragma unroll
  for (int u = 0; u < V; u++) v.v[u] = (v.v[u] - sm.v[u]) * dv;
  vstore<T, V>(X + e, v);
  if (zs) {
    vec<T, V> z = vload<T, V>(zb + e);
#pragma unroll
    for (int u = 0; u < V; u++) z.v[u] += v.v[u];
    vstore<T, V>(zs + e, z);
  }
}

template <typename T, int K>
__global__ void k_gs_zero_map_k(int lo, int hi, int all_n, const int *__restrict__ map, const T *__restrict__ dinv,
                                const T *__restrict__ B, T *__restrict__ X) {
  const long long g = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int r = lo + (int)(g / K), c = (int)(g % K);
  if (r >= all_n) return;
  const size_t e = (size_t)map[r] * K + c;
  X[e] = r < hi ? B[e] * dinv[r] : T(0);
}

// one Jacobi sweep of a triangular system; zs: write zb + result instead
template <typename T, int K>
__global__ __launch_bounds__(BLOCK) void k_jacobi_rows_k(int n, const int *__restrict__ ia,
                                                         const int *__restrict__ ja, const T *__restrict__ val,
                                                         const T *__restrict__ dinv, const T *__restrict__ B,
                                                         const T *__restrict__ Xin, T *__restrict__ Xout,
                                                         const T *__restrict__ zb, T *__restrict__ zs) {
  constexpr int V = VW<T, K>::V, LR = VW<T, K>::LR;
  const int i = logical_block() * (BLOCK / LR) + threadIdx.x / LR;
  const int c0 = (threadIdx.x % LR) * V;
  if (i >= n) return;
  const vec<T, V> sm = row_dot_k<T, K, true>(ia[i], ia[i + 1], ja, val, Xin, c0, i);
  const size_t e = (size_t)i * K + c0;
  const T dv = dinv[i];
  vec<T, V> v = vload<T, V>(B + e);
#pragma unroll
  for (int u = 0; u < V; u++) v.v[u] = (v.v[u] - sm.v[u]) * dv;
  if (zs) {
    vec<T, V> z = vload<T, V>(zb + e);
#pragma unroll
    for (int u = 0; u < V; u++) z.v[u] += v.v[u];
    vstore<T, V>(zs + e, z);
  } else {
    vstore<T, V>(Xout + e, v);
  }
}

// X = om * d_i * B (row scaling); X += om * d_i * R
template <typename T, int K>
__global__ void k_row_scale_k(long long ne, const T *__restrict__ d, T om, const T *__restrict__ B,
                              T *__restrict__ X, int add) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const T v = om * d[e / K] * B[e];
  X[e] = add ? X[e] + v : v;
}

// dense Y = M X (M n x n row-major, FP64): one block per row, BLOCK/K
// column slices of M times K right-hand sides, reduced in LDS
template <int K>
__global__ __launch_bounds__(BLOCK) void k_gemm_k(int n, const double *__restrict__ M,
                                                  const double *__restrict__ X, double *__restrict__ Y) {
  constexpr int SL = BLOCK / K;
  __shared__ double red[BLOCK];
  const int i = blockIdx.x;
  const int sl = threadIdx.x / K, c = threadIdx.x % K;
  const double *m = M + (size_t)i * n;
  double acc = 0.0;
  for (int j = sl; j < n; j += SL) acc += m[j] * X[(size_t)j * K + c];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int w = SL / 2; w > 0; w >>= 1) {
    if (sl < w) red[threadIdx.x] += red[threadIdx.x + w * K];
    __syncthreads();
  }
  if (sl == 0) Y[(size_t)i * K + c] = red[c];
}

// ---------------------------------------------------------------------
// BLAS-1 with one scalar set per column (FP64, outer PCG)
// ---------------------------------------------------------------------
// block sums over rows [row0, row1) of a column-partial into part[b*K+c]
template <int K>
__device__ __forceinline__ void col_partial(double v, double *red, double *part, int b) {
  constexpr int SL = BLOCK / K;
  const int sl = threadIdx.x / K, c = threadIdx.x % K;
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = SL / 2; w > 0; w >>= 1) {
    if (sl < w) red[threadIdx.x] += red[threadIdx.x + w * K];
    __syncthreads();
  }
  if (sl == 0) part[(size_t)b * K + c] = red[c];
  __syncthreads();
}

// rows of block b: [b*chunk, min(n, (b+1)*chunk))
template <int K>
__global__ __launch_bounds__(BLOCK) void k_dot_k(int n, int chunk, const double *__restrict__ x,
                                                 const double *__restrict__ y, const double *__restrict__ u,
                                                 const double *__restrict__ v, double *__restrict__ p1,
                                                 double *__restrict__ p2) {
  constexpr int SL = BLOCK / K;
  __shared__ double red[BLOCK];
  const int b = blockIdx.x;
  const int sl = threadIdx.x / K, c = threadIdx.x % K;
  const int i0 = b * chunk, i1 = min(n, i0 + chunk);
  double a = 0.0, w = 0.0;
  for (int i = i0 + sl; i < i1; i += SL) {
    const size_t e = (size_t)i * K + c;
    a += x[e] * y[e];
    if (u) w += u[e] * v[e];
  }
  col_partial<K>(a, red, p1, b);
  if (u) col_partial<K>(w, red, p2, b);
}

// per-column scalar slots (BN_S_COUNT doubles per column)
#define SC(c, slot) s[(size_t)(c) * BN_S_COUNT + (slot)]

// WRITE_S: also store r rounded to FP32 in rs (the FP32 V-cycle's input)
template <int K, bool WRITE_S = false>
__global__ __launch_bounds__(BLOCK) void k_update_xr_k(int n, int chunk, double *__restrict__ x,
                                                       const double *__restrict__ p, double *__restrict__ r,
                                                       const double *__restrict__ q, const double *__restrict__ s,
                                                       double *__restrict__ part, float *__restrict__ rs = nullptr) {
  constexpr int SL = BLOCK / K;
  __shared__ double red[BLOCK];
  const int b = blockIdx.x;
  const int sl = threadIdx.x / K, c = threadIdx.x % K;
  const double al = SC(c, BN_S_ALPHA);
  const int i0 = b * chunk, i1 = min(n, i0 + chunk);
  double a = 0.0;
  for (int i = i0 + sl; i < i1; i += SL) {
    const size_t e = (size_t)i * K + c;
    x[e] = x[e] + p[e] * al;
    const double rn = r[e] + q[e] * (-al);
    r[e] = rn;
    if (WRITE_S) rs[e] = (float)rn;
    a += rn * rn;
  }
  col_partial<K>(a, red, part, b);
}

// k_dot_k(r, z, z, q) with z held in FP32 (widened exactly: the same
// products in the same order)
template <int K>
__global__ __launch_bounds__(BLOCK) void k_dot_rz_k_s(int n, int chunk, const double *__restrict__ r,
                                                      const float *__restrict__ zs, const double *__restrict__ q,
                                                      double *__restrict__ p1, double *__restrict__ p2) {
  constexpr int SL = BLOCK / K;
  __shared__ double red[BLOCK];
  const int b = blockIdx.x;
  const int sl = threadIdx.x / K, c = threadIdx.x % K;
  const int i0 = b * chunk, i1 = min(n, i0 + chunk);
  double a = 0.0, w = 0.0;
  for (int i = i0 + sl; i < i1; i += SL) {
    const size_t e = (size_t)i * K + c;
    const double z = (double)zs[e];
    a += r[e] * z;
    if (q) w += z * q[e];
  }
  col_partial<K>(a, red, p1, b);
  if (q) col_partial<K>(w, red, p2, b);
}

template <int K>
__global__ void k_update_p_k(long long ne, double *__restrict__ p, const double *__restrict__ z,
                             const double *__restrict__ s) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int c = (int)(e % K);
  p[e] = SC(c, BN_S_BETA) * p[e] + z[e];
}

// the same with z held in FP32
template <int K>
__global__ void k_update_p_k_s(long long ne, double *__restrict__ p, const float *__restrict__ zs,
                               const double *__restrict__ s) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ne) return;
  const int c = (int)(e % K);
  p[e] = SC(c, BN_S_BETA) * p[e] + (double)zs[e];
}

// Per-column scalar logic, one block per column (block c sums column c's
// nb partials, then its thread 0 applies the logic).  Modes as
// bn_scalar_finish: 0 start, 1 alpha, 2 fixed beta, 3 ||b||^2, 4 flexible
// beta.  Modes 2/4 also apply the stop test (relres: sqrt(rr)/||b||, else
// rel) against BN_S_TOL and freeze a converged column (alpha = beta = 0
// from then on, BN_S_DONE set, BN_S_ITS = its iteration count).
template <int K>
__global__ __launch_bounds__(BLOCK) void k_finish_k(int which, int nb, const double *__restrict__ p1,
                                                    const double *__restrict__ p2, const double *__restrict__ p3,
                                                    double *__restrict__ s) {
  __shared__ double red[3][BLOCK];
  const int c = blockIdx.x;
  double a = 0.0, b = 0.0, d = 0.0;
  for (int k = threadIdx.x; k < nb; k += BLOCK) {
    if (p1) a += p1[(size_t)k * K + c];
    if (p2) b += p2[(size_t)k * K + c];
    if (p3) d += p3[(size_t)k * K + c];
  }
  red[0][threadIdx.x] = a;
  red[1][threadIdx.x] = b;
  red[2][threadIdx.x] = d;
  __syncthreads();
  for (int w = BLOCK / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      red[0][threadIdx.x] += red[0][threadIdx.x + w];
      red[1][threadIdx.x] += red[1][threadIdx.x + w];
      red[2][threadIdx.x] += red[2][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  a = red[0][0];
  b = red[1][0];
  d = red[2][0];
  const int done = SC(c, BN_S_DONE) != 0.0;
  switch (which) {
    case 0:
      SC(c, BN_S_RZ) = a;
      SC(c, BN_S_RZ0) = a;
      SC(c, BN_S_RR) = b;
      SC(c, BN_S_REL) = 1.0;
      SC(c, BN_S_BETA) = 0.0;
      SC(c, BN_S_ITER) = 0.0;
      SC(c, BN_S_ITS) = 0.0;
      SC(c, BN_S_DONE) = (a == 0.0) ? 1.0 : 0.0; /* r0 = 0: x0 already solves this column */
      if (a == 0.0) SC(c, BN_S_REL) = 0.0;
      break;
    case 1:
      SC(c, BN_S_PAP) = a;
      SC(c, BN_S_ALPHA) = done ? 0.0 : SC(c, BN_S_RZ) / a;
      break;
    case 2:
    case 4: {
      if (done) {
        SC(c, BN_S_BETA) = 0.0;
        break;
      }
      const double it = SC(c, BN_S_ITER) + 1.0;
      SC(c, BN_S_ITER) = it;
      SC(c, BN_S_RZNEW) = a;
      SC(c, BN_S_RR) = b;
      const double rel = sqrt(a / SC(c, BN_S_RZ0));
      SC(c, BN_S_REL) = rel;
      double bn = sqrt(SC(c, BN_S_BB));
      if (bn == 0.0) bn = 1.0;
      const double test = SC(c, BN_S_STOPREL) != 0.0 ? sqrt(b) / bn : rel;
      if (test < SC(c, BN_S_TOL)) {
        SC(c, BN_S_DONE) = 1.0;
        SC(c, BN_S_ITS) = it;
        SC(c, BN_S_BETA) = 0.0;
      } else {
        SC(c, BN_S_BETA) = (which == 4) ? -SC(c, BN_S_ALPHA) * d / SC(c, BN_S_RZ) : a / SC(c, BN_S_RZ);
      }
      SC(c, BN_S_RZOLD) = SC(c, BN_S_RZ);
      SC(c, BN_S_RZ) = a;
      break;
    }
    case 3:
      SC(c, BN_S_BB) = a;
      break;
    default:
      break;
  }
}

// ---------------------------------------------------------------------
// Segmented BLAS-1 for a block-diagonal union of independent systems:
// work item w covers rows [item_lo[w], item_lo[w+1]) of system
// item_seg[w] (items never cross systems), so the system's scalars are
// uniform over a block.  Partials go to part[w]; k_finish_seg sums system
// g's items item_ptr[g] .. item_ptr[g+1]-1 in order.
// ---------------------------------------------------------------------
__device__ __forceinline__ double block_sum256(double v, double *red) {
  red[threadIdx.x] = v;
  __syncthreads();
  for (int w = BLOCK / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  const double t = red[0];
  __syncthreads();
  return t;
}

__global__ __launch_bounds__(BLOCK) void k_dot_seg(const int *__restrict__ item_lo, const double *__restrict__ x,
                                                   const double *__restrict__ y, const double *__restrict__ u,
                                                   const double *__restrict__ v, double *__restrict__ p1,
                                                   double *__restrict__ p2) {
  __shared__ double red[BLOCK];
  const int w = blockIdx.x;
  const int i0 = item_lo[w], i1 = item_lo[w + 1];
  double a = 0.0, c = 0.0;
  for (int i = i0 + threadIdx.x; i < i1; i += BLOCK) {
    a += x[i] * y[i];
    if (u) c += u[i] * v[i];
  }
  a = block_sum256(a, red);
  if (u) c = block_sum256(c, red);
  if (threadIdx.x == 0) {
    p1[w] = a;
    if (u) p2[w] = c;
  }
}

#define SS(g, slot) s[(size_t)(g) * BN_S_COUNT + (slot)]

__global__ __launch_bounds__(BLOCK) void k_update_xr_seg(const int *__restrict__ item_lo,
                                                         const int *__restrict__ item_seg, double *__restrict__ x,
                                                         const double *__restrict__ p, double *__restrict__ r,
                                                         const double *__restrict__ q, const double *__restrict__ s,
                                                         double *__restrict__ part) {
  __shared__ double red[BLOCK];
  const int w = blockIdx.x;
  const int i0 = item_lo[w], i1 = item_lo[w + 1];
  const double al = SS(item_seg[w], BN_S_ALPHA);
  double a = 0.0;
  for (int i = i0 + threadIdx.x; i < i1; i += BLOCK) {
    x[i] = x[i] + p[i] * al;
    const double rn = r[i] + q[i] * (-al);
    r[i] = rn;
    a += rn * rn;
  }
  a = block_sum256(a, red);
  if (threadIdx.x == 0) part[w] = a;
}

__global__ __launch_bounds__(BLOCK) void k_update_p_seg(const int *__restrict__ item_lo,
                                                        const int *__restrict__ item_seg, double *__restrict__ p,
                                                        const double *__restrict__ z, const double *__restrict__ s) {
  const int w = blockIdx.x;
  const int i0 = item_lo[w], i1 = item_lo[w + 1];
  const double be = SS(item_seg[w], BN_S_BETA);
  for (int i = i0 + threadIdx.x; i < i1; i += BLOCK) p[i] = be * p[i] + z[i];
}

// one block per system; the scalar logic of k_finish_k
__global__ __launch_bounds__(BLOCK) void k_finish_seg(int which, const int *__restrict__ item_ptr,
                                                      const double *__restrict__ p1, const double *__restrict__ p2,
                                                      const double *__restrict__ p3, double *__restrict__ s) {
  __shared__ double red[BLOCK];
  const int g = blockIdx.x;
  double a = 0.0, b = 0.0, d = 0.0;
  for (int k = item_ptr[g] + threadIdx.x; k < item_ptr[g + 1]; k += BLOCK) {
    if (p1) a += p1[k];
    if (p2) b += p2[k];
    if (p3) d += p3[k];
  }
  a = block_sum256(a, red);
  b = block_sum256(b, red);
  d = block_sum256(d, red);
  if (threadIdx.x != 0) return;
  const int done = SS(g, BN_S_DONE) != 0.0;
  switch (which) {
    case 0:
      SS(g, BN_S_RZ) = a;
      SS(g, BN_S_RZ0) = a;
      SS(g, BN_S_RR) = b;
      SS(g, BN_S_REL) = (a == 0.0) ? 0.0 : 1.0;
      SS(g, BN_S_BETA) = 0.0;
      SS(g, BN_S_ITER) = 0.0;
      SS(g, BN_S_ITS) = 0.0;
      SS(g, BN_S_DONE) = (a == 0.0) ? 1.0 : 0.0;
      break;
    case 1:
      SS(g, BN_S_PAP) = a;
      SS(g, BN_S_ALPHA) = done ? 0.0 : SS(g, BN_S_RZ) / a;
      break;
    case 2:
    case 4: {
      if (done) {
        SS(g, BN_S_BETA) = 0.0;
        break;
      }
      const double it = SS(g, BN_S_ITER) + 1.0;
      SS(g, BN_S_ITER) = it;
      SS(g, BN_S_RZNEW) = a;
      SS(g, BN_S_RR) = b;
      const double rel = sqrt(a / SS(g, BN_S_RZ0));
      SS(g, BN_S_REL) = rel;
      double bn = sqrt(SS(g, BN_S_BB));
      if (bn == 0.0) bn = 1.0;
      const double test = SS(g, BN_S_STOPREL) != 0.0 ? sqrt(b) / bn : rel;
      if (test < SS(g, BN_S_TOL)) {
        SS(g, BN_S_DONE) = 1.0;
        SS(g, BN_S_ITS) = it;
        SS(g, BN_S_BETA) = 0.0;
      } else {
        SS(g, BN_S_BETA) = (which == 4) ? -SS(g, BN_S_ALPHA) * d / SS(g, BN_S_RZ) : a / SS(g, BN_S_RZ);
      }
      SS(g, BN_S_RZOLD) = SS(g, BN_S_RZ);
      SS(g, BN_S_RZ) = a;
      break;
    }
    case 3:
      SS(g, BN_S_BB) = a;
      break;
    default:
      break;
  }
}

// column-major k x n (each right-hand side contiguous) <-> interleaved
__global__ void k_interleave(int n, int k, int K, const double *__restrict__ src, double *__restrict__ dst,
                             int back) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)n * K) return;
  const int i = (int)(e / K), c = (int)(e % K);
  if (back) {
    if (c < k) dst[(size_t)c * n + i] = src[e];
  } else {
    dst[e] = c < k ? src[(size_t)c * n + i] : 0.0;
  }
}

}  // namespace

// K is a runtime power of two <= 32; the statement is compiled once per K
#define BN_K_DISPATCH(K, ...)                              \
  switch (K) {                                             \
    case 1: { constexpr int KK = 1; __VA_ARGS__; } break;  \
    case 2: { constexpr int KK = 2; __VA_ARGS__; } break;  \
    case 4: { constexpr int KK = 4; __VA_ARGS__; } break;  \
    case 8: { constexpr int KK = 8; __VA_ARGS__; } break;  \
    case 16: { constexpr int KK = 16; __VA_ARGS__; } break;\
    default: { constexpr int KK = 32; __VA_ARGS__; } break;\
  }

template <typename T>
static void spmm_t(int K, int op, int n, const int *ia, const int *ja, const T *val, const T *X, T *Y, const T *B,
                   int long_len, int nlong, const int *rows, const int *seg_ptr, const int *seg_row, int nseg,
                   int nmulti, double *part, hipStream_t st) {
  if (n <= 0) return;
  if (nlong <= 0) long_len = 0x7fffffff;
#define BN_SPMM(OPV)                                                                                           \
  BN_K_DISPATCH(K, {                                                                                           \
    hipLaunchKernelGGL((k_spmm<T, KK, OPV>), grid_for_rows(n, VW<T, KK>::LR), dim3(BLOCK), 0, st, n, ia, ja, val, X, Y, B, \
                       long_len);                                                                              \
    if (nlong > 0 && nseg > 0) {                                                                               \
      hipLaunchKernelGGL((k_spmm_seg<T, KK, OPV>), dim3(nseg), dim3(BLOCK), 0, st, seg_row, seg_ptr, rows, ia,  \
                         ja, val, X, Y, B, part);                                                              \
      if (nmulti > 0)                                                                                          \
        hipLaunchKernelGGL((k_spmm_seg_finish<T, KK, OPV>), g1((long long)nmulti * KK), dim3(BLOCK), 0, st,    \
                           nmulti, seg_ptr, rows, part, Y, B);                                                 \
    }                                                                                                          \
  })
  if (op == BN_OP_SET) {
    BN_SPMM(BN_OP_SET);
  } else if (op == BN_OP_RESID) {
    BN_SPMM(BN_OP_RESID);
  } else {
    BN_SPMM(BN_OP_ADD);
  }
#undef BN_SPMM
}

extern "C" {

void bn_spmm(char vt, int K, int op, int n, const int *ia, const int *ja, const void *val, const void *X, void *Y,
             const void *B, int long_len, int nlong, const int *rows, const int *seg_ptr, const int *seg_row,
             int nseg, int nmulti, double *part, hipStream_t stream) {
  if (vt == 's')
    spmm_t<float>(K, op, n, ia, ja, (const float *)val, (const float *)X, (float *)Y, (const float *)B, long_len,
                  nlong, rows, seg_ptr, seg_row, nseg, nmulti, part, stream);
  else
    spmm_t<double>(K, op, n, ia, ja, (const double *)val, (const double *)X, (double *)Y, (const double *)B,
                   long_len, nlong, rows, seg_ptr, seg_row, nseg, nmulti, part, stream);
}

void bn_gs_color_map_k(char vt, int K, int lo, int hi, const int *map, const int *ia, const int *ja, const void *val,
                       const void *dinv, const void *B, void *X, const void *zb, void *zs, hipStream_t st) {
  const int cnt = hi - lo;
  if (cnt <= 0) return;
  if (vt == 's') {
    BN_K_DISPATCH(K, hipLaunchKernelGGL((k_gs_map_k<float, KK>), grid_for_rows(cnt, VW<float, KK>::LR), dim3(BLOCK), 0, st, lo,
                                        cnt, map, ia, ja, (const float *)val, (const float *)dinv,
                                        (const float *)B, (float *)X, (const float *)zb, (float *)zs));
  } else {
    BN_K_DISPATCH(K, hipLaunchKernelGGL((k_gs_map_k<double, KK>), grid_for_rows(cnt, VW<double, KK>::LR), dim3(BLOCK), 0, st, lo,
                                        cnt, map, ia, ja, (const double *)val, (const double *)dinv,
                                        (const double *)B, (double *)X, (const double *)zb, (double *)zs));
  }
}

void bn_gs_zero_map_k(char vt, int K, int lo, int hi, int all_n, const int *map, const void *dinv, const void *B,
                      void *X, hipStream_t st) {
  if (all_n < hi) all_n = hi;
  const long long ne = (long long)(all_n - lo) * K;
  if (ne <= 0) return;
  if (vt == 's') {
    BN_K_DISPATCH(K, hipLaunchKernelGGL((k_gs_zero_map_k<float, KK>), g1(ne), dim3(BLOCK), 0, st, lo, hi, all_n,
                                        map, (const float *)dinv, (const float *)B, (float *)X));
  } else {
    BN_K_DISPATCH(K, hipLaunchKernelGGL((k_gs_zero_map_k<double, KK>), g1(ne), dim3(BLOCK), 0, st, lo, hi, all_n,
                                        map, (const double *)dinv, (const double *)B, (double *)X));
  }
}

void bn_jacobi_rows_k(char vt, int K, int n, const int *ia, const int *ja, const void *val, const void *dinv,
                      const void *B, const void *Xin, void *Xout, const void *zb, void *zs, hipStream_t st) {
  if (n <= 0) return;
  if (vt == 's') {
    BN_K_DISPATCH(K, hipLaunchKernelGGL((k_jacobi_rows_k<float, KK>), grid_for_rows(n, VW<float, KK>::LR), dim3(BLOCK), 0, st, n,
                                        ia, ja, (const float *)val, (const float *)dinv, (const float *)B,
                                        (const float *)Xin, (float *)Xout, (const float *)zb, (float *)zs));
  } else {
    BN_K_DISPATCH(K, hipLaunchKernelGGL((k_jacobi_rows_k<double, KK>), grid_for_rows(n, VW<double, KK>::LR), dim3(BLOCK), 0, st, n,
                                        ia, ja, (const double *)val, (const double *)dinv, (const double *)B,
                                        (const double *)Xin, (double *)Xout, (const double *)zb, (double *)zs));
  }
}

void bn_row_scale_k(char vt, int K, int n, const void *d, double om, const void *B, void *X, int add,
                    hipStream_t st) {
  const long long ne = (long long)n * K;
  if (ne <= 0) return;
  if (vt == 's') {
    BN_K_DISPATCH(K, hipLaunchKernelGGL((k_row_scale_k<float, KK>), g1(ne), dim3(BLOCK), 0, st, ne,
                                        (const float *)d, (float)om, (const float *)B, (float *)X, add));
  } else {
    BN_K_DISPATCH(K, hipLaunchKernelGGL((k_row_scale_k<double, KK>), g1(ne), dim3(BLOCK), 0, st, ne,
                                        (const double *)d, om, (const double *)B, (double *)X, add));
  }
}

void bn_gemm_k(int K, int n, const double *M, const double *X, double *Y, hipStream_t st) {
  if (n <= 0) return;
  BN_K_DISPATCH(K, hipLaunchKernelGGL((k_gemm_k<KK>), dim3(n), dim3(BLOCK), 0, st, n, M, X, Y));
}

/* per-column reductions use nb = bn_block_nb(n, K) blocks of chunk rows */
int bn_block_nb(int n) {
  int nb = (n + 1023) / 1024;
  if (nb > BN_BLOCK_RED) nb = BN_BLOCK_RED;
  return nb > 0 ? nb : 1;
}

void bn_dot_k(int K, int n, const double *x, const double *y, const double *u, const double *v, double *p1,
              double *p2, hipStream_t st) {
  const int nb = bn_block_nb(n), chunk = (n + nb - 1) / nb;
  BN_K_DISPATCH(K, hipLaunchKernelGGL((k_dot_k<KK>), dim3(nb), dim3(BLOCK), 0, st, n, chunk, x, y, u, v, p1, p2));
}

void bn_update_xr_k(int K, int n, double *x, const double *p, double *r, const double *q, const double *s,
                    double *part, hipStream_t st) {
  const int nb = bn_block_nb(n), chunk = (n + nb - 1) / nb;
  BN_K_DISPATCH(K, hipLaunchKernelGGL((k_update_xr_k<KK>), dim3(nb), dim3(BLOCK), 0, st, n, chunk, x, p, r, q, s,
                                      part));
}

void bn_update_xr_k_s(int K, int n, double *x, const double *p, double *r, const double *q, const double *s,
                      double *part, float *rs, hipStream_t st) {
  const int nb = bn_block_nb(n), chunk = (n + nb - 1) / nb;
  BN_K_DISPATCH(K, hipLaunchKernelGGL((k_update_xr_k<KK, true>), dim3(nb), dim3(BLOCK), 0, st, n, chunk, x, p, r, q,
                                      s, part, rs));
}

void bn_dot_rz_k_s(int K, int n, const double *r, const float *zs, const double *q, double *p1, double *p2,
                   hipStream_t st) {
  const int nb = bn_block_nb(n), chunk = (n + nb - 1) / nb;
  BN_K_DISPATCH(K, hipLaunchKernelGGL((k_dot_rz_k_s<KK>), dim3(nb), dim3(BLOCK), 0, st, n, chunk, r, zs, q, p1, p2));
}

void bn_update_p_k_s(int K, int n, double *p, const float *zs, const double *s, hipStream_t st) {
  const long long ne = (long long)n * K;
  if (ne <= 0) return;
  BN_K_DISPATCH(K, hipLaunchKernelGGL((k_update_p_k_s<KK>), g1(ne), dim3(BLOCK), 0, st, ne, p, zs, s));
}

void bn_update_p_k(int K, int n, double *p, const double *z, const double *s, hipStream_t st) {
  const long long ne = (long long)n * K;
  if (ne <= 0) return;
  BN_K_DISPATCH(K, hipLaunchKernelGGL((k_update_p_k<KK>), g1(ne), dim3(BLOCK), 0, st, ne, p, z, s));
}

void bn_finish_k(int K, int which, int n, const double *p1, const double *p2, const double *p3, double *s,
                 hipStream_t st) {
  const int nb = bn_block_nb(n);
  BN_K_DISPATCH(K, hipLaunchKernelGGL((k_finish_k<KK>), dim3(KK), dim3(BLOCK), 0, st, which, nb, p1, p2, p3, s));
}

void bn_dot_seg(int nitem, const int *item_lo, const double *x, const double *y, const double *u, const double *v,
                double *p1, double *p2, hipStream_t st) {
  if (nitem <= 0) return;
  hipLaunchKernelGGL(k_dot_seg, dim3(nitem), dim3(BLOCK), 0, st, item_lo, x, y, u, v, p1, p2);
}

void bn_update_xr_seg(int nitem, const int *item_lo, const int *item_seg, double *x, const double *p, double *r,
                      const double *q, const double *s, double *part, hipStream_t st) {
  if (nitem <= 0) return;
  hipLaunchKernelGGL(k_update_xr_seg, dim3(nitem), dim3(BLOCK), 0, st, item_lo, item_seg, x, p, r, q, s, part);
}

void bn_update_p_seg(int nitem, const int *item_lo, const int *item_seg, double *p, const double *z, const double *s,
                     hipStream_t st) {
  if (nitem <= 0) return;
  hipLaunchKernelGGL(k_update_p_seg, dim3(nitem), dim3(BLOCK), 0, st, item_lo, item_seg, p, z, s);
}

void bn_finish_seg(int nsys, int which, const int *item_ptr, const double *p1, const double *p2, const double *p3,
                   double *s, hipStream_t st) {
  if (nsys <= 0) return;
  hipLaunchKernelGGL(k_finish_seg, dim3(nsys), dim3(BLOCK), 0, st, which, item_ptr, p1, p2, p3, s);
}

void bn_interleave(int n, int k, int K, const double *src, double *dst, int back, hipStream_t st) {
  const long long ne = (long long)n * K;
  if (ne <= 0) return;
  hipLaunchKernelGGL(k_interleave, g1(ne), dim3(BLOCK), 0, st, n, k, K, src, dst, back);
}

}  // extern "C"
