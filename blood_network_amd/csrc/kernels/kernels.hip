// kernels.hip - gfx950 kernels for the AMG-preconditioned PCG solve.
//
// Every kernel here is HBM-bandwidth or latency bound (CSR gathers,
// triangular sweeps, BLAS-1): there is no dense contraction to put on
// MFMA.  Design rules used throughout:
//  * 256-thread blocks (4 wave64s); `LANES` consecutive lanes share one
//    CSR row so the ja/val streams of a row are read as one contiguous
//    segment, and the row sum is finished with DPP/xor shuffles.
//  * Streaming kernels use a capped grid (<= BN_RED_BLOCKS blocks, a
//    multiple of 8) whose logical block id is remapped so that each of the
//    8 XCDs owns one contiguous slice of rows: x-gathers of neighbouring
//    rows then hit the same XCD's L2.
//  * Reductions are deterministic: per-block partials in fixed order,
//    finished by one block in fixed order (no floating-point atomics).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <math.h>

#include "bn_kernels.h"

#define BLOCK 256

// A kernel timer armed by bn_time_next_launch: the next launch of this
// thread records its own start and end into the two events (the dispatch
// packet's timestamps, as a profiler reads them) and disarms it.
static thread_local hipEvent_t g_time_start = nullptr, g_time_stop = nullptr;

#define BN_LAUNCH(K, G, B, S, ST, ...)                                            \
  do {                                                                            \
    if (g_time_start) {                                                           \
      hipEvent_t t0_ = g_time_start, t1_ = g_time_stop;                           \
      g_time_start = g_time_stop = nullptr;                                       \
      hipExtLaunchKernelGGL(K, G, B, S, ST, t0_, t1_, 0, __VA_ARGS__);            \
    } else {                                                                      \
      hipLaunchKernelGGL(K, G, B, S, ST, __VA_ARGS__);                            \
    }                                                                             \
  } while (0)

namespace {

__device__ __forceinline__ int logical_block() {
  const int g = gridDim.x, b = blockIdx.x;
  if ((g & 7) != 0) return b;
  return (b & 7) * (g >> 3) + (b >> 3);
}

template <int W, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int o = W / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, W);
  return v;
}

// block-wide sum of one double per thread; result valid in thread 0
__device__ __forceinline__ double block_sum(double v) {
  __shared__ double wsum[BLOCK / 64];
  v = group_sum<64>(v);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) wsum[wid] = v;
  __syncthreads();
  double t = 0.0;
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < BLOCK / 64; k++) t += wsum[k];
  }
  __syncthreads();
  return t;
}

// The sentinel of values not written yet (the one-launch walks' y, the
// long-row segments' partial sums): a signalling-NaN bit pattern, which
// arithmetic never produces (it only ever produces quiet NaNs).
template <typename T> struct walk_word;
template <> struct walk_word<double> {
  typedef unsigned long long U;
  static constexpr U none = 0x7ff4dead0000beefULL;
};
template <> struct walk_word<float> {
  typedef unsigned int U;
  static constexpr U none = 0x7fa0beefU;
};

// Partial-sum buffers always hold BN_RED_BLOCKS slots; the finishing
// kernel sums all of them, so logical block 0 zeroes the slots past the
// grid (instead of a memset node, which is unsafe inside captured graphs).
__device__ __forceinline__ void clear_tail(double *partial) {
  for (int k = gridDim.x + threadIdx.x; k < BN_RED_BLOCKS; k += BLOCK) partial[k] = 0.0;
}

// Sum of val[q] * x[col[q]] over one CSR row segment, `LANES` lanes
// striding through it.  Every round issues R strides of loads before
// using any (the last, partial round with predicated loads), so a lane
// keeps its gathers in flight together instead of forming a chain of
// dependent load rounds.  R = 8 for two lanes per row (a 7-entry fine
// row then finishes in one round), 4 otherwise (one lane: the extra
// registers and masked loads of R = 8 measured slower).  SKIP_DIAG drops
// the entry whose column equals `self`.  `gather(c, j)` returns x at
// stored column word c and sets j to the column it names (the plain form:
// x[c], j = c), so a caller can keep flags in the column words or read x
// from more than one place with the same sums.
template <typename T, int LANES, bool SKIP_DIAG, int R = (LANES == 2 ? 8 : 4), typename G>
__device__ __forceinline__ T row_dot_g(int q, const int e, const int *__restrict__ ja,
                                       const T *__restrict__ val, G gather, int self) {
  T acc[4] = {0, 0, 0, 0};
  for (; q + (R - 1) * LANES < e; q += R * LANES) {
    int j[R];
    T a[R];
#pragma unroll
    for (int u = 0; u < R; u++) {
      j[u] = ja[q + u * LANES];
      a[u] = val[q + u * LANES];
    }
#pragma unroll
    for (int u = 0; u < R; u++) {
      // select the product, not the gathered value: x[self] may be stale
      // (a triangular solve's own unknown) and is never used
      int jd;
      const T xv = gather(j[u], jd);
      const T p = a[u] * xv;
      acc[u & 3] += (SKIP_DIAG && jd == self) ? T(0) : p;
    }
  }
  if (q < e) {
    // the tail loads every slot unconditionally (a slot past the row end
    // re-reads the row's last entry and is then zeroed), so all its loads
    // and gathers are in flight at once instead of one branch at a time
    int j[R - 1];
    T a[R - 1];
#pragma unroll
    for (int u = 0; u < R - 1; u++) {
      const int k = q + u * LANES < e ? q + u * LANES : e - 1;
      j[u] = ja[k];
      a[u] = val[k];
    }
#pragma unroll
    for (int u = 0; u < R - 1; u++) {
      // a padding slot's coefficient is scaled to 0 (x there is a real
      // neighbour's value, finite); only x[self], which may be stale, is
      // selected away - on the product, since a load feeding a select
      // becomes a branch around the load
      const T hm = q + u * LANES < e ? T(1) : T(0);
      int jd;
      const T xv = gather(j[u], jd);
      const T p = (a[u] * hm) * xv;
      acc[u & 3] += (SKIP_DIAG && jd == self) ? T(0) : p;
    }
  }
  return (acc[0] + acc[1]) + (acc[2] + acc[3]);
}

template <typename T, int LANES, bool SKIP_DIAG, int R = (LANES == 2 ? 8 : 4)>
__device__ __forceinline__ T row_dot(int q, const int e, const int *__restrict__ ja,
                                     const T *__restrict__ val, const T *__restrict__ x, int self) {
  return row_dot_g<T, LANES, SKIP_DIAG, R>(
      q, e, ja, val,
      [x](int c, int &j) {
        j = c;
        return x[c];
      },
      self);
}

// Block-wide sum over one long row: each thread keeps 16 strides of
// loads in flight per round, so a row of tens of thousands of entries
// takes a few memory round trips instead of one per 4 strides.
// `gather(c)` returns x at column c.
template <typename T, typename G>
__device__ __forceinline__ T row_dot_deep_g(int q, const int e, const int *__restrict__ ja,
                                            const T *__restrict__ val, G gather) {
  constexpr int U = 16;
  T acc[4] = {0, 0, 0, 0};
  for (; q + (U - 1) * BLOCK < e; q += U * BLOCK) {
    int j[U];
    T a[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
      j[u] = ja[q + u * BLOCK];
      a[u] = val[q + u * BLOCK];
    }
#pragma unroll
    for (int u = 0; u < U; u++) acc[u & 3] += a[u] * gather(j[u]);
  }
  if (q < e) {
    // at most U - 1 strides left: one round, padding slots re-read the
    // row's last entry with a zero coefficient (no branch per load)
    int j[U - 1];
    T a[U - 1];
#pragma unroll
    for (int u = 0; u < U - 1; u++) {
      const int k = q + u * BLOCK < e ? q + u * BLOCK : e - 1;
      j[u] = ja[k];
      a[u] = val[k];
    }
#pragma unroll
    for (int u = 0; u < U - 1; u++) {
      const T hm = q + u * BLOCK < e ? T(1) : T(0);
      acc[u & 3] += (a[u] * hm) * gather(j[u]);
    }
  }
  return (acc[0] + acc[1]) + (acc[2] + acc[3]);
}

template <typename T>
__device__ __forceinline__ T row_dot_deep(int q, const int e, const int *__restrict__ ja,
                                          const T *__restrict__ val, const T *__restrict__ x) {
  return row_dot_deep_g<T>(q, e, ja, val, [x](int c) { return x[c]; });
}

// Grid for one-row-block-per-block kernels: rounded up to a multiple of
// 8 (when large) so logical_block() can give each XCD a contiguous slice.
inline int grid_rows(long long nrb) {
  if (nrb >= 64) return (int)((nrb + 7) & ~7LL);
  return (int)(nrb > 0 ? nrb : 1);
}

inline int grid_for(long long nrb) {
  long long g = nrb;
  if (g > BN_RED_BLOCKS) g = BN_RED_BLOCKS;
  if (g >= 8) g &= ~7LL;
  if (g < 1) g = 1;
  return (int)g;
}

// ---------------------------------------------------------------------
// The Jacobi work a CSR launch can do on the rows it finishes, so that a
// Jacobi-smoothed level needs no launch of its own for it:
//  * SET with xz: also xz_r = om y_r w_r, the first Jacobi step from zero
//    on the result (k_jacobi_zero's product), written beside y; with a
//    colour array only on the rows of colour c and zero elsewhere (om = 1,
//    w = 1/a_ii: k_diag_scale_colour's first Gauss-Seidel colour from
//    zero);
//  * JAC: y_r = x_r + om (b_r - (A x)_r) w_r, one Jacobi step - the
//    residual rounded as the RESID launch stores it, then
//    k_jacobi_update's update - into y, which must not be x.
// ---------------------------------------------------------------------
template <typename T>
struct jacobi_epi {
  T *xz;
  const T *w;
  T om;
  const unsigned char *colour;  // SET: xz_r only where colour_r == c, zero elsewhere
  int c;
};

template <typename T, int OP>
__device__ __forceinline__ T csr_finish(T s, T base, const T *__restrict__ x, int row, const jacobi_epi<T> &e) {
  if (OP == BN_OP_SET) return s;
  if (OP == BN_OP_RESID) return base - s;
  if (OP == BN_OP_JAC) {
    const T r = base - s;
    return x[row] + e.om * r * e.w[row];
  }
  return base + s;
}

template <typename T, int OP>
__device__ __forceinline__ void csr_finish_store(T *__restrict__ y, int row, T out, const jacobi_epi<T> &e) {
  y[row] = out;
  if (OP == BN_OP_SET && e.xz) {
    const T v = e.om * out * e.w[row];
    e.xz[row] = !e.colour || e.colour[row] == e.c ? v : T(0);
  }
}

// ---------------------------------------------------------------------
// CSR apply: y = A x | y = b - A x | y += A x | a Jacobi step (JAC),
// optional fused w.y dot
// ---------------------------------------------------------------------
template <typename T, int LANES, int OP, bool DOT, int R = (LANES == 2 ? 8 : 4)>
__global__ __launch_bounds__(BLOCK) void k_csr(int n, int chunk, const int *__restrict__ ia,
                                               const int *__restrict__ ja, const T *__restrict__ val,
                                               const T *__restrict__ x, T *__restrict__ y,
                                               const T *__restrict__ b, const T *__restrict__ w,
                                               double *__restrict__ partial, int long_len, jacobi_epi<T> e) {
  constexpr int RPB = BLOCK / LANES;
  const int lb = logical_block();
  const int sub = threadIdx.x % LANES;
  const int nrb = (n + RPB - 1) / RPB;
  const int rb0 = lb * chunk;
  int rb1 = rb0 + chunk;
  if (rb1 > nrb) rb1 = nrb;
  double acc = 0.0;
  for (int rb = rb0; rb < rb1; rb++) {
    const int row = rb * RPB + threadIdx.x / LANES;
    if (row < n) {
      const int q0 = ia[row], q1 = ia[row + 1];
      if (q1 - q0 > long_len) continue; /* handled by k_csr_long */
      // b / y ahead of the row's gathers: one memory round trip fewer
      const T base = (OP == BN_OP_RESID || OP == BN_OP_JAC) ? b[row] : OP == BN_OP_ADD ? y[row] : T(0);
      T s = row_dot<T, LANES, false, R>(q0 + sub, q1, ja, val, x, row);
      if (LANES > 1) s = group_sum<LANES>(s);
      if (sub == 0) {
        const T out = csr_finish<T, OP>(s, base, x, row, e);
        csr_finish_store<T, OP>(y, row, out, e);
        if (DOT) acc += (double)w[row] * (double)out;
      }
    }
  }
  if (DOT) {
    double t = block_sum(acc);
    if (threadIdx.x == 0) partial[lb] = t;
    if (lb == 0) clear_tail(partial);
  }
}

template <typename T, int LANES, int R = (LANES == 2 ? 8 : 4)>
void launch_csr(int op, int n, const int *ia, const int *ja, const T *val, const T *x, T *y,
                const T *b, const T *w, double *partial, int long_len, jacobi_epi<T> e, hipStream_t st) {
  constexpr int RPB = BLOCK / LANES;
  const long long nrb = ((long long)n + RPB - 1) / RPB;
#define BN_CSR_LAUNCH(OPV, DOTV, G, CH) \
  BN_LAUNCH((k_csr<T, LANES, OPV, DOTV, R>), dim3(G), dim3(BLOCK), 0, st, n, CH, ia, ja, val, x, y, b, w, partial, \
            long_len, e)
  // fused dot: a capped grid so the partials fit BN_RED_BLOCKS slots.
  // Past that many row blocks each block would walk several of them in
  // turn, which measured slower than the one-row-block-per-block launch
  // followed by a separate w.y pass: FP64 takes that route
  const bool split = partial && nrb > BN_RED_BLOCKS && sizeof(T) == sizeof(double);
  if (partial && !split) {
    const int g = grid_for(nrb);
    const int chunk = (int)((nrb + g - 1) / g);
    if (op == BN_OP_SET) BN_CSR_LAUNCH(BN_OP_SET, true, g, chunk);
    else if (op == BN_OP_RESID) BN_CSR_LAUNCH(BN_OP_RESID, true, g, chunk);
    else BN_CSR_LAUNCH(BN_OP_ADD, true, g, chunk);
  } else {
    // one row-block per block: the hardware keeps every CU full
    const int g = grid_rows(nrb);
    if (op == BN_OP_SET) BN_CSR_LAUNCH(BN_OP_SET, false, g, 1);
    else if (op == BN_OP_RESID) BN_CSR_LAUNCH(BN_OP_RESID, false, g, 1);
    else if (op == BN_OP_JAC) BN_CSR_LAUNCH(BN_OP_JAC, false, g, 1);
    else BN_CSR_LAUNCH(BN_OP_ADD, false, g, 1);
    if (split) bn_dot_partial(n, (const double *)w, (const double *)y, NULL, NULL, partial, NULL, st);
  }
#undef BN_CSR_LAUNCH
}

// lanes = 1..64, optionally with a BN_LANES_DEEP* bit
template <typename T, int LANES>
void launch_csr_rounds(int deep, int op, int n, const int *ia, const int *ja, const T *val, const T *x, T *y,
                       const T *b, const T *w, double *partial, int long_len, jacobi_epi<T> e, hipStream_t st) {
  if (deep == BN_LANES_DEEP16)
    launch_csr<T, LANES, 16>(op, n, ia, ja, val, x, y, b, w, partial, long_len, e, st);
  else if (deep == BN_LANES_DEEP8)
    launch_csr<T, LANES, 8>(op, n, ia, ja, val, x, y, b, w, partial, long_len, e, st);
  else
    launch_csr<T, LANES>(op, n, ia, ja, val, x, y, b, w, partial, long_len, e, st);
}

template <typename T>
void dispatch_csr(int op, int n, const int *ia, const int *ja, const T *val, const T *x, T *y,
                  const T *b, int lanes_in, const T *w, double *partial, int long_len, hipStream_t st,
                  jacobi_epi<T> e = jacobi_epi<T>{nullptr, nullptr, T(0), nullptr, 0}) {
  const int deep = lanes_in & (BN_LANES_DEEP8 | BN_LANES_DEEP16);
  const int lanes = lanes_in & BN_LANES_MASK;
  if (deep) {
    switch (lanes) {
      case 1: launch_csr_rounds<T, 1>(deep, op, n, ia, ja, val, x, y, b, w, partial, long_len, e, st); return;
      case 2: launch_csr_rounds<T, 2>(deep, op, n, ia, ja, val, x, y, b, w, partial, long_len, e, st); return;
      case 4: launch_csr_rounds<T, 4>(deep, op, n, ia, ja, val, x, y, b, w, partial, long_len, e, st); return;
      case 8: launch_csr_rounds<T, 8>(deep, op, n, ia, ja, val, x, y, b, w, partial, long_len, e, st); return;
      case 16: launch_csr_rounds<T, 16>(deep, op, n, ia, ja, val, x, y, b, w, partial, long_len, e, st); return;
      case 32: launch_csr_rounds<T, 32>(deep, op, n, ia, ja, val, x, y, b, w, partial, long_len, e, st); return;
      default: launch_csr_rounds<T, 64>(deep, op, n, ia, ja, val, x, y, b, w, partial, long_len, e, st); return;
    }
  }
  switch (lanes) {
    case 1: launch_csr<T, 1>(op, n, ia, ja, val, x, y, b, w, partial, long_len, e, st); break;
    case 2: launch_csr<T, 2>(op, n, ia, ja, val, x, y, b, w, partial, long_len, e, st); break;
    case 4: launch_csr<T, 4>(op, n, ia, ja, val, x, y, b, w, partial, long_len, e, st); break;
    case 8: launch_csr<T, 8>(op, n, ia, ja, val, x, y, b, w, partial, long_len, e, st); break;
    case 16: launch_csr<T, 16>(op, n, ia, ja, val, x, y, b, w, partial, long_len, e, st); break;
    case 32: launch_csr<T, 32>(op, n, ia, ja, val, x, y, b, w, partial, long_len, e, st); break;
    default: launch_csr<T, 64>(op, n, ia, ja, val, x, y, b, w, partial, long_len, e, st); break;
  }
}

// Long-row bin: one 256-thread block per listed row (rows longer than
// the main kernel's limit), finished with a block-wide reduction.
template <typename T, int OP>
__global__ __launch_bounds__(BLOCK) void k_csr_long(const int *__restrict__ rows, const int *__restrict__ ia,
                                                    const int *__restrict__ ja, const T *__restrict__ val,
                                                    const T *__restrict__ x, T *__restrict__ y,
                                                    const T *__restrict__ b) {
  const int row = rows[blockIdx.x];
  const T s = row_dot_deep<T>(ia[row] + threadIdx.x, ia[row + 1], ja, val, x);
  const double t = block_sum((double)s);
  if (threadIdx.x == 0) {
    if (OP == BN_OP_SET) y[row] = (T)t;
    else if (OP == BN_OP_RESID) y[row] = b[row] - (T)t;
    else y[row] = y[row] + (T)t;
  }
}

// Segmented long-row bin: long row t (rows[t], longest first) is cut
// into segments of BN_SEG_NNZ entries, segments seg_ptr[t] ..
// seg_ptr[t+1]-1, one block each.  A one-segment row is finished by its
// block.  A longer row's blocks publish their partial sums (agent-scope
// stores into part, which holds the sentinel between launches) and count
// themselves in cnt[t]; the block that completes the count adds the
// partials in segment order - the fixed order, whichever block it is -
// finishes the row, and puts back the sentinel and the zero count.  A
// partial is read with agent-scope loads until it is no longer the
// sentinel: its block stored it before counting itself, so the wait ends
// at once, and it is bounded (BN_SEG_SPINS polls) in any case.
#define BN_SEG_SPINS 4096
template <typename T, int OP>
__global__ __launch_bounds__(BLOCK) void k_csr_seg(const int *__restrict__ seg_row, const int *__restrict__ seg_ptr,
                                                   const int *__restrict__ rows, const int *__restrict__ ia,
                                                   const int *__restrict__ ja, const T *__restrict__ val,
                                                   const T *__restrict__ x, T *__restrict__ y,
                                                   const T *__restrict__ b, double *part, int *cnt,
                                                   jacobi_epi<T> e) {
  typedef unsigned long long U;
  const int sg = blockIdx.x;
  const int t = seg_row[sg];
  const int row = rows[t];
  const int s0 = seg_ptr[t], s1 = seg_ptr[t + 1];
  const int beg = ia[row] + (sg - s0) * BN_SEG_NNZ;
  const int end = min(beg + BN_SEG_NNZ, ia[row + 1]);
  const T v = row_dot<T, BLOCK, false>(beg + threadIdx.x, end, ja, val, x, row);
  double tsum = block_sum((double)v);
  if (s1 - s0 > 1) {
    __shared__ int last;
    __shared__ double ps[BLOCK];
    if (threadIdx.x == 0) {
      __hip_atomic_store((U *)(part + sg), __builtin_bit_cast(U, tsum), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");  // the store completes before the count
      last = __hip_atomic_fetch_add(cnt + t, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == s1 - s0 - 1;
    }
    __syncthreads();
    if (!last) return;
    // the row's partials, BLOCK at a time loaded together, added in order
    tsum = 0.0;
    for (int k0 = s0; k0 < s1; k0 += BLOCK) {
      const int k = k0 + threadIdx.x;
      if (k < s1) {
        U u = __hip_atomic_load((U *)(part + k), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        for (int spin = 0; u == walk_word<double>::none && spin < BN_SEG_SPINS; spin++) {
          __builtin_amdgcn_s_sleep(1);
          u = __hip_atomic_load((U *)(part + k), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
        ps[threadIdx.x] = __builtin_bit_cast(double, u);
        __hip_atomic_store((U *)(part + k), walk_word<double>::none, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      }
      __syncthreads();
      if (threadIdx.x == 0)
        for (int q = 0; q < BLOCK && k0 + q < s1; q++) tsum += ps[q];
      __syncthreads();
    }
    if (threadIdx.x == 0) __hip_atomic_store(cnt + t, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  if (threadIdx.x != 0) return;
  const T base = (OP == BN_OP_RESID || OP == BN_OP_JAC) ? b[row] : OP == BN_OP_ADD ? y[row] : T(0);
  csr_finish_store<T, OP>(y, row, csr_finish<T, OP>((T)tsum, base, x, row, e), e);
}

template <typename T>
void launch_seg(int op, int nlong, const int *rows, const int *seg_ptr, const int *seg_row, int nseg,
                double *part, int *cnt, const int *ia, const int *ja, const T *val, const T *x, T *y, const T *b,
                jacobi_epi<T> e, hipStream_t st) {
  if (nlong <= 0 || nseg <= 0) return;
#define BN_SEG(OPV)                                                                                          \
  BN_LAUNCH((k_csr_seg<T, OPV>), dim3(nseg), dim3(BLOCK), 0, st, seg_row, seg_ptr, rows, ia, ja, val, x, \
                     y, b, part, cnt, e)
  if (op == BN_OP_SET) {
    BN_SEG(BN_OP_SET);
  } else if (op == BN_OP_RESID) {
    BN_SEG(BN_OP_RESID);
  } else if (op == BN_OP_JAC) {
    BN_SEG(BN_OP_JAC);
  } else {
    BN_SEG(BN_OP_ADD);
  }
#undef BN_SEG
}

template <typename T>
void launch_long(int op, int nlong, const int *rows, const int *ia, const int *ja, const T *val, const T *x,
                 T *y, const T *b, hipStream_t st) {
  if (nlong <= 0) return;
  if (op == BN_OP_SET)
    BN_LAUNCH((k_csr_long<T, BN_OP_SET>), dim3(nlong), dim3(BLOCK), 0, st, rows, ia, ja, val, x, y, b);
  else if (op == BN_OP_RESID)
    BN_LAUNCH((k_csr_long<T, BN_OP_RESID>), dim3(nlong), dim3(BLOCK), 0, st, rows, ia, ja, val, x, y, b);
  else
    BN_LAUNCH((k_csr_long<T, BN_OP_ADD>), dim3(nlong), dim3(BLOCK), 0, st, rows, ia, ja, val, x, y, b);
}

// ---------------------------------------------------------------------
// Row-block CSR apply (adaptive): each block owns a run of consecutive
// rows holding <= BN_ROWBLOCK_NNZ entries (stream mode) or one longer row
// (vector mode).  Stream mode reads the run's ja/val as one contiguous,
// fully coalesced segment, stages val * x[ja] in LDS, then sums each row
// with tpr = pow2 <= 256 / rows threads.  Fixed summation order per row.
// ---------------------------------------------------------------------
// The block's work over an access policy `io`: io.gx(j) gathers x_j,
// io.base(r) is the value the row's sum is combined with (b_r for RESID,
// y_r for ADD), io.dsc(r) the SCALED factor and io.out(r, s, base, dsc)
// forms and stores the result and returns its term of a fused dot (0 when
// the policy has none).  prod: BN_ROWBLOCK_NNZ + 1 words of LDS.  Returns
// the sum of this thread's dot terms.
template <typename T, int OP, typename IO>
__device__ __forceinline__ double csr_block(int bid, const int *__restrict__ blocks, const int *__restrict__ blkq,
                                          const int *__restrict__ ia, const int *__restrict__ ja,
                                          const T *__restrict__ val, const IO &io, T *prod) {
  constexpr int PER = BN_ROWBLOCK_NNZ / BLOCK;
  const int r0 = blocks[bid], r1 = blocks[bid + 1];
  const int q0 = blkq ? blkq[bid] : ia[r0];
  const int q1 = blkq ? blkq[bid + 1] : ia[r1];
  const int cnt = q1 - q0;
  if (cnt > BN_ROWBLOCK_NNZ) {
    // vector mode: one long row (r1 == r0 + 1)
    const T base = io.base(r0);
    const T s = row_dot_deep_g<T>(q0 + threadIdx.x, q1, ja, val, [&io](int c) { return io.gx(c); });
    const double t = block_sum((double)s);
    return threadIdx.x == 0 ? io.out(r0, (T)t, base, io.dsc(r0)) : 0.0;
  }
  // tpr = largest power of two <= 256 / rows, at most one wave; with at
  // most 256 rows a block (bn_row_blocks' limit) one pass reduces them all
  const int nr = r1 - r0;
  int lg = nr >= BLOCK ? 0 : 31 - __clz(BLOCK / nr);
  if (lg > 6) lg = 6;
  const int tpr = 1 << lg;
  const int sub = threadIdx.x & (tpr - 1);
  // the first pass's row, its bounds and right-hand side, loaded now so
  // they arrive with the entries (a clamped row for threads without one)
  const int rf = r0 + (threadIdx.x >> lg);
  const int rc = rf < r1 ? rf : r1 - 1;
  const int e0 = ia[rc] - q0, e1 = ia[rc + 1] - q0;
  const T base0 = io.base(rc);
  const T dsc0 = io.dsc(rc);
  if (cnt > 0) {
    // every slot loads (past the block's entries: its last entry, with a
    // zero coefficient), so all loads and gathers are in flight together
    int jj[PER];
    T aa[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const int k = threadIdx.x + u * BLOCK;
      const int kk = k < cnt ? k : cnt - 1;
      jj[u] = ja[q0 + kk];
      aa[u] = val[q0 + kk];
    }
#pragma unroll
    for (int u = 0; u < PER; u++) {
      // an unconditional store (padding slots to the spare word, never
      // read) keeps every gather out of a branch
      const int k = threadIdx.x + u * BLOCK;
      const T hm = k < cnt ? T(1) : T(0);
      prod[k < cnt ? k : BN_ROWBLOCK_NNZ] = (aa[u] * hm) * io.gx(jj[u]);
    }
  }
  __syncthreads();
  double acc = 0.0;
  for (int rb = r0; rb < r1; rb += BLOCK >> lg) {
    const int r = rb + (threadIdx.x >> lg);
    const bool first = (rb == r0);
    int ks = e0, ke = e1;
    if (!first && r < r1) {
      ks = ia[r] - q0;
      ke = ia[r + 1] - q0;
    }
    T s = 0;
    if (r < r1)
      for (int k = ks + sub; k < ke; k += tpr) s += prod[k];
    for (int o = tpr >> 1; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (sub == 0 && r < r1) acc += io.out(r, s, first ? base0 : io.base(r), first ? dsc0 : io.dsc(r));
  }
  return acc;
}

// plain loads and stores: y = A x | y = b - A x | y += A x, and with
// SCALED (RESID only) also ys_r = y_r dsc_r, the first Jacobi step from
// zero of a triangle whose right-hand side is this residual (the product
// k_diag_scale would form); with DOT, out returns w_r y_r (in double) for
// the fused w.y dot
template <typename T, int OP, bool SCALED, bool DOT = false>
struct blocks_io {
  const T *__restrict__ x;
  T *__restrict__ y;
  const T *__restrict__ b;
  const T *__restrict__ dsc_;
  T *__restrict__ ys;
  jacobi_epi<T> e;
  const T *__restrict__ w;
  __device__ __forceinline__ T gx(int j) const { return x[j]; }
  __device__ __forceinline__ T base(int r) const {
    return (OP == BN_OP_RESID || OP == BN_OP_JAC) ? b[r] : OP == BN_OP_ADD ? y[r] : T(0);
  }
  __device__ __forceinline__ T dsc(int r) const { return SCALED ? dsc_[r] : T(0); }
  __device__ __forceinline__ double out(int r, T s, T base, T d) const {
    T o;
    if (OP == BN_OP_SET) {
      o = s;
      csr_finish_store<T, OP>(y, r, s, e);
    } else if (OP == BN_OP_RESID) {
      o = base - s;
      y[r] = o;
      if (SCALED) ys[r] = o * d;
    } else if (OP == BN_OP_JAC) {
      o = csr_finish<T, OP>(s, base, x, r, e);
      y[r] = o;
    } else {
      o = base + s;
      y[r] = o;
    }
    return DOT ? (double)w[r] * (double)o : 0.0;
  }
};

// Latency: a block's work is three dependent memory round trips.  (1) The
// block's row bounds and, from blkq (the entry offsets ia[blocks[k]],
// stored beside them), its entry bounds; (2) the entries' columns and
// values, and - issued with them - the bounds of the row this thread
// reduces and its b (or y) and dsc; (3) the x gathers.  Without blkq the
// entry bounds come from ia, one round trip more.  The reduction reads
// LDS only.
// With DOT, block k also writes partial[k] = sum of w_r y_r over its rows
// (a fixed order: each thread's rows in turn, then the block sum).
template <typename T, int OP, bool SCALED, bool DOT>
__device__ __forceinline__ void csr_blocks_body(int nblk, const int *__restrict__ blocks,
                                                const int *__restrict__ blkq, const int *__restrict__ ia,
                                                const int *__restrict__ ja, const T *__restrict__ val,
                                                const T *__restrict__ x, T *__restrict__ y,
                                                const T *__restrict__ b, const T *__restrict__ dsc,
                                                T *__restrict__ ys, jacobi_epi<T> e, const T *__restrict__ w,
                                                double *__restrict__ partial) {
  __shared__ T prod[BN_ROWBLOCK_NNZ + 1];  // + a spare word padding slots write
  const int bid = logical_block();
  if (bid >= nblk) return;
  const blocks_io<T, OP, SCALED, DOT> io{x, y, b, dsc, ys, e, w};
  const double acc = csr_block<T, OP>(bid, blocks, blkq, ia, ja, val, io, prod);
  if (DOT) {
    const double t = block_sum(acc);
    if (threadIdx.x == 0) partial[bid] = t;
  }
}

template <typename T, int OP, bool SCALED = false>
__global__ __launch_bounds__(BLOCK) void k_csr_blocks(int nblk, const int *__restrict__ blocks,
                                                      const int *__restrict__ blkq,
                                                      const int *__restrict__ ia, const int *__restrict__ ja,
                                                      const T *__restrict__ val, const T *__restrict__ x,
                                                      T *__restrict__ y, const T *__restrict__ b,
                                                      const T *__restrict__ dsc = nullptr,
                                                      T *__restrict__ ys = nullptr,
                                                      jacobi_epi<T> e = jacobi_epi<T>{nullptr, nullptr, T(0), nullptr, 0}) {
  csr_blocks_body<T, OP, SCALED, false>(nblk, blocks, blkq, ia, ja, val, x, y, b, dsc, ys, e, nullptr, nullptr);
}

// y = A x with the w.y dot fused: one partial per row block
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_csr_blocks_dot(int nblk, const int *__restrict__ blocks,
                                                          const int *__restrict__ blkq, const int *__restrict__ ia,
                                                          const int *__restrict__ ja, const T *__restrict__ val,
                                                          const T *__restrict__ x, T *__restrict__ y,
                                                          const T *__restrict__ w, double *__restrict__ partial) {
  csr_blocks_body<T, BN_OP_SET, false, true>(nblk, blocks, blkq, ia, ja, val, x, y, nullptr, nullptr, nullptr,
                                             jacobi_epi<T>{nullptr, nullptr, T(0), nullptr, 0}, w, partial);
}

template <typename T>
void launch_blocks(int op, int nblk, const int *blocks, const int *blkq, const int *ia, const int *ja, const T *val,
                   const T *x, T *y, const T *b, jacobi_epi<T> e, hipStream_t st) {
  if (nblk <= 0) return;
  const int g = grid_rows(nblk);
  const T *no_dsc = nullptr;
  T *no_ys = nullptr;
  if (op == BN_OP_SET)
    BN_LAUNCH((k_csr_blocks<T, BN_OP_SET>), dim3(g), dim3(BLOCK), 0, st, nblk, blocks, blkq, ia, ja, val, x, y, b,
              no_dsc, no_ys, e);
  else if (op == BN_OP_RESID)
    BN_LAUNCH((k_csr_blocks<T, BN_OP_RESID>), dim3(g), dim3(BLOCK), 0, st, nblk, blocks, blkq, ia, ja, val, x, y, b,
              no_dsc, no_ys, e);
  else if (op == BN_OP_JAC)
    BN_LAUNCH((k_csr_blocks<T, BN_OP_JAC>), dim3(g), dim3(BLOCK), 0, st, nblk, blocks, blkq, ia, ja, val, x, y, b,
              no_dsc, no_ys, e);
  else
    BN_LAUNCH((k_csr_blocks<T, BN_OP_ADD>), dim3(g), dim3(BLOCK), 0, st, nblk, blocks, blkq, ia, ja, val, x, y, b,
              no_dsc, no_ys, e);
}

// ---------------------------------------------------------------------
// One multicolour Gauss-Seidel colour by row blocks.  The colour's rows
// are rows lo..hi-1 of a colour-ordered copy of the level operator that
// holds only the off-diagonal entries (row r updates level row map[r]);
// blocks b0 .. b0 + nb - 1 of `blocks`/`blkq` cover exactly those rows.
// Same three round trips as k_csr_blocks: block bounds, then the entries
// with the row's map, b, dinv (and zbase), then the x gathers; each row
// is summed from LDS in a fixed order.
//   x_i = (b_i - sum_j a_rj x_j) dinv_r,   i = map[r]
// ---------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_gs_blocks(int b0, int nb, const int *__restrict__ blocks,
                                                     const int *__restrict__ blkq, const int *__restrict__ ja,
                                                     const T *__restrict__ val, const int *__restrict__ ia,
                                                     const int *__restrict__ map, const T *__restrict__ dinv,
                                                     const T *__restrict__ b, T *__restrict__ x,
                                                     const T *__restrict__ zbase, T *__restrict__ zsum) {
  constexpr int PER = BN_ROWBLOCK_NNZ / BLOCK;
  __shared__ T prod[BN_ROWBLOCK_NNZ + 1];
  const int lb = logical_block();
  if (lb >= nb) return;
  const int bid = b0 + lb;
  const int r0 = blocks[bid], r1 = blocks[bid + 1];
  const int q0 = blkq[bid], q1 = blkq[bid + 1];
  const int cnt = q1 - q0;
  if (cnt > BN_ROWBLOCK_NNZ) {
    // one long row
    const int i = map[r0];
    const T bi = b[i], dr = dinv[r0];
    const T s = row_dot_deep<T>(q0 + threadIdx.x, q1, ja, val, x);
    const double t = block_sum((double)s);
    if (threadIdx.x == 0) {
      const T v = (bi - (T)t) * dr;
      x[i] = v;
      if (zsum) zsum[i] = zbase[i] + v;
    }
    return;
  }
  const int nr = r1 - r0;
  int lg = nr >= BLOCK ? 0 : 31 - __clz(BLOCK / nr);
  if (lg > 6) lg = 6;
  const int tpr = 1 << lg;
  const int sub = threadIdx.x & (tpr - 1);
  const int rf = r0 + (threadIdx.x >> lg);
  const int rc = rf < r1 ? rf : r1 - 1;
  const int e0 = ia[rc] - q0, e1 = ia[rc + 1] - q0;
  const int i0 = map[rc];
  const T b0v = b[i0], d0v = dinv[rc];
  const T z0v = zsum ? zbase[i0] : T(0);
  if (cnt > 0) {
    int jj[PER];
    T aa[PER];
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const int k = threadIdx.x + u * BLOCK;
      const int kk = k < cnt ? k : cnt - 1;
      jj[u] = ja[q0 + kk];
      aa[u] = val[q0 + kk];
    }
#pragma unroll
    for (int u = 0; u < PER; u++) {
      const int k = threadIdx.x + u * BLOCK;
      const T hm = k < cnt ? T(1) : T(0);
      prod[k < cnt ? k : BN_ROWBLOCK_NNZ] = (aa[u] * hm) * x[jj[u]];
    }
  }
  // every gather of this colour is done before any of its rows is written
  // (rows of one colour are uncoupled, so no row reads another's result)
  __syncthreads();
  for (int rb = r0; rb < r1; rb += BLOCK >> lg) {
    const int r = rb + (threadIdx.x >> lg);
    const bool first = (rb == r0);
    int ks = e0, ke = e1;
    if (!first && r < r1) {
      ks = ia[r] - q0;
      ke = ia[r + 1] - q0;
    }
    T s = 0;
    if (r < r1)
      for (int k = ks + sub; k < ke; k += tpr) s += prod[k];
    for (int o = tpr >> 1; o > 0; o >>= 1) s += __shfl_xor(s, o);
    if (sub == 0 && r < r1) {
      const int i = first ? i0 : map[r];
      const T v = ((first ? b0v : b[i]) - s) * (first ? d0v : dinv[r]);
      x[i] = v;
      if (zsum) zsum[i] = (first ? z0v : zbase[i]) + v;
    }
  }
}

// ---------------------------------------------------------------------
// Gauss-Seidel colour / triangular level.  Row r of the stored matrix
// updates vector entry i:
//   RS_RANGE  r = i = lo + k            (rows stored in sweep order)
//   RS_LIST   r = i = rows[k]           (level-scheduled triangle)
//   RS_MAP    r = lo + k, i = rows[r]   (rows copied in colour order,
//                                        columns and vectors natural)
// y_i = (b_i - sum_{j != i} a_rj y_j) dinv_r   (dinv_r indexed like r)
// ---------------------------------------------------------------------
enum { RS_RANGE = 0, RS_LIST = 1, RS_MAP = 2 };

template <typename T, int LANES, int MODE>
__global__ __launch_bounds__(BLOCK) void k_rowsolve(int lo, int cnt, const int *__restrict__ rows,
                                                    const int *__restrict__ ia, const int *__restrict__ ja,
                                                    const T *__restrict__ val, const T *__restrict__ dinv,
                                                    const T *__restrict__ b, T *__restrict__ y,
                                                    const T *__restrict__ zbase, T *__restrict__ zsum) {
  constexpr int RPB = BLOCK / LANES;
  const int k = logical_block() * RPB + threadIdx.x / LANES;
  const int sub = threadIdx.x % LANES;
  if (k >= cnt) return;
  const int r = MODE == RS_LIST ? rows[k] : lo + k;
  const int i = MODE == RS_MAP ? rows[r] : r;
  // b_i and dinv_r ahead of the row's gathers: one memory round trip fewer
  const T bi = b[i], dr = dinv[r];
  T s = row_dot<T, LANES, true>(ia[r] + sub, ia[r + 1], ja, val, y, i);
  if (LANES > 1) s = group_sum<LANES>(s);
  if (sub == 0) {
    const T v = (bi - s) * dr;
    y[i] = v;
    if (zsum) zsum[i] = zbase[i] + v;
  }
}

template <typename T, int MODE>
void dispatch_rowsolve(int lo, int cnt, const int *rows, const int *ia, const int *ja, const T *val,
                       const T *dinv, const T *b, T *y, int lanes, const T *zb, T *zs, hipStream_t st) {
  if (cnt <= 0) return;
#define BN_RS(L)                                                                                   \
  BN_LAUNCH((k_rowsolve<T, L, MODE>), dim3(grid_rows((cnt + BLOCK / L - 1) / (BLOCK / L))), dim3(BLOCK), \
                     0, st, lo, cnt, rows, ia, ja, val, dinv, b, y, zb, zs)
  switch (lanes) {
    case 1: BN_RS(1); break;
    case 2: BN_RS(2); break;
    case 4: BN_RS(4); break;
    case 8: BN_RS(8); break;
    case 16: BN_RS(16); break;
    case 32: BN_RS(32); break;
    default: BN_RS(64); break;
  }
#undef BN_RS
}

// ---------------------------------------------------------------------
// Exact triangular levels from a level-ordered ELL copy.  Thread k of a
// level updates row i = rows[k]; slot s of its row sits at s * cnt + k, so
// the slot loads are issued with the row index and 1/t_ii instead of after
// them: one round trip for those, then the gathers of y (two dependent
// rounds per level against four through the CSR rows).  The sums are
// k_rowsolve<T, 1, RS_LIST>'s: the row's t-th entry goes to accumulator
// t & 3, the diagonal and the padding (which names the row itself)
// contribute zero, and the accumulators add as (0 + 1) + (2 + 3), so
// every y_i is the same to the last bit.
// ---------------------------------------------------------------------
// the row sum of W slots whose y values are already gathered (shared by
// the launch-per-level kernel and the one-launch walk, so both form the
// same expression and round alike)
template <typename T, int W>
__device__ __forceinline__ T ell_sum(const int *j, const T *a, const T *yv, int i) {
  T acc[4] = {0, 0, 0, 0};
#pragma unroll
  for (int s = 0; s < W; s++) {
    // select the product, not the gathered value: y_i itself may be stale
    const T p = a[s] * yv[s];
    acc[s & 3] += (j[s] == i) ? T(0) : p;
  }
  return (acc[0] + acc[1]) + (acc[2] + acc[3]);
}

template <typename T, int W>
__device__ __forceinline__ T ell_row(int k, int cnt, int i, const int *__restrict__ ja, const T *__restrict__ val,
                                     const T *y) {
  int j[W];
  T a[W], yv[W];
#pragma unroll
  for (int s = 0; s < W; s++) {
    j[s] = ja[s * cnt + k];
    a[s] = val[s * cnt + k];
  }
#pragma unroll
  for (int s = 0; s < W; s++) yv[s] = y[j[s]];
  return ell_sum<T, W>(j, a, yv, i);
}

template <typename T, int W>
__global__ __launch_bounds__(BLOCK) void k_trsv_ell(int cnt, const int *__restrict__ rows,
                                                    const int *__restrict__ ja, const T *__restrict__ val,
                                                    const T *__restrict__ dl, const T *__restrict__ b, T *y,
                                                    const T *__restrict__ zbase, T *__restrict__ zsum) {
  const int k = logical_block() * BLOCK + threadIdx.x;
  if (k >= cnt) return;
  const int i = rows[k];
  const T d = dl[k];
  const T s = ell_row<T, W>(k, cnt, i, ja, val, y);
  const T v = (b[i] - s) * d;
  y[i] = v;
  if (zsum) zsum[i] = zbase[i] + v;
}

// ---------------------------------------------------------------------
// A run of exact triangular levels l0 .. l1-1 of the ELL copy in ONE
// launch (each level of the run has 1 <= w <= BN_ELL_MAX_W slots per row).
// The rows are cut into np contiguous ranges ("parts"), one workgroup
// each; a workgroup takes its part's rows of each level in turn (kb gives
// their positions in the level's row list, which is ascending), with a
// workgroup barrier between levels.  A row's predecessors of its own part
// were written before that barrier; those of other parts may still be on
// their way.  Before the launch y holds the sentinel walk_word<T>::none on
// every row of the triangle, every y_i is written once, with an
// agent-scope atomic store, and never holds anything else while the launch
// runs (caches hold nothing older: the launch starts with them invalid).
// So a predecessor's value is first read with a plain load (for its own
// part from the CU's caches), and only if that still reads the sentinel
// does the row wait for it with agent-scope atomic loads, which are
// coherent across the XCDs' L2 caches.  The value is its own flag, so no
// ordering between two locations is needed and no cache is written back
// or invalidated as a whole.  With contiguous parts most predecessors of a
// lattice are in the same part, and the dependency chain crosses parts
// about np times instead of once per level.  No deadlock: a row waits only
// for rows of earlier levels, each part walks the levels in order, and
// the grid (np <= one workgroup per CU) is resident at once.  Every wait is
// bounded: past `limit` ticks of the constant-rate clock a thread raises
// *ctl, every other thread stops at its next check (still meeting every
// barrier), and the host re-runs the solve with a launch per level.  The
// sums are ell_sum<T, w>'s, those of the launch-per-level kernel, so y is
// the same to the last bit.
// ---------------------------------------------------------------------

template <typename T>
__device__ __forceinline__ typename walk_word<T>::U walk_load(const T *p) {
  return __hip_atomic_load((typename walk_word<T>::U *)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ bool walk_stopped(const double *ctl) {
  return __hip_atomic_load((unsigned long long *)ctl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0ULL;
}

// waits until the W values y_j of row i are computed (yv gets them), with
// coherent loads; false when this thread or another gave up
template <typename T, int W>
__device__ __forceinline__ bool walk_wait(const int *j, int i, const T *y, double *ctl, long long limit, T *yv) {
  typedef typename walk_word<T>::U U;
  long long t0 = 0;
  for (int spin = 0;; spin++) {
    bool ready = true;
#pragma unroll
    for (int s = 0; s < W; s++) {
      // y_i itself (the diagonal slot, padding) is never waited for
      const U u = walk_load(y + j[s]);
      yv[s] = __builtin_bit_cast(T, u);
      ready &= (j[s] == i) | (u != walk_word<T>::none);
    }
    if (ready) return true;
    // every 16th poll: has another thread given up, and is the limit past
    // (limit < 0 gives up at the first check: the host forces the fallback)
    if ((spin & 15) == 15) {
      if (walk_stopped(ctl)) return false;
      const long long t = (long long)wall_clock64();
      if (spin == 15) t0 = t;
      if (limit < 0 || (spin > 15 && t - t0 > limit)) {
        __hip_atomic_store((unsigned long long *)ctl, 0x3ff0000000000000ULL /* 1.0 */, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
        return false;
      }
    }
    __builtin_amdgcn_s_sleep(1);
  }
}

// the row's predecessors: plain loads first, the coherent wait only when
// one of them still reads the sentinel
template <typename T, int W>
__device__ __forceinline__ bool part_row(const int *j, const T *a, int i, const T *y, double *ctl, long long limit,
                                         T &sum) {
  typedef typename walk_word<T>::U U;
  T yv[W];
  bool ready = true;
#pragma unroll
  for (int s = 0; s < W; s++) {
    const U u = ((const U *)y)[j[s]];
    yv[s] = __builtin_bit_cast(T, u);
    ready &= (j[s] == i) | (u != walk_word<T>::none);
  }
  if (!ready && !walk_wait<T, W>(j, i, y, ctl, limit, yv)) return false;
  sum = ell_sum<T, W>(j, a, yv, i);
  return true;
}

// one row of a part's step: its index, 1/t_ii and slots (nothing computed
// in the launch, so they can be loaded a step ahead)
template <typename T>
struct part_item {
  int i;
  T d;
  int j[BN_ELL_MAX_W];
  T a[BN_ELL_MAX_W];
};

// step descriptor S = {first entry of the level's row list, rows, level
// rows * 16 + slots per row, first slot}: row k < S.y of the step is
// rows[S.x + k], slot s of it at S.w + s * (S.z >> 4) + k
template <typename T>
__device__ __forceinline__ void part_fetch(part_item<T> &it, const int4 S, int k, const int *__restrict__ rows,
                                           const int *__restrict__ ja, const T *__restrict__ val,
                                           const T *__restrict__ dl) {
  const int w = S.z & 15, cnt = S.z >> 4;
  it.i = rows[S.x + k];
  it.d = dl[S.x + k];
#pragma unroll
  for (int s = 0; s < BN_ELL_MAX_W; s++) {
    if (s < w) {
      it.j[s] = ja[S.w + s * cnt + k];
      it.a[s] = val[S.w + s * cnt + k];
    }
  }
}

template <typename T>
__device__ __forceinline__ bool part_solve(const part_item<T> &it, int w, const T *__restrict__ b, T *y,
                                           const T *__restrict__ zbase, T *__restrict__ zsum, double *ctl,
                                           long long limit) {
  typedef typename walk_word<T>::U U;
  const T bi = b[it.i];
  T sum = T(0);
  bool ok;
  switch (w) {
    case 1: ok = part_row<T, 1>(it.j, it.a, it.i, y, ctl, limit, sum); break;
    case 2: ok = part_row<T, 2>(it.j, it.a, it.i, y, ctl, limit, sum); break;
    case 3: ok = part_row<T, 3>(it.j, it.a, it.i, y, ctl, limit, sum); break;
    case 4: ok = part_row<T, 4>(it.j, it.a, it.i, y, ctl, limit, sum); break;
    case 5: ok = part_row<T, 5>(it.j, it.a, it.i, y, ctl, limit, sum); break;
    case 6: ok = part_row<T, 6>(it.j, it.a, it.i, y, ctl, limit, sum); break;
    case 7: ok = part_row<T, 7>(it.j, it.a, it.i, y, ctl, limit, sum); break;
    default: ok = part_row<T, 8>(it.j, it.a, it.i, y, ctl, limit, sum); break;
  }
  if (!ok) return false;
  const T v = (bi - sum) * it.d;
  __hip_atomic_store((U *)(y + it.i), __builtin_bit_cast(U, v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (zsum) zsum[it.i] = zbase[it.i] + v;
  return true;
}

// Part p's steps are steps[sp[p] .. sp[p+1]), its rows of each level of the
// run in level order; the first BN_PART_BLOCK rows of the next step are
// loaded while the current step waits for its predecessors.
template <typename T>
__global__ __launch_bounds__(BN_PART_BLOCK) void k_trsv_parts(const int *__restrict__ sp,
                                                              const int4 *__restrict__ steps,
                                                              const int *__restrict__ rows,
                                                              const int *__restrict__ ja, const T *__restrict__ val,
                                                              const T *__restrict__ dl, const T *__restrict__ b, T *y,
                                                              const T *__restrict__ zbase, T *__restrict__ zsum,
                                                              double *ctl, long long limit) {
  const int p = blockIdx.x, tid = threadIdx.x;
  const int t0 = sp[p], t1 = sp[p + 1];
  if (t0 >= t1) return;
  bool live = !walk_stopped(ctl);
  int4 S = steps[t0];
  part_item<T> cur, nxt;
  part_fetch(cur, S, tid < S.y ? tid : 0, rows, ja, val, dl);
  for (int t = t0; t < t1; t++) {
    const int4 Sn = steps[t + 1 < t1 ? t + 1 : t];
    if (t + 1 < t1) part_fetch(nxt, Sn, tid < Sn.y ? tid : 0, rows, ja, val, dl);
    const int w = S.z & 15;
    if (live && tid < S.y) live = part_solve(cur, w, b, y, zbase, zsum, ctl, limit);
    for (int k = tid + BN_PART_BLOCK; live && k < S.y; k += BN_PART_BLOCK) {
      part_item<T> more;
      part_fetch(more, S, k, rows, ja, val, dl);
      live = part_solve(more, w, b, y, zbase, zsum, ctl, limit);
    }
    __syncthreads();  // a failed thread still meets every barrier
    S = Sn;
    cur = nxt;
  }
}

// ---------------------------------------------------------------------
// One whole multicolour Gauss-Seidel sweep of a level in ONE launch (the
// colour launches' k_rowsolve<T, LANES, RS_MAP> rows, same sums).  The
// level's rows are cut into contiguous parts (level numbering), one
// workgroup each; a part takes its rows of each colour in sweep order
// (backward: colours in reverse), with a workgroup barrier between
// colours.  Before the launch x holds the sentinel on every row and xold
// the sweep's input (ZERO: the input is zero and xold is not read).  A row
// reads a neighbour of an earlier colour of the sweep from x (its new
// value, waited for while it still holds the sentinel) and one of a later
// colour from xold, which nothing writes during the sweep - exactly the
// values the colour launches read, without the launch between colours.
// The column words of the colour-ordered copy carry bit 31 when the
// neighbour's colour precedes the row's (jx, forward order).  Waits,
// bounds, the give-up flag and the deadlock argument are k_trsv_parts's.
// ---------------------------------------------------------------------
template <typename T, int LANES, bool ZERO, bool COHERENT>
__device__ __forceinline__ T gs_walk_sum(int r, int i, const int *__restrict__ ia, const int *__restrict__ jx,
                                         const T *__restrict__ val, const T *x, const T *__restrict__ xold,
                                         int backward, int sub, bool &bad) {
  typedef typename walk_word<T>::U U;
  bool b0 = false;
  const T s = row_dot_g<T, LANES, true>(
      ia[r] + sub, ia[r + 1], jx, val,
      [&](int c, int &j) -> T {
        j = c & 0x7fffffff;
        // the row's own entry (an operator with its diagonal) is never
        // waited for: its product is dropped
        const bool early = ((c < 0) != (backward != 0)) & (j != i);
        U u;
        if (ZERO) {
          // a later colour's input is zero: read x (the sentinel there
          // reads as "not early", so it never counts)
          u = COHERENT ? walk_load(x + j) : ((const U *)x)[j];
          u = early ? u : U(0);
        } else {
          const T *src = early ? x : xold;
          u = COHERENT ? walk_load(src + j) : ((const U *)src)[j];
        }
        b0 |= early & (u == walk_word<T>::none);
        return __builtin_bit_cast(T, u);
      },
      i);
  bad = b0;
  return s;
}

// One row of a sweep in a walk: the sum from plain loads, then, while an
// earlier row's value still reads the sentinel, the whole row again with
// coherent loads; writes x_i (and zsum).  false when this thread or another
// gave up (the row's lanes agree).
template <typename T, int LANES, bool ZERO>
__device__ __forceinline__ bool walk_row_solve(int r, int i, const int *__restrict__ ia, const int *__restrict__ jx,
                                               const T *__restrict__ val, const T *__restrict__ dinv,
                                               const T *__restrict__ b, T *x, const T *__restrict__ xold,
                                               const T *__restrict__ zbase, T *__restrict__ zsum, int backward,
                                               int sub, double *ctl, long long limit) {
  typedef typename walk_word<T>::U U;
  const T bi = b[i], dr = dinv[r];
  bool bad;
  T s = gs_walk_sum<T, LANES, ZERO, false>(r, i, ia, jx, val, x, xold, backward, sub, bad);
  if (LANES > 1) s = group_sum<LANES>(s);
  int nbad = group_sum<LANES>(bad ? 1 : 0);
  long long t0 = 0;
  for (int spin = 0; nbad; spin++) {
    // every 16th retry: has another thread given up, and is the limit past
    // (limit < 0 gives up at the first retry: the host forces the fallback)
    if (limit < 0 || (spin & 15) == 15) {
      const long long tc = (long long)wall_clock64();
      if (spin == 15) t0 = tc;
      if (walk_stopped(ctl) || limit < 0 || (spin > 15 && tc - t0 > limit)) {
        if (!walk_stopped(ctl))
          __hip_atomic_store((unsigned long long *)ctl, 0x3ff0000000000000ULL /* 1.0 */, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
        return false;
      }
    }
    __builtin_amdgcn_s_sleep(1);
    s = gs_walk_sum<T, LANES, ZERO, true>(r, i, ia, jx, val, x, xold, backward, sub, bad);
    if (LANES > 1) s = group_sum<LANES>(s);
    nbad = group_sum<LANES>(bad ? 1 : 0);
  }
  if (sub == 0) {
    const T v = (bi - s) * dr;
    __hip_atomic_store((U *)(x + i), __builtin_bit_cast(U, v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (zsum) zsum[i] = zbase[i] + v;
  }
  return true;
}

template <typename T, int LANES, bool ZERO>
__global__ __launch_bounds__(BN_GSW_BLOCK) void k_gs_walk(int ncol, int backward, const int2 *__restrict__ steps,
                                                          const int *__restrict__ ia, const int *__restrict__ jx,
                                                          const T *__restrict__ val, const int *__restrict__ map,
                                                          const T *__restrict__ dinvc, const T *__restrict__ b, T *x,
                                                          const T *__restrict__ xold, const T *__restrict__ zbase,
                                                          T *__restrict__ zsum, double *ctl, long long limit) {
  constexpr int RPP = BN_GSW_BLOCK / LANES;
  const int p = blockIdx.x, sub = threadIdx.x % LANES, slot = threadIdx.x / LANES;
  bool live = !walk_stopped(ctl);
  for (int t = 0; t < ncol; t++) {
    const int c = backward ? ncol - 1 - t : t;
    const int2 S = steps[p * ncol + c];  // {first row of the colour-ordered copy, rows}
    if (S.y == 0) continue;              // the same for the whole workgroup
    for (int k = slot; live && k < S.y; k += RPP) {
      const int r = S.x + k;
      live = walk_row_solve<T, LANES, ZERO>(r, map[r], ia, jx, val, dinvc, b, x, xold, zbase, zsum, backward, sub,
                                            ctl, limit);
    }
    __syncthreads();  // a failed thread still meets every barrier
  }
}

// ---------------------------------------------------------------------
// An exact-order Gauss-Seidel sweep (the sequential sweep's rows grouped by
// dependency level) in ONE launch: the level operator's rows are cut into
// contiguous parts, one workgroup each, and a part walks its rows of each
// dependency level in turn (steps[sp[p] .. sp[p+1]): {first entry of the
// level's row list, rows}), with a barrier between levels.  A row reads a
// neighbour the sequential sweep visits earlier from x (its new value,
// waited for while it holds the sentinel) and one it visits later from
// xold, the sweep's input: bit 31 of the column word (jx) marks the
// neighbours visited earlier in the forward order.  The sums are
// k_rowsolve<T, LANES, RS_LIST>'s, to the last bit.
// ---------------------------------------------------------------------
template <typename T, int LANES>
__global__ __launch_bounds__(BN_PART_BLOCK) void k_sweep_parts(int backward, const int *__restrict__ sp,
                                                               const int2 *__restrict__ steps,
                                                               const int *__restrict__ rows,
                                                               const int *__restrict__ ia, const int *__restrict__ jx,
                                                               const T *__restrict__ val, const T *__restrict__ dinv,
                                                               const T *__restrict__ b, T *x,
                                                               const T *__restrict__ xold, double *ctl,
                                                               long long limit) {
  constexpr int RPP = BN_PART_BLOCK / LANES;
  const int p = blockIdx.x, sub = threadIdx.x % LANES, slot = threadIdx.x / LANES;
  const int t0 = sp[p], t1 = sp[p + 1];
  if (t0 >= t1) return;
  bool live = !walk_stopped(ctl);
  for (int t = t0; t < t1; t++) {
    const int2 S = steps[t];
    for (int k = slot; live && k < S.y; k += RPP) {
      const int i = rows[S.x + k];
      live = walk_row_solve<T, LANES, false>(i, i, ia, jx, val, dinv, b, x, xold, nullptr, nullptr, backward, sub,
                                             ctl, limit);
    }
    __syncthreads();  // a failed thread still meets every barrier
  }
}

// before a sweep: xold = x (unless the input is zero), x = the sentinel
template <typename T>
__global__ void k_gs_prep(int n, T *__restrict__ x, T *__restrict__ xold) {
  typedef typename walk_word<T>::U U;
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (xold) xold[i] = x[i];
  ((U *)x)[i] = walk_word<T>::none;
}

template <typename U>
__global__ void k_fill_word(int n, U *__restrict__ x, U v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}

__global__ void k_ell_gather(long long n, const int *__restrict__ src, const double *__restrict__ from,
                             double *__restrict__ out, float *__restrict__ outs) {
  const long long q = (long long)blockIdx.x * BLOCK + threadIdx.x;
  if (q >= n) return;
  const int k = src[q];
  const double v = k >= 0 ? from[k] : 0.0;
  out[q] = v;
  outs[q] = (float)v;
}

// ---------------------------------------------------------------------
// Block Gauss-Seidel (l1 hybrid): one workgroup per chunk of consecutive
// rows.  The chunk's iterate lives in LDS for the whole sweep; the chunk's
// local colours are swept in order with workgroup barriers between them,
// so couplings inside the chunk see Gauss-Seidel values and couplings to
// other chunks read the previous iterate xold (never written here), which
// makes the sweep one launch and independent of workgroup timing.  Row
// update in residual form, x_i += (b_i - sum_j a_ij x~_j) * dinv_i with
// dinv_i = 1 / (a_ii + sum of |a_ij| over columns outside the chunk), so a
// sweep is x + M^{-1} (b - A x) with M = D1 + L_in (forward) or D1 + U_in
// (backward) = M_forward^T: the V-cycle stays symmetric.
// ---------------------------------------------------------------------
#define GS_BLOCK 512

template <typename T, bool ZERO>
__device__ __forceinline__ T chunk_x(int j, const T *xs, const T *__restrict__ xold, int lo, unsigned S) {
  const unsigned jl = (unsigned)(j - lo);
  if (jl < S) return xs[jl];
  return ZERO ? T(0) : xold[j];
}

template <typename T, int LANES, bool ZERO>
__device__ __forceinline__ T chunk_row_dot(int q, const int e, const int *__restrict__ ja,
                                           const T *__restrict__ val, const T *xs, const T *__restrict__ xold,
                                           int lo, unsigned S) {
  T s0 = 0, s1 = 0;
  for (; q + 3 * LANES < e; q += 4 * LANES) {
    const int j0 = ja[q], j1 = ja[q + LANES], j2 = ja[q + 2 * LANES], j3 = ja[q + 3 * LANES];
    const T a0 = val[q], a1 = val[q + LANES], a2 = val[q + 2 * LANES], a3 = val[q + 3 * LANES];
    const T x0 = chunk_x<T, ZERO>(j0, xs, xold, lo, S);
    const T x1 = chunk_x<T, ZERO>(j1, xs, xold, lo, S);
    const T x2 = chunk_x<T, ZERO>(j2, xs, xold, lo, S);
    const T x3 = chunk_x<T, ZERO>(j3, xs, xold, lo, S);
    s0 += a0 * x0 + a1 * x1;
    s1 += a2 * x2 + a3 * x3;
  }
  for (; q < e; q += LANES) s0 += val[q] * chunk_x<T, ZERO>(ja[q], xs, xold, lo, S);
  return s0 + s1;
}

template <typename T, int LANES, bool BACKWARD, bool ZERO>
__global__ __launch_bounds__(GS_BLOCK) void k_block_gs(const int *__restrict__ chunk, const int *__restrict__ ccol,
                                                       const int *__restrict__ cbnd, const int *__restrict__ ia,
                                                       const int *__restrict__ ja, const T *__restrict__ val,
                                                       const T *__restrict__ dinv, const T *__restrict__ b,
                                                       const T *__restrict__ xold, T *__restrict__ xnew,
                                                       const T *__restrict__ zbase, T *__restrict__ zsum) {
  extern __shared__ double bn_smem[]; /* 8-byte aligned for either T */
  T *xs = reinterpret_cast<T *>(bn_smem);
  const int k = logical_block();
  const int lo = chunk[k], hi = chunk[k + 1];
  const unsigned S = (unsigned)(hi - lo);
  for (int i = threadIdx.x; i < (int)S; i += GS_BLOCK) xs[i] = ZERO ? T(0) : xold[lo + i];
  __syncthreads();
  const int c0 = ccol[k], nc = ccol[k + 1] - c0 - 1;
  const int sub = threadIdx.x % LANES;
  for (int t = 0; t < nc; t++) {
    const int c = BACKWARD ? nc - 1 - t : t;
    const int r1 = cbnd[c0 + c + 1];
    for (int r = cbnd[c0 + c] + (int)threadIdx.x / LANES; r < r1; r += GS_BLOCK / LANES) {
      // row sum over the whole row, the diagonal on the row's old value
      T s = chunk_row_dot<T, LANES, ZERO>(ia[r] + sub, ia[r + 1], ja, val, xs, xold, lo, S);
      if (LANES > 1) s = group_sum<LANES>(s);
      if (sub == 0) xs[r - lo] += (b[r] - s) * dinv[r];
    }
    __syncthreads();
  }
  for (int i = threadIdx.x; i < (int)S; i += GS_BLOCK) {
    const T v = xs[i];
    xnew[lo + i] = v;
    if (zsum) zsum[lo + i] = zbase[lo + i] + v;
  }
}

template <typename T, int LANES>
void launch_block_gs(int nchunk, int max_rows, const int *chunk, const int *ccol, const int *cbnd, const int *ia,
                     const int *ja, const T *val, const T *dinv, const T *b, const T *xold, T *xnew, int backward,
                     const T *zb, T *zs, hipStream_t st) {
  const size_t sh = (size_t)max_rows * sizeof(T);
#define BN_BGS(BW, Z)                                                                                        \
  BN_LAUNCH((k_block_gs<T, LANES, BW, Z>), dim3(nchunk), dim3(GS_BLOCK), sh, st, chunk, ccol, cbnd, ia, \
                     ja, val, dinv, b, xold, xnew, zb, zs)
  if (backward) {
    if (xold) BN_BGS(true, false);
    else BN_BGS(true, true);
  } else {
    if (xold) BN_BGS(false, false);
    else BN_BGS(false, true);
  }
#undef BN_BGS
}

template <typename T>
void dispatch_block_gs(int nchunk, int max_rows, const int *chunk, const int *ccol, const int *cbnd, const int *ia,
                       const int *ja, const T *val, const T *dinv, const T *b, const T *xold, T *xnew,
                       int backward, int lanes, const T *zb, T *zs, hipStream_t st) {
  if (nchunk <= 0) return;
  switch (lanes) {
    case 1: launch_block_gs<T, 1>(nchunk, max_rows, chunk, ccol, cbnd, ia, ja, val, dinv, b, xold, xnew, backward, zb, zs, st); break;
    case 2: launch_block_gs<T, 2>(nchunk, max_rows, chunk, ccol, cbnd, ia, ja, val, dinv, b, xold, xnew, backward, zb, zs, st); break;
    case 4: launch_block_gs<T, 4>(nchunk, max_rows, chunk, ccol, cbnd, ia, ja, val, dinv, b, xold, xnew, backward, zb, zs, st); break;
    case 8: launch_block_gs<T, 8>(nchunk, max_rows, chunk, ccol, cbnd, ia, ja, val, dinv, b, xold, xnew, backward, zb, zs, st); break;
    case 16: launch_block_gs<T, 16>(nchunk, max_rows, chunk, ccol, cbnd, ia, ja, val, dinv, b, xold, xnew, backward, zb, zs, st); break;
    case 32: launch_block_gs<T, 32>(nchunk, max_rows, chunk, ccol, cbnd, ia, ja, val, dinv, b, xold, xnew, backward, zb, zs, st); break;
    default: launch_block_gs<T, 64>(nchunk, max_rows, chunk, ccol, cbnd, ia, ja, val, dinv, b, xold, xnew, backward, zb, zs, st); break;
  }
}

// SCALED: also write xs_i = x_i dinv_i, the next triangle's first Jacobi
// step from zero when its right-hand side is this sweep's x and both
// triangles share the diagonal (the product k_diag_scale would form)
template <typename T, int LANES, bool SCALED = false>
__global__ __launch_bounds__(BLOCK) void k_jacobi_rows(int n, const int *__restrict__ ia,
                                                       const int *__restrict__ ja, const T *__restrict__ val,
                                                       const T *__restrict__ dinv, const T *__restrict__ b,
                                                       const T *__restrict__ xin, T *__restrict__ xout,
                                                       const T *__restrict__ zbase, T *__restrict__ zsum,
                                                       T *__restrict__ xs = nullptr) {
  constexpr int RPB = BLOCK / LANES;
  const int i = logical_block() * RPB + threadIdx.x / LANES;
  const int sub = threadIdx.x % LANES;
  if (i >= n) return;
  const T bi = b[i], di = dinv[i];
  T s = row_dot<T, LANES, true>(ia[i] + sub, ia[i + 1], ja, val, xin, i);
  if (LANES > 1) s = group_sum<LANES>(s);
  if (sub == 0) {
    const T v = (bi - s) * di;
    if (zsum) zsum[i] = zbase[i] + v;
    else xout[i] = v;
    if (SCALED) xs[i] = v * di;
  }
}

template <typename T>
void dispatch_jacobi_rows(int n, const int *ia, const int *ja, const T *val, const T *dinv, const T *b,
                          const T *xin, T *xout, int lanes, const T *zb, T *zs, hipStream_t st) {
#define BN_JR(L)                                                                                       \
  BN_LAUNCH((k_jacobi_rows<T, L>), dim3(grid_rows((n + BLOCK / L - 1) / (BLOCK / L))), dim3(BLOCK), 0, st, \
                     n, ia, ja, val, dinv, b, xin, xout, zb, zs, (T *)nullptr)
  switch (lanes) {
    case 1: BN_JR(1); break;
    case 2: BN_JR(2); break;
    case 4: BN_JR(4); break;
    case 8: BN_JR(8); break;
    case 16: BN_JR(16); break;
    case 32: BN_JR(32); break;
    default: BN_JR(64); break;
  }
#undef BN_JR
}

template <typename T>
void dispatch_jacobi_rows_scaled(int n, const int *ia, const int *ja, const T *val, const T *dinv, const T *b,
                                 const T *xin, T *xout, T *xs, int lanes, hipStream_t st) {
#define BN_JRS(L)                                                                                        \
  BN_LAUNCH((k_jacobi_rows<T, L, true>), dim3(grid_rows((n + BLOCK / L - 1) / (BLOCK / L))), dim3(BLOCK), \
                     0, st, n, ia, ja, val, dinv, b, xin, xout, (const T *)nullptr, (T *)nullptr, xs)
  switch (lanes) {
    case 1: BN_JRS(1); break;
    case 2: BN_JRS(2); break;
    case 4: BN_JRS(4); break;
    case 8: BN_JRS(8); break;
    case 16: BN_JRS(16); break;
    case 32: BN_JRS(32); break;
    default: BN_JRS(64); break;
  }
#undef BN_JRS
}

// first colour from zero in RS_MAP layout: x_i = b_i dinv_r, i = map[r],
// for r < hi; with all_n > hi the remaining rows (later colours) are
// zeroed in the same pass, which replaces a separate fill of x.
template <typename T>
__global__ void k_diag_scale_map(int lo, int hi, int all_n, const int *__restrict__ map,
                                 const T *__restrict__ dinv, const T *__restrict__ b, T *__restrict__ x) {
  const int r = lo + blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= all_n) return;
  const int i = map[r];
  x[i] = r < hi ? b[i] * dinv[r] : T(0);
}

// The same first colour from zero in level-row order: row i of colour c
// gets b_i / a_ii, every other row 0, all accesses coalesced (dinv holds
// the same 1/a_ii values as the colour-ordered copy)
template <typename T>
__global__ void k_diag_scale_colour(int n, int c, const unsigned char *__restrict__ colour,
                                    const T *__restrict__ dinv, const T *__restrict__ b, T *__restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const T v = b[i] * dinv[i];
  x[i] = colour[i] == c ? v : T(0);
}

template <typename T>
__global__ void k_diag_scale(int lo, int cnt, const T *__restrict__ dinv, const T *__restrict__ b,
                             T *__restrict__ x) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k < cnt) x[lo + k] = b[lo + k] * dinv[lo + k];
}

template <typename T>
__global__ void k_jacobi_zero(int n, const T *__restrict__ b, const T *__restrict__ d, T omega,
                              T *__restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = omega * b[i] * d[i];
}

template <typename T>
__global__ void k_jacobi_update(int n, const T *__restrict__ r, const T *__restrict__ d, T omega,
                                T *__restrict__ x) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] += omega * r[i] * d[i];
}

// ---------------------------------------------------------------------
// dense y = M x: one block per row, coalesced row read
// ---------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_gemv(int n, const T *__restrict__ M, const T *__restrict__ x,
                                                T *__restrict__ y) {
  const int row = blockIdx.x;
  const T *m = M + (size_t)row * (size_t)n;
  double s = 0.0;
  for (int j = threadIdx.x; j < n; j += BLOCK) s += (double)m[j] * (double)x[j];
  s = block_sum(s);
  if (threadIdx.x == 0) y[row] = (T)s;
}

// the FP32 V-cycle's coarsest solve: the FP64 inverse applied to an FP32
// right-hand side, the FP64 result rounded to FP32 - the widening, the
// FP64 product and the rounding that three launches form, in one
__global__ __launch_bounds__(BLOCK) void k_gemv_ds(int n, const double *__restrict__ M, const float *__restrict__ x,
                                                   float *__restrict__ y) {
  const int row = blockIdx.x;
  const double *m = M + (size_t)row * (size_t)n;
  double s = 0.0;
  for (int j = threadIdx.x; j < n; j += BLOCK) s += m[j] * (double)x[j];
  s = block_sum(s);
  if (threadIdx.x == 0) y[row] = (float)s;
}

// ---------------------------------------------------------------------
// BLAS-1
// ---------------------------------------------------------------------
template <typename T>
__global__ void k_fill(int n, T *x, T v) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) x[i] = v;
}
template <typename T>
__global__ void k_copy(int n, const T *__restrict__ x, T *__restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = x[i];
}
template <typename A, typename B>
__global__ void k_convert(int n, const A *__restrict__ x, B *__restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = (B)x[i];
}
template <typename T>
__global__ void k_axpy1(int n, const T *__restrict__ x, T *__restrict__ y) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] += x[i];
}

__global__ __launch_bounds__(BLOCK) void k_dot2(int n, int chunk, const double *__restrict__ x,
                                                const double *__restrict__ y, const double *__restrict__ u,
                                                const double *__restrict__ v, double *__restrict__ p1,
                                                double *__restrict__ p2) {
  const int lb = logical_block();
  const long long i0 = (long long)lb * chunk * BLOCK;
  long long i1 = i0 + (long long)chunk * BLOCK;
  if (i1 > n) i1 = n;
  double a = 0.0, c = 0.0;
  for (long long i = i0 + threadIdx.x; i < i1; i += BLOCK) {
    a += x[i] * y[i];
    if (u) c += u[i] * v[i];
  }
  a = block_sum(a);
  if (u) c = block_sum(c);
  if (threadIdx.x == 0) {
    p1[lb] = a;
    if (u) p2[lb] = c;
  }
  if (lb == 0) {
    clear_tail(p1);
    if (u) clear_tail(p2);
  }
}

// partials of r.z (and of z.q when q != NULL) with z held in FP32: the
// products and their order are k_dot2's over the widened z
__global__ __launch_bounds__(BLOCK) void k_dot_rz_s(int n, int chunk, const double *__restrict__ r,
                                                    const float *__restrict__ zs, const double *__restrict__ q,
                                                    double *__restrict__ p1, double *__restrict__ p2) {
  const int lb = logical_block();
  const long long i0 = (long long)lb * chunk * BLOCK;
  long long i1 = i0 + (long long)chunk * BLOCK;
  if (i1 > n) i1 = n;
  double a = 0.0, c = 0.0;
  for (long long i = i0 + threadIdx.x; i < i1; i += BLOCK) {
    const double z = (double)zs[i];
    a += r[i] * z;
    if (q) c += z * q[i];
  }
  a = block_sum(a);
  if (q) c = block_sum(c);
  if (threadIdx.x == 0) {
    p1[lb] = a;
    if (q) p2[lb] = c;
  }
  if (lb == 0) {
    clear_tail(p1);
    if (q) clear_tail(p2);
  }
}

// WRITE_S: also store r's FP32 rounding into rs (the V-cycle's input)
// WRITE_S: also r rounded to FP32 into rs, and when x0 != NULL the first
// colour of a multicolour sweep from zero with rs as right-hand side
// (k_diag_scale_colour's x0_i = rs_i dinv_i on colour c, zero elsewhere)
template <bool WRITE_S>
__global__ __launch_bounds__(BLOCK) void k_update_xr(int n, int chunk, double *__restrict__ x,
                                                     const double *__restrict__ p, double *__restrict__ r,
                                                     const double *__restrict__ q, const double *__restrict__ s,
                                                     double *__restrict__ part, float *__restrict__ rs,
                                                     const unsigned char *__restrict__ colour, int c,
                                                     const float *__restrict__ dinv, float *__restrict__ x0) {
  const int lb = logical_block();
  const double alpha = s[BN_S_ALPHA];
  const long long i0 = (long long)lb * chunk * BLOCK;
  long long i1 = i0 + (long long)chunk * BLOCK;
  if (i1 > n) i1 = n;
  double a = 0.0;
  for (long long i = i0 + threadIdx.x; i < i1; i += BLOCK) {
    x[i] = x[i] + p[i] * alpha;
    const double rn = r[i] + q[i] * (-alpha);
    r[i] = rn;
    if (WRITE_S) {
      const float rf = (float)rn;
      rs[i] = rf;
      if (x0) {
        const float v = rf * dinv[i];
        x0[i] = colour[i] == c ? v : 0.0f;
      }
    }
    a += rn * rn;
  }
  a = block_sum(a);
  if (threadIdx.x == 0) part[lb] = a;
  if (lb == 0) clear_tail(part);
}

__global__ void k_update_p(int n, double *__restrict__ p, const double *__restrict__ z,
                           const double *__restrict__ s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double beta = s[BN_S_BETA];
  if (i < n) p[i] = beta * p[i] + z[i];
}

// p = beta p + z with z held in FP32 (widening is exact: the same sums as
// k_update_p over the widened copy)
__global__ void k_update_p_s(int n, double *__restrict__ p, const float *__restrict__ zs,
                             const double *__restrict__ s) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const double beta = s[BN_S_BETA];
  if (i < n) p[i] = beta * p[i] + (double)zs[i];
}

// hs != NULL: the scalars are also copied to hs (host memory), so the host
// reads them without a copy of its own after the step
__global__ __launch_bounds__(BLOCK) void k_scalar_finish(int which, const double *p1, const double *p2,
                                                         const double *p3, double *s, int slot, double *hs,
                                                         int n1) {
  // fixed-order finish of BN_RED_BLOCKS partials per array by one block:
  // each thread sums its strided slots in ascending order, then the block
  // sum; the three arrays' loads are issued in one pass
  // (an absent array reads p1's slots instead, so no load sits behind a
  // branch; its sum is not used).  n1 > 0: p1 alone, with n1 slots (one
  // per row block of a fused product), 64 loads in flight per thread (a
  // 4.4M-row product's 17k slots in two rounds) and each thread's slots
  // summed in ascending order
  const double *q1 = p1 ? p1 : (p2 ? p2 : p3);
  const double *q2 = p2 ? p2 : q1, *q3 = p3 ? p3 : q1;
  double va = 0.0, vb = 0.0, vc = 0.0;
  if (n1 > 0) {
    for (int k0 = threadIdx.x; k0 < n1; k0 += 64 * BLOCK) {
      double t[64];
#pragma unroll
      for (int u = 0; u < 64; u++) {
        const int k = k0 + u * BLOCK;
        t[u] = k < n1 ? p1[k] : 0.0;
      }
#pragma unroll
      for (int u = 0; u < 64; u++) va += t[u];
    }
  } else if (q1) {
#pragma unroll
    for (int k = threadIdx.x; k < BN_RED_BLOCKS; k += BLOCK) {
      va += q1[k];
      vb += q2[k];
      vc += q3[k];
    }
  }
  const double a = p1 ? block_sum(va) : 0.0;
  const double b = p2 ? block_sum(vb) : 0.0;
  const double c = p3 ? block_sum(vc) : 0.0;
  if (threadIdx.x != 0) return;
  switch (which) {
    case 0:
      s[BN_S_RZ] = a;
      s[BN_S_RZ0] = a;
      s[BN_S_RR] = b;
      s[BN_S_REL] = 1.0;
      s[BN_S_BETA] = 0.0; /* first p = z + 0 p with p zeroed */
      break;
    case 1:
      s[BN_S_PAP] = a;
      s[BN_S_ALPHA] = s[BN_S_RZ] / a;
      break;
    case 2:
    case 4: {
      s[BN_S_RZNEW] = a;
      s[BN_S_RR] = b;
      s[BN_S_REL] = sqrt(a / s[BN_S_RZ0]);
      s[BN_S_BETA] = (which == 4) ? -s[BN_S_ALPHA] * c / s[BN_S_RZ] : a / s[BN_S_RZ];
      s[BN_S_RZOLD] = s[BN_S_RZ];
      s[BN_S_RZ] = a;
      break;
    }
    case 3:
      s[BN_S_BB] = a;
      break;
    default:
      s[slot] = a;
      break;
  }
  if (hs)
    for (int k = 0; k < BN_S_COUNT; k++) hs[k] = s[k];
}

__global__ void k_permute(int n, const int *__restrict__ perm, const double *__restrict__ x,
                          double *__restrict__ y, int scatter) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= n) return;
  if (scatter) y[perm[k]] = x[k];
  else y[k] = x[perm[k]];
}

__global__ void k_edge_grad(int m, const int *__restrict__ eu, const int *__restrict__ ev,
                            const double *__restrict__ g, const double *__restrict__ p,
                            const double *__restrict__ lam, const double *__restrict__ pdir,
                            double *__restrict__ dg, double *__restrict__ gr) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= m) return;
  const int u = eu[e], v = ev[e];
  const double pu = u >= 0 ? p[u] : pdir[-1 - u];
  const double pv = v >= 0 ? p[v] : pdir[-1 - v];
  const double lu = u >= 0 ? lam[u] : 0.0;
  const double lv = v >= 0 ? lam[v] : 0.0;
  const double d = -(lu - lv) * (pu - pv);
  if (dg) dg[e] = d;
  if (gr) gr[e] = 4.0 * g[e] * d;
}

inline dim3 g1(long long n) { return dim3((unsigned)((n + BLOCK - 1) / BLOCK)); }

template <typename T>
void trsv_ell_t(int w, int cnt, const int *rows, const int *ja, const T *val, const T *dl, const T *b, T *y,
                       const T *zb, T *zs, hipStream_t st) {
  const dim3 g(grid_rows((cnt + BLOCK - 1) / BLOCK));
#define BN_TE(W) BN_LAUNCH((k_trsv_ell<T, W>), g, dim3(BLOCK), 0, st, cnt, rows, ja, val, dl, b, y, zb, zs)
  switch (w) {
    case 1: BN_TE(1); break;
    case 2: BN_TE(2); break;
    case 3: BN_TE(3); break;
    case 4: BN_TE(4); break;
    case 5: BN_TE(5); break;
    case 6: BN_TE(6); break;
    case 7: BN_TE(7); break;
    default: BN_TE(8); break;
  }
#undef BN_TE
}

}  // namespace

// =====================================================================
// C launchers
// =====================================================================
extern "C" {

void bn_time_next_launch(hipEvent_t start, hipEvent_t stop) {
  g_time_start = start;
  g_time_stop = stop;
}

void bn_csr_apply(char vt, int op, int n, const int *ia, const int *ja, const void *val, const void *x,
                  void *y, const void *b, int lanes, const void *dotw, double *partial,
                  hipStream_t stream) {
  if (n <= 0) {
    if (partial) bn_fill('d', BN_RED_BLOCKS, partial, 0.0, stream);
    return;
  }
  if (vt == 's')
    dispatch_csr<float>(op, n, ia, ja, (const float *)val, (const float *)x, (float *)y,
                        (const float *)b, lanes, (const float *)dotw, partial, 0x7fffffff, stream);
  else
    dispatch_csr<double>(op, n, ia, ja, (const double *)val, (const double *)x, (double *)y,
                         (const double *)b, lanes, (const double *)dotw, partial, 0x7fffffff, stream);
}

void bn_csr_apply_binned(char vt, int op, int n, const int *ia, const int *ja, const void *val,
                         const void *x, void *y, const void *b, int lanes, int long_len, int nlong,
                         const int *long_rows, hipStream_t stream) {
  if (n <= 0) return;
  if (nlong <= 0) long_len = 0x7fffffff;
  if (vt == 's') {
    dispatch_csr<float>(op, n, ia, ja, (const float *)val, (const float *)x, (float *)y, (const float *)b,
                        lanes, nullptr, nullptr, long_len, stream);
    launch_long<float>(op, nlong, long_rows, ia, ja, (const float *)val, (const float *)x, (float *)y,
                       (const float *)b, stream);
  } else {
    dispatch_csr<double>(op, n, ia, ja, (const double *)val, (const double *)x, (double *)y,
                         (const double *)b, lanes, nullptr, nullptr, long_len, stream);
    launch_long<double>(op, nlong, long_rows, ia, ja, (const double *)val, (const double *)x, (double *)y,
                        (const double *)b, stream);
  }
}

void bn_csr_resid_blocks_scaled(char vt, int nblk, const int *blocks, const int *blkq, const int *ia, const int *ja,
                                const void *val, const void *x, void *y, const void *b, const void *dsc, void *ys,
                                hipStream_t stream) {
  if (nblk <= 0) return;
  const int g = grid_rows(nblk);
  if (vt == 's')
    BN_LAUNCH((k_csr_blocks<float, BN_OP_RESID, true>), dim3(g), dim3(BLOCK), 0, stream, nblk, blocks, blkq,
                       ia, ja, (const float *)val, (const float *)x, (float *)y, (const float *)b, (const float *)dsc,
                       (float *)ys, jacobi_epi<float>{nullptr, nullptr, 0.0f, nullptr, 0});
  else
    BN_LAUNCH((k_csr_blocks<double, BN_OP_RESID, true>), dim3(g), dim3(BLOCK), 0, stream, nblk, blocks, blkq,
                       ia, ja, (const double *)val, (const double *)x, (double *)y, (const double *)b,
                       (const double *)dsc, (double *)ys, jacobi_epi<double>{nullptr, nullptr, 0.0, nullptr, 0});
}

void bn_csr_apply_blocks_jac(char vt, int op, int nblk, const int *blocks, const int *blkq, const int *ia,
                             const int *ja, const void *val, const void *x, void *y, const void *b, void *xz,
                             const void *w, double om, const unsigned char *colour, int c, hipStream_t stream) {
  if (vt == 's')
    launch_blocks<float>(op, nblk, blocks, blkq, ia, ja, (const float *)val, (const float *)x, (float *)y,
                         (const float *)b, jacobi_epi<float>{(float *)xz, (const float *)w, (float)om, colour, c},
                         stream);
  else
    launch_blocks<double>(op, nblk, blocks, blkq, ia, ja, (const double *)val, (const double *)x, (double *)y,
                          (const double *)b, jacobi_epi<double>{(double *)xz, (const double *)w, om, colour, c},
                          stream);
}

void bn_csr_apply_blocks(char vt, int op, int nblk, const int *blocks, const int *blkq, const int *ia, const int *ja,
                         const void *val, const void *x, void *y, const void *b, hipStream_t stream) {
  bn_csr_apply_blocks_jac(vt, op, nblk, blocks, blkq, ia, ja, val, x, y, b, NULL, NULL, 0.0, NULL, 0, stream);
}

void bn_csr_apply_blocks_dot(int nblk, const int *blocks, const int *blkq, const int *ia, const int *ja,
                             const double *val, const double *x, double *y, const double *w, double *partial,
                             hipStream_t stream) {
  if (nblk <= 0) return;
  BN_LAUNCH(k_csr_blocks_dot<double>, dim3(grid_rows(nblk)), dim3(BLOCK), 0, stream, nblk, blocks, blkq, ia, ja, val,
            x, y, w, partial);
}

void bn_block_gs(char vt, int nchunk, int max_rows, const int *chunk, const int *ccol, const int *cbnd,
                 const int *ia, const int *ja, const void *val, const void *dinv, const void *b,
                 const void *xold, void *xnew, int backward, int lanes, const void *zbase, void *zsum,
                 hipStream_t stream) {
  if (max_rows > BN_CHUNK_MAX_ROWS) return; /* the solver never builds such chunks */
  if (vt == 's')
    dispatch_block_gs<float>(nchunk, max_rows, chunk, ccol, cbnd, ia, ja, (const float *)val, (const float *)dinv,
                             (const float *)b, (const float *)xold, (float *)xnew, backward, lanes,
                             (const float *)zbase, (float *)zsum, stream);
  else
    dispatch_block_gs<double>(nchunk, max_rows, chunk, ccol, cbnd, ia, ja, (const double *)val,
                              (const double *)dinv, (const double *)b, (const double *)xold, (double *)xnew,
                              backward, lanes, (const double *)zbase, (double *)zsum, stream);
}

void bn_csr_apply_seg_jac(char vt, int op, int n, const int *ia, const int *ja, const void *val, const void *x,
                          void *y, const void *b, int lanes, int long_len, int nlong, const int *long_rows,
                          const int *seg_ptr, const int *seg_row, int nseg, double *part, int *cnt, void *xz,
                          const void *w, double om, const unsigned char *colour, int c, hipStream_t stream) {
  if (n <= 0) return;
  if (nlong <= 0) long_len = 0x7fffffff;
  if (vt == 's') {
    const jacobi_epi<float> e{(float *)xz, (const float *)w, (float)om, colour, c};
    dispatch_csr<float>(op, n, ia, ja, (const float *)val, (const float *)x, (float *)y, (const float *)b,
                        lanes, nullptr, nullptr, long_len, stream, e);
    launch_seg<float>(op, nlong, long_rows, seg_ptr, seg_row, nseg, part, cnt, ia, ja, (const float *)val,
                      (const float *)x, (float *)y, (const float *)b, e, stream);
  } else {
    const jacobi_epi<double> e{(double *)xz, (const double *)w, om, colour, c};
    dispatch_csr<double>(op, n, ia, ja, (const double *)val, (const double *)x, (double *)y,
                         (const double *)b, lanes, nullptr, nullptr, long_len, stream, e);
    launch_seg<double>(op, nlong, long_rows, seg_ptr, seg_row, nseg, part, cnt, ia, ja, (const double *)val,
                       (const double *)x, (double *)y, (const double *)b, e, stream);
  }
}

void bn_csr_apply_seg(char vt, int op, int n, const int *ia, const int *ja, const void *val, const void *x,
                      void *y, const void *b, int lanes, int long_len, int nlong, const int *long_rows,
                      const int *seg_ptr, const int *seg_row, int nseg, double *part, int *cnt,
                      hipStream_t stream) {
  bn_csr_apply_seg_jac(vt, op, n, ia, ja, val, x, y, b, lanes, long_len, nlong, long_rows, seg_ptr, seg_row, nseg,
                       part, cnt, NULL, NULL, 0.0, NULL, 0, stream);
}

void bn_seg_reset(int nslot, double *part, int nmulti, int *cnt, hipStream_t stream) {
  if (nslot > 0)
    BN_LAUNCH(k_fill_word<unsigned long long>, g1(nslot), dim3(BLOCK), 0, stream, nslot,
              (unsigned long long *)part, walk_word<double>::none);
  if (nmulti > 0) BN_LAUNCH(k_fill<int>, g1(nmulti), dim3(BLOCK), 0, stream, nmulti, cnt, 0);
}

void bn_gs_color(char vt, int lo, int hi, const int *ia, const int *ja, const void *val,
                 const void *dinv, const void *b, void *x, int lanes, const void *zbase, void *zsum,
                 hipStream_t stream) {
  if (vt == 's')
    dispatch_rowsolve<float, RS_RANGE>(lo, hi - lo, nullptr, ia, ja, (const float *)val, (const float *)dinv,
                                    (const float *)b, (float *)x, lanes, (const float *)zbase,
                                    (float *)zsum, stream);
  else
    dispatch_rowsolve<double, RS_RANGE>(lo, hi - lo, nullptr, ia, ja, (const double *)val,
                                     (const double *)dinv, (const double *)b, (double *)x, lanes,
                                     (const double *)zbase, (double *)zsum, stream);
}

void bn_gs_color_blocks(char vt, int b0, int nb, const int *blocks, const int *blkq, const int *ia, const int *ja,
                        const void *val, const int *map, const void *dinv, const void *b, void *x, const void *zbase,
                        void *zsum, hipStream_t stream) {
  if (nb <= 0) return;
  const int g = grid_rows(nb);
  if (vt == 's')
    BN_LAUNCH(k_gs_blocks<float>, dim3(g), dim3(BLOCK), 0, stream, b0, nb, blocks, blkq, ja,
                       (const float *)val, ia, map, (const float *)dinv, (const float *)b, (float *)x,
                       (const float *)zbase, (float *)zsum);
  else
    BN_LAUNCH(k_gs_blocks<double>, dim3(g), dim3(BLOCK), 0, stream, b0, nb, blocks, blkq, ja,
                       (const double *)val, ia, map, (const double *)dinv, (const double *)b, (double *)x,
                       (const double *)zbase, (double *)zsum);
}

void bn_gs_color_map(char vt, int lo, int hi, const int *map, const int *ia, const int *ja, const void *val,
                     const void *dinv, const void *b, void *x, int lanes, const void *zbase, void *zsum,
                     hipStream_t stream) {
  if (vt == 's')
    dispatch_rowsolve<float, RS_MAP>(lo, hi - lo, map, ia, ja, (const float *)val, (const float *)dinv,
                                     (const float *)b, (float *)x, lanes, (const float *)zbase, (float *)zsum,
                                     stream);
  else
    dispatch_rowsolve<double, RS_MAP>(lo, hi - lo, map, ia, ja, (const double *)val, (const double *)dinv,
                                      (const double *)b, (double *)x, lanes, (const double *)zbase,
                                      (double *)zsum, stream);
}

void bn_gs_color_zero_map(char vt, int lo, int hi, int all_n, const int *map, const void *dinv, const void *b,
                          void *x, hipStream_t stream) {
  if (all_n < hi) all_n = hi;
  const int cnt = all_n - lo;
  if (cnt <= 0) return;
  if (vt == 's')
    BN_LAUNCH(k_diag_scale_map<float>, g1(cnt), dim3(BLOCK), 0, stream, lo, hi, all_n, map,
                       (const float *)dinv, (const float *)b, (float *)x);
  else
    BN_LAUNCH(k_diag_scale_map<double>, g1(cnt), dim3(BLOCK), 0, stream, lo, hi, all_n, map,
                       (const double *)dinv, (const double *)b, (double *)x);
}

void bn_gs_color_zero_colour(char vt, int n, int c, const unsigned char *colour, const void *dinv, const void *b,
                             void *x, hipStream_t stream) {
  if (n <= 0) return;
  if (vt == 's')
    BN_LAUNCH(k_diag_scale_colour<float>, g1(n), dim3(BLOCK), 0, stream, n, c, colour,
                       (const float *)dinv, (const float *)b, (float *)x);
  else
    BN_LAUNCH(k_diag_scale_colour<double>, g1(n), dim3(BLOCK), 0, stream, n, c, colour,
                       (const double *)dinv, (const double *)b, (double *)x);
}

void bn_gs_color_zero(char vt, int lo, int hi, const void *dinv, const void *b, void *x,
                      hipStream_t stream) {
  const int cnt = hi - lo;
  if (cnt <= 0) return;
  if (vt == 's')
    BN_LAUNCH(k_diag_scale<float>, g1(cnt), dim3(BLOCK), 0, stream, lo, cnt, (const float *)dinv,
                       (const float *)b, (float *)x);
  else
    BN_LAUNCH(k_diag_scale<double>, g1(cnt), dim3(BLOCK), 0, stream, lo, cnt,
                       (const double *)dinv, (const double *)b, (double *)x);
}

void bn_jacobi_rows(char vt, int n, const int *ia, const int *ja, const void *val, const void *dinv,
                    const void *b, const void *xin, void *xout, int lanes, const void *zbase, void *zsum,
                    hipStream_t stream) {
  if (n <= 0) return;
  if (!xin) {
    bn_gs_color_zero(vt, 0, n, dinv, b, xout, stream);
    return;
  }
  if (vt == 's')
    dispatch_jacobi_rows<float>(n, ia, ja, (const float *)val, (const float *)dinv, (const float *)b,
                                (const float *)xin, (float *)xout, lanes, (const float *)zbase, (float *)zsum,
                                stream);
  else
    dispatch_jacobi_rows<double>(n, ia, ja, (const double *)val, (const double *)dinv, (const double *)b,
                                 (const double *)xin, (double *)xout, lanes, (const double *)zbase,
                                 (double *)zsum, stream);
}

void bn_jacobi_rows_scaled(char vt, int n, const int *ia, const int *ja, const void *val, const void *dinv,
                           const void *b, const void *xin, void *xout, void *xs, int lanes, hipStream_t stream) {
  if (n <= 0) return;
  if (vt == 's')
    dispatch_jacobi_rows_scaled<float>(n, ia, ja, (const float *)val, (const float *)dinv, (const float *)b,
                                       (const float *)xin, (float *)xout, (float *)xs, lanes, stream);
  else
    dispatch_jacobi_rows_scaled<double>(n, ia, ja, (const double *)val, (const double *)dinv, (const double *)b,
                                        (const double *)xin, (double *)xout, (double *)xs, lanes, stream);
}

void bn_jacobi_update(char vt, int n, const void *r, const void *dscale, double omega, void *x,
                      hipStream_t stream) {
  if (n <= 0) return;
  if (vt == 's')
    BN_LAUNCH(k_jacobi_update<float>, g1(n), dim3(BLOCK), 0, stream, n, (const float *)r,
                       (const float *)dscale, (float)omega, (float *)x);
  else
    BN_LAUNCH(k_jacobi_update<double>, g1(n), dim3(BLOCK), 0, stream, n, (const double *)r,
                       (const double *)dscale, omega, (double *)x);
}

void bn_jacobi_zero(char vt, int n, const void *b, const void *dscale, double omega, void *x,
                    hipStream_t stream) {
  if (n <= 0) return;
  if (vt == 's')
    BN_LAUNCH(k_jacobi_zero<float>, g1(n), dim3(BLOCK), 0, stream, n, (const float *)b,
                       (const float *)dscale, (float)omega, (float *)x);
  else
    BN_LAUNCH(k_jacobi_zero<double>, g1(n), dim3(BLOCK), 0, stream, n, (const double *)b,
                       (const double *)dscale, omega, (double *)x);
}

void bn_trsv_level(char vt, int cnt, const int *rows, const int *ia, const int *ja, const void *val,
                   const void *dinv, const void *b, void *y, int lanes, const void *zbase, void *zsum,
                   hipStream_t stream) {
  if (vt == 's')
    dispatch_rowsolve<float, RS_LIST>(0, cnt, rows, ia, ja, (const float *)val, (const float *)dinv,
                                   (const float *)b, (float *)y, lanes, (const float *)zbase,
                                   (float *)zsum, stream);
  else
    dispatch_rowsolve<double, RS_LIST>(0, cnt, rows, ia, ja, (const double *)val, (const double *)dinv,
                                    (const double *)b, (double *)y, lanes, (const double *)zbase,
                                    (double *)zsum, stream);
}

void bn_trsv_ell(char vt, int w, int cnt, const int *rows, const int *ja, const void *val, const void *dl,
                 const void *b, void *y, const void *zbase, void *zsum, hipStream_t stream) {
  if (cnt <= 0 || w <= 0 || w > BN_ELL_MAX_W) return;
  if (vt == 's')
    trsv_ell_t<float>(w, cnt, rows, ja, (const float *)val, (const float *)dl, (const float *)b, (float *)y,
                      (const float *)zbase, (float *)zsum, stream);
  else
    trsv_ell_t<double>(w, cnt, rows, ja, (const double *)val, (const double *)dl, (const double *)b, (double *)y,
                       (const double *)zbase, (double *)zsum, stream);
}

void bn_trsv_parts(char vt, int np, const int *sp, const int *steps, const int *rows, const int *ja,
                   const void *val, const void *dl, const void *b, void *y, const void *zbase, void *zsum,
                   double *ctl, long long limit, hipStream_t stream) {
  if (np <= 0) return;
  const int4 *S = (const int4 *)steps;
  if (vt == 's')
    BN_LAUNCH(k_trsv_parts<float>, dim3(np), dim3(BN_PART_BLOCK), 0, stream, sp, S, rows, ja, (const float *)val,
              (const float *)dl, (const float *)b, (float *)y, (const float *)zbase, (float *)zsum, ctl, limit);
  else
    BN_LAUNCH(k_trsv_parts<double>, dim3(np), dim3(BN_PART_BLOCK), 0, stream, sp, S, rows, ja,
              (const double *)val, (const double *)dl, (const double *)b, (double *)y, (const double *)zbase,
              (double *)zsum, ctl, limit);
}

int bn_trsv_parts_per_cu(char vt) {
  int nb = 0;
  const hipError_t e = vt == 's'
                           ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_trsv_parts<float>, BN_PART_BLOCK, 0)
                           : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_trsv_parts<double>, BN_PART_BLOCK, 0);
  return e == hipSuccess ? nb : 0;
}

int bn_gs_walk_per_cu(char vt, int lanes) {
  int nb = 0, best = 1 << 30;
  for (int z = 0; z < 2; z++) {
    hipError_t e = hipErrorInvalidValue;
#define BN_OCC(T, L)                                                                                         \
  e = z ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_gs_walk<T, L, true>, BN_GSW_BLOCK, 0)        \
        : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_gs_walk<T, L, false>, BN_GSW_BLOCK, 0)
    if (vt == 's') {
      switch (lanes) {
        case 1: BN_OCC(float, 1); break;
        case 2: BN_OCC(float, 2); break;
        case 4: BN_OCC(float, 4); break;
        default: BN_OCC(float, 8); break;
      }
    } else {
      switch (lanes) {
        case 1: BN_OCC(double, 1); break;
        case 2: BN_OCC(double, 2); break;
        case 4: BN_OCC(double, 4); break;
        default: BN_OCC(double, 8); break;
      }
    }
#undef BN_OCC
    if (e != hipSuccess) return 0;
    if (nb < best) best = nb;
  }
  return best;
}

void bn_gs_walk(char vt, int lanes, int np, int ncol, int backward, const int *steps, const int *ia, const int *jx,
                const void *val, const int *map, const void *dinvc, const void *b, void *x, void *xold, int zero,
                const void *zbase, void *zsum, double *ctl, long long limit, hipStream_t stream) {
  if (np <= 0 || ncol <= 0) return;
  const int2 *S = (const int2 *)steps;
#define BN_GW(T, L, Z)                                                                                            \
  BN_LAUNCH((k_gs_walk<T, L, Z>), dim3(np), dim3(BN_GSW_BLOCK), 0, stream, ncol, backward, S, ia, jx, (const T *)val, \
            map, (const T *)dinvc, (const T *)b, (T *)x, (const T *)xold, (const T *)zbase, (T *)zsum, ctl, limit)
#define BN_GWL(T, Z)              \
  switch (lanes) {                \
    case 1: BN_GW(T, 1, Z); break; \
    case 2: BN_GW(T, 2, Z); break; \
    case 4: BN_GW(T, 4, Z); break; \
    default: BN_GW(T, 8, Z); break; \
  }
  if (vt == 's') {
    if (zero) { BN_GWL(float, true) } else { BN_GWL(float, false) }
  } else {
    if (zero) { BN_GWL(double, true) } else { BN_GWL(double, false) }
  }
#undef BN_GWL
#undef BN_GW
}

int bn_sweep_parts_per_cu(char vt) {
  int best = 1 << 30;
  for (int k = 0; k < 7; k++) {
    int nb = 0;
    hipError_t e = hipErrorInvalidValue;
#define BN_OCC(T, L) e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_sweep_parts<T, L>, BN_PART_BLOCK, 0)
#define BN_OCCL(T)                 \
  switch (k) {                     \
    case 0: BN_OCC(T, 1); break;   \
    case 1: BN_OCC(T, 2); break;   \
    case 2: BN_OCC(T, 4); break;   \
    case 3: BN_OCC(T, 8); break;   \
    case 4: BN_OCC(T, 16); break;  \
    case 5: BN_OCC(T, 32); break;  \
    default: BN_OCC(T, 64); break; \
  }
    if (vt == 's') { BN_OCCL(float) } else { BN_OCCL(double) }
#undef BN_OCCL
#undef BN_OCC
    if (e != hipSuccess) return 0;
    if (nb < best) best = nb;
  }
  return best;
}

void bn_sweep_parts(char vt, int lanes, int np, int backward, const int *sp, const int *steps, const int *rows,
                    const int *ia, const int *jx, const void *val, const void *dinv, const void *b, void *x,
                    const void *xold, double *ctl, long long limit, hipStream_t stream) {
  if (np <= 0) return;
  const int2 *S = (const int2 *)steps;
#define BN_SP(T, L)                                                                                                \
  BN_LAUNCH((k_sweep_parts<T, L>), dim3(np), dim3(BN_PART_BLOCK), 0, stream, backward, sp, S, rows, ia, jx,         \
            (const T *)val, (const T *)dinv, (const T *)b, (T *)x, (const T *)xold, ctl, limit)
#define BN_SPL(T)                    \
  switch (lanes) {                   \
    case 1: BN_SP(T, 1); break;      \
    case 2: BN_SP(T, 2); break;      \
    case 4: BN_SP(T, 4); break;      \
    case 8: BN_SP(T, 8); break;      \
    case 16: BN_SP(T, 16); break;    \
    case 32: BN_SP(T, 32); break;    \
    default: BN_SP(T, 64); break;    \
  }
  if (vt == 's') { BN_SPL(float) } else { BN_SPL(double) }
#undef BN_SPL
#undef BN_SP
}

void bn_gs_prep(char vt, int n, void *x, void *xold, hipStream_t stream) {
  if (n <= 0) return;
  if (vt == 's')
    BN_LAUNCH(k_gs_prep<float>, g1(n), dim3(BLOCK), 0, stream, n, (float *)x, (float *)xold);
  else
    BN_LAUNCH(k_gs_prep<double>, g1(n), dim3(BLOCK), 0, stream, n, (double *)x, (double *)xold);
}

void bn_walk_fill(char vt, int n, void *y, hipStream_t stream) {
  if (n <= 0) return;
  if (vt == 's')
    BN_LAUNCH(k_fill_word<unsigned int>, g1(n), dim3(BLOCK), 0, stream, n, (unsigned int *)y,
              walk_word<float>::none);
  else
    BN_LAUNCH(k_fill_word<unsigned long long>, g1(n), dim3(BLOCK), 0, stream, n, (unsigned long long *)y,
              walk_word<double>::none);
}

void bn_ell_gather(long long n, const int *src, const double *from, double *out, float *outs, hipStream_t stream) {
  if (n <= 0) return;
  BN_LAUNCH(k_ell_gather, dim3((unsigned)((n + BLOCK - 1) / BLOCK)), dim3(BLOCK), 0, stream, n, src, from, out, outs);
}

void bn_dense_gemv(char vt, int n, const void *M, const void *x, void *y, hipStream_t stream) {
  if (n <= 0) return;
  if (vt == 's')
    BN_LAUNCH(k_gemv<float>, dim3(n), dim3(BLOCK), 0, stream, n, (const float *)M,
                       (const float *)x, (float *)y);
  else
    BN_LAUNCH(k_gemv<double>, dim3(n), dim3(BLOCK), 0, stream, n, (const double *)M,
                       (const double *)x, (double *)y);
}

void bn_dense_gemv_ds(int n, const double *M, const float *x, float *y, hipStream_t stream) {
  if (n <= 0) return;
  BN_LAUNCH(k_gemv_ds, dim3(n), dim3(BLOCK), 0, stream, n, M, x, y);
}

void bn_fill(char vt, int n, void *x, double v, hipStream_t stream) {
  if (n <= 0) return;
  if (vt == 's')
    BN_LAUNCH(k_fill<float>, g1(n), dim3(BLOCK), 0, stream, n, (float *)x, (float)v);
  else
    BN_LAUNCH(k_fill<double>, g1(n), dim3(BLOCK), 0, stream, n, (double *)x, v);
}

void bn_copy(char vt, int n, const void *x, void *y, hipStream_t stream) {
  if (n <= 0) return;
  if (vt == 's')
    BN_LAUNCH(k_copy<float>, g1(n), dim3(BLOCK), 0, stream, n, (const float *)x, (float *)y);
  else
    BN_LAUNCH(k_copy<double>, g1(n), dim3(BLOCK), 0, stream, n, (const double *)x, (double *)y);
}

void bn_convert(char vt_in, char vt_out, int n, const void *x, void *y, hipStream_t stream) {
  if (n <= 0) return;
  if (vt_in == 'd' && vt_out == 's')
    BN_LAUNCH((k_convert<double, float>), g1(n), dim3(BLOCK), 0, stream, n, (const double *)x,
                       (float *)y);
  else if (vt_in == 's' && vt_out == 'd')
    BN_LAUNCH((k_convert<float, double>), g1(n), dim3(BLOCK), 0, stream, n, (const float *)x,
                       (double *)y);
  else
    bn_copy(vt_in, n, x, y, stream);
}

void bn_axpy1(char vt, int n, const void *x, void *y, hipStream_t stream) {
  if (n <= 0) return;
  if (vt == 's')
    BN_LAUNCH(k_axpy1<float>, g1(n), dim3(BLOCK), 0, stream, n, (const float *)x, (float *)y);
  else
    BN_LAUNCH(k_axpy1<double>, g1(n), dim3(BLOCK), 0, stream, n, (const double *)x, (double *)y);
}

void bn_dot_partial(int n, const double *x, const double *y, const double *u, const double *v,
                    double *partial, double *partial2, hipStream_t stream) {
  if (n <= 0) {
    bn_fill('d', BN_RED_BLOCKS, partial, 0.0, stream);
    if (u) bn_fill('d', BN_RED_BLOCKS, partial2, 0.0, stream);
    return;
  }
  const long long nb = ((long long)n + BLOCK - 1) / BLOCK;
  const int g = grid_for(nb);
  const int chunk = (int)((nb + g - 1) / g);
  BN_LAUNCH(k_dot2, dim3(g), dim3(BLOCK), 0, stream, n, chunk, x, y, u, v, partial, partial2);
}

void bn_pcg_update_xr(int n, double *x, const double *p, double *r, const double *q, const double *s,
                      double *partial, hipStream_t stream) {
  if (n <= 0) {
    bn_fill('d', BN_RED_BLOCKS, partial, 0.0, stream);
    return;
  }
  const long long nb = ((long long)n + BLOCK - 1) / BLOCK;
  const int g = grid_for(nb);
  const int chunk = (int)((nb + g - 1) / g);
  BN_LAUNCH(k_update_xr<false>, dim3(g), dim3(BLOCK), 0, stream, n, chunk, x, p, r, q, s, partial,
                     (float *)NULL, (const unsigned char *)nullptr, 0, (const float *)nullptr, (float *)nullptr);
}

void bn_pcg_update_xr_s(int n, double *x, const double *p, double *r, const double *q, const double *s,
                        double *partial, float *rs, hipStream_t stream) {
  if (n <= 0) {
    bn_fill('d', BN_RED_BLOCKS, partial, 0.0, stream);
    return;
  }
  const long long nb = ((long long)n + BLOCK - 1) / BLOCK;
  const int g = grid_for(nb);
  const int chunk = (int)((nb + g - 1) / g);
  BN_LAUNCH(k_update_xr<true>, dim3(g), dim3(BLOCK), 0, stream, n, chunk, x, p, r, q, s, partial, rs,
            (const unsigned char *)nullptr, 0, (const float *)nullptr, (float *)nullptr);
}

void bn_pcg_update_xr_s0(int n, double *x, const double *p, double *r, const double *q, const double *s,
                         double *partial, float *rs, const unsigned char *colour, int c, const float *dinv, float *x0,
                         hipStream_t stream) {
  if (n <= 0) {
    bn_fill('d', BN_RED_BLOCKS, partial, 0.0, stream);
    return;
  }
  const long long nb = ((long long)n + BLOCK - 1) / BLOCK;
  const int g = grid_for(nb);
  const int chunk = (int)((nb + g - 1) / g);
  BN_LAUNCH(k_update_xr<true>, dim3(g), dim3(BLOCK), 0, stream, n, chunk, x, p, r, q, s, partial, rs, colour, c, dinv,
            x0);
}

void bn_dot_rz_s(int n, const double *r, const float *zs, const double *q, double *partial, double *partial2,
                 hipStream_t stream) {
  if (n <= 0) {
    bn_fill('d', BN_RED_BLOCKS, partial, 0.0, stream);
    if (q) bn_fill('d', BN_RED_BLOCKS, partial2, 0.0, stream);
    return;
  }
  const long long nb = ((long long)n + BLOCK - 1) / BLOCK;
  const int g = grid_for(nb);
  const int chunk = (int)((nb + g - 1) / g);
  BN_LAUNCH(k_dot_rz_s, dim3(g), dim3(BLOCK), 0, stream, n, chunk, r, zs, q, partial, partial2);
}

void bn_pcg_update_p_s(int n, double *p, const float *zs, const double *s, hipStream_t stream) {
  if (n <= 0) return;
  BN_LAUNCH(k_update_p_s, g1(n), dim3(BLOCK), 0, stream, n, p, zs, s);
}

void bn_pcg_update_p(int n, double *p, const double *z, const double *s, hipStream_t stream) {
  if (n <= 0) return;
  BN_LAUNCH(k_update_p, g1(n), dim3(BLOCK), 0, stream, n, p, z, s);
}

void bn_scalar_finish(int which, const double *p1, const double *p2, const double *p3, double *s, int slot,
                      hipStream_t stream) {
  BN_LAUNCH(k_scalar_finish, dim3(1), dim3(BLOCK), 0, stream, which, p1, p2, p3, s, slot, (double *)nullptr, 0);
}

void bn_scalar_finish_pap(const double *part, int nslot, double *s, hipStream_t stream) {
  BN_LAUNCH(k_scalar_finish, dim3(1), dim3(BLOCK), 0, stream, 1, part, (const double *)nullptr,
            (const double *)nullptr, s, 0, (double *)nullptr, nslot);
}

void bn_scalar_finish_host(int which, const double *p1, const double *p2, const double *p3, double *s, double *hs,
                           hipStream_t stream) {
  BN_LAUNCH(k_scalar_finish, dim3(1), dim3(BLOCK), 0, stream, which, p1, p2, p3, s, 0, hs, 0);
}

void bn_permute(int n, const int *perm, const double *x, double *y, int scatter, hipStream_t stream) {
  if (n <= 0) return;
  BN_LAUNCH(k_permute, g1(n), dim3(BLOCK), 0, stream, n, perm, x, y, scatter);
}

void bn_edge_grad(int m, const int *eu, const int *ev, const double *g, const double *p, const double *lam,
                  const double *pdir, double *dg, double *grad_logr, hipStream_t stream) {
  if (m <= 0) return;
  BN_LAUNCH(k_edge_grad, g1(m), dim3(BLOCK), 0, stream, m, eu, ev, g, p, lam, pdir, dg, grad_logr);
}

}  // extern "C"
