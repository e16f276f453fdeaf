// devsort.hip - device sorting and scans for the set-up kernels (gfx950):
//   * exclusive prefix sums of int32 and int64 arrays;
//   * a stable least-significant-digit radix sort of (u64 key, int32 value)
//     pairs over the key's low `end_bit` bits.
//
// Scan: each block sums a tile of SCAN_TILE items; the tile sums are
// scanned (recursively while more than one tile of them); each block then
// scans its tile from its tile offset.
//
// Radix sort: 8-bit digits, one pass per digit.  A pass (1) counts each
// digit per tile, (2) scans the counts digit-major (digit 0 of every tile,
// then digit 1, ...), which gives every (digit, tile) its first output
// slot, and (3) re-reads each tile and writes every item to that slot plus
// its rank among the tile's earlier items of the same digit.  The rank is
// stable: a tile is read 256 items at a time in input order; inside a
// wavefront the lanes holding one digit find each other with eight
// ballots (one per digit bit), count the lanes below them, and the
// wavefronts of a group add the counts of the wavefronts before them and
// of the groups before.  Items therefore keep their input order inside a
// digit, and LSD passes over stable ranks sort stably.
#include <hip/hip_runtime.h>

#include "bn_kernels.h"

namespace {

constexpr int SB = 256;            // threads per block
constexpr int SCAN_IPT = 8;        // scan items per thread
constexpr int SCAN_TILE = SB * SCAN_IPT;
constexpr int RADIX = 256;
constexpr int SORT_GROUPS = 16;    // 256-item groups per sort tile
constexpr int SORT_TILE = SB * SORT_GROUPS;
typedef unsigned long long u64;

template <typename T>
__device__ T block_exclusive(T v, T *total) {
  __shared__ T part[SB / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  // inclusive scan in the wavefront
  T x = v;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const T y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) part[wid] = x;
  __syncthreads();
  T before = 0, all = 0;
#pragma unroll
  for (int w = 0; w < SB / 64; w++) {
    const T p = part[w];
    if (w < wid) before += p;
    all += p;
  }
  __syncthreads();
  if (total) *total = all;
  return before + x - v;
}

template <typename T>
__global__ __launch_bounds__(SB) void k_tile_sums(long long n, const T *__restrict__ in, T *__restrict__ sums) {
  const long long base = (long long)blockIdx.x * SCAN_TILE;
  T s = 0;
#pragma unroll
  for (int u = 0; u < SCAN_IPT; u++) {
    const long long i = base + (long long)u * SB + threadIdx.x;
    if (i < n) s += in[i];
  }
  T tot;
  block_exclusive<T>(s, &tot);
  if (threadIdx.x == 0) sums[blockIdx.x] = tot;
}

// exclusive scan of a tile (thread t owns items t*IPT .. t*IPT+IPT-1), plus
// the tile's offset
template <typename T>
__global__ __launch_bounds__(SB) void k_tile_scan(long long n, const T *__restrict__ in, T *__restrict__ out,
                                                  const T *__restrict__ offs) {
  const long long base = (long long)blockIdx.x * SCAN_TILE + (long long)threadIdx.x * SCAN_IPT;
  T v[SCAN_IPT];
  T s = 0;
#pragma unroll
  for (int u = 0; u < SCAN_IPT; u++) {
    const long long i = base + u;
    v[u] = i < n ? in[i] : T(0);
    s += v[u];
  }
  T run = block_exclusive<T>(s, nullptr) + (offs ? offs[blockIdx.x] : T(0));
#pragma unroll
  for (int u = 0; u < SCAN_IPT; u++) {
    const long long i = base + u;
    if (i < n) out[i] = run;
    run += v[u];
  }
}

template <typename T>
int scan_rec(const T *in, T *out, long long n, hipStream_t st) {
  if (n <= 0) return 0;
  const long long tiles = (n + SCAN_TILE - 1) / SCAN_TILE;
  if (tiles == 1) {
    hipLaunchKernelGGL(k_tile_scan<T>, dim3(1), dim3(SB), 0, st, n, in, out, (const T *)nullptr);
    return hipGetLastError() == hipSuccess ? 0 : -1;
  }
  T *sums = nullptr, *offs = nullptr;
  if (hipMalloc((void **)&sums, (size_t)tiles * sizeof(T)) != hipSuccess) return -1;
  if (hipMalloc((void **)&offs, (size_t)tiles * sizeof(T)) != hipSuccess) {
    (void)hipFree(sums);
    return -1;
  }
  hipLaunchKernelGGL(k_tile_sums<T>, dim3((unsigned)tiles), dim3(SB), 0, st, n, in, sums);
  int rc = scan_rec<T>(sums, offs, tiles, st);
  if (!rc) hipLaunchKernelGGL(k_tile_scan<T>, dim3((unsigned)tiles), dim3(SB), 0, st, n, in, out, (const T *)offs);
  if (hipStreamSynchronize(st) != hipSuccess) rc = -1;
  (void)hipFree(sums);
  (void)hipFree(offs);
  return rc;
}

// ---------------------------------------------------------------------
// radix sort
// ---------------------------------------------------------------------
// the digit of key bits shift .. shift+7, bits at or past end_bit read as 0
__device__ __forceinline__ int digit_of(u64 k, int shift, int dmask) { return (int)((k >> shift) & dmask); }

// counts[d * ntiles + tile] = items of digit d in the tile
__global__ __launch_bounds__(SB) void k_digit_counts(int n, int shift, int dmask, int ntiles,
                                                     const u64 *__restrict__ key, int *__restrict__ counts) {
  __shared__ int h[RADIX];
  h[threadIdx.x] = 0;
  __syncthreads();
  const long long base = (long long)blockIdx.x * SORT_TILE;
  for (int g = 0; g < SORT_GROUPS; g++) {
    const long long i = base + (long long)g * SB + threadIdx.x;
    if (i < n) atomicAdd(&h[digit_of(key[i], shift, dmask)], 1);
  }
  __syncthreads();
  counts[(long long)threadIdx.x * ntiles + blockIdx.x] = h[threadIdx.x];
}

// lanes of this wavefront whose digit equals mine (eight ballots)
__device__ __forceinline__ u64 same_digit_lanes(int d, bool valid) {
  u64 m = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; b++) {
    const u64 set = __ballot(valid && ((d >> b) & 1));
    m &= ((d >> b) & 1) ? set : ~set;
  }
  return m;
}

__global__ __launch_bounds__(SB) void k_digit_scatter(int n, int shift, int dmask, int ntiles, const u64 *__restrict__ kin,
                                                      const int *__restrict__ vin, u64 *__restrict__ kout,
                                                      int *__restrict__ vout, const int *__restrict__ offs) {
  __shared__ int base_d[RADIX];          // this tile's next slot per digit
  __shared__ int wcnt[SB / 64][RADIX];   // per-wavefront digit counts of the group
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  base_d[threadIdx.x] = offs[(long long)threadIdx.x * ntiles + blockIdx.x];
#pragma unroll
  for (int w = 0; w < SB / 64; w++) wcnt[w][threadIdx.x] = 0;
  __syncthreads();
  const long long tile0 = (long long)blockIdx.x * SORT_TILE;
  const u64 below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (int g = 0; g < SORT_GROUPS; g++) {
    const long long i = tile0 + (long long)g * SB + threadIdx.x;
    const bool valid = i < n;
    const u64 k = valid ? kin[i] : 0ull;
    const int v = valid ? vin[i] : 0;
    const int d = digit_of(k, shift, dmask);
    const u64 peers = same_digit_lanes(d, valid);
    const int rank_w = __popcll(peers & below);
    // the lowest lane of each digit's peers records the wavefront's count
    if (valid && rank_w == 0) wcnt[wid][d] = __popcll(peers);
    __syncthreads();
    if (valid) {
      int r = base_d[d] + rank_w;
      for (int w = 0; w < wid; w++) r += wcnt[w][d];
      kout[r] = k;
      vout[r] = v;
    }
    __syncthreads();
    int add = 0;
#pragma unroll
    for (int w = 0; w < SB / 64; w++) {
      add += wcnt[w][threadIdx.x];
      wcnt[w][threadIdx.x] = 0;
    }
    base_d[threadIdx.x] += add;
    __syncthreads();
  }
}

}  // namespace

extern "C" {

int bn_dev_scan_i32(const int *in, int *out, long long n, hipStream_t st) { return scan_rec<int>(in, out, n, st); }

int bn_dev_scan_i64(const long long *in, long long *out, long long n, hipStream_t st) {
  return scan_rec<long long>(in, out, n, st);
}

int bn_dev_sort_pairs(const unsigned long long *kin, unsigned long long *kout, const int *vin, int *vout, int n,
                      int end_bit, hipStream_t st) {
  if (n <= 0) return 0;
  if (end_bit < 1) end_bit = 1;
  if (end_bit > 64) end_bit = 64;
  const int passes = (end_bit + 7) / 8;
  const int ntiles = (int)(((long long)n + SORT_TILE - 1) / SORT_TILE);
  const long long nc = (long long)RADIX * ntiles;
  u64 *kb = nullptr;
  int *vb = nullptr, *cnt = nullptr, *off = nullptr;
  int rc = 0;
  if (hipMalloc((void **)&kb, (size_t)n * sizeof(u64)) != hipSuccess ||
      hipMalloc((void **)&vb, (size_t)n * sizeof(int)) != hipSuccess ||
      hipMalloc((void **)&cnt, (size_t)nc * sizeof(int)) != hipSuccess ||
      hipMalloc((void **)&off, (size_t)nc * sizeof(int)) != hipSuccess)
    rc = -1;
  // ping-pong so that the last pass writes kout / vout
  const u64 *ksrc = kin;
  const int *vsrc = vin;
  for (int p = 0; p < passes && !rc; p++) {
    const bool to_out = ((passes - 1 - p) % 2) == 0;
    u64 *kdst = to_out ? kout : kb;
    int *vdst = to_out ? vout : vb;
    const int bits = end_bit - 8 * p < 8 ? end_bit - 8 * p : 8;
    const int dmask = (1 << bits) - 1;
    hipLaunchKernelGGL(k_digit_counts, dim3(ntiles), dim3(SB), 0, st, n, 8 * p, dmask, ntiles, ksrc, cnt);
    rc = scan_rec<int>(cnt, off, nc, st);
    if (!rc)
      hipLaunchKernelGGL(k_digit_scatter, dim3(ntiles), dim3(SB), 0, st, n, 8 * p, dmask, ntiles, ksrc, vsrc, kdst,
                         vdst, off);
    ksrc = kdst;
    vsrc = vdst;
  }
  if (!rc && hipStreamSynchronize(st) != hipSuccess) rc = -1;
  (void)hipFree(kb);
  (void)hipFree(vb);
  (void)hipFree(cnt);
  (void)hipFree(off);
  return rc;
}

}  // extern "C"
