// galerkin.hip - the set-up's Galerkin product on the device, for gfx950:
//   A_c = (RAP + RAP^T) / 2,  R = P^T,
// with the host product's layout and values, entry for entry (setup.c's
// dcsr_rap, a row-by-row SMMP product, then bn_csr_symmetrize):
//  - row ic of RAP holds its diagonal first, then its other columns in the
//    order the traversal "R row ic, then A row, then P row" first meets
//    them;
//  - an entry's value is the sum of its terms (r a) p in traversal order,
//    starting from its first term (the diagonal from +0.0);
//  - A_c keeps RAP's layout, each value (v_ij + v_ji) * 0.5.
// Every term (R entry q1, A entry q2, P entry q3) becomes an item keyed by
// (row, column); a stable radix sort (devsort.hip) groups the items of an entry in
// traversal order, and a per-row virtual item ahead of all terms stands
// for the diagonal the host reserves.  An entry's first item then gives
// its place in the row.  Compiled without contraction.
#include <hip/hip_runtime.h>

#include <stdlib.h>
#include <string.h>

#include "amg.h"
#include "bn_kernels.h"

#pragma clang fp contract(off)

extern "C" {
int bn_device_ok(void);
void *bn_xmalloc(size_t bytes);
}

namespace {

constexpr int B = 256;
typedef unsigned long long u64;

inline unsigned nb(long long n) { return (unsigned)((n + B - 1) / B); }

__global__ void k_g_ap(int nf, const int *__restrict__ aia, const int *__restrict__ aja, const int *__restrict__ pia,
                       long long *__restrict__ ap) {
  const int i = blockIdx.x * B + threadIdx.x;
  if (i >= nf) return;
  long long t = 0;
  for (int q = aia[i]; q < aia[i + 1]; q++) t += pia[aja[q] + 1] - pia[aja[q]];
  ap[i] = t;
}

__global__ void k_g_w(int nnzr, const int *__restrict__ rja, const long long *__restrict__ ap,
                      long long *__restrict__ w) {
  const int q = blockIdx.x * B + threadIdx.x;
  if (q < nnzr) w[q] = ap[rja[q]];
}

// row r of a CSR row pointer with ia[r] <= q < ia[r + 1]
__device__ int row_of(const int *__restrict__ ia, int n, int q) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (ia[mid] <= q) lo = mid;
    else hi = mid;
  }
  return lo;
}

// the terms of R entry q1, in traversal order, at items nc + off[q1] ...
__global__ void k_g_gen(int nnzr, int nc, int cb, const int *__restrict__ ria, const int *__restrict__ rja,
                        const int *__restrict__ aia, const int *__restrict__ aja, const int *__restrict__ pia,
                        const int *__restrict__ pja, const long long *__restrict__ off, u64 *__restrict__ key,
                        int *__restrict__ idx, int *__restrict__ Q1, int *__restrict__ Q2, int *__restrict__ Q3) {
  const int q1 = blockIdx.x * B + threadIdx.x;
  if (q1 >= nnzr) return;
  const int ic = row_of(ria, nc, q1), i1 = rja[q1];
  const u64 rk = (u64)ic << cb;
  long long t = off[q1];
  for (int q2 = aia[i1]; q2 < aia[i1 + 1]; q2++) {
    const int i2 = aja[q2];
    for (int q3 = pia[i2]; q3 < pia[i2 + 1]; q3++, t++) {
      key[nc + t] = rk | (u64)pja[q3];
      idx[nc + t] = (int)(nc + t);
      Q1[t] = q1;
      Q2[t] = q2;
      Q3[t] = q3;
    }
  }
}

// items 0 .. nc-1: the diagonal of each row, ahead of every term
__global__ void k_g_virtual(int nc, int cb, u64 *__restrict__ key, int *__restrict__ idx) {
  const int ic = blockIdx.x * B + threadIdx.x;
  if (ic >= nc) return;
  key[ic] = ((u64)ic << cb) | (u64)ic;
  idx[ic] = ic;
}

__global__ void k_g_heads(int N, const u64 *__restrict__ key, int *__restrict__ flag) {
  const int k = blockIdx.x * B + threadIdx.x;
  if (k < N) flag[k] = (k == 0 || key[k] != key[k - 1]) ? 1 : 0;
}

// entry e (the e-th group of equal keys): its first item, column and its
// place key (row, first item), which orders a row as the host lays it out
__global__ void k_g_entries(int N, int cb, const u64 *__restrict__ key, const int *__restrict__ idx,
                            const int *__restrict__ flag, const int *__restrict__ eid, int *__restrict__ start,
                            u64 *__restrict__ ekey, int *__restrict__ ecol, int *__restrict__ epay) {
  const int k = blockIdx.x * B + threadIdx.x;
  if (k >= N || !flag[k]) return;
  const int e = eid[k];
  const u64 row = key[k] >> cb;
  start[e] = k;
  ecol[e] = (int)(key[k] & ((1ull << cb) - 1));
  ekey[e] = (row << 31) | (u64)idx[k];
  epay[e] = e;
}

__global__ void k_g_layout(int E, int nc, const u64 *__restrict__ ekey, const int *__restrict__ order,
                           const int *__restrict__ ecol, int *__restrict__ rowptr, int *__restrict__ ja,
                           int *__restrict__ pos, int *__restrict__ prow) {
  const int p = blockIdx.x * B + threadIdx.x;
  if (p >= E) return;
  const int e = order[p];
  const int row = (int)(ekey[p] >> 31);
  ja[p] = ecol[e];
  pos[e] = p;
  prow[p] = row;
  if (p == 0 || row != (int)(ekey[p - 1] >> 31)) rowptr[row] = p;
  if (p == E - 1) rowptr[nc] = E;
}

__device__ __forceinline__ double g_term(int it, int nc, const int *__restrict__ Q1, const int *__restrict__ Q2,
                                         const int *__restrict__ Q3, const double *__restrict__ rv,
                                         const double *__restrict__ av, const double *__restrict__ pv) {
  const int t = it - nc;
  const double ra = rv[Q1[t]] * av[Q2[t]];
  return ra * pv[Q3[t]];
}

// value of every entry with at most BN_RB_LONG items; longer ones are
// listed for k_g_vals_wave
__global__ void k_g_vals(int E, int nc, const int *__restrict__ start, const int *__restrict__ idx,
                         const int *__restrict__ pos, const int *__restrict__ Q1, const int *__restrict__ Q2,
                         const int *__restrict__ Q3, const double *__restrict__ rv, const double *__restrict__ av,
                         const double *__restrict__ pv, double *__restrict__ val, int *__restrict__ nlong,
                         int *__restrict__ elist) {
  const int e = blockIdx.x * B + threadIdx.x;
  if (e >= E) return;
  int k = start[e];
  const int hi = start[e + 1];
  if (hi - k > BN_RB_LONG) {
    elist[atomicAdd(nlong, 1)] = e;
    return;
  }
  double acc = -0.0;  // -0.0 + first term = first term
  if (idx[k] < nc) {
    acc = 0.0;  // the diagonal starts from +0.0
    k++;
  }
  for (; k < hi; k++) acc += g_term(idx[k], nc, Q1, Q2, Q3, rv, av, pv);
  val[pos[e]] = acc;
}

__device__ __forceinline__ double lane_value(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)b, l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (unsigned)lo);
}

// one wavefront per long entry: 64 terms at a time, summed in lane order
__global__ void __launch_bounds__(B) k_g_vals_wave(int nlong, const int *__restrict__ elist, int nc,
                                                  const int *__restrict__ start, const int *__restrict__ idx,
                                                  const int *__restrict__ pos, const int *__restrict__ Q1,
                                                  const int *__restrict__ Q2, const int *__restrict__ Q3,
                                                  const double *__restrict__ rv, const double *__restrict__ av,
                                                  const double *__restrict__ pv, double *__restrict__ val) {
  const int w = __builtin_amdgcn_readfirstlane((int)(blockIdx.x * (B / 64) + threadIdx.x / 64));
  const int lane = threadIdx.x & 63;
  if (w >= nlong) return;
  const int e = __builtin_amdgcn_readfirstlane(elist[w]);
  const int lo = __builtin_amdgcn_readfirstlane(start[e]), hi = __builtin_amdgcn_readfirstlane(start[e + 1]);
  double acc = idx[lo] < nc ? 0.0 : -0.0;
  for (int base = lo; base < hi; base += 64) {
    const int k = base + lane;
    double c = -0.0;  // the diagonal's virtual item adds nothing
    if (k < hi && idx[k] >= nc) c = g_term(idx[k], nc, Q1, Q2, Q3, rv, av, pv);
    const int cnt = hi - base < 64 ? hi - base : 64;
    for (int l = 0; l < cnt; l++) acc += lane_value(c, l);
  }
  if (lane == 0) val[pos[e]] = acc;
}

__global__ void k_g_tkeys(int E, int cb, const int *__restrict__ prow, const int *__restrict__ ja,
                          u64 *__restrict__ k1, u64 *__restrict__ k2, int *__restrict__ pay) {
  const int p = blockIdx.x * B + threadIdx.x;
  if (p >= E) return;
  const u64 r = (u64)prow[p], c = (u64)ja[p];
  k1[p] = (r << cb) | c;
  k2[p] = (c << cb) | r;
  pay[p] = p;
}

// with the entries ranked by (row, col) and by (col, row), rank k holds an
// entry and its transpose; a mismatch means the pattern is not symmetric
__global__ void k_g_tpos(int E, const u64 *__restrict__ k1, const u64 *__restrict__ k2, const int *__restrict__ p1,
                         const int *__restrict__ p2, int *__restrict__ tpos, int *__restrict__ bad) {
  const int k = blockIdx.x * B + threadIdx.x;
  if (k >= E) return;
  if (k1[k] != k2[k]) *bad = 1;
  tpos[p2[k]] = p1[k];
}

__global__ void k_g_sym(int E, const int *__restrict__ tpos, const double *__restrict__ v, double *__restrict__ out) {
  const int p = blockIdx.x * B + threadIdx.x;
  if (p < E) out[p] = (v[p] + v[tpos[p]]) * 0.5;
}

int bits_for(long long n) {
  int b = 1;
  while ((1ll << b) < n) b++;
  return b;
}

// device allocations of one product, released together
struct Pool {
  void *p[64];
  int n = 0;
  template <typename T>
  T *get(size_t count) {
    void *q = nullptr;
    if (n >= 64 || hipMalloc(&q, (count ? count : 1) * sizeof(T)) != hipSuccess) return nullptr;
    p[n++] = q;
    return (T *)q;
  }
  ~Pool() {
    for (int i = 0; i < n; i++) (void)hipFree(p[i]);
  }
};

template <typename T>
T *upload(Pool &pool, const T *h, size_t count) {
  T *d = pool.get<T>(count);
  if (d && count && hipMemcpy(d, h, count * sizeof(T), hipMemcpyHostToDevice) != hipSuccess) return nullptr;
  return d;
}

#define GCHECK(x)                 \
  do {                            \
    if ((x) != hipSuccess) return -1; \
  } while (0)
#define GNULL(p)         \
  do {                   \
    if (!(p)) return -1; \
  } while (0)

int sort_pairs(Pool &pool, const u64 *kin, u64 *kout, const int *vin, int *vout, int n, int end_bit) {
  (void)pool;
  return bn_dev_sort_pairs(kin, kout, vin, vout, n, end_bit, (hipStream_t)0);
}

int galerkin(const dcsr_mat *A, const dcsr_mat *P, const dcsr_mat *R, dcsr_mat *Ac) {
  Pool pool;
  const int nf = A->row, nc = R->row, nnzr = R->nnz;
  if (nc <= 0 || nnzr <= 0 || P->row != nf || R->col != nf) return -1;
  const int *aia = upload(pool, A->ia, (size_t)nf + 1), *aja = upload(pool, A->ja, (size_t)A->nnz);
  const double *av = upload(pool, A->val, (size_t)A->nnz);
  const int *pia = upload(pool, P->ia, (size_t)nf + 1), *pja = upload(pool, P->ja, (size_t)P->nnz);
  const double *pv = upload(pool, P->val, (size_t)P->nnz);
  const int *ria = upload(pool, R->ia, (size_t)nc + 1), *rja = upload(pool, R->ja, (size_t)nnzr);
  const double *rv = upload(pool, R->val, (size_t)nnzr);
  long long *ap = pool.get<long long>((size_t)nf), *w = pool.get<long long>((size_t)nnzr),
            *off = pool.get<long long>((size_t)nnzr);
  GNULL(aia && aja && av && pia && pja && pv && ria && rja && rv && ap && w && off);
  hipLaunchKernelGGL(k_g_ap, dim3(nb(nf)), dim3(B), 0, 0, nf, aia, aja, pia, ap);
  hipLaunchKernelGGL(k_g_w, dim3(nb(nnzr)), dim3(B), 0, 0, nnzr, rja, ap, w);
  if (bn_dev_scan_i64(w, off, nnzr, (hipStream_t)0)) return -1;
  long long last[2];
  GCHECK(hipMemcpy(&last[0], off + nnzr - 1, sizeof(long long), hipMemcpyDeviceToHost));
  GCHECK(hipMemcpy(&last[1], w + nnzr - 1, sizeof(long long), hipMemcpyDeviceToHost));
  const long long T = last[0] + last[1];
  if (T + nc >= (1ll << 31) - 1) return -1;  // item indices are int32
  const int N = (int)(T + nc), cb = bits_for((long long)nc + 1);
  if (2 * cb > 62 || cb + 31 > 64) return -1;

  u64 *key = pool.get<u64>((size_t)N), *key_s = pool.get<u64>((size_t)N);
  int *idx = pool.get<int>((size_t)N), *idx_s = pool.get<int>((size_t)N);
  int *Q1 = pool.get<int>((size_t)T), *Q2 = pool.get<int>((size_t)T), *Q3 = pool.get<int>((size_t)T);
  GNULL(key && key_s && idx && idx_s && Q1 && Q2 && Q3);
  hipLaunchKernelGGL(k_g_virtual, dim3(nb(nc)), dim3(B), 0, 0, nc, cb, key, idx);
  hipLaunchKernelGGL(k_g_gen, dim3(nb(nnzr)), dim3(B), 0, 0, nnzr, nc, cb, ria, rja, aia, aja, pia, pja, off, key, idx,
                     Q1, Q2, Q3);
  if (sort_pairs(pool, key, key_s, idx, idx_s, N, 2 * cb)) return -1;

  // entries: groups of equal (row, col)
  int *flag = pool.get<int>((size_t)N), *eid = pool.get<int>((size_t)N);
  GNULL(flag && eid);
  hipLaunchKernelGGL(k_g_heads, dim3(nb(N)), dim3(B), 0, 0, N, key_s, flag);
  if (bn_dev_scan_i32(flag, eid, N, (hipStream_t)0)) return -1;
  int lastf[2];
  GCHECK(hipMemcpy(&lastf[0], eid + N - 1, sizeof(int), hipMemcpyDeviceToHost));
  GCHECK(hipMemcpy(&lastf[1], flag + N - 1, sizeof(int), hipMemcpyDeviceToHost));
  const int E = lastf[0] + lastf[1];
  int *start = pool.get<int>((size_t)E + 1), *ecol = pool.get<int>((size_t)E), *epay = pool.get<int>((size_t)E),
      *order = pool.get<int>((size_t)E);
  u64 *ekey = pool.get<u64>((size_t)E), *ekey_s = pool.get<u64>((size_t)E);
  GNULL(start && ecol && epay && order && ekey && ekey_s);
  hipLaunchKernelGGL(k_g_entries, dim3(nb(N)), dim3(B), 0, 0, N, cb, key_s, idx_s, flag, eid, start, ekey, ecol, epay);
  GCHECK(hipMemcpy(start + E, &N, sizeof(int), hipMemcpyHostToDevice));
  if (sort_pairs(pool, ekey, ekey_s, epay, order, E, 31 + cb)) return -1;

  // layout, values
  int *rowptr = pool.get<int>((size_t)nc + 1), *cja = pool.get<int>((size_t)E), *pos = pool.get<int>((size_t)E),
      *prow = pool.get<int>((size_t)E), *nlong = pool.get<int>(1), *elist = pool.get<int>((size_t)E);
  double *rapv = pool.get<double>((size_t)E), *cval = pool.get<double>((size_t)E);
  GNULL(rowptr && cja && pos && prow && nlong && elist && rapv && cval);
  hipLaunchKernelGGL(k_g_layout, dim3(nb(E)), dim3(B), 0, 0, E, nc, ekey_s, order, ecol, rowptr, cja, pos, prow);
  const int zero = 0;
  GCHECK(hipMemcpy(nlong, &zero, sizeof(int), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_g_vals, dim3(nb(E)), dim3(B), 0, 0, E, nc, start, idx_s, pos, Q1, Q2, Q3, rv, av, pv, rapv,
                     nlong, elist);
  int nl = 0;
  GCHECK(hipMemcpy(&nl, nlong, sizeof(int), hipMemcpyDeviceToHost));
  if (nl > 0)
    hipLaunchKernelGGL(k_g_vals_wave, dim3((unsigned)((nl + B / 64 - 1) / (B / 64))), dim3(B), 0, 0, nl, elist, nc,
                       start, idx_s, pos, Q1, Q2, Q3, rv, av, pv, rapv);

  // (RAP + RAP^T) / 2 on RAP's layout
  u64 *k1 = pool.get<u64>((size_t)E), *k2 = pool.get<u64>((size_t)E), *k1s = pool.get<u64>((size_t)E),
      *k2s = pool.get<u64>((size_t)E);
  int *pay = pool.get<int>((size_t)E), *p1 = pool.get<int>((size_t)E), *p2 = pool.get<int>((size_t)E),
      *tpos = pool.get<int>((size_t)E), *bad = pool.get<int>(1);
  GNULL(k1 && k2 && k1s && k2s && pay && p1 && p2 && tpos && bad);
  hipLaunchKernelGGL(k_g_tkeys, dim3(nb(E)), dim3(B), 0, 0, E, cb, prow, cja, k1, k2, pay);
  if (sort_pairs(pool, k1, k1s, pay, p1, E, 2 * cb) || sort_pairs(pool, k2, k2s, pay, p2, E, 2 * cb)) return -1;
  GCHECK(hipMemcpy(bad, &zero, sizeof(int), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_g_tpos, dim3(nb(E)), dim3(B), 0, 0, E, k1s, k2s, p1, p2, tpos, bad);
  hipLaunchKernelGGL(k_g_sym, dim3(nb(E)), dim3(B), 0, 0, E, tpos, rapv, cval);
  int isbad = 0;
  GCHECK(hipMemcpy(&isbad, bad, sizeof(int), hipMemcpyDeviceToHost));
  if (isbad) return -1;  // not structurally symmetric: the host forms it

  Ac->row = nc;
  Ac->col = nc;
  Ac->nnz = E;
  Ac->ia = (int *)bn_xmalloc(((size_t)nc + 1) * sizeof(int));
  Ac->ja = (int *)bn_xmalloc((size_t)(E > 0 ? E : 1) * sizeof(int));
  Ac->val = (double *)bn_xmalloc((size_t)(E > 0 ? E : 1) * sizeof(double));
  hipError_t e = hipMemcpy(Ac->ia, rowptr, ((size_t)nc + 1) * sizeof(int), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(Ac->ja, cja, (size_t)E * sizeof(int), hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(Ac->val, cval, (size_t)E * sizeof(double), hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    free(Ac->ia);
    free(Ac->ja);
    free(Ac->val);
    memset(Ac, 0, sizeof(*Ac));
    return -1;
  }
  return 0;
}

}  // namespace

extern "C" {

/* A_c = (R A P + (R A P)^T) / 2 of the set-up (setup.c's galerkin_level)
 * on the device, identical to the host product; R = P^T as the host
 * transposes it.  Returns 0 with Ac allocated on the host, or -1 (no
 * device, BN_HOST_GALERKIN set, fewer than BN_DEVICE_GALERKIN_MIN operator
 * entries, sizes beyond int32 items, a pattern that is not symmetric, a
 * device failure): the caller then forms it on the host. */
int bn_galerkin_device(const dcsr_mat *A, const dcsr_mat *P, const dcsr_mat *R, dcsr_mat *Ac) {
  const char *env = getenv("BN_HOST_GALERKIN");
  if ((env && *env && *env != '0') || !bn_device_ok()) return -1;
  /* below ~4M operator entries the device's fixed cost (allocation, sort
   * set-up, copies: ~25 ms) exceeds the host product's time
   * (BN_DEVICE_GALERKIN_MIN overrides) */
  long long min_nnz = 4000000;
  const char *mn = getenv("BN_DEVICE_GALERKIN_MIN");
  if (mn && *mn) min_nnz = atoll(mn);
  if (A->nnz < min_nnz) return -1;
  const int rc = galerkin(A, P, R, Ac);
  if (rc) (void)hipGetLastError();
  return rc;
}

}  // extern "C"
