// setup_dev.hip - the per-row stages of one level of the classical AMG
// set-up on the device (gfx950), with the host routines' results entry for
// entry (setup.c cites the reference lines each one follows):
//   * strength (C_amg1/src/rs_classical.c:30-88): per row the threshold
//     theta * max(0, max_{j != i} -a_ij), then the strong entries in row
//     order;
//   * the PMIS independent-set rounds (rs_classical.c:305-349) on the
//     symmetric strength graph S u S^T: a round's select phase reads only
//     the previous round's state and its F phase only reacts to C marks,
//     so both are order-free;
//   * standard interpolation (rs_classical.c:643-805): the pattern of F row
//     i is its strong C neighbours in S order, and each weight is formed by
//     one thread in the host's order of operations.
// The greedy RS / HMIS passes and the promotion of F points without a
// strong C neighbour are sequential and stay on the host, as do the
// tie-break draws of PMIS (the C library's rand() sequence).
// Compiled without contraction, so every FP64 operation is the host's.
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
constexpr int CPT = 1, FPT = -1;
// Interpolation runs on the device only where its per-row threads stay
// short: each weight's lookups scan the strong row, and a row visits the
// operator rows of its strong F neighbours.  Galerkin levels with hub rows
// of thousands of entries would leave single threads walking them for
// seconds (the host's threads take milliseconds there).
constexpr int INTERP_MAX_S = 64;       // longest strong row
constexpr int INTERP_MAX_ROW = 1024;   // longest operator row
constexpr double INTERP_MAX_MEAN = 40; // mean operator row

inline unsigned nb(long long n) { return (unsigned)((n + B - 1) / B); }

__device__ __forceinline__ bool strong(double a, double t, int clamp) {
  const double v = -a;
  return clamp ? ((v < 0.0 ? 0.0 : v) >= t) : (v >= t);
}

// thr[i] = theta * max(0, max_{j != i} -a_ij); cnt[i] = strong entries
__global__ void k_s_count(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                          const double *__restrict__ val, double theta, int clamp, double *__restrict__ thr,
                          int *__restrict__ cnt) {
  const int i = blockIdx.x * B + threadIdx.x;
  if (i >= n) return;
  const int q0 = ia[i], q1 = ia[i + 1];
  double mx = 0.0;
  for (int k = q0; k < q1; k++) {
    if (ja[k] == i) continue;
    const double v = -val[k];
    if (v > mx) mx = v;
  }
  const double t = theta * mx;
  int c = 0;
  for (int k = q0; k < q1; k++)
    if (ja[k] != i && strong(val[k], t, clamp)) c++;
  thr[i] = t;
  cnt[i] = c;
}

__global__ void k_s_fill(int n, const int *__restrict__ ia, const int *__restrict__ ja,
                         const double *__restrict__ val, const double *__restrict__ thr, int clamp,
                         const int *__restrict__ sia, int *__restrict__ sja, double *__restrict__ sval) {
  const int i = blockIdx.x * B + threadIdx.x;
  if (i >= n) return;
  const double t = thr[i];
  int p = sia[i];
  for (int k = ia[i]; k < ia[i + 1]; k++)
    if (ja[k] != i && strong(val[k], t, clamp)) {
      sja[p] = ja[k];
      sval[p] = val[k];
      p++;
    }
}

// column counts of S (integer atomics: the counts do not depend on order)
__global__ void k_col_count(int n, const int *__restrict__ sia, const int *__restrict__ sja, int *__restrict__ tcnt) {
  const int i = blockIdx.x * B + threadIdx.x;
  if (i >= n) return;
  for (int k = sia[i]; k < sia[i + 1]; k++) atomicAdd(&tcnt[sja[k]], 1);
}

// S^T's rows as sets (the rounds only test membership, so the order the
// atomics give each row does not matter)
__global__ void k_col_fill(int n, const int *__restrict__ sia, const int *__restrict__ sja, int *__restrict__ tnext,
                           int *__restrict__ tja) {
  const int i = blockIdx.x * B + threadIdx.x;
  if (i >= n) return;
  for (int k = sia[i]; k < sia[i + 1]; k++) tja[atomicAdd(&tnext[sja[k]], 1)] = i;
}

__device__ __forceinline__ int block_count(int v) {
  __shared__ int part[B / 64];
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = v;
  __syncthreads();
  int t = 0;
  if (threadIdx.x == 0)
    for (int w = 0; w < B / 64; w++) t += part[w];
  return t;
}

// select: an undecided node with no undecided neighbour of larger weight
__global__ void k_select(int n, const int *__restrict__ sia, const int *__restrict__ sja, const int *__restrict__ tia,
                         const int *__restrict__ tja, const double *__restrict__ w, const int *__restrict__ cf,
                         unsigned char *__restrict__ pick, int *__restrict__ counts) {
  const int i = blockIdx.x * B + threadIdx.x;
  int top = 0;
  if (i < n) {
    pick[i] = 0;
    if (cf[i] == 0) {
      top = 1;
      const double wi = w[i];
      for (int k = sia[i]; k < sia[i + 1] && top; k++) {
        const int j = sja[k];
        if (cf[j] == 0 && w[j] > wi) top = 0;
      }
      for (int k = tia[i]; k < tia[i + 1] && top; k++) {
        const int j = tja[k];
        if (cf[j] == 0 && w[j] > wi) top = 0;
      }
      if (top) pick[i] = 1;
    }
  }
  const int t = block_count(top);
  if (threadIdx.x == 0 && t) atomicAdd(&counts[0], t);
}

__global__ void k_apply(int n, const unsigned char *__restrict__ pick, int *__restrict__ cf) {
  const int i = blockIdx.x * B + threadIdx.x;
  if (i < n && pick[i]) cf[i] = CPT;
}

// F phase: an undecided node with a C neighbour becomes F (only C marks
// are read, and F marks never create one, so the order is free)
__global__ void k_fmark(int n, const int *__restrict__ sia, const int *__restrict__ sja, const int *__restrict__ tia,
                        const int *__restrict__ tja, int *__restrict__ cf, int *__restrict__ counts) {
  const int i = blockIdx.x * B + threadIdx.x;
  int f = 0;
  if (i < n && cf[i] == 0) {
    for (int k = sia[i]; k < sia[i + 1] && !f; k++) f = (cf[sja[k]] == CPT);
    for (int k = tia[i]; k < tia[i + 1] && !f; k++) f = (cf[tja[k]] == CPT);
    if (f) cf[i] = FPT;
  }
  const int t = block_count(f);
  if (threadIdx.x == 0 && t) atomicAdd(&counts[1], t);
}

__global__ void k_rest_c(int n, int *__restrict__ cf) {
  const int i = blockIdx.x * B + threadIdx.x;
  if (i < n && cf[i] == 0) cf[i] = CPT;
}

// interpolation pattern: 1 entry for a C row, its strong C neighbours for
// an F row; cflag[i] = 1 for C (scanned into coarse indices)
__global__ void k_p_count(int n, const int *__restrict__ sia, const int *__restrict__ sja, const int *__restrict__ cf,
                          int *__restrict__ cnt, int *__restrict__ cflag) {
  const int i = blockIdx.x * B + threadIdx.x;
  if (i >= n) return;
  if (cf[i] == CPT) {
    cnt[i] = 1;
    cflag[i] = 1;
    return;
  }
  int c = 0;
  for (int k = sia[i]; k < sia[i + 1]; k++) c += (cf[sja[k]] == CPT);
  cnt[i] = c;
  cflag[i] = 0;
}

__global__ void k_p_fill(int n, const int *__restrict__ sia, const int *__restrict__ sja, const int *__restrict__ cf,
                         const int *__restrict__ cidx, const int *__restrict__ pia, int *__restrict__ pja) {
  const int i = blockIdx.x * B + threadIdx.x;
  if (i >= n) return;
  int p = pia[i];
  if (cf[i] == CPT) {
    pja[p] = cidx[i];
    return;
  }
  for (int k = sia[i]; k < sia[i + 1]; k++)
    if (cf[sja[k]] == CPT) pja[p++] = cidx[sja[k]];
}

// j in S_i: returns 0 (not in S_i), 1 (a strong F neighbour) or 2 (a
// strong C neighbour, *slot its P entry).  A column listed twice counts by
// its last copy, as the host's markers do.
__device__ __forceinline__ int in_row(int j, const int *__restrict__ sja, int s0, int s1, const int *__restrict__ cf,
                                      int p0, int *slot) {
  int kind = 0, c = 0;
  for (int k = s0; k < s1; k++) {
    const int m = sja[k];
    const int isc = (cf[m] == CPT);
    if (m == j) {
      kind = isc ? 2 : 1;
      if (isc) *slot = p0 + c;
    }
    c += isc;
  }
  return kind;
}

// weights of F row i, in the host's order: P starts from the strong C
// entries a_ij, weak entries lump into the diagonal, each strong F
// neighbour f distributes a_if a_fj / den_f over row f's entries in C_i^s
// (or lumps a_if when |den_f| < 1e-15), and the row is scaled by -1/dhat
// when |dhat| > 1e-15
__global__ void k_p_vals(int n, const int *__restrict__ aia, const int *__restrict__ aja,
                         const double *__restrict__ aval, const int *__restrict__ sia, const int *__restrict__ sja,
                         const int *__restrict__ cf, const int *__restrict__ pia, double *__restrict__ pval) {
  const int i = blockIdx.x * B + threadIdx.x;
  if (i >= n) return;
  const int p0 = pia[i], p1 = pia[i + 1];
  if (cf[i] == CPT) {
    pval[p0] = 1.0;
    return;
  }
  const int s0 = sia[i], s1 = sia[i + 1];
  for (int p = p0; p < p1; p++) pval[p] = 0.0;
  double diag = 0.0, weak = 0.0;
  for (int k = aia[i]; k < aia[i + 1]; k++) {
    const int j = aja[k];
    const double a = aval[k];
    if (j == i) {
      diag = a;
      continue;
    }
    int slot = -1;
    const int kind = in_row(j, sja, s0, s1, cf, p0, &slot);
    if (kind == 2) pval[slot] += a;
    else if (kind == 1 && cf[j] == FPT) continue;
    else weak += a;
  }
  double dhat = diag + weak;
  for (int k = aia[i]; k < aia[i + 1]; k++) {
    const int f = aja[k];
    if (f == i) continue;
    int slot = -1;
    if (in_row(f, sja, s0, s1, cf, p0, &slot) != 1 || cf[f] != FPT) continue;
    const double aik = aval[k];
    double den = 0.0;
    for (int q = aia[f]; q < aia[f + 1]; q++)
      if (in_row(aja[q], sja, s0, s1, cf, p0, &slot) == 2) den += aval[q];
    if (fabs(den) < 1e-15) {
      dhat += aik;
    } else {
      for (int q = aia[f]; q < aia[f + 1]; q++)
        if (in_row(aja[q], sja, s0, s1, cf, p0, &slot) == 2) pval[slot] += aik * aval[q] / den;
    }
  }
  if (fabs(dhat) > 1e-15)
    for (int p = p0; p < p1; p++) pval[p] /= -dhat;
}

struct dbuf {
  void *p[32];
  int k;
  template <typename T>
  T *get(size_t count) {
    if (k >= 32) return nullptr;
    void *q = nullptr;
    if (hipMalloc(&q, (count ? count : 1) * sizeof(T)) != hipSuccess) return nullptr;
    p[k++] = q;
    return (T *)q;
  }
  void release() {
    for (int t = 0; t < k; t++) (void)hipFree(p[t]);
    k = 0;
  }
};

}  // namespace

// one level's device state: A, then S, S^T and the split as they form
struct bn_dsetup {
  int n, nnz;
  int *aia, *aja;
  double *aval;
  int snnz, smax, amax;
  int *sia, *sja, *tia, *tja;
  double *sval;
  int *cf;
  dbuf buf;
};

#define DS_CHECK(x)                         \
  do {                                      \
    if ((x) != hipSuccess) {                \
      (void)hipGetLastError();              \
      return -1;                            \
    }                                       \
  } while (0)

extern "C" {

/* Upload A for the device stages of its level; NULL when no device is
 * present or an allocation fails (the host routines then run). */
bn_dsetup *bn_dsetup_begin(const dcsr_mat *A) {
  if (!bn_device_ok() || !A || A->row <= 0 || A->nnz <= 0) return nullptr;
  bn_dsetup *c = (bn_dsetup *)calloc(1, sizeof(bn_dsetup));
  if (!c) return nullptr;
  c->n = A->row;
  c->nnz = A->nnz;
  for (int i = 0; i < A->row; i++)
    if (A->ia[i + 1] - A->ia[i] > c->amax) c->amax = A->ia[i + 1] - A->ia[i];
  c->aia = c->buf.get<int>((size_t)c->n + 1);
  c->aja = c->buf.get<int>((size_t)c->nnz);
  c->aval = c->buf.get<double>((size_t)c->nnz);
  c->cf = c->buf.get<int>((size_t)c->n);
  if (!c->aia || !c->aja || !c->aval || !c->cf ||
      hipMemcpy(c->aia, A->ia, ((size_t)c->n + 1) * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->aja, A->ja, (size_t)c->nnz * sizeof(int), hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->aval, A->val, (size_t)c->nnz * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
    (void)hipGetLastError();
    c->buf.release();
    free(c);
    return nullptr;
  }
  return c;
}

void bn_dsetup_end(bn_dsetup *c) {
  if (!c) return;
  c->buf.release();
  free(c);
}

/* S (host copy in *S, allocated here) with clamp = 1 for the MEX rule;
 * colcount (n ints, may be NULL) receives |S^T_i|, the column counts. */
int bn_dsetup_strength(bn_dsetup *c, double theta, int clamp, dcsr_mat *S, int *colcount) {
  const int n = c->n;
  double *thr = c->buf.get<double>((size_t)n);
  int *cnt = c->buf.get<int>((size_t)n + 1);
  c->sia = c->buf.get<int>((size_t)n + 1);
  if (!thr || !cnt || !c->sia) return -1;
  hipLaunchKernelGGL(k_s_count, dim3(nb(n)), dim3(B), 0, 0, n, c->aia, c->aja, c->aval, theta, clamp, thr, cnt);
  DS_CHECK(hipMemsetAsync(cnt + n, 0, sizeof(int), 0));
  if (bn_dev_scan_i32(cnt, c->sia, (long long)n + 1, 0)) return -1;
  int snnz = 0;
  DS_CHECK(hipMemcpy(&snnz, c->sia + n, sizeof(int), hipMemcpyDeviceToHost));
  c->snnz = snnz;
  c->sja = c->buf.get<int>((size_t)snnz);
  c->sval = c->buf.get<double>((size_t)snnz);
  if (!c->sja || !c->sval) return -1;
  hipLaunchKernelGGL(k_s_fill, dim3(nb(n)), dim3(B), 0, 0, n, c->aia, c->aja, c->aval, thr, clamp, c->sia, c->sja,
                     c->sval);
  // S^T as sets, for the rounds, and its row lengths
  int *tcnt = c->buf.get<int>((size_t)n + 1);
  c->tia = c->buf.get<int>((size_t)n + 1);
  c->tja = c->buf.get<int>((size_t)snnz);
  if (!tcnt || !c->tia || !c->tja) return -1;
  DS_CHECK(hipMemsetAsync(tcnt, 0, ((size_t)n + 1) * sizeof(int), 0));
  hipLaunchKernelGGL(k_col_count, dim3(nb(n)), dim3(B), 0, 0, n, c->sia, c->sja, tcnt);
  if (bn_dev_scan_i32(tcnt, c->tia, (long long)n + 1, 0)) return -1;
  int *tnext = (int *)thr;  // thr is spent: its n doubles serve as the n fill cursors
  DS_CHECK(hipMemcpyAsync(tnext, c->tia, (size_t)n * sizeof(int), hipMemcpyDeviceToDevice, 0));
  hipLaunchKernelGGL(k_col_fill, dim3(nb(n)), dim3(B), 0, 0, n, c->sia, c->sja, tnext, c->tja);
  DS_CHECK(hipGetLastError());
  // host copies
  memset(S, 0, sizeof(*S));
  S->row = n;
  S->col = n;
  S->nnz = snnz;
  S->ia = (int *)bn_xmalloc(((size_t)n + 1) * sizeof(int));
  S->ja = (int *)bn_xmalloc((size_t)(snnz > 0 ? snnz : 1) * sizeof(int));
  S->val = (double *)bn_xmalloc((size_t)(snnz > 0 ? snnz : 1) * sizeof(double));
  hipError_t e = hipMemcpy(S->ia, c->sia, ((size_t)n + 1) * sizeof(int), hipMemcpyDeviceToHost);
  if (e == hipSuccess && snnz > 0) e = hipMemcpy(S->ja, c->sja, (size_t)snnz * sizeof(int), hipMemcpyDeviceToHost);
  if (e == hipSuccess && snnz > 0)
    e = hipMemcpy(S->val, c->sval, (size_t)snnz * sizeof(double), hipMemcpyDeviceToHost);
  if (e == hipSuccess && colcount) e = hipMemcpy(colcount, tcnt, (size_t)n * sizeof(int), hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    free(S->ia);
    free(S->ja);
    free(S->val);
    memset(S, 0, sizeof(*S));
    return -1;
  }
  int smax = 0;
  for (int i = 0; i < n; i++)
    if (S->ia[i + 1] - S->ia[i] > smax) smax = S->ia[i + 1] - S->ia[i];
  c->smax = smax;
  return 0;
}

/* Independent-set rounds on S u S^T from the host split cf (0 undecided)
 * and weights w until no node is undecided; cf is updated in place.  The
 * sequence of rounds is the host's: a round with no pick makes every
 * undecided node C and ends the rounds. */
int bn_dsetup_rounds(bn_dsetup *c, const double *w, int *cf, int undecided) {
  const int n = c->n;
  if (!c->tja) return -1;
  double *dw = c->buf.get<double>((size_t)n);
  unsigned char *pick = c->buf.get<unsigned char>((size_t)n);
  int *counts = c->buf.get<int>(2);
  if (!dw || !pick || !counts) return -1;
  DS_CHECK(hipMemcpy(dw, w, (size_t)n * sizeof(double), hipMemcpyHostToDevice));
  DS_CHECK(hipMemcpy(c->cf, cf, (size_t)n * sizeof(int), hipMemcpyHostToDevice));
  long long left = undecided;
  while (left > 0) {
    int h[2] = {0, 0};
    DS_CHECK(hipMemsetAsync(counts, 0, 2 * sizeof(int), 0));
    hipLaunchKernelGGL(k_select, dim3(nb(n)), dim3(B), 0, 0, n, c->sia, c->sja, c->tia, c->tja, dw, c->cf, pick,
                       counts);
    hipLaunchKernelGGL(k_apply, dim3(nb(n)), dim3(B), 0, 0, n, pick, c->cf);
    // with no pick there is no new C mark and the F phase changes nothing
    hipLaunchKernelGGL(k_fmark, dim3(nb(n)), dim3(B), 0, 0, n, c->sia, c->sja, c->tia, c->tja, c->cf, counts);
    DS_CHECK(hipMemcpy(h, counts, 2 * sizeof(int), hipMemcpyDeviceToHost));
    if (h[0] == 0) {
      hipLaunchKernelGGL(k_rest_c, dim3(nb(n)), dim3(B), 0, 0, n, c->cf);
      break;
    }
    left -= (long long)h[0] + (long long)h[1];
  }
  DS_CHECK(hipMemcpy(cf, c->cf, (size_t)n * sizeof(int), hipMemcpyDeviceToHost));
  return 0;
}

/* Standard interpolation of the final split cf (host copy in *P, allocated
 * here); -1 when the level's rows are too long for the per-row kernel. */
int bn_dsetup_interp(bn_dsetup *c, const int *cf, dcsr_mat *P) {
  const int n = c->n;
  if (!c->sja || c->smax > INTERP_MAX_S || c->amax > INTERP_MAX_ROW || (double)c->nnz > INTERP_MAX_MEAN * n)
    return -1;
  int *cnt = c->buf.get<int>((size_t)n + 1), *cflag = c->buf.get<int>((size_t)n + 1);
  int *pia = c->buf.get<int>((size_t)n + 1), *cidx = c->buf.get<int>((size_t)n + 1);
  if (!cnt || !cflag || !pia || !cidx) return -1;
  DS_CHECK(hipMemcpy(c->cf, cf, (size_t)n * sizeof(int), hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_p_count, dim3(nb(n)), dim3(B), 0, 0, n, c->sia, c->sja, c->cf, cnt, cflag);
  DS_CHECK(hipMemsetAsync(cnt + n, 0, sizeof(int), 0));
  DS_CHECK(hipMemsetAsync(cflag + n, 0, sizeof(int), 0));
  if (bn_dev_scan_i32(cnt, pia, (long long)n + 1, 0) || bn_dev_scan_i32(cflag, cidx, (long long)n + 1, 0)) return -1;
  int h[2] = {0, 0};
  DS_CHECK(hipMemcpy(&h[0], pia + n, sizeof(int), hipMemcpyDeviceToHost));
  DS_CHECK(hipMemcpy(&h[1], cidx + n, sizeof(int), hipMemcpyDeviceToHost));
  const int pnnz = h[0], nc = h[1];
  int *pja = c->buf.get<int>((size_t)pnnz);
  double *pval = c->buf.get<double>((size_t)pnnz);
  if (!pja || !pval) return -1;
  hipLaunchKernelGGL(k_p_fill, dim3(nb(n)), dim3(B), 0, 0, n, c->sia, c->sja, c->cf, cidx, pia, pja);
  hipLaunchKernelGGL(k_p_vals, dim3(nb(n)), dim3(B), 0, 0, n, c->aia, c->aja, c->aval, c->sia, c->sja, c->cf, pia,
                     pval);
  DS_CHECK(hipGetLastError());
  memset(P, 0, sizeof(*P));
  P->row = n;
  P->col = nc;
  P->nnz = pnnz;
  P->ia = (int *)bn_xmalloc(((size_t)n + 1) * sizeof(int));
  P->ja = (int *)bn_xmalloc((size_t)(pnnz > 0 ? pnnz : 1) * sizeof(int));
  P->val = (double *)bn_xmalloc((size_t)(pnnz > 0 ? pnnz : 1) * sizeof(double));
  hipError_t e = hipMemcpy(P->ia, pia, ((size_t)n + 1) * sizeof(int), hipMemcpyDeviceToHost);
  if (e == hipSuccess && pnnz > 0) e = hipMemcpy(P->ja, pja, (size_t)pnnz * sizeof(int), hipMemcpyDeviceToHost);
  if (e == hipSuccess && pnnz > 0) e = hipMemcpy(P->val, pval, (size_t)pnnz * sizeof(double), hipMemcpyDeviceToHost);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    free(P->ia);
    free(P->ja);
    free(P->val);
    memset(P, 0, sizeof(*P));
    return -1;
  }
  return 0;
}

}  // extern "C"
