// HIP kernels of the Fock-build engine (gfx950).
//
//   k_shell_norm     shell-block infinity norms of the density
//                    (compute_shellblock_norm, libint2_direct.cpp:43-59)
//   k_task_count /   two-pass construction of the screened canonical quartet
//   k_task_write     list, grouped by class (the loop + test of
//                    libint2_direct.cpp:357-411)
//   k_symmetrize     J = (J+J^T)/4, K = (K+K^T)/2 (libint2_direct.cpp:594-625)
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "eri_core.hpp"
#include "kernels.hpp"
#include "screen.hpp"

namespace qfe {

// ------------------------------------------------------------ density norms
// Pn is kept in float rounded towards +inf, so the screening test can only
// keep more quartets than the double-precision test of the reference
__global__ void k_shell_norm(const double* __restrict__ P, int nbf, const int* __restrict__ sh_bf0,
                             const int* __restrict__ sh_nf, int nsh, float* __restrict__ Pn,
                             double* __restrict__ Pnd, int* __restrict__ bad) {
  const int s2 = blockIdx.x * blockDim.x + threadIdx.x;
  const int s1 = blockIdx.y;
  if (s2 >= nsh) return;
  const int r0 = sh_bf0[s1], c0 = sh_bf0[s2], nr = sh_nf[s1], nc = sh_nf[s2];
  double m = 0.0;
  for (int i = 0; i < nr; ++i)
    for (int j = 0; j < nc; ++j) {
      const double v = fabs(P[(size_t)(r0 + i) * nbf + c0 + j]);
      if (!(v <= 1.0e300)) *bad = 1;  // NaN or Inf
      m = fmax(m, v);
    }
  Pn[(size_t)s1 * nsh + s2] = __double2float_ru(m);
  Pnd[(size_t)s1 * nsh + s2] = m;
}

__global__ void k_max_reduce(const double* __restrict__ x, size_t n, double* __restrict__ out) {
  __shared__ double sm[256];
  double m = 0.0;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
    m = fmax(m, x[i]);
  sm[threadIdx.x] = m;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) sm[threadIdx.x] = fmax(sm[threadIdx.x], sm[threadIdx.x + s]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    // values are non-negative: compare the bit patterns as unsigned integers
    atomicMax((unsigned long long*)out, (unsigned long long)__double_as_longlong(sm[0]));
  }
}

// ------------------------------------------------------------ task list
// work item = (bra pair p, ket bucket b); one wavefront scans bucket b.
__device__ __forceinline__ bool quartet_passes(int2 P1, double Q1, int2 P2, double Q2,
                                               const float* __restrict__ Pn, int nsh, double thr) {
  const size_t a = P1.x, b = P1.y, c = P2.x, d = P2.y;
  float m = Pn[a * nsh + b];
  m = fmaxf(m, Pn[c * nsh + d]);
  m = fmaxf(m, Pn[a * nsh + c]);
  m = fmaxf(m, Pn[a * nsh + d]);
  m = fmaxf(m, Pn[b * nsh + c]);
  m = fmaxf(m, Pn[b * nsh + d]);
  return !((double)m * Q1 * Q2 < thr);
}

// counts/offsets are [2][nitems]: the FP64 part, then the FP32 part (quartets
// whose Schwarz bound Q_ab Q_cd is below mix_thr when mixed precision is on).
template <bool WRITE>
__global__ __launch_bounds__(256) void k_task_scan(const int2* __restrict__ items, long long nitems,
                                                   const int2* __restrict__ buckets,
                                                   const int* __restrict__ pair_bucket,
                                                   const int2* __restrict__ psh,
                                                   const double* __restrict__ Q,
                                                   const float* __restrict__ Pn, int nsh,
                                                   const double* __restrict__ pmax_ptr, double thr,
                                                   double mix_thr, int* __restrict__ counts,
                                                   const long long* __restrict__ offsets,
                                                   int2* __restrict__ tasks) {
  const long long item = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / 64;
  const int lane = threadIdx.x & 63;
  if (item >= nitems) return;
  if (WRITE && counts[item] == 0 && counts[nitems + item] == 0) return;
  const int2 it = items[item];
  const int p = it.x;
  const int2 bk = buckets[it.y];
  int end = bk.y;
  if (pair_bucket[p] == it.y) end = p + 1;  // same bucket: q <= p
  const double Qp = Q[p];
  const int2 Pp = psh[p];
  const double cut = thr / (*pmax_ptr);
  const long long base64 = WRITE ? offsets[item] : 0;
  const long long base32 = WRITE ? offsets[nitems + item] : 0;
  int c64 = 0, c32 = 0;
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int q0 = bk.x; q0 < end; q0 += 64) {
    // sorted by Q descending: if the first ket fails the Schwarz bound, all later do
    if (Qp * Q[q0] < cut) break;
    const int q = q0 + lane;
    bool ok = false, lo = false;
    if (q < end) {
      const double Qq = Q[q];
      ok = quartet_passes(Pp, Qp, psh[q], Qq, Pn, nsh, thr);
      lo = Qp * Qq < mix_thr;
    }
    const unsigned long long m64 = __ballot(ok && !lo);
    const unsigned long long m32 = __ballot(ok && lo);
    if (WRITE && ok) {
      if (lo)
        tasks[base32 + c32 + __popcll(m32 & below)] = make_int2(p, q);
      else
        tasks[base64 + c64 + __popcll(m64 & below)] = make_int2(p, q);
    }
    c64 += __popcll(m64);
    c32 += __popcll(m32);
  }
  if (!WRITE && lane == 0) {
    counts[item] = c64;
    counts[nitems + item] = c32;
  }
}

// per class pair: surviving shell quartets and their primitive quartets
// (block-level LDS reduction, one global atomic per block and class pair)
constexpr int kMaxCP = 64;
__global__ __launch_bounds__(256) void k_class_stats(const int* __restrict__ counts, const int2* __restrict__ items,
                                                     long long nitems, const int* __restrict__ item_cp,
                                                     const int2* __restrict__ buckets,
                                                     const PairHdr* __restrict__ pairs, int ncp,
                                                     double* __restrict__ q_out, double* __restrict__ p_out) {
  __shared__ double sq[kMaxCP], sp[kMaxCP];
  for (int k = threadIdx.x; k < kMaxCP; k += blockDim.x) {
    sq[k] = 0.0;
    sp[k] = 0.0;
  }
  __syncthreads();
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < nitems) {
    const int c = counts[i];
    if (c != 0) {
      const int2 it = items[i];
      const double np = (double)pairs[it.x].nprim * (double)pairs[buckets[it.y].x].nprim;
      atomicAdd(&sq[item_cp[i]], (double)c);
      atomicAdd(&sp[item_cp[i]], (double)c * np);
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < ncp && k < kMaxCP; k += blockDim.x)
    if (sq[k] != 0.0) {
      unsafeAtomicAdd(&q_out[k], sq[k]);
      unsafeAtomicAdd(&p_out[k], sp[k]);
    }
}

__global__ void k_gather_offsets(const long long* __restrict__ offsets, const long long* __restrict__ idx,
                                 int n, long long* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = offsets[idx[i]];
}

// ------------------------------------------------------------ symmetrisation
__global__ void k_symmetrize(double* __restrict__ J, double* __restrict__ K, int n) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j > i || i >= n) return;
  const size_t a = (size_t)i * n + j, b = (size_t)j * n + i;
  if (J) {
    const double v = 0.25 * (J[a] + J[b]);
    J[a] = v;
    J[b] = v;
  }
  if (K) {
    const double v = 0.5 * (K[a] + K[b]);
    K[a] = v;
    K[b] = v;
  }
}

// ------------------------------------------------------------ host wrappers
void launch_shell_norm(const double* P, int nbf, const int* sh_bf0, const int* sh_nf, int nsh,
                       float* Pn, double* Pnd, int* bad, double* pmax, hipStream_t s) {
  dim3 block(64), grid((nsh + 63) / 64, nsh);
  hipLaunchKernelGGL(k_shell_norm, grid, block, 0, s, P, nbf, sh_bf0, sh_nf, nsh, Pn, Pnd, bad);
  (void)hipMemsetAsync(pmax, 0, sizeof(double), s);
  size_t n = (size_t)nsh * nsh;
  int g = (int)((n + 255) / 256);
  if (g > 1024) g = 1024;
  hipLaunchKernelGGL(k_max_reduce, dim3(g), dim3(256), 0, s, Pnd, n, pmax);
}

void launch_task_count(const int2* items, long long nitems, const int2* buckets, const int* pair_bucket,
                       const int2* pairs, const double* Q, const float* Pn, int nsh,
                       const double* pmax, double thr, double mix_thr, int* counts, hipStream_t s) {
  if (nitems <= 0) return;
  const long long threads = nitems * 64;
  const long long grid = (threads + 255) / 256;
  hipLaunchKernelGGL((k_task_scan<false>), dim3((unsigned)grid), dim3(256), 0, s, items, nitems, buckets,
                     pair_bucket, pairs, Q, Pn, nsh, pmax, thr, mix_thr, counts, (const long long*)nullptr,
                     (int2*)nullptr);
}

void launch_task_write(const int2* items, long long nitems, const int2* buckets, const int* pair_bucket,
                       const int2* pairs, const double* Q, const float* Pn, int nsh,
                       const double* pmax, double thr, double mix_thr, const int* counts,
                       const long long* offsets, int2* tasks, hipStream_t s) {
  if (nitems <= 0) return;
  const long long threads = nitems * 64;
  const long long grid = (threads + 255) / 256;
  hipLaunchKernelGGL((k_task_scan<true>), dim3((unsigned)grid), dim3(256), 0, s, items, nitems, buckets,
                     pair_bucket, pairs, Q, Pn, nsh, pmax, thr, mix_thr, (int*)counts, offsets, tasks);
}

size_t scan_temp_bytes(long long n) {
  size_t bytes = 0;
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, bytes, (const int*)nullptr, (long long*)nullptr, (int)n);
  return bytes;
}

void launch_exclusive_scan(void* temp, size_t temp_bytes, const int* in, long long* out, long long n,
                           hipStream_t s) {
  (void)hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, in, out, (int)n, s);
}

void launch_gather_offsets(const long long* offsets, const long long* idx, int n, long long* out,
                           hipStream_t s) {
  hipLaunchKernelGGL(k_gather_offsets, dim3((n + 63) / 64), dim3(64), 0, s, offsets, idx, n, out);
}

void launch_class_stats(const int* counts, const int2* items, long long nitems, const int* item_cp,
                        const int2* buckets, const PairHdr* pairs, int ncp, double* q_out, double* p_out,
                        hipStream_t s) {
  if (nitems <= 0) return;
  hipLaunchKernelGGL(k_class_stats, dim3((unsigned)((nitems + 255) / 256)), dim3(256), 0, s, counts, items,
                     nitems, item_cp, buckets, pairs, ncp, q_out, p_out);
}

void launch_symmetrize(double* J, double* K, int n, hipStream_t s) {
  dim3 block(256), grid((n + 255) / 256, n);
  hipLaunchKernelGGL(k_symmetrize, grid, block, 0, s, J, K, n);
}

// ------------------------------------------------------------ J_cd by ket pair
__global__ void k_run_bounds(const unsigned* __restrict__ keys, long long n, int* __restrict__ run_start,
                             int* __restrict__ run_end) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const unsigned k = keys[i];
  if (i == 0 || keys[i - 1] != k) run_start[k] = (int)i;
  if (i == n - 1 || keys[i + 1] != k) run_end[k] = (int)(i + 1);
}

__device__ __forceinline__ double wave_sum64(double v) {
  v += __shfl_xor(v, 32);
  v += __shfl_xor(v, 16);
  v += __shfl_xor(v, 8);
  v += __shfl_xor(v, 4);
  v += __shfl_xor(v, 2);
  v += __shfl_xor(v, 1);
  return v;
}

// one wave per ket pair of one pair class (NK functions per pair): each lane
// adds whole shares of the run (NK contiguous doubles), then the wave sums
// them and adds one atomic per J element
template <int NK>
__global__ __launch_bounds__(256) void k_jcd_sum(const PairHdr* __restrict__ pairs, int q0, int q1, int nsb,
                                                 const int* __restrict__ run_start, const int* __restrict__ run_end,
                                                 const unsigned* __restrict__ vals, const double* __restrict__ jbuf,
                                                 double* __restrict__ J, int nbf) {
  const int q = q0 + (int)((blockIdx.x * blockDim.x + threadIdx.x) / 64);
  const int lane = (int)(threadIdx.x & 63);
  if (q >= q1) return;  // whole waves
  const int i0 = run_start[q], i1 = run_end[q];
  if (i0 >= i1) return;
  double acc[NK];
#pragma unroll
  for (int f = 0; f < NK; ++f) acc[f] = 0.0;
  for (int i = i0 + lane; i < i1; i += 64) {
    const double* sh = jbuf + vals[i];
#pragma unroll
    for (int f = 0; f < NK; ++f) acc[f] += sh[f];
  }
  const PairHdr h = pairs[q];
#pragma unroll
  for (int f = 0; f < NK; ++f) {
    const double v = wave_sum64(acc[f]);
    if (lane == 0 && v != 0.0) unsafeAtomicAdd(&J[(size_t)(h.bfa + f / nsb) * nbf + h.bfb + f % nsb], v);
  }
}

template <int NK>
static void jcd_sum_class(const PairHdr* pairs, int q0, int q1, int nsb, const int* run_start, const int* run_end,
                          const unsigned* vals, const double* jbuf, double* J, int nbf, hipStream_t s) {
  const long long threads = (long long)(q1 - q0) * 64;
  hipLaunchKernelGGL((k_jcd_sum<NK>), dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, pairs, q0, q1, nsb,
                     run_start, run_end, vals, jbuf, J, nbf);
}

size_t jcd_sort_temp_bytes(long long n, int bits) {
  size_t bytes = 0;
  hipcub::DoubleBuffer<unsigned> k(nullptr, nullptr), v(nullptr, nullptr);
  (void)hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, k, v, (int)n, 0, bits);
  return bytes;
}

void launch_jcd_reduce(void* temp, size_t temp_bytes, unsigned* keys0, unsigned* keys1, unsigned* vals0,
                       unsigned* vals1, long long n, int bits, const PairHdr* pairs, int npair,
                       const int* cls_begin, int ncls, int* run_start, int* run_end, const double* jbuf, double* J,
                       int nbf, hipStream_t s) {
  if (n <= 0) return;
  hipcub::DoubleBuffer<unsigned> k(keys0, keys1), v(vals0, vals1);
  (void)hipcub::DeviceRadixSort::SortPairs(temp, temp_bytes, k, v, (int)n, 0, bits, s);
  (void)hipMemsetAsync(run_start, 0, sizeof(int) * npair, s);
  (void)hipMemsetAsync(run_end, 0, sizeof(int) * npair, s);
  hipLaunchKernelGGL(k_run_bounds, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, k.Current(), n, run_start,
                     run_end);
  // pairs are sorted by class: one launch per pair class, templated on its functions
  for (int c = 0; c < ncls; ++c) {
    const int q0 = cls_begin[c], q1 = cls_begin[c + 1];
    if (q0 >= q1) continue;
    int la = 0;
    while ((la + 1) * (la + 2) / 2 <= c) ++la;
    const int lb = c - la * (la + 1) / 2;
    const int nsa = la >= 2 ? 2 * la + 1 : (la + 1) * (la + 2) / 2;
    const int nsb = lb >= 2 ? 2 * lb + 1 : (lb + 1) * (lb + 2) / 2;
    const unsigned* vs = v.Current();
    switch (nsa * nsb) {
      case 1: jcd_sum_class<1>(pairs, q0, q1, nsb, run_start, run_end, vs, jbuf, J, nbf, s); break;
      case 3: jcd_sum_class<3>(pairs, q0, q1, nsb, run_start, run_end, vs, jbuf, J, nbf, s); break;
      case 5: jcd_sum_class<5>(pairs, q0, q1, nsb, run_start, run_end, vs, jbuf, J, nbf, s); break;
      case 7: jcd_sum_class<7>(pairs, q0, q1, nsb, run_start, run_end, vs, jbuf, J, nbf, s); break;
      case 9: jcd_sum_class<9>(pairs, q0, q1, nsb, run_start, run_end, vs, jbuf, J, nbf, s); break;
      case 15: jcd_sum_class<15>(pairs, q0, q1, nsb, run_start, run_end, vs, jbuf, J, nbf, s); break;
      case 21: jcd_sum_class<21>(pairs, q0, q1, nsb, run_start, run_end, vs, jbuf, J, nbf, s); break;
      case 25: jcd_sum_class<25>(pairs, q0, q1, nsb, run_start, run_end, vs, jbuf, J, nbf, s); break;
      case 35: jcd_sum_class<35>(pairs, q0, q1, nsb, run_start, run_end, vs, jbuf, J, nbf, s); break;
      case 49: jcd_sum_class<49>(pairs, q0, q1, nsb, run_start, run_end, vs, jbuf, J, nbf, s); break;
      default: break;  // no other pure pair sizes up to f
    }
  }
}

__global__ __launch_bounds__(256) void k_copy_words(const unsigned long long* __restrict__ src,
                                                    unsigned long long* __restrict__ dst, size_t n) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) dst[i] = src[i];
}

void launch_copy_words(const unsigned long long* src, unsigned long long* dst, size_t n, hipStream_t s) {
  if (n == 0) return;
  const size_t want = (n + 255) / 256;
  const unsigned grid = (unsigned)(want < 8192 ? want : 8192);
  hipLaunchKernelGGL(k_copy_words, dim3(grid), dim3(256), 0, s, src, dst, n);
}

}  // namespace qfe
