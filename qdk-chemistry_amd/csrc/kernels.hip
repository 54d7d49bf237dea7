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
#include "fixed_point.hpp"
#include "kernels.hpp"
#include "screen.hpp"

namespace qfe {

// ------------------------------------------------------------ density norms
// Pn[s1][s2] = max |P_block(s1, s2)| in double, the norms the reference
// screens with (compute_shellblock_norm, libint2_direct.cpp:43-59)
__global__ void k_shell_norm(const double* __restrict__ P, int nbf, const int* __restrict__ sh_bf0,
                             const int* __restrict__ sh_nf, int nsh, double* __restrict__ Pn,
                             int* __restrict__ bad) {
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
  Pn[(size_t)s1 * nsh + s2] = m;
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
//
// counts/offsets are [2][nitems]: the FP64 part, then the FP32 part (quartets
// whose Schwarz bound Q_ab Q_cd is below mix_thr when mixed precision is on).
// The ket pairs q of an item with Q_p Q_q >= thr / pmax form a prefix of its
// bucket range (the bucket is sorted by Q descending); only they can pass the
// density test.  The count pass also stores nslot[item] = prefix length x nk
// (nk = functions per ket pair, cp_nk[item_cp[item]]): the item's J_cd share
// slots, and nmask[item] = 64-bit words of its survivor mask (one per 64
// prefix positions).  With jmask, the write pass stores those words (bit j
// of word k: position 64 k + j survived), so the J_cd reduction reads only
// the slots that the class kernels wrote.
//
// Schwarz-prefix cut: thr / pmax lowered by 1e-12 relative, so that every
// quartet it drops fails the exact density test whatever the rounding
__device__ __forceinline__ double prefix_cut(double thr, double pmax) { return thr / pmax * (1.0 - 1.0e-12); }

// One wave scans one item.  pn(c, d) returns max(Pn[a][c], Pn[a][d],
// Pn[b][c], Pn[b][d]) for the bra shells (a, b); Pcd[q] = Pn of ket pair q.
//
// The density test is the reference's own, in double precision and with its
// operand order: the reference's loop visits a quartet with the pair holding
// the larger shell index (then the larger second index) as (s1, s2) and skips
// it when max6 * Q(s1, s2) * Q(s3, s4) < threshold (libint2_direct.cpp:
// 357-402), so the product is formed in that order here too.  The Schwarz
// prefix cut uses thr / pmax lowered by 1e-12 relative: a quartet beyond the
// prefix then fails the exact test in either product order.
template <bool WRITE, class PN>
__device__ __forceinline__ void scan_item(long long item, int lane, const int2* __restrict__ items, long long nitems,
                                          const int2* __restrict__ buckets, const int* __restrict__ pair_bucket,
                                          const int2* __restrict__ psh, const double* __restrict__ Q,
                                          const double* __restrict__ Pcd, double pab, int2 ab, const PN& pn,
                                          double cut, double thr, double mix_thr, int* __restrict__ counts,
                                          const long long* __restrict__ offsets, int2* __restrict__ tasks,
                                          int nk, int* __restrict__ nslot, int* __restrict__ nmask,
                                          const long long* __restrict__ mbase,
                                          unsigned long long* __restrict__ jmask) {
  const bool masks = WRITE && jmask != nullptr;
  if (WRITE && !masks && counts[item] == 0 && counts[nitems + item] == 0) return;
  const int2 it = items[item];
  const int p = it.x;
  const int2 bk = buckets[it.y];
  int end = bk.y;
  if (pair_bucket[p] == it.y) end = p + 1;  // same bucket: q <= p
  const double Qp = Q[p];
  // the reference's loop order of the bra pair: (larger shell, smaller shell)
  const int pa1 = ab.x > ab.y ? ab.x : ab.y, pa2 = ab.x > ab.y ? ab.y : ab.x;
  const long long base64 = WRITE ? offsets[item] : 0;
  const long long base32 = WRITE ? offsets[nitems + item] : 0;
  unsigned long long* const mw = masks ? jmask + mbase[item] : nullptr;
  int c64 = 0, c32 = 0, npre = 0;
  const unsigned long long below = (1ull << lane) - 1ull;
  for (int q0 = bk.x; q0 < end; q0 += 64) {
    const int q = q0 + lane;
    const double Qq = q < end ? Q[q] : 0.0;
    // sorted by Q descending: if the first ket (lane 0, q0 < end) fails the
    // Schwarz bound, all later do
    if (Qp * __shfl(Qq, 0) < cut) break;
    bool ok = false, lo = false, pre = false;
    if (q < end) {
      pre = !(Qp * Qq < cut);
      if (pre) {
        const int2 cd = psh[q];
        const double m = fmax(fmax(pab, Pcd[q]), pn(cd.x, cd.y));
        const int qc1 = cd.x > cd.y ? cd.x : cd.y, qc2 = cd.x > cd.y ? cd.y : cd.x;
        const bool bra_first = pa1 > qc1 || (pa1 == qc1 && pa2 >= qc2);
        const double Q12 = bra_first ? Qp : Qq, Q34 = bra_first ? Qq : Qp;
        ok = !(m * Q12 * Q34 < thr);  // libint2_direct.cpp:400-402
      }
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
    if (masks && lane == 0) mw[(q0 - bk.x) >> 6] = m64 | m32;
    c64 += __popcll(m64);
    c32 += __popcll(m32);
    npre += __popcll(__ballot(pre));
  }
  if (!WRITE && lane == 0) {
    counts[item] = c64;
    counts[nitems + item] = c32;
    nslot[item] = npre * nk;
    nmask[item] = (npre + 63) >> 6;
  }
}

// One block per segment (the items of one bra pair (a, b) in one class
// pair): the density-norm rows of a and b are staged in LDS, so the four
// bra-ket norms of a candidate are LDS reads and only Pcd[q] comes from HBM
// (coalesced, like Q and psh).  Waves take the segment's items in turn.
template <bool WRITE>
__global__ __launch_bounds__(256) void k_task_scan_seg(const int2* __restrict__ segs, const int2* __restrict__ items,
                                                       long long nitems, const int2* __restrict__ buckets,
                                                       const int* __restrict__ pair_bucket,
                                                       const int2* __restrict__ psh, const double* __restrict__ Q,
                                                       const double* __restrict__ Pn, const double* __restrict__ Pcd,
                                                       int nsh, const double* __restrict__ pmax_ptr, double thr,
                                                       double mix_thr, int* __restrict__ counts,
                                                       const long long* __restrict__ offsets,
                                                       int2* __restrict__ tasks, const int* __restrict__ item_cp,
                                                       const int* __restrict__ cp_nk, int* __restrict__ nslot,
                                                       int* __restrict__ nmask, const long long* __restrict__ mbase,
                                                       unsigned long long* __restrict__ jmask) {
  extern __shared__ double prow[];  // [2][nsh]: rows a and b of Pn
  const int2 sg = segs[blockIdx.x];
  const int2 ab = psh[items[sg.x].x];
  for (int c = threadIdx.x; c < nsh; c += blockDim.x) {
    prow[c] = Pn[(size_t)ab.x * nsh + c];
    prow[nsh + c] = Pn[(size_t)ab.y * nsh + c];
  }
  __syncthreads();
  const double pab = prow[ab.y];
  const double cut = prefix_cut(thr, *pmax_ptr);
  const int nk = cp_nk[item_cp[sg.x]];
  const int lane = (int)(threadIdx.x & 63);
  auto pn = [&](int c, int d) { return fmax(fmax(prow[c], prow[d]), fmax(prow[nsh + c], prow[nsh + d])); };
  for (int item = sg.x + (int)(threadIdx.x >> 6); item < sg.y; item += (int)(blockDim.x >> 6))
    scan_item<WRITE>(item, lane, items, nitems, buckets, pair_bucket, psh, Q, Pcd, pab, ab, pn, cut, thr, mix_thr,
                     counts, offsets, tasks, nk, nslot, nmask, mbase, jmask);
}

// The same scan with one wave per item and the norms gathered from HBM (for
// shell counts whose two rows do not fit the LDS).
template <bool WRITE>
__global__ __launch_bounds__(256) void k_task_scan(const int2* __restrict__ items, long long nitems,
                                                   const int2* __restrict__ buckets,
                                                   const int* __restrict__ pair_bucket,
                                                   const int2* __restrict__ psh, const double* __restrict__ Q,
                                                   const double* __restrict__ Pn, const double* __restrict__ Pcd,
                                                   int nsh, const double* __restrict__ pmax_ptr, double thr,
                                                   double mix_thr, int* __restrict__ counts,
                                                   const long long* __restrict__ offsets,
                                                   int2* __restrict__ tasks, const int* __restrict__ item_cp,
                                                   const int* __restrict__ cp_nk, int* __restrict__ nslot,
                                                   int* __restrict__ nmask, const long long* __restrict__ mbase,
                                                   unsigned long long* __restrict__ jmask) {
  const long long item = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / 64;
  const int lane = threadIdx.x & 63;
  if (item >= nitems) return;
  const int2 ab = psh[items[item].x];
  const double* ra = Pn + (size_t)ab.x * nsh;
  const double* rb = Pn + (size_t)ab.y * nsh;
  auto pn = [&](int c, int d) { return fmax(fmax(ra[c], ra[d]), fmax(rb[c], rb[d])); };
  scan_item<WRITE>(item, lane, items, nitems, buckets, pair_bucket, psh, Q, Pcd, ra[ab.y], ab, pn,
                   prefix_cut(thr, *pmax_ptr), thr, mix_thr, counts, offsets, tasks, cp_nk[item_cp[item]], nslot,
                   nmask, mbase, jmask);
}

// Pcd[q] = Pn of the shell pair of pair q
__global__ void k_pair_norm(const int2* __restrict__ psh, int npair, const double* __restrict__ Pn, int nsh,
                            double* __restrict__ Pcd) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q < npair) Pcd[q] = Pn[(size_t)psh[q].x * nsh + psh[q].y];
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

// J and K of a deterministic build from their fixed-point sums, symmetrised
// like k_symmetrize: J = (J + J^T) / 4, K = kq (K + K^T) / 2 (kq, the
// exchange scale (alpha + beta) / 4, is applied here rather than per update)
__global__ void k_fixed_finish(const unsigned long long* __restrict__ Jx, const unsigned long long* __restrict__ Kx,
                               double* __restrict__ J, double* __restrict__ K, int n, double kq) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  const int i = blockIdx.y;
  if (j > i || i >= n) return;
  const size_t a = (size_t)i * n + j, b = (size_t)j * n + i;
  if (J) {
    const double v = 0.25 * (fx_value((long long)Jx[2 * a], (long long)Jx[2 * a + 1]) +
                             fx_value((long long)Jx[2 * b], (long long)Jx[2 * b + 1]));
    J[a] = v;
    J[b] = v;
  }
  if (K) {
    const double v = 0.5 * kq * (fx_value((long long)Kx[2 * a], (long long)Kx[2 * a + 1]) +
                                 fx_value((long long)Kx[2 * b], (long long)Kx[2 * b + 1]));
    K[a] = v;
    K[b] = v;
  }
}

// ------------------------------------------------------------ host wrappers
void launch_shell_norm(const double* P, int nbf, const int* sh_bf0, const int* sh_nf, int nsh, double* Pn, int* bad,
                       double* pmax, hipStream_t s) {
  dim3 block(64), grid((nsh + 63) / 64, nsh);
  hipLaunchKernelGGL(k_shell_norm, grid, block, 0, s, P, nbf, sh_bf0, sh_nf, nsh, Pn, bad);
  (void)hipMemsetAsync(pmax, 0, sizeof(double), s);
  size_t n = (size_t)nsh * nsh;
  int g = (int)((n + 255) / 256);
  if (g > 1024) g = 1024;
  hipLaunchKernelGGL(k_max_reduce, dim3(g), dim3(256), 0, s, Pn, n, pmax);
}

void launch_pair_norm(const int2* psh, int npair, const double* Pn, int nsh, double* Pcd, hipStream_t s) {
  if (npair <= 0) return;
  hipLaunchKernelGGL(k_pair_norm, dim3((npair + 255) / 256), dim3(256), 0, s, psh, npair, Pn, nsh, Pcd);
}

template <bool WRITE>
static void task_scan(const ScanArgs& a, int* counts, const long long* offsets, int2* tasks, int* nslot,
                      int* nmask, const long long* mbase, unsigned long long* jmask, hipStream_t s) {
  if (a.nitems <= 0) return;
  const size_t lds = 2 * sizeof(double) * (size_t)a.nsh;
  if (a.segs && a.nseg > 0 && lds <= (size_t)a.lds_max) {
    auto kern = k_task_scan_seg<WRITE>;
    if (lds > 65536) (void)hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    hipLaunchKernelGGL(kern, dim3((unsigned)a.nseg), dim3(256), lds, s, a.segs, a.items, a.nitems, a.buckets,
                       a.pair_bucket, a.psh, a.Q, a.Pn, a.Pcd, a.nsh, a.pmax, a.thr, a.mix_thr, counts, offsets,
                       tasks, a.item_cp, a.cp_nk, nslot, nmask, mbase, jmask);
  } else {
    const long long threads = a.nitems * 64;
    hipLaunchKernelGGL((k_task_scan<WRITE>), dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, a.items,
                       a.nitems, a.buckets, a.pair_bucket, a.psh, a.Q, a.Pn, a.Pcd, a.nsh, a.pmax, a.thr, a.mix_thr,
                       counts, offsets, tasks, a.item_cp, a.cp_nk, nslot, nmask, mbase, jmask);
  }
}

void launch_task_count(const ScanArgs& a, int* counts, int* nslot, int* nmask, hipStream_t s) {
  task_scan<false>(a, counts, nullptr, nullptr, nslot, nmask, nullptr, nullptr, s);
}

void launch_task_write(const ScanArgs& a, const int* counts, const long long* offsets, int2* tasks,
                       const long long* mbase, unsigned long long* jmask, hipStream_t s) {
  task_scan<true>(a, (int*)counts, offsets, tasks, nullptr, nullptr, mbase, jmask, s);
}

// The scans and sums read the int counts widened to 64 bits: hipcub
// accumulates in the input's value type, and the J_cd slot total of a large
// system exceeds 2^31.
struct WidenInt {
  __host__ __device__ long long operator()(int x) const { return (long long)x; }
};
using WideInts = rocprim::transform_iterator<const int*, WidenInt, long long>;

size_t scan_temp_bytes(long long n) {
  size_t a = 0, b = 0;
  const WideInts it(nullptr, WidenInt());
  (void)hipcub::DeviceScan::ExclusiveSum(nullptr, a, it, (long long*)nullptr, (int)n);
  (void)hipcub::DeviceReduce::Sum(nullptr, b, it, (long long*)nullptr, (int)n);
  return a > b ? a : b;
}

void launch_exclusive_scan(void* temp, size_t temp_bytes, const int* in, long long* out, long long n,
                           hipStream_t s) {
  (void)hipcub::DeviceScan::ExclusiveSum(temp, temp_bytes, WideInts(in, WidenInt()), out, (int)n, s);
}

void launch_sum(void* temp, size_t temp_bytes, const int* in, long long* out, long long n, hipStream_t s) {
  (void)hipcub::DeviceReduce::Sum(temp, temp_bytes, WideInts(in, WidenInt()), out, (int)n, s);
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

void launch_fixed_finish(const unsigned long long* Jx, const unsigned long long* Kx, double* J, double* K, int n,
                         double kq, hipStream_t s) {
  if (!J && !K) return;
  dim3 block(256), grid((n + 255) / 256, n);
  hipLaunchKernelGGL(k_fixed_finish, grid, block, 0, s, Jx, Kx, J, K, n, kq);
}

// ------------------------------------------------------------ J_cd by ket pair
// Share slots: item (p, kb) owns nslot[item] / NK positions j (ket pair
// q = first pair of bucket kb + j), NK doubles each, from jbuf + jslot[item].
// A tile is up to 256 consecutive bra pairs of one bra bucket (items
// item0 + i * stride, i < np) against positions [j0, j0 + kJcdTileKet) of one
// ket bucket whose first pair is q0.  One wave per tile walks the tile's
// items (their prefix lengths, slot bases and mask bases are loaded once, four
// per lane); for each 64-position block it gathers every item's survivor word
// (one per lane), visits only the items with a survivor there (two at a time,
// their bases broadcast with readlane), reads the surviving shares of the
// block as contiguous runs (lane l holds elements l, l + 64, ... of the
// block's 64 * NK doubles) and adds the sums to J.
template <int NK>
__global__ __launch_bounds__(256) void k_jcd_tiles(const int4* __restrict__ tiles, int ntiles,
                                                   const int* __restrict__ nslot,
                                                   const long long* __restrict__ jslot,
                                                   const long long* __restrict__ mbase,
                                                   const unsigned long long* __restrict__ jmask,
                                                   const double* __restrict__ jbuf,
                                                   const PairHdr* __restrict__ pairs, int nsb,
                                                   double* __restrict__ J, int nbf,
                                                   unsigned long long* __restrict__ Jx, int* __restrict__ over) {
  const int w = (int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6);
  const int lane = (int)(threadIdx.x & 63);
  if (w >= ntiles) return;  // whole waves
  const int4 tl = tiles[w];
  const int stride = tl.y >> 16, np = tl.y & 0xffff;
  int nl[4];
  long long bl[4], ml[4];
  int nmax = 0;
#pragma unroll
  for (int k = 0; k < 4; ++k) {
    const int i = lane + 64 * k;
    nl[k] = 0;
    bl[k] = 0;
    ml[k] = 0;
    if (i < np) {
      const int item = tl.x + i * stride;
      nl[k] = nslot[item] / NK;
      bl[k] = jslot[item];
      ml[k] = mbase[item];
    }
    nmax = nl[k] > nmax ? nl[k] : nmax;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    const int v = __shfl_xor(nmax, o);
    nmax = v > nmax ? v : nmax;
  }
  const int jend = tl.w + kJcdTileKet < nmax ? tl.w + kJcdTileKet : nmax;
  // lanes own flattened elements of the 64-position block: element
  // e = f * 64 + lane is function e % NK of position jc + e / NK, so a wave's
  // share loads are contiguous 512-byte runs of the item's slots
  for (int jc = tl.w; jc < jend; jc += 64) {
    double acc[NK];
#pragma unroll
    for (int f = 0; f < NK; ++f) acc[f] = 0.0;
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      // each lane gathers its own item's survivor word for these 64 positions;
      // the wave then visits only the items that have a survivor here
      unsigned long long wk = 0ull;
      if (nl[k] > jc) wk = jmask[ml[k] + (jc >> 6)];
      unsigned long long act = __ballot(wk != 0ull);
      while (act) {
        // two items per step, so that their share loads overlap (one for the
        // largest ket pairs, whose shares would not fit the registers twice)
        constexpr bool kTwo = NK <= 25;
        const int l0 = __builtin_ctzll(act);
        act &= act - 1ull;
        const int l1 = (kTwo && act) ? __builtin_ctzll(act) : l0;
        if (kTwo && act) act &= act - 1ull;
        const unsigned long long w0 = ((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)(wk >> 32), l0) << 32) |
                                      (unsigned)__builtin_amdgcn_readlane((int)(wk & 0xffffffffull), l0);
        const unsigned long long w1 = ((unsigned long long)(unsigned)__builtin_amdgcn_readlane((int)(wk >> 32), l1) << 32) |
                                      (unsigned)__builtin_amdgcn_readlane((int)(wk & 0xffffffffull), l1);
        const long long b0 = ((long long)__builtin_amdgcn_readlane((int)(bl[k] >> 32), l0) << 32) |
                             (unsigned)__builtin_amdgcn_readlane((int)(bl[k] & 0xffffffffLL), l0);
        const long long b1 = ((long long)__builtin_amdgcn_readlane((int)(bl[k] >> 32), l1) << 32) |
                             (unsigned)__builtin_amdgcn_readlane((int)(bl[k] & 0xffffffffLL), l1);
        const double* s0 = jbuf + b0 + (long long)jc * NK;
        const double* s1 = jbuf + b1 + (long long)jc * NK;
        if constexpr (kTwo) {
          double v0[NK], v1[NK];
#pragma unroll
          for (int f = 0; f < NK; ++f) {
            const int e = f * 64 + lane, jj = e / NK;
            const bool t0 = (w0 >> jj) & 1ull;
            const bool t1 = l1 != l0 && ((w1 >> jj) & 1ull);
            v0[f] = t0 ? s0[e] : 0.0;
            v1[f] = t1 ? s1[e] : 0.0;
          }
#pragma unroll
          for (int f = 0; f < NK; ++f) acc[f] += v0[f] + v1[f];
        } else {
#pragma unroll
          for (int f = 0; f < NK; ++f) {
            const int e = f * 64 + lane;
            if ((w0 >> (e / NK)) & 1ull) acc[f] += s0[e];
          }
        }
      }
    }
    int bad = 0;
#pragma unroll
    for (int f = 0; f < NK; ++f) {
      const int e = f * 64 + lane, j = jc + e / NK, fn = e % NK;
      if (j < nmax && acc[f] != 0.0) {
        const PairHdr* h = pairs + tl.z + j;
        const size_t e = (size_t)(h->bfa + fn / nsb) * nbf + h->bfb + fn % nsb;
        if (Jx)
          fx_global_add(Jx + 2 * e, acc[f], bad);
        else
          unsafeAtomicAdd(&J[e], acc[f]);
      }
    }
    if (bad) *over = 1;
  }
}

template <int NK>
static void jcd_tiles_launch(const int4* tiles, int ntiles, const int* nslot, const long long* jslot,
                             const long long* mbase, const unsigned long long* jmask, const double* jbuf,
                             const PairHdr* pairs, int nsb, double* J, int nbf, unsigned long long* Jx, int* over,
                             hipStream_t s) {
  const long long threads = (long long)ntiles * 64;
  hipLaunchKernelGGL((k_jcd_tiles<NK>), dim3((unsigned)((threads + 255) / 256)), dim3(256), 0, s, tiles, ntiles,
                     nslot, jslot, mbase, jmask, jbuf, pairs, nsb, J, nbf, Jx, over);
}

void launch_jcd_tiles(const int4* tiles, int ntiles, int nk, int nsb, const int* nslot, const long long* jslot,
                      const long long* mbase, const unsigned long long* jmask, const double* jbuf,
                      const PairHdr* pairs, double* J, int nbf, unsigned long long* Jx, int* over, hipStream_t s) {
  if (ntiles <= 0) return;
  switch (nk) {
    case 1: jcd_tiles_launch<1>(tiles, ntiles, nslot, jslot, mbase, jmask, jbuf, pairs, nsb, J, nbf, Jx, over, s); break;
    case 3: jcd_tiles_launch<3>(tiles, ntiles, nslot, jslot, mbase, jmask, jbuf, pairs, nsb, J, nbf, Jx, over, s); break;
    case 5: jcd_tiles_launch<5>(tiles, ntiles, nslot, jslot, mbase, jmask, jbuf, pairs, nsb, J, nbf, Jx, over, s); break;
    case 7: jcd_tiles_launch<7>(tiles, ntiles, nslot, jslot, mbase, jmask, jbuf, pairs, nsb, J, nbf, Jx, over, s); break;
    case 9: jcd_tiles_launch<9>(tiles, ntiles, nslot, jslot, mbase, jmask, jbuf, pairs, nsb, J, nbf, Jx, over, s); break;
    case 15: jcd_tiles_launch<15>(tiles, ntiles, nslot, jslot, mbase, jmask, jbuf, pairs, nsb, J, nbf, Jx, over, s); break;
    case 21: jcd_tiles_launch<21>(tiles, ntiles, nslot, jslot, mbase, jmask, jbuf, pairs, nsb, J, nbf, Jx, over, s); break;
    case 25: jcd_tiles_launch<25>(tiles, ntiles, nslot, jslot, mbase, jmask, jbuf, pairs, nsb, J, nbf, Jx, over, s); break;
    case 35: jcd_tiles_launch<35>(tiles, ntiles, nslot, jslot, mbase, jmask, jbuf, pairs, nsb, J, nbf, Jx, over, s); break;
    case 49: jcd_tiles_launch<49>(tiles, ntiles, nslot, jslot, mbase, jmask, jbuf, pairs, nsb, J, nbf, Jx, over, s); break;
    default: break;  // no other pure pair sizes up to f
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


// ------------------------------------------------------------ profiling
template <typename R>
__global__ __launch_bounds__(256) void k_count_prims(const int2* __restrict__ tasks, long long ntask,
                                                     const PairHdr* __restrict__ pairs, const R* __restrict__ sig,
                                                     const double* __restrict__ pmax,
                                                     int prim_screen, unsigned long long* __restrict__ count) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  unsigned long long n = 0;
  if (t < ntask) {
    const int2 tk = tasks[t];
    const PairHdr bra = pairs[tk.x], ket = pairs[tk.y];
    const R lnp = prim_ln_cut<R>(pmax, prim_screen);
    const R sb0 = sig[bra.prim0];
    for (int ik = ket.prim0; ik < ket.prim0 + ket.nprim; ++ik) {
      const R sk = sig[ik];
      if (sk + sb0 < lnp) break;
      const R blim = lnp - sk;
      for (int ib = bra.prim0; ib < bra.prim0 + bra.nprim; ++ib) {
        if (sig[ib] < blim) break;
        ++n;
      }
    }
  }
  // one atomic per wave
  for (int o = 32; o > 0; o >>= 1) n += __shfl_xor(n, o);
  if ((threadIdx.x & 63) == 0 && n) atomicAdd(count, n);
}

void launch_count_prims(const int2* tasks, long long ntask, const PairHdr* pairs, const double* sig,
                        const float* sig32, const double* pmax, int prim_screen, int fp32,
                        unsigned long long* count, hipStream_t s) {
  if (ntask <= 0) return;
  const dim3 grid((unsigned)((ntask + 255) / 256));
  if (fp32)
    hipLaunchKernelGGL(k_count_prims<float>, grid, dim3(256), 0, s, tasks, ntask, pairs, sig32, pmax,
                       prim_screen, count);
  else
    hipLaunchKernelGGL(k_count_prims<double>, grid, dim3(256), 0, s, tasks, ntask, pairs, sig, pmax,
                       prim_screen, count);
}

}  // namespace qfe
