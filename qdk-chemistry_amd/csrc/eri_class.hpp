// Per-class ERI kernels (templates), instantiated one class per translation
// unit by the build (gen/cls_*.hip) to keep compile times bounded.
//   k_eri_jk_seg  Rys ERIs of one class chunk for the tasks of one bra pair
//                 + J/K digestion (libint2_direct.cpp:414-567)
//   k_schwarz  sum of squares of (ab|ab) for Q_ab = sqrt ||(ab|ab)||_F
//              (schwarz.cpp:26-78)
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>

#include "eri_core.hpp"
#include "kernels.hpp"

namespace qfe {

// ------------------------------------------------------------ host sfor
template <typename F, int... Is>
inline void hsfor_impl(F&& f, std::integer_sequence<int, Is...>) {
  (f(std::integral_constant<int, Is>{}), ...);
}
template <int N, typename F>
inline void hsfor(F&& f) {
  hsfor_impl(f, std::make_integer_sequence<int, N>{});
}

// LDS rows a segment kernel holds for bra functions [q0, q1): the a functions
// touched, then the b functions (all of them once the range spans two a's)
template <class C>
constexpr int seg_rows(int q0, int q1) {
  const int nma = (q1 - 1) / C::NSB - q0 / C::NSB + 1;
  return nma + (nma > 1 ? C::NSB : q1 - q0);
}
constexpr int kSegRowsMax = 6;  // K and P rows of nbf doubles: 6 x 16 B x nbf <= 160 KB up to nbf = 1706
template <class C>
constexpr int seg_bch(int hi) {
  for (int b = hi; b > 1; --b) {
    int worst = 0;
    for (int q0 = 0; q0 < C::NBRA; q0 += b) {
      const int r = seg_rows<C>(q0, q0 + b < C::NBRA ? q0 + b : C::NBRA);
      worst = r > worst ? r : worst;
    }
    if (worst <= kSegRowsMax) return b;
  }
  return 1;
}

// Accumulators per thread: bra functions x ket cartesians of one chunk.  48
// for classes up to d shells (their whole ket block fits); classes with an f
// shell hold fewer and are also chunked over the ket cartesians, so that no
// kernel instance spills registers (24, and 16 from L = 11 on).
constexpr int kAccMax = 48;
// A class whose kernels spilled with the default is compiled with its own
// QFE_JK_ACC_F (defined by its generated source).
template <class C>
constexpr int jk_acc() {
  if (C::LA <= 2 && C::LB <= 2 && C::LC <= 2 && C::LD <= 2) return kAccMax;
#ifdef QFE_JK_ACC_F
  return QFE_JK_ACC_F;
#else
  return C::L >= 11 ? 16 : 24;
#endif
}
template <class C>
struct JKChunks {
  static constexpr bool F = C::LA > 2 || C::LB > 2 || C::LC > 2 || C::LD > 2;
  static constexpr int ACC = jk_acc<C>();
  static constexpr int KCH = C::NKETC <= ACC ? C::NKETC : ACC;  // ket cartesians per chunk
  static constexpr int NKCH = (C::NKETC + KCH - 1) / KCH;
  // bra functions per chunk: as many as the accumulators allow, then fewer
  // until no chunk's segment kernel needs more than kSegRowsMax LDS rows
  static constexpr int BCH_ = ACC / KCH < 1 ? 1 : (ACC / KCH > C::NBRA ? C::NBRA : ACC / KCH);
  static constexpr int BCH = seg_bch<C>(BCH_);
  static constexpr int NBU = (C::NBRA + BCH - 1) / BCH;  // uniform bra chunks of BCH functions
  // or row-aligned: each a function's NSB bra functions in MP near-equal parts
  // of at most kSegRowsMax - 1 (one a row plus its b rows), when that gives
  // fewer chunks (wide b shells, where uniform chunks would straddle rows)
  static constexpr int WA = BCH_ < C::NSB ? (BCH_ < kSegRowsMax - 1 ? BCH_ : kSegRowsMax - 1)
                                          : (C::NSB < kSegRowsMax - 1 ? C::NSB : kSegRowsMax - 1);
  static constexpr int MP = (C::NSB + WA - 1) / WA;
  static constexpr bool ALIGNED = C::NSA * MP < NBU;
  static constexpr int NBCH = ALIGNED ? C::NSA * MP : NBU;
  static constexpr int NCH = NBCH * NKCH;  // chunk c: bra chunk c / NKCH, ket chunk c % NKCH
  template <int CI>
  struct At {
    static constexpr int b = CI / NKCH, k = CI % NKCH;
    static constexpr int Q0 = ALIGNED ? (b / MP) * C::NSB + (b % MP) * C::NSB / MP : b * BCH;
    static constexpr int Q1 = ALIGNED ? (b / MP) * C::NSB + (b % MP + 1) * C::NSB / MP
                                      : ((Q0 + BCH < C::NBRA) ? Q0 + BCH : C::NBRA);
    static constexpr int K0 = k * KCH, K1 = (K0 + KCH < C::NKETC) ? K0 + KCH : C::NKETC;
    static_assert(seg_rows<C>(Q0, Q1) <= kSegRowsMax, "bra chunk needs too many LDS rows");
  };
};

// sum over the 64 lanes of a full wave (result valid on every lane)
QFE_DI double wave_sum(double v) {
  v += __shfl_xor(v, 32);
  v += __shfl_xor(v, 16);
  v += __shfl_xor(v, 8);
  v += __shfl_xor(v, 4);
  v += __shfl_xor(v, 2);
  v += __shfl_xor(v, 1);
  return v;
}

// ------------------------------------------------------------ segment J/K kernel
// One block per segment: the tasks of one bra pair (a, b) within one class
// pair.  Every exchange update of a canonical quartet lands in a row of a or
// of b (K_ac, K_ad -> rows of a; K_bc, K_bd -> rows of b; the final
// K <- (K + K^T)/2 makes the placement immaterial), so the block keeps those
// rows in LDS, adds into them with LDS atomics, and flushes each non-zero
// entry once with one global atomic.  J_ab is reduced over the block; J_cd
// (one entry per task) stays a global atomic.  With PROW the density rows of
// a and b are staged in LDS beside the K rows.
template <class C, int Q0, int Q1>
struct SegRows {
  static constexpr int MA0 = Q0 / C::NSB, MA1 = (Q1 - 1) / C::NSB;
  static constexpr int NMA = MA1 - MA0 + 1;
  static constexpr int MB_LO = (NMA > 1) ? 0 : (Q0 % C::NSB);
  static constexpr int MB_HI = (NMA > 1) ? C::NSB - 1 : ((Q1 - 1) % C::NSB);
  static constexpr int NMB = MB_HI - MB_LO + 1;
  static constexpr int NR = NMA + NMB;  // LDS rows: a functions, then b functions
};

template <class C, int Q0, int Q1, typename R, bool PROW, int K0, int K1, int BLK>
__global__ __launch_bounds__(BLK) void k_eri_jk_seg(const int2* __restrict__ segs,
                                                    const long long* __restrict__ offs,
                                                    const int2* __restrict__ tasks,
                                                    const PairHdr* __restrict__ pairs,
                                                    const typename Vec4<R>::type* __restrict__ ppos,
                                                    const R* __restrict__ pk, RysArgsT<R> ra,
                                                    const double* __restrict__ P, double* __restrict__ J,
                                                    double* __restrict__ K, int nbf, double kscale,
                                                    double* __restrict__ jbuf, const long long* __restrict__ jslot,
                                                    const int2* __restrict__ qpos, int jfirst, int nsplit) {
  using S = SegRows<C, Q0, Q1>;
  extern __shared__ double lds[];
  // block = (segment, part): a long segment may be split over nsplit blocks,
  // each with its own LDS rows, when there are too few segments for the GPU
  const int2 sg = segs[blockIdx.x / nsplit];
  const int part = (int)(blockIdx.x % nsplit);
  const long long s0 = offs[sg.x], s1 = offs[sg.y];
  const long long t0 = s0 + (s1 - s0) * part / nsplit, t1 = s0 + (s1 - s0) * (part + 1) / nsplit;
  if (t0 >= t1) return;  // the same for every thread of the block
  const PairHdr bra = pairs[tasks[t0].x];
  constexpr int NQ = Q1 - Q0;
  const size_t n = (size_t)nbf;
  const int i0 = bra.bfa, j0 = bra.bfb;
  double* Ks = lds;
  const double* Pa = PROW ? lds + (size_t)S::NR * nbf : nullptr;
  __shared__ double sjab[NQ];  // J_ab of the block's bra functions
  if (threadIdx.x < NQ) sjab[threadIdx.x] = 0.0;
  for (int r = 0; r < S::NR; ++r) {
    const size_t ao = r < S::NMA ? (size_t)(i0 + S::MA0 + r) : (size_t)(j0 + S::MB_LO + (r - S::NMA));
    for (int c = threadIdx.x; c < nbf; c += blockDim.x) {
      Ks[(size_t)r * nbf + c] = 0.0;
      if constexpr (PROW) lds[(size_t)(S::NR + r) * nbf + c] = P[ao * n + c];
    }
  }
  __syncthreads();
  // density row of bra function (row r of the LDS layout)
  auto prow = [&](int r, int col) -> double {
    if constexpr (PROW) {
      return Pa[(size_t)r * nbf + col];
    } else {
      const size_t ao = r < S::NMA ? (size_t)(i0 + S::MA0 + r) : (size_t)(j0 + S::MB_LO + (r - S::NMA));
      return P[ao * n + col];
    }
  };
  double pij[NQ], jij[NQ];
  sfor<NQ>([&](auto qq) {
    constexpr int qi = Q0 + qq;
    pij[qq] = P[(size_t)(i0 + qi / C::NSB) * n + j0 + qi % C::NSB];
    jij[qq] = 0.0;
  });
  const double dbra = bra.sa == bra.sb ? 1.0 : 2.0;
  for (long long t = t0 + threadIdx.x; t < t1; t += blockDim.x) {
    const int2 tk = tasks[t];
    const PairHdr ket = pairs[tk.y];
    R acc[NQ][K1 - K0];
    contract_chunk<C, Q0, Q1, R, K0, K1>(bra, ket, ppos, pk, ra, acc);
    const double deg = dbra * (ket.sa == ket.sb ? 1.0 : 2.0) * (tk.x == tk.y ? 1.0 : 2.0);
    const int k0 = ket.bfa, l0 = ket.bfb;
    double Pkl[C::NKETS], jkl[C::NKETS];
    sfor<C::NSC>([&](auto c) {
      sfor<C::NSD>([&](auto d) {
        Pkl[c * C::NSD + d] = P[(size_t)(k0 + c) * n + l0 + d];
        jkl[c * C::NSD + d] = 0.0;
      });
    });
    double kac[S::NMA][C::NSC], kad[S::NMA][C::NSD], kbc[S::NMB][C::NSC], kbd[S::NMB][C::NSD];
    sfor<S::NMA>([&](auto a) {
      sfor<C::NSC>([&](auto c) { kac[a][c] = 0.0; });
      sfor<C::NSD>([&](auto d) { kad[a][d] = 0.0; });
    });
    sfor<S::NMB>([&](auto b) {
      sfor<C::NSC>([&](auto c) { kbc[b][c] = 0.0; });
      sfor<C::NSD>([&](auto d) { kbd[b][d] = 0.0; });
    });
    sfor<NQ>([&](auto qq) {
      constexpr int qi = Q0 + qq;
      constexpr int ia = qi / C::NSB - S::MA0, ib = qi % C::NSB - S::MB_LO;
      double v[C::NKETS];
      ket_to_sph<C, R, K0, K1>(acc[qq], v);
      sfor<C::NSC>([&](auto c) {
        const double Pac = prow(ia, k0 + c), Pbc = prow(S::NMA + ib, k0 + c);
        sfor<C::NSD>([&](auto d) {
          const double x = v[c * C::NSD + d] * deg;
          const double Pbd = prow(S::NMA + ib, l0 + d), Pad = prow(ia, l0 + d);
          jij[qq] = fma(Pkl[c * C::NSD + d], x, jij[qq]);
          jkl[c * C::NSD + d] = fma(pij[qq], x, jkl[c * C::NSD + d]);
          kac[ia][c] = fma(Pbd, x, kac[ia][c]);
          kbc[ib][c] = fma(Pad, x, kbc[ib][c]);
          kad[ia][d] = fma(Pbc, x, kad[ia][d]);
          kbd[ib][d] = fma(Pac, x, kbd[ib][d]);
        });
      });
    });
    if (jbuf) {
      // J_cd share of this task in its item's slots, summed per ket pair
      // after the class kernels
      const int2 qp = qpos[tk.y];
      double* jt = jbuf + jslot[sg.x + qp.x] + (long long)qp.y * C::NKETS;
      if (jfirst) {
        sfor<C::NKETS>([&](auto f) { jt[f] = jkl[f]; });
      } else {
        sfor<C::NKETS>([&](auto f) { jt[f] += jkl[f]; });
      }
    } else {
      sfor<C::NSC>([&](auto c) {
        sfor<C::NSD>([&](auto d) { unsafeAtomicAdd(&J[(size_t)(k0 + c) * n + l0 + d], jkl[c * C::NSD + d]); });
      });
    }
    if (kscale != 0.0) {
      sfor<S::NMA>([&](auto a) {
        double* row = Ks + (size_t)a * nbf;
        sfor<C::NSC>([&](auto c) { atomicAdd(&row[k0 + c], kac[a][c]); });
        sfor<C::NSD>([&](auto d) { atomicAdd(&row[l0 + d], kad[a][d]); });
      });
      sfor<S::NMB>([&](auto b) {
        double* row = Ks + (size_t)(S::NMA + b) * nbf;
        sfor<C::NSC>([&](auto c) { atomicAdd(&row[k0 + c], kbc[b][c]); });
        sfor<C::NSD>([&](auto d) { atomicAdd(&row[l0 + d], kbd[b][d]); });
      });
    }
  }
  // J_ab: every thread of the block holds partial sums of the same entries;
  // waves reduce in registers, then through LDS to one global atomic each
  const int lane = (int)(threadIdx.x & 63);
  sfor<NQ>([&](auto qq) {
    const double tot = wave_sum(jij[qq]);
    if (lane == 0 && tot != 0.0) atomicAdd(&sjab[qq], tot);
  });
  __syncthreads();
  if (threadIdx.x < NQ) {
    const int qi = Q0 + (int)threadIdx.x;
    const double v = sjab[threadIdx.x];
    if (v != 0.0) unsafeAtomicAdd(&J[(size_t)(i0 + qi / C::NSB) * n + j0 + qi % C::NSB], v);
  }
  if (kscale != 0.0) {
    const double kq = 0.25 * kscale;
    for (int r = 0; r < S::NR; ++r) {
      const size_t ao = r < S::NMA ? (size_t)(i0 + S::MA0 + r) : (size_t)(j0 + S::MB_LO + (r - S::NMA));
      for (int c = threadIdx.x; c < nbf; c += blockDim.x) {
        const double v = Ks[(size_t)r * nbf + c];
        if (v != 0.0) unsafeAtomicAdd(&K[ao * n + c], kq * v);
      }
    }
  }
}

// ------------------------------------------------------------ Schwarz kernel
// q2[pair] += sum of squares of this chunk of (ab|ab)
template <class C, int Q0, int Q1>
__global__ __launch_bounds__(256) void k_schwarz(const int* __restrict__ plist, int npair,
                                                 const PairHdr* __restrict__ pairs,
                                                 const double4* __restrict__ ppos,
                                                 const double* __restrict__ pk, RysArgs ra,
                                                 double* __restrict__ q2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npair) return;
  const int p = plist[t];
  const PairHdr pr = pairs[p];
  double acc[Q1 - Q0][C::NKETC];
  contract_chunk<C, Q0, Q1, double>(pr, pr, ppos, pk, ra, acc);
  double s = 0.0;
  sfor<Q1 - Q0>([&](auto qq) {
    double v[C::NKETS];
    ket_to_sph<C, double>(acc[qq], v);
    sfor<C::NKETS>([&](auto x) { s = fma(v[x], v[x], s); });
  });
  unsafeAtomicAdd(&q2[p], s);
}

// Schwarz of classes with f shells, in two steps: each chunk kernel writes its
// share (bra functions x ket cartesians) of the (ab|ab) block to blk, then the
// fold transforms every bra function's ket row to output functions and adds
// its sum of squares to q2 (the square needs the whole row, so it cannot be
// taken per chunk).
template <class C, int Q0, int Q1, int K0, int K1>
__global__ __launch_bounds__(256) void k_schwarz_part(const int* __restrict__ plist, int npair,
                                                      const PairHdr* __restrict__ pairs,
                                                      const double4* __restrict__ ppos,
                                                      const double* __restrict__ pk, RysArgs ra,
                                                      double* __restrict__ blk) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= npair) return;
  const PairHdr pr = pairs[plist[t]];
  double acc[Q1 - Q0][K1 - K0];
  contract_chunk<C, Q0, Q1, double, K0, K1>(pr, pr, ppos, pk, ra, acc);
  sfor<Q1 - Q0>([&](auto qq) {
    double* row = blk + ((size_t)t * C::NBRA + Q0 + qq) * C::NKETC + K0;
    sfor<K1 - K0>([&](auto kk) { row[kk] = acc[qq][kk]; });
  });
}

template <class C>
__global__ __launch_bounds__(256) void k_schwarz_fold(const int* __restrict__ plist, int npair,
                                                      const double* __restrict__ blk, double* __restrict__ q2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)npair * C::NBRA) return;
  const double* row = blk + (size_t)i * C::NKETC;
  double s = 0.0;
  sfor<C::NSC>([&](auto mc) {
    sfor<C::NSD>([&](auto md) {
      double v = 0.0;
      sfor<C::NKETC>([&](auto k) {
        constexpr double c = sph_coef(C::LC, C::PURE, mc, k / C::NCD) * sph_coef(C::LD, C::PURE, md, k % C::NCD);
        if constexpr (c != 0.0) v = fma(c, row[k], v);
      });
      s = fma(v, v, s);
    });
  });
  unsafeAtomicAdd(&q2[plist[i / C::NBRA]], s);
}

// ------------------------------------------------------------ launchers
template <int LA, int LB, int LC, int LD, bool PURE>
struct ClassOps {
  using C = Cls<LA, LB, LC, LD, PURE>;
  using K = JKChunks<C>;
  static constexpr int NCH = K::NCH;
  // threads per segment block: the LDS rows of one bra pair are shared by
  // all of them, so light classes use large blocks to keep enough waves per
  // CU resident (the register budget per thread shrinks with the block)
  static constexpr int SEG_BLK =
      K::F ? 256 : ((C::L <= 2 && C::NBRA <= 5) ? 1024 : (C::L <= 3 ? 512 : 256));
  // doubles of the Schwarz staging block per pair (classes with f shells)
  static constexpr int SCHWARZ_BLOCK = K::F ? C::NBRA * C::NKETC : 0;

  // dynamic LDS bytes of the segment kernel (K rows, and P rows with prow),
  // plus its static J_ab slots when checked against the LDS limit
  template <int Q0, int Q1>
  static size_t seg_bytes(int nbf, bool prow) {
    return (size_t)SegRows<C, Q0, Q1>::NR * nbf * sizeof(double) * (prow ? 2 : 1);
  }
  template <int Q0, int Q1>
  static size_t seg_total(int nbf, bool prow) {
    return seg_bytes<Q0, Q1>(nbf, prow) + (size_t)(Q1 - Q0) * sizeof(double);
  }

  template <int Q0, int Q1, int K0, int K1, typename R, bool PROW>
  static bool launch_seg(const JKLaunch& a, hipStream_t s, size_t bytes) {
    auto kern = k_eri_jk_seg<C, Q0, Q1, R, PROW, K0, K1, SEG_BLK>;
    if (bytes > 65536 &&
        hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    const int jfirst = (Q0 == 0 && K0 == 0) ? 1 : 0;  // chunk 0 of the class runs first
    // split segments when there are fewer than ~2048 blocks and the average
    // segment gives every thread several tasks
    const long long per = a.ntask / std::max(a.nseg, 1);
    const long long want = (2048 + a.nseg - 1) / std::max(a.nseg, 1);
    const int nsplit = (int)std::max(1LL, std::min({want, per / (2 * SEG_BLK), 16LL}));
    const dim3 grid((unsigned)((long long)a.nseg * nsplit));
    if constexpr (sizeof(R) == 4)
      hipLaunchKernelGGL(kern, grid, dim3(SEG_BLK), bytes, s, a.segs, a.offs, a.tasks_all, a.pairs, a.ppos32, a.pk32,
                         *a.ra32, a.P, a.J, a.K, a.nbf, a.kscale, a.jbuf, a.jslot, a.qpos, jfirst, nsplit);
    else
      hipLaunchKernelGGL(kern, grid, dim3(SEG_BLK), bytes, s, a.segs, a.offs, a.tasks_all, a.pairs, a.ppos, a.pk,
                         *a.ra, a.P, a.J, a.K, a.nbf, a.kscale, a.jbuf, a.jslot, a.qpos, jfirst,
                         nsplit);
    return true;
  }

  // segment kernel with K and P rows in LDS when they fit, else K rows only
  // (P from L2); a negative lds_max forces the K-only variant
  // waves of this kernel one CU can hold when a block needs `bytes` of LDS
  // (the CU's LDS equals the per-block maximum on gfx950; 32 wave slots)
  static int waves_per_cu(size_t bytes, size_t lds) {
    const size_t blocks_lds = bytes ? lds / bytes : 32;
    const size_t blocks_slots = (size_t)32 / (SEG_BLK / 64);
    return (int)std::min(blocks_lds, blocks_slots) * (SEG_BLK / 64);
  }

  template <int Q0, int Q1, int K0, int K1, typename R, bool BOTH>
  static void seg(const JKLaunch& a, hipStream_t s) {
    static_assert(SegRows<C, Q0, Q1>::NR <= kSegRowsMax, "chunk needs more LDS rows than planned");
    if (a.nseg <= 0 || a.segs == nullptr) throw std::runtime_error("segment list missing");
    const size_t lim = (size_t)(a.lds_max < 0 ? -a.lds_max : a.lds_max);
    const size_t with_p = seg_bytes<Q0, Q1>(a.nbf, true), k_only = seg_bytes<Q0, Q1>(a.nbf, false);
    const size_t tot_p = seg_total<Q0, Q1>(a.nbf, true), tot_k = seg_total<Q0, Q1>(a.nbf, false);
    // density rows in LDS unless that leaves fewer than seg_min_waves waves
    // per CU and dropping them lets more reside
    const int wp = waves_per_cu(tot_p, lim), wk = waves_per_cu(tot_k, lim);
    const bool prow = a.lds_max > 0 && tot_p <= lim && (!BOTH || wp >= wk || wp >= a.seg_min_waves);
    if (prow && launch_seg<Q0, Q1, K0, K1, R, true>(a, s, with_p)) return;
    if constexpr (BOTH)
      if (tot_k <= lim && launch_seg<Q0, Q1, K0, K1, R, false>(a, s, k_only)) return;
    throw std::runtime_error("the exchange rows of one bra pair exceed the LDS of a block (nbf too large)");
  }

  template <int C0, int C1, bool LAST>
  static void jk_range(const JKLaunch& a, hipStream_t s) {
    static_assert(C0 < C1 && C1 <= NCH && (!LAST || C1 == NCH), "chunk range does not match the class");
    static_assert(!K::F, "classes with f shells use jk_seg_only_range");
    if (a.ntask <= 0) return;
    hsfor<C1 - C0>([&](auto cc) {
      using A = typename K::template At<C0 + cc>;
      if (a.fp32)
        seg<A::Q0, A::Q1, A::K0, A::K1, float, true>(a, s);
      else
        seg<A::Q0, A::Q1, A::K0, A::K1, double, true>(a, s);
    });
  }

  // FP64 segment kernels only (FP32 tasks are evaluated in FP64): the variant
  // compiled for classes with f shells, where the per-task kernel would double
  // the compile time
  template <int C0, int C1, bool LAST>
  static void jk_seg_only_range(const JKLaunch& a, hipStream_t s) {
    static_assert(C0 < C1 && C1 <= NCH && (!LAST || C1 == NCH), "chunk range does not match the class");
    if (a.ntask <= 0) return;
    hsfor<C1 - C0>([&](auto cc) {
      using A = typename K::template At<C0 + cc>;
      seg<A::Q0, A::Q1, A::K0, A::K1, double, false>(a, s);
    });
  }

  template <int C0, int C1, bool LAST>
  static void schwarz_range(const SchwarzLaunch& a, hipStream_t s) {
    static_assert(C0 < C1 && C1 <= NCH && (!LAST || C1 == NCH), "chunk range does not match the class");
    if (a.npair <= 0) return;
    if (K::F && a.blk == nullptr) throw std::runtime_error("Schwarz staging block missing for an f-shell class");
    const int block = 256;
    const int grid = (a.npair + block - 1) / block;
    hsfor<C1 - C0>([&](auto cc) {
      using A = typename K::template At<C0 + cc>;
      if constexpr (K::F)
        hipLaunchKernelGGL((k_schwarz_part<C, A::Q0, A::Q1, A::K0, A::K1>), dim3(grid), dim3(block), 0, s, a.plist,
                           a.npair, a.pairs, a.ppos, a.pk, *a.ra, a.blk);
      else
        hipLaunchKernelGGL((k_schwarz<C, A::Q0, A::Q1>), dim3(grid), dim3(block), 0, s, a.plist, a.npair, a.pairs,
                           a.ppos, a.pk, *a.ra, a.q2);
    });
  }

  // after every chunk range: fold the staged blocks (classes with f shells)
  static void schwarz_fold(const SchwarzLaunch& a, hipStream_t s) {
    if constexpr (K::F) {
      if (a.npair <= 0) return;
      const long long n = (long long)a.npair * C::NBRA;
      hipLaunchKernelGGL((k_schwarz_fold<C>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, a.plist, a.npair,
                         a.blk, a.q2);
    }
  }
};

}  // namespace qfe
