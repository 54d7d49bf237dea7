// Per-class ERI kernels (templates), instantiated one class per translation
// unit by the build (gen/cls_*.hip) to keep compile times bounded.
//   k_eri_jk   Rys ERIs of one class (one bra chunk) + J/K digestion
//              (libint2_direct.cpp:414-567)
//   k_schwarz  sum of squares of (ab|ab) for Q_ab = sqrt ||(ab|ab)||_F
//              (schwarz.cpp:26-78)
#pragma once

#include <hip/hip_runtime.h>

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

constexpr int kAccMax = 48;  // register budget (doubles) for one thread's accumulators
template <class C>
constexpr int chunk_size() {
  int ch = kAccMax / C::NKETC;
  if (ch < 1) ch = 1;
  if (ch > C::NBRA) ch = C::NBRA;
  return ch;
}

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

// ------------------------------------------------------------ J/K kernel
template <class C, int Q0, int Q1, typename R>
__global__ __launch_bounds__(256) void k_eri_jk(const int2* __restrict__ tasks, long long ntask,
                                                const PairHdr* __restrict__ pairs,
                                                const typename Vec4<R>::type* __restrict__ ppos,
                                                const R* __restrict__ pk, RysArgsT<R> ra,
                                                const double* __restrict__ P, double* __restrict__ J,
                                                double* __restrict__ K, int nbf, double kscale) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntask) return;
  const int2 tk = tasks[t];
  const PairHdr bra = pairs[tk.x];
  const PairHdr ket = pairs[tk.y];
  constexpr int NQ = Q1 - Q0;
  R acc[NQ][C::NKETC];
  contract_chunk<C, Q0, Q1, R>(bra, ket, ppos, pk, ra, acc);

  const double deg = (bra.sa == bra.sb ? 1.0 : 2.0) * (ket.sa == ket.sb ? 1.0 : 2.0) *
                     (tk.x == tk.y ? 1.0 : 2.0);
  // is the bra pair the same on every active lane of this wave?
  const int lane = (int)(threadIdx.x & 63);
  const unsigned long long active = __ballot(1);
  const int bra0 = __shfl(tk.x, __ffsll((long long)active) - 1);
  const bool uniform_bra = (active == ~0ull) && (__ballot(tk.x != bra0) == 0ull);
  const int i0 = bra.bfa, j0 = bra.bfb, k0 = ket.bfa, l0 = ket.bfb;
  const size_t n = (size_t)nbf;
  constexpr int MA0 = Q0 / C::NSB, MA1 = (Q1 - 1) / C::NSB;
  constexpr int NMA = MA1 - MA0 + 1;
  // density blocks needed by the digestion
  double Pkl[C::NKETS];
  sfor<C::NSC>([&](auto c) { sfor<C::NSD>([&](auto d) { Pkl[c * C::NSD + d] = P[(k0 + c) * n + l0 + d]; }); });
  double jkl[C::NKETS];
  sfor<C::NKETS>([&](auto x) { jkl[x] = 0.0; });
  double kik[NMA][C::NSC], kil[NMA][C::NSD], kjk[C::NSB][C::NSC], kjl[C::NSB][C::NSD];
  sfor<NMA>([&](auto a) {
    sfor<C::NSC>([&](auto c) { kik[a][c] = 0.0; });
    sfor<C::NSD>([&](auto d) { kil[a][d] = 0.0; });
  });
  sfor<C::NSB>([&](auto b) {
    sfor<C::NSC>([&](auto c) { kjk[b][c] = 0.0; });
    sfor<C::NSD>([&](auto d) { kjl[b][d] = 0.0; });
  });
  const double kq = 0.25 * kscale;
  sfor<NQ>([&](auto qq) {
    constexpr int qi = Q0 + qq;
    constexpr int ma = qi / C::NSB, mb = qi % C::NSB;
    constexpr int ia = ma - MA0;
    const size_t ri = (size_t)(i0 + ma) * n, rj = (size_t)(j0 + mb) * n;
    double v[C::NKETS];
    ket_to_sph<C, R>(acc[qq], v);
    const double Pij = P[ri + j0 + mb];
    double jij = 0.0;
    sfor<C::NSC>([&](auto c) {
      const double Pik = P[ri + k0 + c], Pjk = P[rj + k0 + c];
      sfor<C::NSD>([&](auto d) {
        const double x = v[c * C::NSD + d] * deg;
        const double Pjl = P[rj + l0 + d], Pil = P[ri + l0 + d];
        jij = fma(Pkl[c * C::NSD + d], x, jij);
        jkl[c * C::NSD + d] = fma(Pij, x, jkl[c * C::NSD + d]);
        kik[ia][c] = fma(Pjl, x, kik[ia][c]);
        kjk[mb][c] = fma(Pil, x, kjk[mb][c]);
        kil[ia][d] = fma(Pjk, x, kil[ia][d]);
        kjl[mb][d] = fma(Pik, x, kjl[mb][d]);
      });
    });
    if (uniform_bra) {
      // every lane holds the same (i, j): one atomic per wave
      const double tot = wave_sum(jij);
      if (lane == 0) unsafeAtomicAdd(&J[ri + j0 + mb], tot);
    } else {
      unsafeAtomicAdd(&J[ri + j0 + mb], jij);
    }
  });
  sfor<C::NSC>([&](auto c) {
    sfor<C::NSD>([&](auto d) { unsafeAtomicAdd(&J[(k0 + c) * n + l0 + d], jkl[c * C::NSD + d]); });
  });
  if (kscale != 0.0) {
    sfor<NMA>([&](auto a) {
      const size_t ri = (size_t)(i0 + MA0 + a) * n;
      sfor<C::NSC>([&](auto c) { unsafeAtomicAdd(&K[ri + k0 + c], kq * kik[a][c]); });
      sfor<C::NSD>([&](auto d) { unsafeAtomicAdd(&K[ri + l0 + d], kq * kil[a][d]); });
    });
    // bra-b rows touched by this chunk
    constexpr int MB_LO = (NMA > 1) ? 0 : (Q0 % C::NSB);
    constexpr int MB_HI = (NMA > 1) ? C::NSB - 1 : ((Q1 - 1) % C::NSB);
    sfor<MB_HI - MB_LO + 1>([&](auto bb) {
      constexpr int b = MB_LO + bb;
      const size_t rj = (size_t)(j0 + b) * n;
      sfor<C::NSC>([&](auto c) { unsafeAtomicAdd(&K[rj + k0 + c], kq * kjk[b][c]); });
      sfor<C::NSD>([&](auto d) { unsafeAtomicAdd(&K[rj + l0 + d], kq * kjl[b][d]); });
    });
  }
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

template <class C, int Q0, int Q1, typename R, bool PROW>
__global__ __launch_bounds__(256) void k_eri_jk_seg(const int2* __restrict__ segs,
                                                    const long long* __restrict__ offs,
                                                    const int2* __restrict__ tasks,
                                                    const PairHdr* __restrict__ pairs,
                                                    const typename Vec4<R>::type* __restrict__ ppos,
                                                    const R* __restrict__ pk, RysArgsT<R> ra,
                                                    const double* __restrict__ P, double* __restrict__ J,
                                                    double* __restrict__ K, int nbf, double kscale) {
  using S = SegRows<C, Q0, Q1>;
  extern __shared__ double lds[];
  const int2 sg = segs[blockIdx.x];
  const long long t0 = offs[sg.x], t1 = offs[sg.y];
  if (t0 >= t1) return;  // the same for every thread of the block
  const PairHdr bra = pairs[tasks[t0].x];
  constexpr int NQ = Q1 - Q0;
  const size_t n = (size_t)nbf;
  const int i0 = bra.bfa, j0 = bra.bfb;
  double* Ks = lds;
  const double* Pa = PROW ? lds + (size_t)S::NR * nbf : nullptr;
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
    R acc[NQ][C::NKETC];
    contract_chunk<C, Q0, Q1, R>(bra, ket, ppos, pk, ra, acc);
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
      ket_to_sph<C, R>(acc[qq], v);
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
    sfor<C::NSC>([&](auto c) {
      sfor<C::NSD>([&](auto d) { unsafeAtomicAdd(&J[(size_t)(k0 + c) * n + l0 + d], jkl[c * C::NSD + d]); });
    });
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
  // J_ab: every thread of the block holds partial sums of the same entries
  const int lane = (int)(threadIdx.x & 63);
  sfor<NQ>([&](auto qq) {
    constexpr int qi = Q0 + qq;
    const double tot = wave_sum(jij[qq]);
    if (lane == 0 && tot != 0.0) unsafeAtomicAdd(&J[(size_t)(i0 + qi / C::NSB) * n + j0 + qi % C::NSB], tot);
  });
  if (kscale != 0.0) {
    __syncthreads();
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

// ------------------------------------------------------------ launchers
template <int LA, int LB, int LC, int LD, bool PURE>
struct ClassOps {
  using C = Cls<LA, LB, LC, LD, PURE>;
  static constexpr int CH = chunk_size<C>();
  static constexpr int NCH = (C::NBRA + CH - 1) / CH;

  // LDS bytes of the segment kernel for chunk c (K rows, and P rows with PROW)
  template <int Q0, int Q1>
  static size_t seg_bytes(int nbf, bool prow) {
    return (size_t)SegRows<C, Q0, Q1>::NR * nbf * sizeof(double) * (prow ? 2 : 1);
  }

  template <int Q0, int Q1, typename R, bool PROW>
  static bool launch_seg(const JKLaunch& a, hipStream_t s, size_t bytes) {
    auto kern = k_eri_jk_seg<C, Q0, Q1, R, PROW>;
    if (bytes > 65536 &&
        hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes) != hipSuccess) {
      (void)hipGetLastError();
      return false;
    }
    if constexpr (sizeof(R) == 4)
      hipLaunchKernelGGL(kern, dim3((unsigned)a.nseg), dim3(256), bytes, s, a.segs, a.offs, a.tasks_all, a.pairs,
                         a.ppos32, a.pk32, *a.ra32, a.P, a.J, a.K, a.nbf, a.kscale);
    else
      hipLaunchKernelGGL(kern, dim3((unsigned)a.nseg), dim3(256), bytes, s, a.segs, a.offs, a.tasks_all, a.pairs,
                         a.ppos, a.pk, *a.ra, a.P, a.J, a.K, a.nbf, a.kscale);
    return true;
  }

  // segment kernel when its K and P rows fit in LDS (else the per-task kernel)
  template <int Q0, int Q1, typename R>
  static bool try_seg(const JKLaunch& a, hipStream_t s) {
    if (a.nseg <= 0 || a.lds_max <= 0 || a.segs == nullptr) return false;
    const size_t bytes = seg_bytes<Q0, Q1>(a.nbf, true);
    return bytes <= (size_t)a.lds_max && launch_seg<Q0, Q1, R, true>(a, s, bytes);
  }

  // Bra-function chunks [C0, C1) of the class; the build may split the chunks
  // of a heavy class over several translation units (LAST marks the range
  // that must end at NCH, so a mismatch with the build fails to compile).
  template <int C0, int C1, bool LAST>
  static void jk_range(const JKLaunch& a, hipStream_t s) {
    static_assert(C0 < C1 && C1 <= NCH && (!LAST || C1 == NCH), "chunk range does not match the class");
    if (a.ntask <= 0) return;
    const int block = 256;
    const long long grid = (a.ntask + block - 1) / block;
    hsfor<C1 - C0>([&](auto cc) {
      constexpr int q0 = (C0 + cc) * CH;
      constexpr int q1 = (q0 + CH < C::NBRA) ? q0 + CH : C::NBRA;
      if (a.fp32 ? try_seg<q0, q1, float>(a, s) : try_seg<q0, q1, double>(a, s)) return;
      if (a.fp32)
        hipLaunchKernelGGL((k_eri_jk<C, q0, q1, float>), dim3((unsigned)grid), dim3(block), 0, s, a.tasks,
                           a.ntask, a.pairs, a.ppos32, a.pk32, *a.ra32, a.P, a.J, a.K, a.nbf, a.kscale);
      else
        hipLaunchKernelGGL((k_eri_jk<C, q0, q1, double>), dim3((unsigned)grid), dim3(block), 0, s, a.tasks,
                           a.ntask, a.pairs, a.ppos, a.pk, *a.ra, a.P, a.J, a.K, a.nbf, a.kscale);
    });
  }

  // FP64 segment kernels only (FP32 tasks are evaluated in FP64): the variant
  // compiled for classes with f shells, where every chunk holds one bra
  // function and the per-task kernel would double the compile time
  template <int C0, int C1, bool LAST>
  static void jk_seg_only_range(const JKLaunch& a, hipStream_t s) {
    static_assert(C0 < C1 && C1 <= NCH && (!LAST || C1 == NCH), "chunk range does not match the class");
    if (a.ntask <= 0) return;
    hsfor<C1 - C0>([&](auto cc) {
      constexpr int q0 = (C0 + cc) * CH;
      constexpr int q1 = (q0 + CH < C::NBRA) ? q0 + CH : C::NBRA;
      if (!try_seg<q0, q1, double>(a, s))
        throw std::runtime_error("exchange rows of this f-shell class exceed the LDS of one block");
    });
  }

  template <int C0, int C1, bool LAST>
  static void schwarz_range(const SchwarzLaunch& a, hipStream_t s) {
    static_assert(C0 < C1 && C1 <= NCH && (!LAST || C1 == NCH), "chunk range does not match the class");
    if (a.npair <= 0) return;
    const int block = 256;
    const int grid = (a.npair + block - 1) / block;
    hsfor<C1 - C0>([&](auto cc) {
      constexpr int q0 = (C0 + cc) * CH;
      constexpr int q1 = (q0 + CH < C::NBRA) ? q0 + CH : C::NBRA;
      hipLaunchKernelGGL((k_schwarz<C, q0, q1>), dim3(grid), dim3(block), 0, s, a.plist, a.npair,
                         a.pairs, a.ppos, a.pk, *a.ra, a.q2);
    });
  }
};

}  // namespace qfe
