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
#include "fixed_point.hpp"
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
#ifdef QFE_JK_ACC_F
  return QFE_JK_ACC_F;
#else
  if (C::LA <= 2 && C::LB <= 2 && C::LC <= 2 && C::LD <= 2) return kAccMax;
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
// (one entry per task) goes to the task's share slot in jbuf, or is a global
// atomic when jbuf is null.  With PROW the density rows of a and b are staged
// in LDS beside the K rows.  A null J skips the Coulomb digestion and
// kscale = 0 the exchange digestion (build_JK with J or K null,
// eri_multiplexer.cpp:66-67).
template <class C, int Q0, int Q1>
struct SegRows {
  static constexpr int MA0 = Q0 / C::NSB, MA1 = (Q1 - 1) / C::NSB;
  static constexpr int NMA = MA1 - MA0 + 1;
  static constexpr int MB_LO = (NMA > 1) ? 0 : (Q0 % C::NSB);
  static constexpr int MB_HI = (NMA > 1) ? C::NSB - 1 : ((Q1 - 1) % C::NSB);
  static constexpr int NMB = MB_HI - MB_LO + 1;
  static constexpr int NR = NMA + NMB;  // LDS rows: a functions, then b functions
};

// The ket pair of a task from its 16-byte record {prim0, nprim, sa, sb} and
// the shell table {x, y, z, first AO} (nsh entries, small enough to stay in
// L2): 16 bytes per task from the pair table instead of the 80-byte header
QFE_DI PairHdr ket_record(const int4* __restrict__ kets, const double4* __restrict__ shells, int p) {
  const int4 r = kets[p];
  const double4 c = shells[r.z], d = shells[r.w];
  PairHdr h;
  h.A[0] = c.x;
  h.A[1] = c.y;
  h.A[2] = c.z;
  h.B[0] = d.x;
  h.B[1] = d.y;
  h.B[2] = d.z;
  h.prim0 = r.x;
  h.nprim = r.y;
  h.sa = r.z;
  h.sb = r.w;
  h.bfa = (int)c.w;
  h.bfb = (int)d.w;
  h.cls = 0;
  h.pad = 0;
  return h;
}

// P[e] for an element index e < 2^29 (nbf < 23170, checked on the host): a
// 32-bit byte offset from the uniform base, which the load can take in its
// scalar-base form (no 64-bit address per lane)
QFE_DI double pload(const double* __restrict__ P, unsigned e) {
  return *reinterpret_cast<const double*>(reinterpret_cast<const char*>(P) + (e << 3));
}

// The one-root classes stage the Boys rows they read (F_0 .. F_7 at every
// grid point, 8 values a row) in LDS after the exchange and density rows:
// their kernels wait on these per-lane gathers more than on anything else.
constexpr int kLdsBoysW = 8;
template <class C>
constexpr bool lds_boys() {
  return C::LA <= 2 && C::LB <= 2 && C::LC <= 2 && C::LD <= 2 && C::NR == 1;
}

template <class C, int Q0, int Q1, typename R, bool PROW, int K0, int K1, int BLK>
__global__ __launch_bounds__(BLK) void k_eri_jk_seg(const int2* __restrict__ segs,
                                                    const long long* __restrict__ offs,
                                                    const int2* __restrict__ tasks,
                                                    const PairHdr* __restrict__ pairs,
                                                    const typename Vec4<R>::type* __restrict__ ppos,
                                                    const R* __restrict__ pk, const R* __restrict__ sig,
                                                    RysArgsT<R> ra,
                                                    const double* __restrict__ P, double* __restrict__ J,
                                                    double* __restrict__ K, int nbf, double kscale,
                                                    double* __restrict__ jbuf, const long long* __restrict__ jslot,
                                                    const int2* __restrict__ qpos, int jfirst, int nsplit,
                                                    SegExtra ex) {
  using S = SegRows<C, Q0, Q1>;
  constexpr bool FCLS = C::LA > 2 || C::LB > 2 || C::LC > 2 || C::LD > 2;  // a class with an f shell
  extern __shared__ double lds[];
  // block = (segment, part, column tile): a long segment may be split over
  // nsplit blocks, each with its own LDS rows, when there are too few
  // segments for the GPU; the exchange rows may be split into column tiles
  // when a whole row does not fit the LDS
  const int tile = (int)(blockIdx.x % (unsigned)ex.ntile);
  const int rest = (int)(blockIdx.x / (unsigned)ex.ntile);
  const int2 sg = segs[rest / nsplit];
  const int part = rest % nsplit;
  const long long s0 = offs[sg.x], s1 = offs[sg.y];
  const long long t0 = s0 + (s1 - s0) * part / nsplit, t1 = s0 + (s1 - s0) * (part + 1) / nsplit;
  if (t0 >= t1) return;  // the same for every thread of the block
  const PairHdr bra = pairs[tasks[t0].x];
  constexpr int NQ = Q1 - Q0;
  const size_t n = (size_t)nbf;
  const int i0 = bra.bfa, j0 = bra.bfb;
  // this block's exchange columns: column tiles only in the FP64 kernel with
  // K rows only (FP32 tasks that need tiles are evaluated by it)
  constexpr bool TILED = !PROW && sizeof(R) == 8;
  const int c0 = TILED ? ex.tile_lo[tile] : 0, W = TILED ? ex.tile_hi[tile] - c0 : nbf;
  const bool det = ex.det != 0;
  const bool doJ = det ? ex.Jx != nullptr : J != nullptr;
  const bool doK = kscale != 0.0;
  // exchange rows: doubles, or (hi, lo) fixed-point pairs in deterministic builds
  const size_t krow = det ? 2 * (size_t)W : (size_t)W;
  double* Ks = lds;
  unsigned long long* Kf = reinterpret_cast<unsigned long long*>(lds);
  const double* Pa = PROW ? lds + (size_t)S::NR * krow : nullptr;
  __shared__ unsigned long long sjab[2 * NQ];  // J_ab of the block's bra functions
  for (int i = threadIdx.x; i < 2 * NQ; i += blockDim.x) sjab[i] = 0ull;
  constexpr bool LBOYS = lds_boys<C>();
  R* btab = nullptr;
  if constexpr (LBOYS) {
    btab = reinterpret_cast<R*>(lds + (size_t)S::NR * krow + (PROW ? (size_t)S::NR * W : 0));
    for (int i = threadIdx.x; i < BOYS_NG * kLdsBoysW; i += blockDim.x)
      btab[i] = ra.boys[(size_t)(i / kLdsBoysW) * BOYS_W + i % kLdsBoysW];
  }
  for (int r = 0; r < S::NR; ++r) {
    const size_t ao = r < S::NMA ? (size_t)(i0 + S::MA0 + r) : (size_t)(j0 + S::MB_LO + (r - S::NMA));
    for (int c = threadIdx.x; c < (int)krow; c += blockDim.x) Ks[(size_t)r * krow + c] = 0.0;
    if constexpr (PROW)
      for (int c = threadIdx.x; c < W; c += blockDim.x) lds[(size_t)S::NR * krow + (size_t)r * W + c] = P[ao * n + c0 + c];
  }
  __syncthreads();
  // density row of bra function (row r of the LDS layout)
  auto prow = [&](int r, int col) -> double {
    if constexpr (PROW) return Pa[(size_t)r * W + col];
    const unsigned ao = r < S::NMA ? (unsigned)(i0 + S::MA0 + r) : (unsigned)(j0 + S::MB_LO + (r - S::NMA));
    return pload(P, ao * (unsigned)nbf + (unsigned)col);
  };
  int bad = 0;  // a fixed-point add out of range (reported after the last add)
  auto kadd = [&](int r, int col, double v) {
    if (det)
      fx_lds_add(Kf + (size_t)r * krow + 2 * (size_t)(col - c0), v, bad);
    else
      atomicAdd(&Ks[(size_t)r * krow + col - c0], v);
  };
  const R lnp = prim_ln_cut<R>(ex.pmax, ex.prim_screen);
  double pij[NQ], jij[NQ];
  sfor<NQ>([&](auto qq) {
    constexpr int qi = Q0 + qq;
    pij[qq] = P[(size_t)(i0 + qi / C::NSB) * n + j0 + qi % C::NSB];
    jij[qq] = 0.0;
  });
  const double dbra = bra.sa == bra.sb ? 1.0 : 2.0;
  for (long long t = t0 + threadIdx.x; t < t1; t += blockDim.x) {
    const int2 tk = tasks[t];
    // classes with an f shell are bound by their integrals, not by this
    // load: they read the full header (the compact record costs them registers)
    const PairHdr ket = FCLS ? pairs[tk.y] : ket_record(ex.kets, ex.shells, tk.y);
    const int k0 = ket.bfa, l0 = ket.bfb;
    // with column tiles a task belongs to the tiles of its ket shells: its
    // c-column updates and its Coulomb terms go with the tile of c, its
    // d-column updates with the tile of d
    const bool inC = !TILED || (unsigned)(k0 - c0) < (unsigned)W;
    const bool inD = !TILED || (unsigned)(l0 - c0) < (unsigned)W;
    if (!inC && !inD) continue;
    R acc[NQ][K1 - K0];
    // the wave-uniform bra loop, except in classes with an f shell (their
    // long primitive bodies hide the per-lane break's wait, and the uniform
    // loop costs them registers)
    contract_chunk<C, Q0, Q1, R, K0, K1, !FCLS, LBOYS ? kLdsBoysW : BOYS_W>(bra, ket, ppos, pk, ra, acc, lnp,
                                                                             sig, btab);
    const double deg = dbra * (ket.sa == ket.sb ? 1.0 : 2.0) * (tk.x == tk.y ? 1.0 : 2.0);
    const double degj = inC ? deg : 0.0;
    double Pkl[C::NKETS], jkl[C::NKETS];
    sfor<C::NSC>([&](auto c) {
      sfor<C::NSD>([&](auto d) {
        Pkl[c * C::NSD + d] = pload(P, (unsigned)(k0 + c) * (unsigned)nbf + (unsigned)(l0 + d));
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
          const double xj = v[c * C::NSD + d] * degj;
          const double Pbd = prow(S::NMA + ib, l0 + d), Pad = prow(ia, l0 + d);
          jij[qq] = fma(Pkl[c * C::NSD + d], xj, jij[qq]);
          jkl[c * C::NSD + d] = fma(pij[qq], x, jkl[c * C::NSD + d]);
          kac[ia][c] = fma(Pbd, x, kac[ia][c]);
          kbc[ib][c] = fma(Pad, x, kbc[ib][c]);
          kad[ia][d] = fma(Pbc, x, kad[ia][d]);
          kbd[ib][d] = fma(Pac, x, kbd[ib][d]);
        });
      });
    });
    if (doJ && inC) {
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
          sfor<C::NSD>([&](auto d) {
            const size_t e = (size_t)(k0 + c) * n + l0 + d;
            if (det)
              fx_global_add(ex.Jx + 2 * e, jkl[c * C::NSD + d], bad);
            else
              unsafeAtomicAdd(&J[e], jkl[c * C::NSD + d]);
          });
        });
      }
    }
    if (doK) {
      sfor<S::NMA>([&](auto a) {
        if (inC) sfor<C::NSC>([&](auto c) { kadd(a, k0 + c, kac[a][c]); });
        if (inD) sfor<C::NSD>([&](auto d) { kadd(a, l0 + d, kad[a][d]); });
      });
      sfor<S::NMB>([&](auto b) {
        if (inC) sfor<C::NSC>([&](auto c) { kadd(S::NMA + b, k0 + c, kbc[b][c]); });
        if (inD) sfor<C::NSD>([&](auto d) { kadd(S::NMA + b, l0 + d, kbd[b][d]); });
      });
    }
  }
  // J_ab: every thread of the block holds partial sums of the same entries;
  // waves reduce in registers (a fixed order), then through LDS to one global
  // atomic each
  const int lane = (int)(threadIdx.x & 63);
  if (doJ) {
    sfor<NQ>([&](auto qq) {
      const double tot = wave_sum(jij[qq]);
      if (lane == 0 && tot != 0.0) {
        if (det)
          fx_lds_add(sjab + 2 * qq, tot, bad);
        else
          atomicAdd(reinterpret_cast<double*>(sjab) + qq, tot);
      }
    });
  }
  if (bad) *ex.fx_over = 1;
  __syncthreads();
  if (doJ && threadIdx.x < NQ) {
    const int qi = Q0 + (int)threadIdx.x;
    const size_t e = (size_t)(i0 + qi / C::NSB) * n + j0 + qi % C::NSB;
    if (det) {
      fx_flush(sjab + 2 * threadIdx.x, ex.Jx + 2 * e);
    } else {
      const double v = reinterpret_cast<const double*>(sjab)[threadIdx.x];
      if (v != 0.0) unsafeAtomicAdd(&J[e], v);
    }
  }
  if (doK) {
    const double kq = 0.25 * kscale;
    for (int r = 0; r < S::NR; ++r) {
      const size_t ao = r < S::NMA ? (size_t)(i0 + S::MA0 + r) : (size_t)(j0 + S::MB_LO + (r - S::NMA));
      for (int c = threadIdx.x; c < W; c += blockDim.x) {
        if (det) {
          // the exchange scale is applied when the fixed-point sums are read
          fx_flush(Kf + (size_t)r * krow + 2 * (size_t)c, ex.Kx + 2 * (ao * n + c0 + c));
        } else {
          const double v = Ks[(size_t)r * krow + c];
          if (v != 0.0) unsafeAtomicAdd(&K[ao * n + c0 + c], kq * v);
        }
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
  // CU resident (the register budget per thread shrinks with the block:
  // 128 VGPRs at 1024 threads, 170 at 768, 256 at 512); 768 for the L = 2
  // classes and the L = 3 classes with an (ss) ket, which fit 170; 512 for
  // the heavier classes with an (ss) or (ps) ket, which fit 256
  static constexpr int SEG_BLK =
      K::F ? 256
           : (C::L <= 1 ? 1024
                        : (C::L <= 3 ? ((C::L == 2 || C::FC == 0) ? 768 : 512) : (C::FC <= 1 ? 512 : 256)));
  // doubles of the Schwarz staging block per pair (classes with f shells)
  static constexpr int SCHWARZ_BLOCK = K::F ? C::NBRA * C::NKETC : 0;

  // dynamic LDS bytes of a segment block over W exchange columns: the K rows
  // (doubles, or fixed-point pairs in deterministic builds) and the P rows
  // with prow; seg_total adds the static J_ab slots
  // bytes of the Boys rows staged in LDS (one-root classes; sized for FP64)
  static constexpr size_t BOYS_LDS = lds_boys<C>() ? (size_t)BOYS_NG * kLdsBoysW * sizeof(double) : 0;
  template <int Q0, int Q1>
  static size_t seg_bytes(int W, bool prow, bool det) {
    return (size_t)SegRows<C, Q0, Q1>::NR * W * sizeof(double) * ((det ? 2 : 1) + (prow ? 1 : 0)) + BOYS_LDS;
  }
  template <int Q0, int Q1>
  static size_t seg_total(int W, bool prow, bool det) {
    return seg_bytes<Q0, Q1>(W, prow, det) + 2 * (size_t)(Q1 - Q0) * sizeof(double);
  }
  // the widest W whose block fits lim bytes
  template <int Q0, int Q1>
  static int seg_width(size_t lim, bool prow, bool det) {
    const size_t fixed = 2 * (size_t)(Q1 - Q0) * sizeof(double) + BOYS_LDS;
    const size_t per = seg_bytes<Q0, Q1>(1, prow, det) - BOYS_LDS;
    return lim > fixed ? (int)std::min<size_t>((lim - fixed) / per, (size_t)1 << 30) : 0;
  }

  // shell-aligned column tiles of at most wmax columns covering [0, nbf)
  static bool make_tiles(const JKLaunch& a, int wmax, SegExtra& ex) {
    ex.ntile = 0;
    int lo = 0;
    while (lo < a.nbf) {
      if (ex.ntile == kMaxColTiles || a.sh_bf0 == nullptr) return false;
      // the last shell boundary within lo + wmax
      const int* e = std::upper_bound(a.sh_bf0, a.sh_bf0 + a.nsh + 1, lo + wmax);
      const int hi = *(e - 1);
      if (hi <= lo) return false;  // one shell wider than a tile
      ex.tile_lo[ex.ntile] = lo;
      ex.tile_hi[ex.ntile] = hi;
      ++ex.ntile;
      lo = hi;
    }
    return true;
  }

  template <int Q0, int Q1, int K0, int K1, typename R, bool PROW>
  static bool launch_seg(const JKLaunch& a, hipStream_t s, size_t bytes, const SegExtra& ex) {
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
    const dim3 grid((unsigned)((long long)a.nseg * nsplit * ex.ntile));
    if constexpr (sizeof(R) == 4)
      hipLaunchKernelGGL(kern, grid, dim3(SEG_BLK), bytes, s, a.segs, a.offs, a.tasks_all, a.pairs, a.ppos32, a.pk32,
                         a.ra32->sig, *a.ra32, a.P, a.J, a.K, a.nbf, a.kscale, a.jbuf, a.jslot, a.qpos, jfirst,
                         nsplit, ex);
    else
      hipLaunchKernelGGL(kern, grid, dim3(SEG_BLK), bytes, s, a.segs, a.offs, a.tasks_all, a.pairs, a.ppos, a.pk,
                         a.ra->sig, *a.ra, a.P, a.J, a.K, a.nbf, a.kscale, a.jbuf, a.jslot, a.qpos, jfirst,
                         nsplit, ex);
    return true;
  }

  // waves of this kernel one CU can hold when a block needs `bytes` of LDS
  // (the CU's LDS equals the per-block maximum on gfx950; 32 wave slots)
  static int waves_per_cu(size_t bytes, size_t lds) {
    const size_t blocks_lds = bytes ? lds / bytes : 32;
    const size_t blocks_slots = (size_t)32 / (SEG_BLK / 64);
    return (int)std::min(blocks_lds, blocks_slots) * (SEG_BLK / 64);
  }

  // The segment kernel with K and P rows in LDS when they fit (compiled when
  // BOTH), else K rows only (P from L2; a negative lds_max forces it); when
  // even whole K rows do not fit, the K rows are split into shell-aligned
  // column tiles, each block evaluating the tasks whose ket shells fall in
  // its tile.  kcol_tile > 0 caps the tile width (to exercise the tiled path
  // on small systems).
  template <int Q0, int Q1, int K0, int K1, typename R, bool BOTH>
  static void seg(const JKLaunch& a, hipStream_t s) {
    static_assert(SegRows<C, Q0, Q1>::NR <= kSegRowsMax, "chunk needs more LDS rows than planned");
    if (a.nseg <= 0 || a.segs == nullptr) throw std::runtime_error("segment list missing");
    const size_t lim = (size_t)(a.lds_max < 0 ? -a.lds_max : a.lds_max);
    const bool det = a.det != 0;
    SegExtra ex{a.pmax, a.prim_screen, a.det, a.Jx, a.Kx, a.fx_over, a.kets, a.shells, 1, {0}, {a.nbf}};
    const int cap = a.kcol_tile > 0 ? a.kcol_tile : a.nbf;
    if (cap >= a.nbf) {
      const size_t tot_k = seg_total<Q0, Q1>(a.nbf, false, det);
      if constexpr (BOTH) {
        // density rows in LDS unless that leaves fewer than seg_min_waves
        // waves per CU and dropping them lets more reside
        const size_t tot_p = seg_total<Q0, Q1>(a.nbf, true, det);
        const int wp = waves_per_cu(tot_p, lim), wk = waves_per_cu(tot_k, lim);
        const bool prow = a.lds_max > 0 && tot_p <= lim && (wp >= wk || wp >= a.seg_min_waves);
        if (prow && launch_seg<Q0, Q1, K0, K1, R, true>(a, s, seg_bytes<Q0, Q1>(a.nbf, true, det), ex)) return;
      }
      if (tot_k <= lim && launch_seg<Q0, Q1, K0, K1, R, false>(a, s, seg_bytes<Q0, Q1>(a.nbf, false, det), ex))
        return;
    }
    // tiles: the FP64 K-row kernel (FP32 tasks too)
    const int w = std::min(cap, seg_width<Q0, Q1>(lim, false, det));
    if (!make_tiles(a, w, ex))
      throw std::runtime_error("the exchange rows of one bra pair need more LDS column tiles than a launch holds "
                               "(nbf too large)");
    int wmax = 0;
    for (int t = 0; t < ex.ntile; ++t) wmax = std::max(wmax, ex.tile_hi[t] - ex.tile_lo[t]);
    if (launch_seg<Q0, Q1, K0, K1, double, false>(a, s, seg_bytes<Q0, Q1>(wmax, false, det), ex)) return;
    throw std::runtime_error("segment kernel launch configuration rejected");
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

  // FP64 segment kernels with K rows only (FP32 tasks are evaluated in FP64):
  // the variant compiled for classes with f shells, where more variants
  // would multiply the compile time; their integral work outweighs reading
  // the density from L2
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
