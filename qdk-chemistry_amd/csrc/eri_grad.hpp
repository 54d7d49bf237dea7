// Adjoint nuclear-gradient kernel: derivative ERIs contracted with the density
// pairs on the fly, never materialised.  Restates the contraction of
// ERI::get_gradients (libint2_direct.cpp:646-890):
//   dJ[x] = 1/2 sum_ordered P_ij P_kl d(ij|kl)/dx
//   dK[x] = -1/4 (alpha+beta) sum_ordered P_ik P_jl d(ij|kl)/dx
// over the canonical quartet list with degeneracy weights, where the ordered
// sum over the 8 index permutations turns the exchange weight into
// deg/2 (P_ik P_jl + P_il P_jk).
//
// Per primitive quartet and Rys root the 1-D integrals are built with one
// extra unit of angular momentum on A, B and C; the Gaussian derivative
// d/dA_x phi_a = 2 alpha phi_{a+1x} - a_x phi_{a-1x} then gives the x, y, z
// derivatives for centres A, B, C, and translation invariance gives D.  The
// cartesian weights W_abcd (density blocks back-transformed from solid
// harmonics) are formed once per shell quartet.  Gradients are summed per
// block in LDS (6 doubles per atom) and flushed with one atomic per entry.
#pragma once

#include "eri_class.hpp"

namespace qfe {

template <class C>
struct GradTraits {
  static constexpr int EA1 = C::EA + 1, FC1 = C::FC + 1;
  static constexpr int NR1 = (C::L + 1) / 2 + 1;
  static constexpr int NBRAC = C::NCA * C::NCB;  // bra cartesian pairs
};

// Weights held per thread (bra cartesians x ket cartesians of one chunk): 12
// keeps every class up to (dd|dd) free of register spills; classes with more
// ket cartesians than that are also chunked over the ket.
constexpr int kGradWeights = 12;
template <class C>
struct GradChunks {
  static constexpr int NBC = C::NCA * C::NCB;
  static constexpr int KCH = C::NKETC <= kGradWeights ? C::NKETC : kGradWeights;  // ket cartesians per chunk
  static constexpr int NKCH = (C::NKETC + KCH - 1) / KCH;
  static constexpr int BCH_ = kGradWeights / KCH;
  static constexpr int BCH = BCH_ < 1 ? 1 : (BCH_ > NBC ? NBC : BCH_);  // bra cartesians per chunk
  static constexpr int NBCH = (NBC + BCH - 1) / BCH;
  static constexpr int NCH = NBCH * NKCH;  // chunk c: bra chunk c / NKCH, ket chunk c % NKCH
};

// density block (sa, sb) back-transformed to cartesians: Pc[x][y] =
// sum_{m,n} S_la[m][x] S_lb[n][y] P[bf_a + m][bf_b + n]
template <int LX, int LY, bool PURE>
QFE_DI void cart_density(const double* __restrict__ P, size_t n, int bx, int by,
                         double (&out)[ncart(LX)][ncart(LY)]) {
  constexpr int NX = nsph(LX, PURE), NY = nsph(LY, PURE);
  double blk[NX][NY];
  sfor<NX>([&](auto m) { sfor<NY>([&](auto k) { blk[m][k] = P[(size_t)(bx + m) * n + by + k]; }); });
  sfor<ncart(LX)>([&](auto x) {
    sfor<ncart(LY)>([&](auto y) {
      double s = 0.0;
      sfor<NX>([&](auto m) {
        constexpr double cx = sph_coef(LX, PURE, m, x);
        if constexpr (cx != 0.0) {
          sfor<NY>([&](auto k) {
            constexpr double cy = sph_coef(LY, PURE, k, y);
            if constexpr (cy != 0.0) s = fma(cx * cy, blk[m][k], s);
          });
        }
      });
      out[x][y] = s;
    });
  });
}

// exchange weights of one spin density: W[q][k] += wk (P_ac P_bd + P_ad P_bc)
// over the cartesian chunk (bra cartesians [Q0, Q1), ket cartesians [K0, K1))
template <class C, int Q0, int Q1, int K0, int K1>
QFE_DI void exchange_weights(const double* __restrict__ P, size_t n, const PairHdr& bra, const PairHdr& ket,
                             double wk, double (&W)[Q1 - Q0][K1 - K0]) {
  double Pac[C::NCA][C::NCC], Pbd[C::NCB][C::NCD], Pad[C::NCA][C::NCD], Pbc[C::NCB][C::NCC];
  cart_density<C::LA, C::LC, C::PURE>(P, n, bra.bfa, ket.bfa, Pac);
  cart_density<C::LB, C::LD, C::PURE>(P, n, bra.bfb, ket.bfb, Pbd);
  cart_density<C::LA, C::LD, C::PURE>(P, n, bra.bfa, ket.bfb, Pad);
  cart_density<C::LB, C::LC, C::PURE>(P, n, bra.bfb, ket.bfa, Pbc);
  sfor<Q1 - Q0>([&](auto qq) {
    constexpr int ac = (Q0 + qq) / C::NCB, bc = (Q0 + qq) % C::NCB;
    sfor<K1 - K0>([&](auto kk) {
      constexpr int cc = (K0 + kk) / C::NCD, dc = (K0 + kk) % C::NCD;
      W[qq][kk] = fma(wk, Pac[ac][cc] * Pbd[bc][dc] + Pad[ac][dc] * Pbc[bc][cc], W[qq][kk]);
    });
  });
}

// Derivatives of one canonical task (bra pair tk.x, ket pair tk.y) by its
// centres A, B and C, contracted with the J and K weights: gj[centre][xyz],
// gk[centre][xyz] (centre D follows by translation invariance).  P is the
// density of the Coulomb term (the total density of a spin-split pair); the
// exchange term sums over Pk (RHF: the total density, scale -1/4 of the
// ordered sum) or over the spin densities Pk and Pk2 (scale -1/2), as
// libint2_direct.cpp:683-692,866-886 do.  Host-callable so the arithmetic
// can be checked on the CPU.
template <class C, int Q0, int Q1, int K0, int K1>
QFE_DI void grad_task(int2 tk, const PairHdr* __restrict__ pairs, const double4* __restrict__ ppos,
                      const double* __restrict__ pk, const double* __restrict__ pa, const RysArgs& ra,
                      const double* __restrict__ P, const double* __restrict__ Pk, const double* __restrict__ Pk2,
                      int nbf, double kscale, double (&gj)[3][3], double (&gk)[3][3]) {
  using G = GradTraits<C>;
  const PairHdr bra = pairs[tk.x];
  const PairHdr ket = pairs[tk.y];
  const double deg = (bra.sa == bra.sb ? 1.0 : 2.0) * (ket.sa == ket.sb ? 1.0 : 2.0) *
                     (tk.x == tk.y ? 1.0 : 2.0);
  const size_t n = (size_t)nbf;
  // cartesian density blocks of the Coulomb weight
  double Pab[C::NCA][C::NCB], Pcd[C::NCC][C::NCD];
  cart_density<C::LA, C::LB, C::PURE>(P, n, bra.bfa, bra.bfb, Pab);
  cart_density<C::LC, C::LD, C::PURE>(P, n, ket.bfa, ket.bfb, Pcd);
  constexpr int NQ = Q1 - Q0, NK = K1 - K0;
  double WJ[NQ][NK], WK[NQ][NK];
  const double wj = 0.5 * deg;
  // the canonical task stands for its deg ordered quartets: -1/4 (RHF) or
  // -1/2 (per spin) of the ordered sum is -deg/8 or -deg/4 per task
  const double wk = (Pk2 ? -0.25 : -0.125) * kscale * deg;
  sfor<NQ>([&](auto qq) {
    constexpr int ac = (Q0 + qq) / C::NCB, bc = (Q0 + qq) % C::NCB;
    sfor<NK>([&](auto kk) {
      constexpr int cc = (K0 + kk) / C::NCD, dc = (K0 + kk) % C::NCD;
      WJ[qq][kk] = wj * Pab[ac][bc] * Pcd[cc][dc];
      WK[qq][kk] = 0.0;
    });
  });
  exchange_weights<C, Q0, Q1, K0, K1>(Pk, n, bra, ket, wk, WK);
  if (Pk2) exchange_weights<C, Q0, Q1, K0, K1>(Pk2, n, bra, ket, wk, WK);
  sfor<3>([&](auto a) { sfor<3>([&](auto b) { gj[a][b] = 0.0; gk[a][b] = 0.0; }); });
  constexpr double TWO_PI_52 = 34.98683665524972;
  const double ABx = bra.A[0] - bra.B[0], ABy = bra.A[1] - bra.B[1], ABz = bra.A[2] - bra.B[2];
  const double CDx = ket.A[0] - ket.B[0], CDy = ket.A[1] - ket.B[1], CDz = ket.A[2] - ket.B[2];
  for (int ik = ket.prim0; ik < ket.prim0 + ket.nprim; ++ik) {
    const double4 Q = ppos[ik];
    const double Kk = pk[ik];
    const double q = Q.x;
    const double tg = 2.0 * pa[ik];  // 2 gamma (exponent on centre C)
    const double QCx = Q.y - ket.A[0], QCy = Q.z - ket.A[1], QCz = Q.w - ket.A[2];
    const double h2q = 0.5 * frcp(q);
    for (int ib = bra.prim0; ib < bra.prim0 + bra.nprim; ++ib) {
      const double4 Pp = ppos[ib];
      const double Kb = pk[ib];
      const double p = Pp.x;
      const double ta = 2.0 * pa[ib], tb = 2.0 * (p - pa[ib]);  // 2 alpha, 2 beta
      const double PAx = Pp.y - bra.A[0], PAy = Pp.z - bra.A[1], PAz = Pp.w - bra.A[2];
      const double PQx = Pp.y - Q.y, PQy = Pp.z - Q.z, PQz = Pp.w - Q.w;
      const double rs = frsq(p + q);
      const double rpq = rs * rs;
      const double T = p * q * rpq * (PQx * PQx + PQy * PQy + PQz * PQz);
      const double pref = TWO_PI_52 * Kb * Kk * rs;
      double u[G::NR1], w[G::NR1];
      rys_roots<G::NR1>(T, ra, u, w);
      const double qr = q * rpq, pr = p * rpq;
      const double h2p = 0.5 * frcp(p);
#pragma unroll 1
      for (int r = 0; r < G::NR1; ++r) {
        const double ur = u[r];
        const double B00 = 0.5 * rpq * ur;
        const double B10 = h2p - h2p * qr * ur;
        const double B01 = h2q - h2q * pr * ur;
        const double xa = qr * ur, xc = pr * ur;
        OneDX<C::LA + 1, C::LB + 1, C::LC + 1, C::LD, double> I[3];
        rys_1d_x<G::EA1, G::FC1, C::LA + 1, C::LB + 1, C::LC + 1, C::LD, double>(
            1.0, PAx - xa * PQx, QCx + xc * PQx, B00, B10, B01, ABx, CDx, I[0]);
        rys_1d_x<G::EA1, G::FC1, C::LA + 1, C::LB + 1, C::LC + 1, C::LD, double>(
            1.0, PAy - xa * PQy, QCy + xc * PQy, B00, B10, B01, ABy, CDy, I[1]);
        rys_1d_x<G::EA1, G::FC1, C::LA + 1, C::LB + 1, C::LC + 1, C::LD, double>(
            w[r] * pref, PAz - xa * PQz, QCz + xc * PQz, B00, B10, B01, ABz, CDz, I[2]);
        sfor<NQ>([&](auto qq) {
          constexpr int ac = (Q0 + qq) / C::NCB, bc = (Q0 + qq) % C::NCB;
          constexpr int ae[3] = {cart_x(C::LA, ac), cart_y(C::LA, ac), cart_z(C::LA, ac)};
          constexpr int be[3] = {cart_x(C::LB, bc), cart_y(C::LB, bc), cart_z(C::LB, bc)};
          sfor<NK>([&](auto kk) {
            constexpr int cc = (K0 + kk) / C::NCD, dc = (K0 + kk) % C::NCD;
            constexpr int ce[3] = {cart_x(C::LC, cc), cart_y(C::LC, cc), cart_z(C::LC, cc)};
            {
              constexpr int de[3] = {cart_x(C::LD, dc), cart_y(C::LD, dc), cart_z(C::LD, dc)};
              const double wJ = WJ[qq][kk], wK = WK[qq][kk];
              double v[3];
              sfor<3>([&](auto x) { v[x] = I[x].v[ae[x]][be[x]][ce[x]][de[x]]; });
              sfor<3>([&](auto x) {
                constexpr int y = (x + 1) % 3, z = (x + 2) % 3;
                const double o = v[y] * v[z];
                const double oj = wJ * o, ok = wK * o;
                // centre A
                double dA = ta * I[x].v[ae[x] + 1][be[x]][ce[x]][de[x]];
                if constexpr (ae[x] > 0) dA -= (double)ae[x] * I[x].v[ae[x] - 1][be[x]][ce[x]][de[x]];
                // centre B
                double dB = tb * I[x].v[ae[x]][be[x] + 1][ce[x]][de[x]];
                if constexpr (be[x] > 0) dB -= (double)be[x] * I[x].v[ae[x]][be[x] - 1][ce[x]][de[x]];
                // centre C
                double dC = tg * I[x].v[ae[x]][be[x]][ce[x] + 1][de[x]];
                if constexpr (ce[x] > 0) dC -= (double)ce[x] * I[x].v[ae[x]][be[x]][ce[x] - 1][de[x]];
                gj[0][x] = fma(dA, oj, gj[0][x]);
                gk[0][x] = fma(dA, ok, gk[0][x]);
                gj[1][x] = fma(dB, oj, gj[1][x]);
                gk[1][x] = fma(dB, ok, gk[1][x]);
                gj[2][x] = fma(dC, oj, gj[2][x]);
                gk[2][x] = fma(dC, ok, gk[2][x]);
              });
            }
          });
        });
      }
    }
  }
}

// One thread per canonical task; bra cartesians [Q0, Q1) and ket cartesians
// [K0, K1) of the class (the chunks keep every instance free of register
// spills).  With lds the block sums its contributions in LDS ([2][3][natoms])
// and flushes one atomic per entry; otherwise each thread adds to global.
template <class C, int Q0, int Q1, int K0, int K1>
__global__ __launch_bounds__(256) void k_eri_grad(const int2* __restrict__ tasks, long long ntask,
                                                  const PairHdr* __restrict__ pairs,
                                                  const double4* __restrict__ ppos, const double* __restrict__ pk,
                                                  const double* __restrict__ pa, RysArgs ra,
                                                  const double* __restrict__ P, const double* __restrict__ Pk,
                                                  const double* __restrict__ Pk2, int nbf,
                                                  const int* __restrict__ sh_atom, int natoms, double kscale,
                                                  int lds, double* __restrict__ gJ, double* __restrict__ gK) {
  extern __shared__ double sg[];
  if (lds) {
    for (int i = threadIdx.x; i < 6 * natoms; i += blockDim.x) sg[i] = 0.0;
    __syncthreads();
  }
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < ntask) {
    const int2 tk = tasks[t];
    double gj[3][3], gk[3][3];
    grad_task<C, Q0, Q1, K0, K1>(tk, pairs, ppos, pk, pa, ra, P, Pk, Pk2, nbf, kscale, gj, gk);
    const PairHdr& bra = pairs[tk.x];
    const PairHdr& ket = pairs[tk.y];
    const int at[4] = {sh_atom[bra.sa], sh_atom[bra.sb], sh_atom[ket.sa], sh_atom[ket.sb]};
    sfor<3>([&](auto x) {
      const double dj[4] = {gj[0][x], gj[1][x], gj[2][x], -(gj[0][x] + gj[1][x] + gj[2][x])};
      const double dk[4] = {gk[0][x], gk[1][x], gk[2][x], -(gk[0][x] + gk[1][x] + gk[2][x])};
      sfor<4>([&](auto c) {
        const int idx = at[c] + x * natoms;
        if (lds) {
          atomicAdd(&sg[idx], dj[c]);
          atomicAdd(&sg[3 * natoms + idx], dk[c]);
        } else {
          unsafeAtomicAdd(&gJ[idx], dj[c]);
          unsafeAtomicAdd(&gK[idx], dk[c]);
        }
      });
    });
  }
  if (lds) {
    __syncthreads();
    for (int i = threadIdx.x; i < 3 * natoms; i += blockDim.x) {
      if (sg[i] != 0.0) unsafeAtomicAdd(&gJ[i], sg[i]);
      if (sg[3 * natoms + i] != 0.0) unsafeAtomicAdd(&gK[i], sg[3 * natoms + i]);
    }
  }
}

template <int LA, int LB, int LC, int LD, bool PURE>
struct GradOps {
  using C = Cls<LA, LB, LC, LD, PURE>;
  using K = GradChunks<C>;
  static constexpr int NCH = K::NCH;

  template <int C0, int C1, bool LAST>
  static void run_range(const GradLaunch& a, hipStream_t s) {
    static_assert(C0 < C1 && C1 <= NCH && (!LAST || C1 == NCH), "chunk range does not match the class");
    if (a.ntask <= 0) return;
    const int block = 256;
    const long long grid = (a.ntask + block - 1) / block;
    const int lds = a.natoms * 6 * (int)sizeof(double) <= 64 * 1024 ? 1 : 0;
    const size_t shm = lds ? (size_t)a.natoms * 6 * sizeof(double) : 0;
    hsfor<C1 - C0>([&](auto cc) {
      constexpr int c = C0 + cc;
      constexpr int b = c / K::NKCH, k = c % K::NKCH;
      constexpr int q0 = b * K::BCH, q1 = (q0 + K::BCH < K::NBC) ? q0 + K::BCH : K::NBC;
      constexpr int k0 = k * K::KCH, k1 = (k0 + K::KCH < C::NKETC) ? k0 + K::KCH : C::NKETC;
      hipLaunchKernelGGL((k_eri_grad<C, q0, q1, k0, k1>), dim3((unsigned)grid), dim3(block), shm, s, a.tasks,
                         a.ntask, a.pairs, a.ppos, a.pk, a.pa, *a.ra, a.P, a.Pk, a.Pk2, a.nbf, a.sh_atom, a.natoms, a.kscale,
                         lds, a.gJ, a.gK);
    });
  }
};

}  // namespace qfe
