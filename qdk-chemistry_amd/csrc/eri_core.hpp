// Device-side core of the Fock-build engine: compile-time-specialised Rys
// quadrature for one shell quartet class (LA LB | LC LD).
//
// One thread evaluates one shell quartet (bra pair, ket pair) -- or, for the
// wide classes, one chunk of bra function pairs of it -- contracting over all
// primitive quartets in registers:
//   per primitive quartet: Gaussian product data from the pair tables, Rys
//   roots/weights (piecewise polynomial table or large-T Hermite limit), the
//   2-D recursion for G(i,k) along x, y, z, 1-D horizontal transfer to
//   I(a,b,c,d), and the assembly of the chunk's (ab|cd) in cartesians on the
//   ket and solid harmonics on the bra.
// Every index below is a compile-time constant (sfor over integer sequences)
// so all arrays live in VGPRs and the compiler removes unused terms.
//
// Recurrences: Rys, Dupuis & King, J. Comput. Chem. 4, 154 (1983);
// horizontal transfer: Head-Gordon & Pople, JCP 89, 5777 (1988), applied per
// cartesian direction.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>
#include <utility>

#include "sph_coef.inc"

namespace qfe {

#define QFE_HD __host__ __device__
// device functions are also host-callable, so the kernel arithmetic can be
// checked on the CPU (the hardware estimates below fall back to exact math)
#define QFE_DI __host__ __device__ __forceinline__

// ---------------------------------------------------------------- indices
QFE_HD constexpr int ncart(int l) { return (l + 1) * (l + 2) / 2; }
QFE_HD constexpr int nsph(int l, bool pure) { return (pure && l >= 2) ? 2 * l + 1 : ncart(l); }
// CCA order inside a shell: lx descending, then ly descending.
QFE_HD constexpr int cart_x(int l, int c) {
  int k = 0;
  while ((k + 1) * (k + 2) / 2 <= c) ++k;
  return l - k;
}
QFE_HD constexpr int cart_z(int l, int c) {
  int k = 0;
  while ((k + 1) * (k + 2) / 2 <= c) ++k;
  return c - k * (k + 1) / 2;
}
QFE_HD constexpr int cart_y(int l, int c) { return l - cart_x(l, c) - cart_z(l, c); }

template <typename F, int... Is>
QFE_DI void sfor_impl(F&& f, std::integer_sequence<int, Is...>) {
  (f(std::integral_constant<int, Is>{}), ...);
}
// compile-time loop: f(integral_constant<int, i>) for i = 0..N-1
template <int N, typename F>
QFE_DI void sfor(F&& f) {
  sfor_impl(f, std::make_integer_sequence<int, N>{});
}

// solid-harmonic coefficient (bra side) or identity for cartesian / l < 2
QFE_HD constexpr double sph_coef(int l, bool pure, int m, int c) {
  if (!(pure && l >= 2)) return m == c ? 1.0 : 0.0;
  return kSph[l][m][c];
}

// ------------------------------------------------------------ device data
struct PairHdr {        // one significant shell pair (80 B)
  double A[3];          // centre of shell a
  double B[3];          // centre of shell b
  int prim0, nprim;     // range in the primitive-pair arrays
  int sa, sb;           // shell indices (AO order of the reference)
  int bfa, bfb;         // first AO of each shell
  int cls, pad;
};

// primitive pair: ppos = (p, Px, Py, Pz); pk = c_a c_b exp(-ab/p |AB|^2) / p
template <typename R>
struct RysArgsT {
  const R* coef[9];   // per root count n: [nint][2n][RYS_DEG+1]
  R tmax[9];
  R hx2[9][8];        // large-T limit: u_i = hx2/T
  R hw[9][8];         //               w_i = hw/sqrt(T)
  const R* boys;      // Boys Taylor table [BOYS_NG][BOYS_W]: F_m(k * BOYS_DT)
  const R* h2;        // 1/(2p) per primitive pair (p = exponent sum), in ppos order
  // primitive-quartet screening factor per primitive pair, in ppos order:
  // sig = ln(|pk| p^-1/4) + ln(2^1/2 pi^5/2) / 2, so that the prefactor of a
  // primitive quartet, 2 pi^5/2 |pk_ab pk_cd| / sqrt(p + q), never exceeds
  // exp(sig_ab + sig_cd) (p + q >= 2 sqrt(pq)); each pair's primitives are
  // sorted by sig, largest first
  const R* sig;
};
using RysArgs = RysArgsT<double>;
using RysArgs32 = RysArgsT<float>;

template <typename R>
struct Vec4;
template <>
struct Vec4<double> {
  using type = double4;
};
template <>
struct Vec4<float> {
  using type = float4;
};
constexpr int RYS_DEG = 11;
constexpr double BOYS_DT = 0.05;
constexpr double BOYS_TSW = 40.0;  // above: F_0 = sqrt(pi/T)/2 (erfc(sqrt 40) < 4e-19) and upward recursion
constexpr int BOYS_NG = 802;
constexpr int BOYS_W = 24;

// 1/x and 1/sqrt(x) for positive normal x: hardware estimate plus two
// Newton steps (quadratic convergence to full double precision).
QFE_DI double frcp(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rcp(x);
  r = fma(fma(-x, r, 1.0), r, r);
  return fma(fma(-x, r, 1.0), r, r);
#else
  return 1.0 / x;
#endif
}
QFE_DI double frsq(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  double r = __builtin_amdgcn_rsq(x);
  const double hx = 0.5 * x;
  r = r * fma(-hx * r, r, 1.5);
  return r * fma(-hx * r, r, 1.5);
#else
  return 1.0 / sqrt(x);
#endif
}
// single precision: the hardware results (v_rcp_f32, v_rsq_f32: 1 ulp for
// normal arguments, the FP32 rounding level), with no Newton step
QFE_DI float frcp(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}
QFE_DI float frsq(float x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rsqf(x);
#else
  return 1.0f / sqrtf(x);
#endif
}
// An empty asm that reads x: x must be in a register at this point, so the
// compiler issues its load before it and cannot sink the load below a later
// branch.  held_s for wave-uniform values (scalar registers), held_v for
// per-lane values.
template <typename T>
QFE_DI void held_s(const T& x) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" ::"s"(x));
#endif
}
template <typename T>
QFE_DI void held_v(const T& x) {
#if defined(__HIP_DEVICE_COMPILE__)
  asm volatile("" ::"v"(x));
#endif
}
QFE_DI double fexp(double x) { return exp(x); }
QFE_DI float fexp(float x) { return expf(x); }

// F_0..F_M at T.  Below BOYS_TSW: Taylor expansion about the nearest grid
// point, |dT| <= 0.025, with NT terms: 7 in double precision (truncation
// < 2e-16 relative), 4 in single precision (the first dropped term is below
// F_{m+4} dT^4 / 4! <= 1.7e-8 F_m, under the FP32 rounding unit); for M <= 2
// every order is expanded directly (no exp), otherwise F_M is expanded and
// the downward recursion F_m = (2T F_{m+1} + e^-T)/(2m+1) fills the rest.
// Above: F_0 from the asymptotic form and the upward recursion (stable for
// T > M; e^-T < 4e-18 is dropped).
// tab: rows of BW values F_m(k * BOYS_DT), m < BW
template <int M, typename R, int BW = BOYS_W>
QFE_DI void boys_fm(R T, const R* __restrict__ tab, R (&F)[M + 1]) {
  constexpr int NT = sizeof(R) == 4 ? 4 : 7;
  static_assert(M + 7 <= BW, "the Taylor series reads at most F_M .. F_{M+6}");
  if (T < R(BOYS_TSW)) {
    const int g = (int)(T * R(1.0 / BOYS_DT) + R(0.5));
    const R d = (R)g * R(BOYS_DT) - T;
    const R* __restrict__ row = tab + (size_t)g * BW;
    if constexpr (M <= 2) {
      sfor<M + 1>([&](auto m) {
        R s = row[m + NT - 1];
        sfor<NT - 1>([&](auto kk) {
          constexpr int k = NT - 2 - kk;
          s = fma(s, d * R(1.0 / (double)(k + 1)), row[m + k]);
        });
        F[m] = s;
      });
    } else {
      const R e = fexp(-T);
      R s = row[M + NT - 1];
      sfor<NT - 1>([&](auto kk) {
        constexpr int k = NT - 2 - kk;
        s = fma(s, d * R(1.0 / (double)(k + 1)), row[M + k]);
      });
      F[M] = s;
      sfor<M>([&](auto mm) {
        constexpr int m = M - 1 - mm;
        F[m] = (R(2) * T * F[m + 1] + e) * R(1.0 / (double)(2 * m + 1));
      });
    }
  } else {
    const R is = frsq(T);
    const R it = is * is;                     // 1/T from 1/sqrt(T): no second reciprocal
    F[0] = R(0.88622692545275801365) * is;  // sqrt(pi)/2 / sqrt(T)
    sfor<M>([&](auto m) { F[m + 1] = R(2 * m + 1) * F[m] * (R(0.5) * it); });
  }
}

// boys: the Boys table (rows of BW values) if not ra.boys
template <int N, typename R, int BW = BOYS_W>
QFE_DI void rys_roots(R T, const RysArgsT<R>& ra, R (&u)[N], R (&w)[N], const R* boys = nullptr) {
  if constexpr (N <= 2) {
    // Gauss rule of the Boys measure from its moments F_0..F_{2N-1}
    R F[2 * N];
    boys_fm<2 * N - 1, R, BW>(T, BW == BOYS_W ? ra.boys : boys, F);
    if constexpr (N == 1) {
      w[0] = F[0];
      u[0] = F[1] * frcp(F[0]);
    } else {
      // monic pi_2(x) = x^2 + a x + b orthogonal to 1 and x
      const R idet = frcp(F[0] * F[2] - F[1] * F[1]);
      const R a = (F[1] * F[2] - F[0] * F[3]) * idet;
      const R b = (F[1] * F[3] - F[2] * F[2]) * idet;
      const R disc = a * a - R(4) * b;
      const R q = R(0.5) * (disc * frsq(disc) - a);  // larger root (a < 0)
      const R x1 = q, x0 = b * frcp(q);
      const R w0 = (F[1] - x1 * F[0]) * frcp(x0 - x1);
      u[0] = x0;
      u[1] = x1;
      w[0] = w0;
      w[1] = F[0] - w0;
    }
  } else {
    if (T >= ra.tmax[N]) {
      const R it = frcp(T);
      const R is = frsq(T);
      sfor<N>([&](auto i) {
        u[i] = ra.hx2[N][i] * it;
        w[i] = ra.hw[N][i] * is;
      });
    } else {
      const int k = (int)T;
      const R x = T - ((R)k + R(0.5));
      const R* __restrict__ c = ra.coef[N] + (size_t)k * (2 * N) * (RYS_DEG + 1);
      sfor<N>([&](auto i) {
        const R* cu = c + i * (RYS_DEG + 1);
        const R* cw = c + (N + i) * (RYS_DEG + 1);
        R vu = cu[RYS_DEG], vw = cw[RYS_DEG];
#pragma unroll
        for (int d = RYS_DEG - 1; d >= 0; --d) {
          vu = fma(vu, x, cu[d]);
          vw = fma(vw, x, cw[d]);
        }
        u[i] = vu;
        w[i] = vw;
      });
    }
  }
}

// ------------------------------------------------------------ class traits
template <int LA_, int LB_, int LC_, int LD_, bool PURE_>
struct Cls {
  static constexpr int LA = LA_, LB = LB_, LC = LC_, LD = LD_;
  static constexpr bool PURE = PURE_;
  static constexpr int EA = LA + LB, FC = LC + LD;
  static constexpr int L = EA + FC;
  static constexpr int NR = L / 2 + 1;
  static constexpr int NCA = ncart(LA), NCB = ncart(LB), NCC = ncart(LC), NCD = ncart(LD);
  static constexpr int NSA = nsph(LA, PURE), NSB = nsph(LB, PURE), NSC = nsph(LC, PURE),
                       NSD = nsph(LD, PURE);
  static constexpr int NBRA = NSA * NSB;   // bra function pairs
  static constexpr int NKETC = NCC * NCD;  // ket cartesian pairs
  static constexpr int NKETS = NSC * NSD;  // ket function pairs
};

// 1-D integrals of one direction for one root: I[a][b][c][d]
template <int AMX, int BMX, int CMX, int DMX, typename R>
struct OneDX {
  R v[AMX + 1][BMX + 1][CMX + 1][DMX + 1];
};
template <class C, typename R>
using OneD = OneDX<C::LA, C::LB, C::LC, C::LD, R>;

// 2-D recursion G(i, k), i <= EA, k <= FC, then 1-D horizontal transfer to
// I(a, b, c, d) for a <= AMX, b <= BMX, c <= CMX, d <= DMX with a + b <= EA
// and c + d <= FC (entries outside those bounds are not written).
template <int EA, int FC, int AMX, int BMX, int CMX, int DMX, typename R>
QFE_DI void rys_1d_x(R s, R C00, R C00p, R B00, R B10, R B01, R AB, R CD,
                     OneDX<AMX, BMX, CMX, DMX, R>& out) {
  R g[EA + 1][FC + 1];
  g[0][0] = s;
  if constexpr (EA >= 1) g[1][0] = C00 * s;
  sfor<(EA >= 2 ? EA - 1 : 0)>([&](auto ii) {
    constexpr int i = ii + 1;
    g[i + 1][0] = C00 * g[i][0] + R(i) * B10 * g[i - 1][0];
  });
  if constexpr (FC >= 1) {
    g[0][1] = C00p * s;
    sfor<EA>([&](auto ii) {
      constexpr int i = ii + 1;
      g[i][1] = C00p * g[i][0] + R(i) * B00 * g[i - 1][0];
    });
    sfor<(FC >= 2 ? FC - 1 : 0)>([&](auto kk) {
      constexpr int k = kk + 1;
      g[0][k + 1] = C00p * g[0][k] + R(k) * B01 * g[0][k - 1];
      sfor<EA>([&](auto ii) {
        constexpr int i = ii + 1;
        g[i][k + 1] = C00p * g[i][k] + R(k) * B01 * g[i][k - 1] + R(i) * B00 * g[i - 1][k];
      });
    });
  }
  // ket transfer: h[i][c][d] from g[i][c + d]: (c, d) = (c+1, d-1) + CD (c, d-1)
  R h[EA + 1][CMX + 1][DMX + 1];
  sfor<EA + 1>([&](auto i) {
    R t[FC + 1][DMX + 1];
    sfor<FC + 1>([&](auto k) { t[k][0] = g[i][k]; });
    sfor<DMX>([&](auto dd) {
      constexpr int d = dd + 1;
      sfor<FC + 1 - d>([&](auto k) { t[k][d] = t[k + 1][d - 1] + CD * t[k][d - 1]; });
    });
    sfor<CMX + 1>([&](auto c) {
      sfor<DMX + 1>([&](auto d) {
        if constexpr (c + d <= FC) h[i][c][d] = t[c][d];
      });
    });
  });
  // bra transfer: I[a][b][c][d] from h[a + b][c][d]
  sfor<CMX + 1>([&](auto c) {
    sfor<DMX + 1>([&](auto d) {
      if constexpr (c + d <= FC) {
        R t[EA + 1][BMX + 1];
        sfor<EA + 1>([&](auto i) { t[i][0] = h[i][c][d]; });
        sfor<BMX>([&](auto bb) {
          constexpr int b = bb + 1;
          sfor<EA + 1 - b>([&](auto i) { t[i][b] = t[i + 1][b - 1] + AB * t[i][b - 1]; });
        });
        sfor<AMX + 1>([&](auto a) {
          sfor<BMX + 1>([&](auto b) {
            if constexpr (a + b <= EA) out.v[a][b][c][d] = t[a][b];
          });
        });
      }
    });
  });
}

template <class C, typename R>
QFE_DI void rys_1d(R s, R C00, R C00p, R B00, R B10, R B01, R AB, R CD, OneD<C, R>& out) {
  rys_1d_x<C::EA, C::FC, C::LA, C::LB, C::LC, C::LD, R>(s, C00, C00p, B00, B10, B01, AB, CD, out);
}

// ln of the primitive-quartet cut.  FP64: the reference's engine precision
// max(eps / max|P|, eps / 1e10) (libint2_direct.cpp:279-287).  FP32 tasks
// (Q_ab Q_cd below the mixed threshold, 1e-6) carry ~1e-13 absolute rounding
// of their own; they cut at 1e-14 / max|P|, clamped the same way.
template <typename R>
QFE_DI R prim_ln_cut(const double* pmax, int prim_screen) {
  if (!prim_screen) return R(-3.0e38);
  const double pm = *pmax;
  const double eps = 2.220446049250313e-16;
  double prec = pm > 0.0 ? fmax(eps / pm, eps * 1e-10) : eps;
  if (sizeof(R) == 4) prec = fmax(prec, pm > 0.0 ? fmax(1e-14 / pm, 1e-24) : 1e-14);
  return (R)log(prec);
}

// Contract all primitive quartets of (bra|ket) into acc[q][k] for the bra
// function pairs q = Q0 .. Q1-1 (row-major over (ma, mb)) and the ket
// cartesian pairs K0 + k, k < K1 - K0 (row-major over (c, d)), in precision R.
// Primitive quartets whose prefactor bound exp(sig_ab + sig_cd) is below
// exp(lnp) are skipped, as the reference's Libint2 engine skips those below
// its precision (ScreeningMethod::Original at eps / max|P|,
// libint2_direct.cpp:279-292); the primitives of a pair are sorted by sig,
// so the kept ones form a prefix of both loops.  lnp = -inf keeps all.
// UNIFORM: the bra loop ends on a wave-wide vote instead of a per-lane break
// (sig must then be a __restrict__ kernel parameter, passed as sigp).  The
// bra primitive index is the same on every lane; with a per-lane break the
// loop exit is divergent and the compiler fetches sig with a vector load
// waited on in every bra iteration, and sig read through the argument struct
// is not known to be unchanged by the kernel's own stores.  With the vote,
// readfirstlane and the restrict pointer the bra data are scalar loads; a
// lane past its cut runs on with a zero prefactor (in SIMT a masked lane
// would cost the same).
// BW < BOYS_W: the one- and two-root classes read the Boys rows from boys
// (rows of BW values, e.g. staged in LDS) instead of ra.boys.
template <class C, int Q0, int Q1, typename R, int K0 = 0, int K1 = C::NKETC, bool UNIFORM = false,
          int BW = BOYS_W>
QFE_DI void contract_chunk(const PairHdr& bra, const PairHdr& ket,
                           const typename Vec4<R>::type* __restrict__ ppos, const R* __restrict__ pk,
                           const RysArgsT<R>& ra, R (&acc)[Q1 - Q0][K1 - K0], R lnp = R(-3.0e38),
                           const R* __restrict__ sigp = nullptr, const R* boys = nullptr) {
  constexpr int NK = K1 - K0;
  const R TWO_PI_52 = R(34.98683665524972);  // 2 pi^(5/2)
  sfor<Q1 - Q0>([&](auto q) { sfor<NK>([&](auto k) { acc[q][k] = R(0); }); });
  const R ABx = R(bra.A[0] - bra.B[0]), ABy = R(bra.A[1] - bra.B[1]), ABz = R(bra.A[2] - bra.B[2]);
  const R CDx = R(ket.A[0] - ket.B[0]), CDy = R(ket.A[1] - ket.B[1]), CDz = R(ket.A[2] - ket.B[2]);
  const R Ax = R(bra.A[0]), Ay = R(bra.A[1]), Az = R(bra.A[2]);
  const R Cx = R(ket.A[0]), Cy = R(ket.A[1]), Cz = R(ket.A[2]);
  const R* __restrict__ sig = UNIFORM ? sigp : ra.sig;
  const R sb0 = sig[bra.prim0];  // the bra's largest factor
  // ket primitives outer (per-lane data), bra primitives inner: the bra pair
  // is usually shared by the whole wave, so the inner loads are broadcasts
  // with UNIFORM, each iteration holds its primitive's data in registers
  // before the cut test, so that the loads share one round trip instead of
  // following each other across the test (a primitive past the cut still
  // lies in its pair's range: the reads stay in bounds)
  using V4 = typename Vec4<R>::type;
  for (int ik = ket.prim0; ik < ket.prim0 + ket.nprim; ++ik) {
    const R sk = sig[ik];
    V4 Q;
    R Kk, h2q;  // h2q: used from a ket (pp) or (ds) on
    if constexpr (UNIFORM) {
      Q = ppos[ik], Kk = pk[ik], h2q = ra.h2[ik];
      held_v(Q.x), held_v(Q.y), held_v(Q.z), held_v(Q.w), held_v(Kk);
      if constexpr (C::FC >= 2) held_v(h2q);
    }
    if (sk + sb0 < lnp) break;  // this and every later ket primitive
    if constexpr (!UNIFORM) Q = ppos[ik], Kk = pk[ik], h2q = ra.h2[ik];
    const R blim = lnp - sk;
    const R q = Q.x;
    const R QCx = Q.y - Cx, QCy = Q.z - Cy, QCz = Q.w - Cz;
    for (int ib0 = bra.prim0; ib0 < bra.prim0 + bra.nprim; ++ib0) {
      int ib = ib0;
      bool on = true;
      V4 P;
      R Kb, h2p;  // h2p: used from a bra (pp) or (ds) on
      if constexpr (UNIFORM) {
        ib = __builtin_amdgcn_readfirstlane(ib0);
        const R sgb = sig[ib];
        P = ppos[ib], Kb = pk[ib], h2p = ra.h2[ib];
        held_s(P.x), held_s(P.y), held_s(P.z), held_s(P.w), held_s(Kb);
        if constexpr (C::EA >= 2) held_s(h2p);
        on = !(sgb < blim);
        if (__ballot(on) == 0ull) break;  // past the cut on every lane
        Kb = on ? Kb : R(0);
      } else {
        if (sig[ib] < blim) break;  // this and every later bra primitive
        P = ppos[ib], Kb = pk[ib];
      }
      const R p = P.x;
      const R PAx = P.y - Ax, PAy = P.z - Ay, PAz = P.w - Az;
      const R PQx = P.y - Q.y, PQy = P.z - Q.z, PQz = P.w - Q.w;
      const R rs = frsq(p + q);
      const R rpq = rs * rs;
      const R rho = p * q * rpq;
      const R T = rho * (PQx * PQx + PQy * PQy + PQz * PQz);
      const R pref = TWO_PI_52 * Kb * Kk * rs;
      if constexpr (C::L == 1 && C::LA == 1) {
        // (ps|ss): one root, and the integral is linear in it; with the
        // weight w = F_0 and w u = F_1 no division by F_0 is needed:
        // (p_i s|ss) = pref (F_0 PA_i - q / (p + q) F_1 PQ_i)
        R F[2];
        boys_fm<1, R, BW>(T, BW == BOYS_W ? ra.boys : boys, F);
        const R a0 = pref * F[0], a1 = pref * F[1] * (q * rpq);
        const R PA[3] = {PAx, PAy, PAz}, PQ[3] = {PQx, PQy, PQz};
        sfor<Q1 - Q0>([&](auto qq) { acc[qq][0] += fma(a0, PA[Q0 + qq], -a1 * PQ[Q0 + qq]); });
        continue;
      }
      R u[C::NR], w[C::NR];
      rys_roots<C::NR, R, BW>(T, ra, u, w, boys);
      const R qr = q * rpq, pr = p * rpq;
      if constexpr (!UNIFORM) h2p = ra.h2[ib];
#pragma unroll 1
      for (int r = 0; r < C::NR; ++r) {
        const R ur = u[r];
        const R B00 = R(0.5) * rpq * ur;
        const R B10 = h2p - h2p * qr * ur;
        const R B01 = h2q - h2q * pr * ur;
        const R xa = qr * ur, xc = pr * ur;
        OneD<C, R> Ix, Iy, Iz;
        rys_1d<C, R>(R(1), PAx - xa * PQx, QCx + xc * PQx, B00, B10, B01, ABx, CDx, Ix);
        rys_1d<C, R>(R(1), PAy - xa * PQy, QCy + xc * PQy, B00, B10, B01, ABy, CDy, Iy);
        rys_1d<C, R>(w[r] * pref, PAz - xa * PQz, QCz + xc * PQz, B00, B10, B01, ABz, CDz, Iz);
        // assemble: bra in output functions, ket in cartesians
        sfor<Q1 - Q0>([&](auto qq) {
          constexpr int qi = Q0 + qq;
          constexpr int ma = qi / C::NSB, mb = qi % C::NSB;
          sfor<NK>([&](auto kk) {
            constexpr int cc = (K0 + kk) / C::NCD, dc = (K0 + kk) % C::NCD;
            constexpr int cx = cart_x(C::LC, cc), cy = cart_y(C::LC, cc), cz = cart_z(C::LC, cc);
            {
              constexpr int dx = cart_x(C::LD, dc), dy = cart_y(C::LD, dc), dz = cart_z(C::LD, dc);
              R s = R(0);
              sfor<C::NCA>([&](auto ac) {
                constexpr double ca = sph_coef(C::LA, C::PURE, ma, ac);
                if constexpr (ca != 0.0) {
                  constexpr int ax = cart_x(C::LA, ac), ay = cart_y(C::LA, ac), az = cart_z(C::LA, ac);
                  sfor<C::NCB>([&](auto bc) {
                    constexpr double cb = sph_coef(C::LB, C::PURE, mb, bc);
                    if constexpr (cb != 0.0) {
                      constexpr int bx = cart_x(C::LB, bc), by = cart_y(C::LB, bc), bz = cart_z(C::LB, bc);
                      const R v = Ix.v[ax][bx][cx][dx] * Iy.v[ay][by][cy][dy] * Iz.v[az][bz][cz][dz];
                      if constexpr (ca * cb == 1.0)
                        s += v;
                      else
                        s = fma(R(ca * cb), v, s);
                    }
                  });
                }
              });
              acc[qq][kk] += s;
            }
          });
        });
      }
    }
  }
}

// ket cartesian -> output functions for one bra function pair; with a ket
// cartesian range [K0, K1) the result is that range's share of every output
// function (the transform is linear, so the shares of all ranges add up)
template <class C, typename R, int K0 = 0, int K1 = C::NKETC>
QFE_DI void ket_to_sph(const R (&in)[K1 - K0], double (&out)[C::NKETS]) {
  sfor<C::NSC>([&](auto mc) {
    sfor<C::NSD>([&](auto md) {
      double s = 0.0;
      sfor<K1 - K0>([&](auto kk) {
        constexpr int cc = (K0 + kk) / C::NCD, dc = (K0 + kk) % C::NCD;
        constexpr double c = sph_coef(C::LC, C::PURE, mc, cc) * sph_coef(C::LD, C::PURE, md, dc);
        if constexpr (c != 0.0) s = fma(c, (double)in[kk], s);
      });
      out[mc * C::NSD + md] = s;
    });
  });
}

}  // namespace qfe
