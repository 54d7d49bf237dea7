// Adjoint nuclear-gradient kernel for the ERI classes with f shells, with
// runtime angular momenta (shells up to f): the same contraction as the
// compile-time k_eri_grad (eri_grad.hpp), the reference's ERI::get_gradients
// (libint2_direct.cpp:646-890) over the canonical task list,
//   dJ[x] += 1/2 deg sum P_ab P_cd d(ab|cd)/dx,
//   dK[x] += wk deg sum_s (P_ac P_bd + P_ad P_bc) d(ab|cd)/dx,
// wk = -1/8 (alpha + beta) for one total density, -1/4 per spin of a pair.
//
// These classes are a small share of any build's quartets (the f shells of
// cc-pVTZ carbon), while compile-time instances of them need hundreds of
// spill-free chunk kernels each; one runtime kernel keeps the build short.
// One thread per task: per primitive quartet and Rys root it builds the 1-D
// integrals I(a, b, c, d) with one extra unit on A, B and C (2-D recursion,
// then the two horizontal transfers), then walks the cartesian components,
// forming the density weights on the fly from the cartesian blocks of P.
#include <hip/hip_runtime.h>

#include <stdexcept>

#include "eri_core.hpp"
#include "kernels.hpp"

namespace qfe {

namespace {

constexpr int kGL = 3;               // shells up to f
constexpr int kGNC = (kGL + 1) * (kGL + 2) / 2;
constexpr int kGE = 2 * kGL + 1;     // bra (ket) recursion extent with the derivative unit

__device__ int ga_ncart(int l) { return (l + 1) * (l + 2) / 2; }

__device__ void ga_xyz(int l, int c, int& x, int& y, int& z) {
  int k = 0;
  while ((k + 1) * (k + 2) / 2 <= c) ++k;
  x = l - k;
  z = c - k * (k + 1) / 2;
  y = l - x - z;
}

// Rys roots and weights for a runtime root count (1..7)
__device__ void ga_rys(int n, double T, const RysArgs& ra, double* u, double* w) {
  switch (n) {
#define QFE_GA_ROOTS(N)          \
  case N: {                      \
    double uu[N], ww[N];         \
    rys_roots<N>(T, ra, uu, ww); \
    for (int i = 0; i < N; ++i) { \
      u[i] = uu[i];              \
      w[i] = ww[i];              \
    }                            \
    break;                       \
  }
    QFE_GA_ROOTS(1)
    QFE_GA_ROOTS(2)
    QFE_GA_ROOTS(3)
    QFE_GA_ROOTS(4)
    QFE_GA_ROOTS(5)
    QFE_GA_ROOTS(6)
    default:
      QFE_GA_ROOTS(7)
#undef QFE_GA_ROOTS
  }
}

// cartesian block of the density: out[x][y] = sum_{m,k} S_lx[m][x] S_ly[k][y]
// P[bx + m][by + k] (solid harmonics for l >= 2, identity below)
__device__ void ga_cart_density(const double* __restrict__ P, size_t n, const double* __restrict__ sph, int lx,
                                int bx, int ly, int by, double (&out)[kGNC][kGNC]) {
  const int ncx = ga_ncart(lx), ncy = ga_ncart(ly);
  const int nsx = lx >= 2 ? 2 * lx + 1 : ncx, nsy = ly >= 2 ? 2 * ly + 1 : ncy;
  for (int x = 0; x < ncx; ++x)
    for (int y = 0; y < ncy; ++y) {
      double s = 0.0;
      for (int m = 0; m < nsx; ++m) {
        const double cx = lx >= 2 ? sph[(lx * 9 + m) * 15 + x] : (m == x ? 1.0 : 0.0);
        if (cx == 0.0) continue;
        for (int k = 0; k < nsy; ++k) {
          const double cy = ly >= 2 ? sph[(ly * 9 + k) * 15 + y] : (k == y ? 1.0 : 0.0);
          if (cy != 0.0) s = fma(cx * cy, P[(size_t)(bx + m) * n + by + k], s);
        }
      }
      out[x][y] = s;
    }
}

}  // namespace

__global__ __launch_bounds__(64) void k_eri_grad_any(const int2* __restrict__ tasks, long long ntask,
                                                     const PairHdr* __restrict__ pairs,
                                                     const double4* __restrict__ ppos, const double* __restrict__ pk,
                                                     const double* __restrict__ pa, RysArgs ra,
                                                     const double* __restrict__ sph, const double* __restrict__ P,
                                                     const double* __restrict__ Pk, const double* __restrict__ Pk2,
                                                     int nbf, const int* __restrict__ sh_atom, int natoms,
                                                     double kscale, int la, int lb, int lc, int ld,
                                                     double* __restrict__ gJ, double* __restrict__ gK) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ntask) return;
  const int2 tk = tasks[t];
  const PairHdr bra = pairs[tk.x];
  const PairHdr ket = pairs[tk.y];
  const size_t n = (size_t)nbf;
  const double deg = (bra.sa == bra.sb ? 1.0 : 2.0) * (ket.sa == ket.sb ? 1.0 : 2.0) * (tk.x == tk.y ? 1.0 : 2.0);
  const double wj = 0.5 * deg;
  const double wk = (Pk2 ? -0.25 : -0.125) * kscale * deg;
  const int na = ga_ncart(la), nb = ga_ncart(lb), nc = ga_ncart(lc), nd = ga_ncart(ld);
  // cartesian density blocks: Coulomb (ab), (cd); exchange (ac), (bd), (ad), (bc) per spin
  double Pab[kGNC][kGNC], Pcd[kGNC][kGNC];
  ga_cart_density(P, n, sph, la, bra.bfa, lb, bra.bfb, Pab);
  ga_cart_density(P, n, sph, lc, ket.bfa, ld, ket.bfb, Pcd);
  double X[2][4][kGNC][kGNC];
  const int nspin = Pk2 ? 2 : 1;
  for (int sp = 0; sp < nspin; ++sp) {
    const double* D = sp ? Pk2 : Pk;
    ga_cart_density(D, n, sph, la, bra.bfa, lc, ket.bfa, X[sp][0]);  // ac
    ga_cart_density(D, n, sph, lb, bra.bfb, ld, ket.bfb, X[sp][1]);  // bd
    ga_cart_density(D, n, sph, la, bra.bfa, ld, ket.bfb, X[sp][2]);  // ad
    ga_cart_density(D, n, sph, lb, bra.bfb, lc, ket.bfa, X[sp][3]);  // bc
  }
  // cartesian exponents, decoded once
  int ea[kGNC][3], eb[kGNC][3], ec[kGNC][3], ed[kGNC][3];
  for (int c = 0; c < na; ++c) ga_xyz(la, c, ea[c][0], ea[c][1], ea[c][2]);
  for (int c = 0; c < nb; ++c) ga_xyz(lb, c, eb[c][0], eb[c][1], eb[c][2]);
  for (int c = 0; c < nc; ++c) ga_xyz(lc, c, ec[c][0], ec[c][1], ec[c][2]);
  for (int c = 0; c < nd; ++c) ga_xyz(ld, c, ed[c][0], ed[c][1], ed[c][2]);
  double gj[3][3] = {}, gk[3][3] = {};
  const double AB[3] = {bra.A[0] - bra.B[0], bra.A[1] - bra.B[1], bra.A[2] - bra.B[2]};
  const double CD[3] = {ket.A[0] - ket.B[0], ket.A[1] - ket.B[1], ket.A[2] - ket.B[2]};
  const int E1 = la + lb + 1, F1 = lc + ld + 1;
  const int nroots = (la + lb + lc + ld + 1) / 2 + 1;
  constexpr double TWO_PI_52 = 34.98683665524972;
  for (int ik = ket.prim0; ik < ket.prim0 + ket.nprim; ++ik) {
    const double4 Q = ppos[ik];
    const double Kk = pk[ik], q = Q.x;
    const double tg = 2.0 * pa[ik];  // 2 gamma
    const double QC[3] = {Q.y - ket.A[0], Q.z - ket.A[1], Q.w - ket.A[2]};
    const double h2q = 0.5 / q;
    for (int ib = bra.prim0; ib < bra.prim0 + bra.nprim; ++ib) {
      const double4 Pp = ppos[ib];
      const double Kb = pk[ib], p = Pp.x;
      const double ta = 2.0 * pa[ib], tb = 2.0 * (p - pa[ib]);  // 2 alpha, 2 beta
      const double PA[3] = {Pp.y - bra.A[0], Pp.z - bra.A[1], Pp.w - bra.A[2]};
      const double PQ[3] = {Pp.y - Q.y, Pp.z - Q.z, Pp.w - Q.w};
      const double rpq = 1.0 / (p + q);
      const double T = p * q * rpq * (PQ[0] * PQ[0] + PQ[1] * PQ[1] + PQ[2] * PQ[2]);
      const double pref = TWO_PI_52 * Kb * Kk * sqrt(rpq);
      double u[7], w[7];
      ga_rys(nroots, T, ra, u, w);
      const double qr = q * rpq, pr = p * rpq, h2p = 0.5 / p;
      for (int r = 0; r < nroots; ++r) {
        const double ur = u[r];
        const double B00 = 0.5 * rpq * ur, B10 = h2p - h2p * qr * ur, B01 = h2q - h2q * pr * ur;
        // I[x][a][b][c][d]: a <= la+1, b <= lb+1, c <= lc+1, d <= ld
        double I[3][kGL + 2][kGL + 2][kGL + 2][kGL + 1];
        for (int x = 0; x < 3; ++x) {
          const double C00 = PA[x] - qr * ur * PQ[x], D00 = QC[x] + pr * ur * PQ[x];
          double G[kGE + 1][kGE + 1];
          G[0][0] = (x == 2) ? w[r] * pref : 1.0;
          for (int i = 0; i < E1; ++i) G[i + 1][0] = C00 * G[i][0] + (i > 0 ? i * B10 * G[i - 1][0] : 0.0);
          for (int k = 0; k < F1; ++k)
            for (int i = 0; i <= E1; ++i) {
              double v = D00 * G[i][k];
              if (k > 0) v += k * B01 * G[i][k - 1];
              if (i > 0) v += i * B00 * G[i - 1][k];
              G[i][k + 1] = v;
            }
          // bra transfer (a, b+1) = (a+1, b) + AB (a, b): H[a][b][k], a + b <= E1
          double H[kGL + 2][kGL + 2][kGE + 1];
          for (int k = 0; k <= F1; ++k) {
            double tt[kGE + 1][kGL + 2];
            for (int i = 0; i <= E1; ++i) tt[i][0] = G[i][k];
            for (int j = 1; j <= lb + 1; ++j)
              for (int i = 0; i <= E1 - j; ++i) tt[i][j] = tt[i + 1][j - 1] + AB[x] * tt[i][j - 1];
            for (int i = 0; i <= la + 1; ++i)
              for (int j = 0; j <= lb + 1 && i + j <= E1; ++j) H[i][j][k] = tt[i][j];
          }
          // ket transfer (c, d+1) = (c+1, d) + CD (c, d): c + d <= F1
          for (int i = 0; i <= la + 1; ++i)
            for (int j = 0; j <= lb + 1 && i + j <= E1; ++j) {
              double tt[kGE + 1][kGL + 1];
              for (int k = 0; k <= F1; ++k) tt[k][0] = H[i][j][k];
              for (int m = 1; m <= ld; ++m)
                for (int k = 0; k <= F1 - m; ++k) tt[k][m] = tt[k + 1][m - 1] + CD[x] * tt[k][m - 1];
              for (int k = 0; k <= lc + 1; ++k)
                for (int m = 0; m <= ld && k + m <= F1; ++m) I[x][i][j][k][m] = tt[k][m];
            }
        }
        for (int ca = 0; ca < na; ++ca)
          for (int cb = 0; cb < nb; ++cb)
            for (int cc = 0; cc < nc; ++cc)
              for (int cd = 0; cd < nd; ++cd) {
                double wK = 0.0;
                for (int sp = 0; sp < nspin; ++sp)
                  wK += X[sp][0][ca][cc] * X[sp][1][cb][cd] + X[sp][2][ca][cd] * X[sp][3][cb][cc];
                const double wJ = wj * Pab[ca][cb] * Pcd[cc][cd];
                wK *= wk;
                double v[3];
                for (int x = 0; x < 3; ++x) v[x] = I[x][ea[ca][x]][eb[cb][x]][ec[cc][x]][ed[cd][x]];
                for (int x = 0; x < 3; ++x) {
                  const int y = (x + 1) % 3, z = (x + 2) % 3;
                  const double o = v[y] * v[z];
                  const int a = ea[ca][x], b = eb[cb][x], c = ec[cc][x], d = ed[cd][x];
                  double dA = ta * I[x][a + 1][b][c][d];
                  if (a > 0) dA -= a * I[x][a - 1][b][c][d];
                  double dB = tb * I[x][a][b + 1][c][d];
                  if (b > 0) dB -= b * I[x][a][b - 1][c][d];
                  double dC = tg * I[x][a][b][c + 1][d];
                  if (c > 0) dC -= c * I[x][a][b][c - 1][d];
                  gj[0][x] = fma(dA, wJ * o, gj[0][x]);
                  gk[0][x] = fma(dA, wK * o, gk[0][x]);
                  gj[1][x] = fma(dB, wJ * o, gj[1][x]);
                  gk[1][x] = fma(dB, wK * o, gk[1][x]);
                  gj[2][x] = fma(dC, wJ * o, gj[2][x]);
                  gk[2][x] = fma(dC, wK * o, gk[2][x]);
                }
              }
      }
    }
  }
  const int at[4] = {sh_atom[bra.sa], sh_atom[bra.sb], sh_atom[ket.sa], sh_atom[ket.sb]};
  for (int x = 0; x < 3; ++x) {
    const double dj[4] = {gj[0][x], gj[1][x], gj[2][x], -(gj[0][x] + gj[1][x] + gj[2][x])};
    const double dk[4] = {gk[0][x], gk[1][x], gk[2][x], -(gk[0][x] + gk[1][x] + gk[2][x])};
    for (int c = 0; c < 4; ++c) {
      unsafeAtomicAdd(&gJ[at[c] + x * natoms], dj[c]);
      unsafeAtomicAdd(&gK[at[c] + x * natoms], dk[c]);
    }
  }
}

void launch_grad_any(const GradLaunch& a, int la, int lb, int lc, int ld, hipStream_t s) {
  if (a.ntask <= 0) return;
  if (la > kGL || lb > kGL || lc > kGL || ld > kGL) throw std::runtime_error("gradient class above f");
  if (a.sph == nullptr) throw std::runtime_error("solid-harmonic table missing for the runtime gradient kernel");
  const int block = 64;
  hipLaunchKernelGGL(k_eri_grad_any, dim3((unsigned)((a.ntask + block - 1) / block)), dim3(block), 0, s, a.tasks,
                     a.ntask, a.pairs, a.ppos, a.pk, a.pa, *a.ra, a.sph, a.P, a.Pk, a.Pk2, a.nbf, a.sh_atom, a.natoms,
                     a.kscale, la, lb, lc, ld, a.gJ, a.gK);
}

}  // namespace qfe
