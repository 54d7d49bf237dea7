// The dense AO two-electron tensor (ij|kl) on the GPU, for the quarter
// transform out[i][j][k][p] = sum_l (ij|kl) C[l][p] that the reference's
// ERI::quarter_trans computes (libint2_direct.cpp:913-1141) for the MO
// transforms of scf/src/eri/moeri.cpp.  The tensor is N^4 doubles, so this
// path is for small systems (N of a few hundred at most; 288 GB of HBM holds
// N = 400 with room to spare).  engine.cpp applies C with one rocBLAS GEMM.
//
// Rys quadrature with runtime angular momenta (shells up to d): per
// primitive quartet and root, the 2-D recursion I(i, k) for i <= la + lb and
// k <= lc + ld, horizontal transfers to (a, b) and (c, d), then the
// solid-harmonic transforms.  One thread per canonical shell quartet stores
// its block at the eight symmetry-equivalent places; every element has one
// canonical quartet, so no atomics are needed.  The work is O(nsh^4 / 8).
#include <hip/hip_runtime.h>

#include "dfj.hpp"
#include "eri_core.hpp"
#include "eri_dense.hpp"

namespace qfe {

namespace {

constexpr double kTwoPi52d = 34.98683665524972;  // 2 pi^(5/2)
constexpr int kL = 2;                             // shells up to d
constexpr int kL2 = 2 * kL;
constexpr int kNC = (kL + 1) * (kL + 2) / 2;
constexpr int kBlk = kNC * kNC * kNC * kNC;

struct Sh {
  int l, n;
  const double* e;
  const double* c;
  const double* O;
};

__device__ int dn_ncart(int l) { return (l + 1) * (l + 2) / 2; }

__device__ void dn_cart_xyz(int l, int c, int& x, int& y, int& z) {
  int k = 0;
  while ((k + 1) * (k + 2) / 2 <= c) ++k;
  x = l - k;
  z = c - k * (k + 1) / 2;
  y = l - x - z;
}

__device__ void dn_rys(int n, double T, const RysArgs& ra, double* u, double* w) {
  switch (n) {
#define QFE_DN_ROOTS(N)          \
  case N: {                      \
    double uu[N], ww[N];         \
    rys_roots<N>(T, ra, uu, ww); \
    for (int i = 0; i < N; ++i) { \
      u[i] = uu[i];              \
      w[i] = ww[i];              \
    }                            \
    break;                       \
  }
    QFE_DN_ROOTS(1)
    QFE_DN_ROOTS(2)
    QFE_DN_ROOTS(3)
    QFE_DN_ROOTS(4)
    default:
      QFE_DN_ROOTS(5)
#undef QFE_DN_ROOTS
  }
}

// cartesian block (a b | c d): out[na][nb][nc][nd]
__device__ void dn_cart(const Sh& a, const Sh& b, const Sh& c, const Sh& d, const RysArgs& ra, double* out) {
  const int la = a.l, lb = b.l, lc = c.l, ld = d.l, lab = la + lb, lcd = lc + ld;
  const int na = dn_ncart(la), nb = dn_ncart(lb), nc = dn_ncart(lc), nd = dn_ncart(ld);
  for (int i = 0; i < na * nb * nc * nd; ++i) out[i] = 0.0;
  double AB[3], CD[3], AB2 = 0.0, CD2 = 0.0;
  for (int x = 0; x < 3; ++x) {
    AB[x] = a.O[x] - b.O[x];
    CD[x] = c.O[x] - d.O[x];
    AB2 += AB[x] * AB[x];
    CD2 += CD[x] * CD[x];
  }
  const int nroots = (lab + lcd) / 2 + 1;
  // cartesian exponents of every component, decoded once per quartet
  int ex[4][3][kNC];
  {
    const int ls[4] = {la, lb, lc, ld};
    for (int q = 0; q < 4; ++q)
      for (int c = 0; c < dn_ncart(ls[q]); ++c) dn_cart_xyz(ls[q], c, ex[q][0][c], ex[q][1][c], ex[q][2][c]);
  }
  for (int ia = 0; ia < a.n; ++ia)
    for (int ib = 0; ib < b.n; ++ib) {
      const double p = a.e[ia] + b.e[ib];
      const double Kab = a.c[ia] * b.c[ib] * exp(-a.e[ia] * b.e[ib] / p * AB2) / p;
      if (Kab == 0.0) continue;
      double P[3], PA[3];
      for (int x = 0; x < 3; ++x) {
        P[x] = (a.e[ia] * a.O[x] + b.e[ib] * b.O[x]) / p;
        PA[x] = P[x] - a.O[x];
      }
      for (int ic = 0; ic < c.n; ++ic)
        for (int id = 0; id < d.n; ++id) {
          const double q = c.e[ic] + d.e[id];
          const double Kcd = c.c[ic] * d.c[id] * exp(-c.e[ic] * d.e[id] / q * CD2) / q;
          if (Kcd == 0.0) continue;
          double Q[3], QC[3], PQ[3], PQ2 = 0.0;
          for (int x = 0; x < 3; ++x) {
            Q[x] = (c.e[ic] * c.O[x] + d.e[id] * d.O[x]) / q;
            QC[x] = Q[x] - c.O[x];
            PQ[x] = P[x] - Q[x];
            PQ2 += PQ[x] * PQ[x];
          }
          const double rpq = 1.0 / (p + q);
          const double T = p * q * rpq * PQ2;
          const double pref = kTwoPi52d * Kab * Kcd * sqrt(rpq);
          double u[5], w[5];
          dn_rys(nroots, T, ra, u, w);
          const double h2p = 0.5 / p, h2q = 0.5 / q, qr = q * rpq, pr = p * rpq;
          for (int r = 0; r < nroots; ++r) {
            const double ur = u[r];
            const double B00 = 0.5 * rpq * ur, B10 = h2p - h2p * qr * ur, B01 = h2q - h2q * pr * ur;
            double I[3][kL + 1][kL + 1][kL + 1][kL + 1];
            for (int x = 0; x < 3; ++x) {
              const double C00 = PA[x] - qr * ur * PQ[x], D00 = QC[x] + pr * ur * PQ[x];
              double G[kL2 + 1][kL2 + 1];
              G[0][0] = (x == 2) ? w[r] * pref : 1.0;
              for (int i = 0; i < lab; ++i) G[i + 1][0] = C00 * G[i][0] + (i > 0 ? i * B10 * G[i - 1][0] : 0.0);
              for (int k = 0; k < lcd; ++k)
                for (int i = 0; i <= lab; ++i) {
                  double v = D00 * G[i][k];
                  if (k > 0) v += k * B01 * G[i][k - 1];
                  if (i > 0) v += i * B00 * G[i - 1][k];
                  G[i][k + 1] = v;
                }
              // bra transfer (a, b+1) = (a+1, b) + AB (a, b), for every k
              double H[kL + 1][kL + 1][kL2 + 1];
              for (int k = 0; k <= lcd; ++k) {
                double t[kL2 + 1][kL + 1];
                for (int i = 0; i <= lab; ++i) t[i][0] = G[i][k];
                for (int j = 1; j <= lb; ++j)
                  for (int i = 0; i <= lab - j; ++i) t[i][j] = t[i + 1][j - 1] + AB[x] * t[i][j - 1];
                for (int i = 0; i <= la; ++i)
                  for (int j = 0; j <= lb; ++j) H[i][j][k] = t[i][j];
              }
              // ket transfer (c, d+1) = (c+1, d) + CD (c, d)
              for (int i = 0; i <= la; ++i)
                for (int j = 0; j <= lb; ++j) {
                  double t[kL2 + 1][kL + 1];
                  for (int k = 0; k <= lcd; ++k) t[k][0] = H[i][j][k];
                  for (int m = 1; m <= ld; ++m)
                    for (int k = 0; k <= lcd - m; ++k) t[k][m] = t[k + 1][m - 1] + CD[x] * t[k][m - 1];
                  for (int k = 0; k <= lc; ++k)
                    for (int m = 0; m <= ld; ++m) I[x][i][j][k][m] = t[k][m];
                }
            }
            for (int ca = 0; ca < na; ++ca)
              for (int cb = 0; cb < nb; ++cb)
                for (int cc = 0; cc < nc; ++cc)
                  for (int cd = 0; cd < nd; ++cd)
                    out[((ca * nb + cb) * nc + cc) * nd + cd] +=
                        I[0][ex[0][0][ca]][ex[1][0][cb]][ex[2][0][cc]][ex[3][0][cd]] *
                        I[1][ex[0][1][ca]][ex[1][1][cb]][ex[2][1][cc]][ex[3][1][cd]] *
                        I[2][ex[0][2][ca]][ex[1][2][cb]][ex[2][2][cc]][ex[3][2][cd]];
          }
        }
    }
}

// one index of a 4-index block to output functions: dims n[4] (cartesian in
// position `pos`), result with that index pure when `pure` and l >= 2
__device__ void dn_transform(const double* sph, const int (&n)[4], int pos, int l, int pure, const double* in,
                             double* out, int& nout) {
  const bool p = pure && l >= 2;
  nout = p ? 2 * l + 1 : n[pos];
  int outer = 1, inner = 1;
  for (int q = 0; q < pos; ++q) outer *= n[q];
  for (int q = pos + 1; q < 4; ++q) inner *= n[q];
  for (int o = 0; o < outer; ++o)
    for (int m = 0; m < nout; ++m)
      for (int i = 0; i < inner; ++i) {
        double s = 0.0;
        if (p)
          for (int c = 0; c < n[pos]; ++c) s += sph[(l * 9 + m) * 15 + c] * in[(o * n[pos] + c) * inner + i];
        else
          s = in[(o * n[pos] + m) * inner + i];
        out[(o * nout + m) * inner + i] = s;
      }
}

__device__ Sh dn_shell(const DFBasisView& v, int s) {
  return Sh{v.l[s], v.po[s + 1] - v.po[s], v.exps + v.po[s], v.coefs + v.po[s], v.cen + 3 * s};
}

}  // namespace

// row r >= column c of the lower triangle holding linear index t
__device__ void dn_tri(long long t, int& r, int& c) {
  int x = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) / 2.0);
  while ((long long)x * (x + 1) / 2 > t) --x;
  while ((long long)(x + 1) * (x + 2) / 2 <= t) ++x;
  r = x;
  c = (int)(t - (long long)x * (x + 1) / 2);
}

// One thread per unordered pair of shell pairs (p12 >= p34 over the list
// `pairs`, which holds every unordered shell pair once, sorted by class and
// contraction length so that a wave's quartets mostly share their class):
// the block is computed once and stored at its eight
// permutation-equivalent places (ab|cd) = (ba|cd) = (ab|dc) = (cd|ab) ...;
// where shells coincide two of the stores hit the same element with the same
// value.
__global__ __launch_bounds__(256) void k_eri_dense(DFBasisView b, RysArgs ra, const double* __restrict__ sph,
                                                   const int2* __restrict__ pairs, double* __restrict__ eri) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long np = (long long)b.nsh * (b.nsh + 1) / 2;
  if (t >= np * (np + 1) / 2) return;
  int p12, p34;
  dn_tri(t, p12, p34);
  const int2 u = pairs[p12], v = pairs[p34];
  const int s1 = u.x, s2 = u.y, s3 = v.x, s4 = v.y;
  const Sh A = dn_shell(b, s1), B = dn_shell(b, s2), C = dn_shell(b, s3), D = dn_shell(b, s4);
  double x0[kBlk], x1[kBlk];
  dn_cart(A, B, C, D, ra, x0);
  int n[4] = {dn_ncart(A.l), dn_ncart(B.l), dn_ncart(C.l), dn_ncart(D.l)};
  const int ls[4] = {A.l, B.l, C.l, D.l};
  double* src = x0;
  double* dst = x1;
  for (int pos = 0; pos < 4; ++pos) {
    int m = 0;
    dn_transform(sph, n, pos, ls[pos], b.pure, src, dst, m);
    n[pos] = m;
    double* tmp = src;
    src = dst;
    dst = tmp;
  }
  const size_t N = (size_t)b.nbf;
  const int f1 = b.bf0[s1], f2 = b.bf0[s2], f3 = b.bf0[s3], f4 = b.bf0[s4];
  for (int i = 0; i < n[0]; ++i)
    for (int j = 0; j < n[1]; ++j) {
      const size_t ab = (size_t)(f1 + i) * N + f2 + j, ba = (size_t)(f2 + j) * N + f1 + i;
      for (int k = 0; k < n[2]; ++k)
        for (int l = 0; l < n[3]; ++l) {
          const size_t cd = (size_t)(f3 + k) * N + f4 + l, dc = (size_t)(f4 + l) * N + f3 + k;
          const double v = src[((i * n[1] + j) * n[2] + k) * n[3] + l];
          const size_t NN = N * N;
          eri[ab * NN + cd] = v;
          eri[ba * NN + cd] = v;
          eri[ab * NN + dc] = v;
          eri[ba * NN + dc] = v;
          eri[cd * NN + ab] = v;
          eri[dc * NN + ab] = v;
          eri[cd * NN + ba] = v;
          eri[dc * NN + ba] = v;
        }
    }
}

void launch_eri_dense(const DFBasisView& b, const RysArgs& ra, const double* sph, const int2* pairs, double* eri,
                      hipStream_t s) {
  const long long np = (long long)b.nsh * (b.nsh + 1) / 2;
  const long long n = np * (np + 1) / 2;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_eri_dense, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, b, ra, sph, pairs, eri);
}

}  // namespace qfe
