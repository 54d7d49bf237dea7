// Density-fitted Coulomb (DF-J) on the GPU: the three-centre integrals
// B[P][mu nu] = (mu nu|P) over an auxiliary basis, mu >= nu only (the packed
// lower triangle: half the HBM and half the bytes a build reads, against the
// full square of the reference's INCORE DF engine,
// scf/src/eri/INCORE/incore_impl_df.cpp:53-80), and the two-centre metric
// V[P][Q] = (P|Q), kept in HBM.  The
// build itself (X = B . vec(D), V c = X, J = B^T c; incore_impl_df.cpp:
// 81-200) runs in engine.cpp with rocBLAS and rocSOLVER.
//
// The integrals are Rys quadrature with runtime angular momenta (orbital
// shells up to d, auxiliary shells up to g): per primitive triple and root,
// 2-D recursions I(i, k) for i <= la + lb on the bra and k <= lc on the
// auxiliary centre, the transfer to (a, b), and the solid-harmonic
// transforms.  An auxiliary function is a ket pair (c, s0) with s0 an s
// function of exponent 0 at the same centre, so q = gamma, Q = C and the
// ket transfer term vanishes.  The work is far below the Fock build's; the
// kernels are plain code (one thread per block).
#include <hip/hip_runtime.h>

#include "dfj.hpp"
#include "eri_core.hpp"

namespace qfe {

namespace {

constexpr double kTwoPi52 = 34.98683665524972;  // 2 pi^(5/2)
constexpr int kMaxLo = 2;                        // orbital shells up to d
constexpr int kMaxLx = 4;                        // auxiliary shells up to g
constexpr int kMaxAB = 2 * kMaxLo;               // bra recursion extent (also an auxiliary shell with s0)
constexpr int kNCo = (kMaxLo + 1) * (kMaxLo + 2) / 2;
constexpr int kNCx = (kMaxLx + 1) * (kMaxLx + 2) / 2;

struct ShellRef {
  int l, n;          // angular momentum, primitives
  const double* e;   // exponents
  const double* c;   // contraction coefficients (normalised)
  const double* O;   // centre
};

__device__ int df_ncart(int l) { return (l + 1) * (l + 2) / 2; }

__device__ void df_cart_xyz(int l, int c, int& x, int& y, int& z) {
  int k = 0;
  while ((k + 1) * (k + 2) / 2 <= c) ++k;
  x = l - k;
  z = c - k * (k + 1) / 2;
  y = l - x - z;
}

__device__ void df_rys(int n, double T, const RysArgs& ra, double* u, double* w) {
  switch (n) {
#define QFE_DF_ROOTS(N)                 \
  case N: {                             \
    double uu[N], ww[N];                \
    rys_roots<N>(T, ra, uu, ww);        \
    for (int i = 0; i < N; ++i) {       \
      u[i] = uu[i];                     \
      w[i] = ww[i];                     \
    }                                   \
    break;                              \
  }
    QFE_DF_ROOTS(1)
    QFE_DF_ROOTS(2)
    QFE_DF_ROOTS(3)
    QFE_DF_ROOTS(4)
    QFE_DF_ROOTS(5)
    default:
      QFE_DF_ROOTS(6)
#undef QFE_DF_ROOTS
  }
}

// cartesian block (a b | c): out[na][nb][nc]
__device__ void df_cart(const ShellRef& a, const ShellRef& b, const ShellRef& c, const RysArgs& ra,
                        double* out) {
  const int la = a.l, lb = b.l, lc = c.l, lab = la + lb;
  const int na = df_ncart(la), nb = df_ncart(lb), nc = df_ncart(lc);
  for (int i = 0; i < na * nb * nc; ++i) out[i] = 0.0;
  double AB[3], AB2 = 0.0;
  for (int d = 0; d < 3; ++d) {
    AB[d] = a.O[d] - b.O[d];
    AB2 += AB[d] * AB[d];
  }
  const int nroots = (lab + lc) / 2 + 1;
  for (int ia = 0; ia < a.n; ++ia)
    for (int ib = 0; ib < b.n; ++ib) {
      const double p = a.e[ia] + b.e[ib];
      const double Kab = a.c[ia] * b.c[ib] * exp(-a.e[ia] * b.e[ib] / p * AB2) / p;
      if (Kab == 0.0) continue;
      double P[3], PA[3];
      for (int d = 0; d < 3; ++d) {
        P[d] = (a.e[ia] * a.O[d] + b.e[ib] * b.O[d]) / p;
        PA[d] = P[d] - a.O[d];
      }
      for (int ic = 0; ic < c.n; ++ic) {
        const double q = c.e[ic];
        const double Kc = c.c[ic] / q;
        double PQ[3], PQ2 = 0.0;
        for (int d = 0; d < 3; ++d) {
          PQ[d] = P[d] - c.O[d];
          PQ2 += PQ[d] * PQ[d];
        }
        const double rpq = 1.0 / (p + q);
        const double T = p * q * rpq * PQ2;
        const double pref = kTwoPi52 * Kab * Kc * sqrt(rpq);
        double u[6], w[6];
        df_rys(nroots, T, ra, u, w);
        const double h2p = 0.5 / p, h2q = 0.5 / q, qr = q * rpq, pr = p * rpq;
        for (int r = 0; r < nroots; ++r) {
          const double ur = u[r];
          const double B00 = 0.5 * rpq * ur, B10 = h2p - h2p * qr * ur, B01 = h2q - h2q * pr * ur;
          double I[3][kMaxAB + 1][kMaxLo + 1][kMaxLx + 1];  // first bra index up to la <= lab
          for (int d = 0; d < 3; ++d) {
            const double C00 = PA[d] - qr * ur * PQ[d], D00 = pr * ur * PQ[d];
            double G[kMaxAB + 1][kMaxLx + 1];
            G[0][0] = (d == 2) ? w[r] * pref : 1.0;
            for (int i = 0; i < lab; ++i) G[i + 1][0] = C00 * G[i][0] + (i > 0 ? i * B10 * G[i - 1][0] : 0.0);
            for (int k = 0; k < lc; ++k)
              for (int i = 0; i <= lab; ++i) {
                double v = D00 * G[i][k];
                if (k > 0) v += k * B01 * G[i][k - 1];
                if (i > 0) v += i * B00 * G[i - 1][k];
                G[i][k + 1] = v;
              }
            // transfer to the second bra centre: (a, b+1) = (a+1, b) + AB (a, b)
            for (int k = 0; k <= lc; ++k) {
              double t[kMaxAB + 1][kMaxLo + 1];
              for (int i = 0; i <= lab; ++i) t[i][0] = G[i][k];
              for (int j = 1; j <= lb; ++j)
                for (int i = 0; i <= lab - j; ++i) t[i][j] = t[i + 1][j - 1] + AB[d] * t[i][j - 1];
              for (int i = 0; i <= la; ++i)
                for (int j = 0; j <= lb; ++j) I[d][i][j][k] = t[i][j];
            }
          }
          for (int x = 0; x < na; ++x) {
            int ax, ay, az;
            df_cart_xyz(la, x, ax, ay, az);
            for (int y = 0; y < nb; ++y) {
              int bx, by, bz;
              df_cart_xyz(lb, y, bx, by, bz);
              for (int z = 0; z < nc; ++z) {
                int cx, cy, cz;
                df_cart_xyz(lc, z, cx, cy, cz);
                out[(x * nb + y) * nc + z] += I[0][ax][bx][cx] * I[1][ay][by][cy] * I[2][az][bz][cz];
              }
            }
          }
        }
      }
    }
}

// output functions of a shell: solid harmonics from d on when pure
__device__ int df_nfun(int l, int pure) { return (pure && l >= 2) ? 2 * l + 1 : df_ncart(l); }

// block [na][nb][nc] cartesian -> [ma][mb][mc] output functions
__device__ void df_to_out(const double* sph, int la, int lb, int lc, int pure_o, int pure_x, const double* in,
                          double* out) {
  const int na = df_ncart(la), nb = df_ncart(lb), nc = df_ncart(lc);
  const int ma = df_nfun(la, pure_o), mb = df_nfun(lb, pure_o), mc = df_nfun(lc, pure_x);
  const bool pa = pure_o && la >= 2, pb = pure_o && lb >= 2, pc = pure_x && lc >= 2;
  double t1[kNCo * kNCo * kNCx];
  // c index
  for (int x = 0; x < na * nb; ++x)
    for (int m = 0; m < mc; ++m) {
      double s = 0.0;
      if (pc)
        for (int z = 0; z < nc; ++z) s += sph[(lc * 9 + m) * 15 + z] * in[x * nc + z];
      else
        s = in[x * nc + m];
      t1[x * mc + m] = s;
    }
  // b index: t1[na][nb][mc] -> t2[na][mb][mc]
  double t2[kNCo * kNCo * kNCx];
  for (int x = 0; x < na; ++x)
    for (int m = 0; m < mb; ++m)
      for (int z = 0; z < mc; ++z) {
        double s = 0.0;
        if (pb)
          for (int y = 0; y < nb; ++y) s += sph[(lb * 9 + m) * 15 + y] * t1[(x * nb + y) * mc + z];
        else
          s = t1[(x * nb + m) * mc + z];
        t2[(x * mb + m) * mc + z] = s;
      }
  // a index
  for (int m = 0; m < ma; ++m)
    for (int y = 0; y < mb; ++y)
      for (int z = 0; z < mc; ++z) {
        double s = 0.0;
        if (pa)
          for (int x = 0; x < na; ++x) s += sph[(la * 9 + m) * 15 + x] * t2[(x * mb + y) * mc + z];
        else
          s = t2[(m * mb + y) * mc + z];
        out[(m * mb + y) * mc + z] = s;
      }
}

__device__ ShellRef df_shell(const DFBasisView& v, int s) {
  return ShellRef{v.l[s], v.po[s + 1] - v.po[s], v.exps + v.po[s], v.coefs + v.po[s], v.cen + 3 * s};
}

}  // namespace

// B[P][mu >= nu] (packed lower triangle, element mu (mu + 1) / 2 + nu) for one
// (orbital shell pair s1 >= s2, auxiliary shell) per thread
__global__ __launch_bounds__(256) void k_df_three_centre(DFBasisView o, DFBasisView x, RysArgs ra,
                                                         const double* __restrict__ sph, double* __restrict__ B) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long npair = (long long)o.nsh * (o.nsh + 1) / 2;
  if (t >= npair * x.nsh) return;
  const int sc = (int)(t / npair);
  const long long pq = t % npair;
  int s1 = (int)((sqrt(8.0 * (double)pq + 1.0) - 1.0) / 2.0);
  while ((long long)s1 * (s1 + 1) / 2 > pq) --s1;
  while ((long long)(s1 + 1) * (s1 + 2) / 2 <= pq) ++s1;
  const int s2 = (int)(pq - (long long)s1 * (s1 + 1) / 2);
  const ShellRef a = df_shell(o, s1), b = df_shell(o, s2), c = df_shell(x, sc);
  double cart[kNCo * kNCo * kNCx], blk[kNCo * kNCo * kNCx];
  df_cart(a, b, c, ra, cart);
  df_to_out(sph, a.l, b.l, c.l, o.pure, x.pure, cart, blk);
  const int ma = df_nfun(a.l, o.pure), mb = df_nfun(b.l, o.pure), mc = df_nfun(c.l, x.pure);
  const size_t n = (size_t)o.nbf, np = n * (n + 1) / 2;
  for (int i = 0; i < ma; ++i)
    for (int j = 0; j < mb; ++j) {
      const size_t mu = (size_t)o.bf0[s1] + i, nu = (size_t)o.bf0[s2] + j;
      const size_t hi = mu > nu ? mu : nu, lo = mu > nu ? nu : mu;
      const size_t e = hi * (hi + 1) / 2 + lo;
      for (int k = 0; k < mc; ++k) B[(size_t)(x.bf0[sc] + k) * np + e] = blk[(i * mb + j) * mc + k];
    }
}

// V[P][Q] for one auxiliary shell pair c1 >= c2 per thread: (c1 s0 | c2)
__global__ __launch_bounds__(256) void k_df_metric(DFBasisView x, RysArgs ra, const double* __restrict__ sph,
                                                   double* __restrict__ V) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long npair = (long long)x.nsh * (x.nsh + 1) / 2;
  if (t >= npair) return;
  int c1 = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) / 2.0);
  while ((long long)c1 * (c1 + 1) / 2 > t) --c1;
  while ((long long)(c1 + 1) * (c1 + 2) / 2 <= t) ++c1;
  const int c2 = (int)(t - (long long)c1 * (c1 + 1) / 2);
  const double zero = 0.0, one = 1.0;
  const ShellRef a = df_shell(x, c1), c = df_shell(x, c2);
  const ShellRef s0{0, 1, &zero, &one, c.O};
  // the bra holds an auxiliary shell (up to g): use the auxiliary limits for it
  double cart[kNCx * kNCx], blk[kNCx * kNCx];
  df_cart(c, s0, a, ra, cart);  // (c2 s0 | c1): c2 on the bra, c1 on the auxiliary side
  df_to_out(sph, c.l, 0, a.l, x.pure, x.pure, cart, blk);
  const int m2 = df_nfun(c.l, x.pure), m1 = df_nfun(a.l, x.pure);
  const size_t n = (size_t)x.nbf;
  for (int j = 0; j < m2; ++j)
    for (int i = 0; i < m1; ++i) {
      const double v = blk[j * m1 + i];
      V[(size_t)(x.bf0[c1] + i) * n + x.bf0[c2] + j] = v;
      V[(size_t)(x.bf0[c2] + j) * n + x.bf0[c1] + i] = v;
    }
}

// d[mu (mu + 1) / 2 + nu] = D[mu][nu] + D[nu][mu] (mu > nu), D[mu][mu] on the
// diagonal: the density in the packed layout of B (one thread per element)
__global__ __launch_bounds__(256) void k_df_pack(const double* __restrict__ D, double* __restrict__ d, int n) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)n * (n + 1) / 2) return;
  long long mu = (long long)((sqrt(8.0 * (double)e + 1.0) - 1.0) / 2.0);
  while (mu * (mu + 1) / 2 > e) --mu;
  while ((mu + 1) * (mu + 2) / 2 <= e) ++mu;
  const long long nu = e - mu * (mu + 1) / 2;
  d[e] = mu == nu ? D[mu * n + mu] : D[mu * n + nu] + D[nu * n + mu];
}

// J[mu][nu] = J[nu][mu] = j[mu (mu + 1) / 2 + nu]: one thread per element of
// the full matrix, so the stores are coalesced
__global__ __launch_bounds__(256) void k_df_unpack(const double* __restrict__ j, double* __restrict__ J, int n) {
  const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long long)n * n) return;
  const long long r = e / n, c = e % n;
  const long long hi = r > c ? r : c, lo = r > c ? c : r;
  J[e] = j[hi * (hi + 1) / 2 + lo];
}

void launch_df_pack(const double* D, double* d, int n, hipStream_t s) {
  const long long m = (long long)n * (n + 1) / 2;
  if (m > 0) hipLaunchKernelGGL(k_df_pack, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, D, d, n);
}

void launch_df_unpack(const double* j, double* J, int n, hipStream_t s) {
  const long long m = (long long)n * n;
  if (m > 0) hipLaunchKernelGGL(k_df_unpack, dim3((unsigned)((m + 255) / 256)), dim3(256), 0, s, j, J, n);
}

void launch_df_three_centre(const DFBasisView& o, const DFBasisView& x, const RysArgs& ra, const double* sph,
                            double* B, hipStream_t s) {
  const long long n = (long long)o.nsh * (o.nsh + 1) / 2 * x.nsh;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_df_three_centre, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, o, x, ra, sph, B);
}

void launch_df_metric(const DFBasisView& x, const RysArgs& ra, const double* sph, double* V, hipStream_t s) {
  const long long n = (long long)x.nsh * (x.nsh + 1) / 2;
  if (n <= 0) return;
  hipLaunchKernelGGL(k_df_metric, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, x, ra, sph, V);
}

}  // namespace qfe
