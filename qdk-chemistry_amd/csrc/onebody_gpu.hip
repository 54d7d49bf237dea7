// One-electron integrals on the GPU for the SCF harness: overlap S, kinetic
// T, nuclear attraction V over all shell pairs, and the density-contracted
// nuclear derivatives sum_mn W_mn dO_mn/dR.  The reference evaluates these
// with Libint2 in scf/src/util/int1e.cpp (S :518, T and V and their
// derivatives :473-977).  The recursions are those of the host path
// (onebody.cpp): Obara-Saika 1-D overlaps for S and T (T from the
// second-derivative identity), Rys quadrature for V with the engine's root
// tables (roots of the Boys measure at T = p |PC|^2, C00 = PA - u PC,
// B10 = (1 - u)/2p, horizontal transfer to the second centre).  One thread
// per shell pair (per shell pair and nucleus for the V derivative); the work
// is O(nsh^2), far below the Fock build, so the kernels favour plain code
// over tuning.
#include <hip/hip_runtime.h>

#include "eri_core.hpp"
#include "onebody_gpu.hpp"

namespace qfe {

namespace {

constexpr double kPi1 = 3.14159265358979323846;
constexpr int kLD = 8;      // 1-D overlap table stride: i <= la <= 4, j <= lb + 2 <= 5
constexpr int kMaxC = 15;   // cartesian functions up to l = 4 (derivative of f)
constexpr int kMaxF = 10;   // output functions of one shell up to l = 3 (cartesian f)

__device__ int ob_ncart(int l) { return (l + 1) * (l + 2) / 2; }

__device__ void ob_cart_xyz(int l, int c, int& x, int& y, int& z) {
  int k = 0;
  while ((k + 1) * (k + 2) / 2 <= c) ++k;
  x = l - k;
  z = c - k * (k + 1) / 2;
  y = l - x - z;
}

__device__ int ob_nfun(const OneBodyView& b, int l) { return (b.pure && l >= 2) ? 2 * l + 1 : ob_ncart(l); }

// 1-D overlap S[i][j] without the sqrt(pi/p) exp prefactor, i <= im, j <= jm
__device__ void ob_os1d(double PA, double PB, double p, int im, int jm, double* S) {
  const double o2p = 0.5 / p;
  for (int i = 0; i <= im; ++i)
    for (int j = 0; j <= jm; ++j) {
      double v;
      if (i == 0 && j == 0)
        v = 1.0;
      else if (j == 0) {
        v = PA * S[(i - 1) * kLD];
        if (i >= 2) v += (i - 1) * o2p * S[(i - 2) * kLD];
      } else {
        v = PB * S[i * kLD + j - 1];
        if (i >= 1) v += i * o2p * S[(i - 1) * kLD + j - 1];
        if (j >= 2) v += (j - 1) * o2p * S[i * kLD + j - 2];
      }
      S[i * kLD + j] = v;
    }
}

__device__ void ob_rys(int n, double T, const RysArgs& ra, double* u, double* w) {
  switch (n) {
    case 1: {
      double uu[1], ww[1];
      rys_roots<1>(T, ra, uu, ww);
      u[0] = uu[0];
      w[0] = ww[0];
      break;
    }
    case 2: {
      double uu[2], ww[2];
      rys_roots<2>(T, ra, uu, ww);
      for (int i = 0; i < 2; ++i) u[i] = uu[i], w[i] = ww[i];
      break;
    }
    case 3: {
      double uu[3], ww[3];
      rys_roots<3>(T, ra, uu, ww);
      for (int i = 0; i < 3; ++i) u[i] = uu[i], w[i] = ww[i];
      break;
    }
    default: {
      double uu[4], ww[4];
      rys_roots<4>(T, ra, uu, ww);
      for (int i = 0; i < 4; ++i) u[i] = uu[i], w[i] = ww[i];
      break;
    }
  }
}

// Cartesian block of shell pair (s1 with la = l(s1) + dl, s2): primitive sum
// weighted by (2 alpha_a)^wpow.  op 0 = S, 1 = T, 2 = V (nucleus nuc, or all
// nuclei when nuc < 0).  out[na][nb].
__device__ void ob_cart_block(const OneBodyView& b, const RysArgs& ra, int s1, int s2, int dl, int wpow, int op,
                              int nuc, double* out) {
  const int la = b.l[s1] + dl, lb = b.l[s2];
  const int na = la >= 0 ? ob_ncart(la) : 0, nb = ob_ncart(lb);
  for (int i = 0; i < na * nb; ++i) out[i] = 0.0;
  if (la < 0) return;
  const double* A = b.cen + 3 * s1;
  const double* B = b.cen + 3 * s2;
  double AB[3], AB2 = 0.0;
  for (int d = 0; d < 3; ++d) {
    AB[d] = A[d] - B[d];
    AB2 += AB[d] * AB[d];
  }
  for (int ia = b.po[s1]; ia < b.po[s1 + 1]; ++ia)
    for (int ib = b.po[s2]; ib < b.po[s2 + 1]; ++ib) {
      const double a = b.exps[ia], bb = b.exps[ib], p = a + bb;
      double cc = b.coefs[ia] * b.coefs[ib];
      for (int k = 0; k < wpow; ++k) cc *= 2.0 * a;
      const double Kab = exp(-a * bb / p * AB2) * cc;
      if (Kab == 0.0) continue;
      double P[3], PA[3], PB[3];
      for (int d = 0; d < 3; ++d) {
        P[d] = (a * A[d] + bb * B[d]) / p;
        PA[d] = P[d] - A[d];
        PB[d] = P[d] - B[d];
      }
      if (op <= 1) {
        double S[3][kLD * kLD];
        for (int d = 0; d < 3; ++d) ob_os1d(PA[d], PB[d], p, la, lb + 2, S[d]);
        const double s00 = pow(kPi1 / p, 1.5) * Kab;
        for (int i = 0; i < na; ++i) {
          int ea[3];
          ob_cart_xyz(la, i, ea[0], ea[1], ea[2]);
          for (int j = 0; j < nb; ++j) {
            int eb[3];
            ob_cart_xyz(lb, j, eb[0], eb[1], eb[2]);
            double v;
            if (op == 0) {
              v = S[0][ea[0] * kLD + eb[0]] * S[1][ea[1] * kLD + eb[1]] * S[2][ea[2] * kLD + eb[2]];
            } else {
              double Sd[3], Td[3];
              for (int d = 0; d < 3; ++d) {
                const int ii = ea[d], jj = eb[d];
                Sd[d] = S[d][ii * kLD + jj];
                double t = -2.0 * bb * (2 * jj + 1) * S[d][ii * kLD + jj] + 4.0 * bb * bb * S[d][ii * kLD + jj + 2];
                if (jj >= 2) t += jj * (jj - 1) * S[d][ii * kLD + jj - 2];
                Td[d] = -0.5 * t;
              }
              v = Td[0] * Sd[1] * Sd[2] + Sd[0] * Td[1] * Sd[2] + Sd[0] * Sd[1] * Td[2];
            }
            out[i * nb + j] += s00 * v;
          }
        }
      } else {
        const int L = la + lb, nr = L / 2 + 1;
        const int c0 = nuc < 0 ? 0 : nuc, c1 = nuc < 0 ? b.natoms : nuc + 1;
        for (int cn = c0; cn < c1; ++cn) {
          const double* C = b.R + 3 * cn;
          double PC[3], PC2 = 0.0;
          for (int d = 0; d < 3; ++d) {
            PC[d] = P[d] - C[d];
            PC2 += PC[d] * PC[d];
          }
          double u[4], w[4];
          ob_rys(nr, p * PC2, ra, u, w);
          const double pref = -b.Z[cn] * 2.0 * kPi1 / p * Kab;
          for (int r = 0; r < nr; ++r) {
            double I[3][5][4];
            for (int d = 0; d < 3; ++d) {
              double G[8];
              const double C00 = PA[d] - u[r] * PC[d], B10 = (1.0 - u[r]) / (2.0 * p);
              G[0] = (d == 2) ? w[r] * pref : 1.0;
              if (L >= 1) G[1] = C00 * G[0];
              for (int i = 1; i < L; ++i) G[i + 1] = C00 * G[i] + i * B10 * G[i - 1];
              // horizontal transfer (a, b+1) = (a+1, b) + AB (a, b)
              double t[8][4];
              for (int i = 0; i <= L; ++i) t[i][0] = G[i];
              for (int jj = 1; jj <= lb; ++jj)
                for (int i = 0; i <= L - jj; ++i) t[i][jj] = t[i + 1][jj - 1] + AB[d] * t[i][jj - 1];
              for (int i = 0; i <= la; ++i)
                for (int jj = 0; jj <= lb; ++jj) I[d][i][jj] = t[i][jj];
            }
            for (int i = 0; i < na; ++i) {
              int ax, ay, az;
              ob_cart_xyz(la, i, ax, ay, az);
              for (int j = 0; j < nb; ++j) {
                int bx, by, bz;
                ob_cart_xyz(lb, j, bx, by, bz);
                out[i * nb + j] += I[0][ax][bx] * I[1][ay][by] * I[2][az][bz];
              }
            }
          }
        }
      }
    }
}

// cartesian block [n1c][n2c] -> output functions [n1][n2]
__device__ void ob_to_out(const OneBodyView& b, int l1, int l2, const double* in, double* out) {
  const int n1c = ob_ncart(l1), n2c = ob_ncart(l2);
  const bool p1 = b.pure && l1 >= 2, p2 = b.pure && l2 >= 2;
  const int n1 = p1 ? 2 * l1 + 1 : n1c, n2 = p2 ? 2 * l2 + 1 : n2c;
  double t[kMaxC * kMaxF];
  for (int m = 0; m < n1; ++m)
    for (int j = 0; j < n2c; ++j) {
      double s = 0.0;
      if (p1)
        for (int c = 0; c < n1c; ++c) s += b.sph[(l1 * 9 + m) * 15 + c] * in[c * n2c + j];
      else
        s = in[m * n2c + j];
      t[m * n2c + j] = s;
    }
  for (int m = 0; m < n1; ++m)
    for (int n = 0; n < n2; ++n) {
      double s = 0.0;
      if (p2)
        for (int c = 0; c < n2c; ++c) s += b.sph[(l2 * 9 + n) * 15 + c] * t[m * n2c + c];
      else
        s = t[m * n2c + n];
      out[m * n2 + n] = s;
    }
}

// d/dA_x of the cartesian block (l1, l2) -> [3][n1c][n2c]
__device__ void ob_deriv_A(const OneBodyView& b, const RysArgs& ra, int s1, int s2, int op, int nuc, double* out) {
  const int l1 = b.l[s1], l2 = b.l[s2];
  const int n1 = ob_ncart(l1), n2 = ob_ncart(l2);
  double up[kMaxC * kMaxF], dn[kMaxC * kMaxF];
  ob_cart_block(b, ra, s1, s2, +1, 1, op, nuc, up);
  if (l1 > 0) ob_cart_block(b, ra, s1, s2, -1, 0, op, nuc, dn);
  for (int i = 0; i < n1; ++i) {
    int e[3];
    ob_cart_xyz(l1, i, e[0], e[1], e[2]);
    for (int d = 0; d < 3; ++d) {
      int eu[3] = {e[0], e[1], e[2]};
      eu[d]++;
      const int lu = l1 + 1;
      const int iu = ((lu - eu[0]) * (lu - eu[0] + 1)) / 2 + eu[2];
      int idn = -1;
      if (e[d] > 0) {
        int ed[3] = {e[0], e[1], e[2]};
        ed[d]--;
        const int ll = l1 - 1;
        idn = ((ll - ed[0]) * (ll - ed[0] + 1)) / 2 + ed[2];
      }
      for (int j = 0; j < n2; ++j) {
        double v = up[iu * n2 + j];
        if (idn >= 0) v -= e[d] * dn[idn * n2 + j];
        out[(d * n1 + i) * n2 + j] = v;
      }
    }
  }
}

// 2 sum_ij W[bf(s1) + i][bf(s2) + j] o_x[i][j] for the three components of d
__device__ void ob_contract(const OneBodyView& b, int s1, int s2, const double* d, const double* W, double (&g)[3]) {
  const int l1 = b.l[s1], l2 = b.l[s2];
  const int n1c = ob_ncart(l1), n2c = ob_ncart(l2);
  const int n1 = ob_nfun(b, l1), n2 = ob_nfun(b, l2);
  const size_t n = (size_t)b.nbf;
  double o[kMaxF * kMaxF];
  for (int x = 0; x < 3; ++x) {
    ob_to_out(b, l1, l2, d + x * n1c * n2c, o);
    double s = 0.0;
    for (int i = 0; i < n1; ++i)
      for (int j = 0; j < n2; ++j) s += W[(size_t)(b.bf0[s1] + i) * n + b.bf0[s2] + j] * o[i * n2 + j];
    g[x] = 2.0 * s;
  }
}

// true when every primitive pair of (s1, s2) has |c_a c_b| exp(-ab/p |AB|^2)
// below thr: such a pair's S, T, V and derivative blocks are below
// thr times factors of order (pi/p)^(3/2), Z and the angular polynomials
__device__ bool ob_pair_negligible(const OneBodyView& b, int s1, int s2, double thr) {
  const double* A = b.cen + 3 * s1;
  const double* B = b.cen + 3 * s2;
  double AB2 = 0.0;
  for (int d = 0; d < 3; ++d) AB2 += (A[d] - B[d]) * (A[d] - B[d]);
  for (int ia = b.po[s1]; ia < b.po[s1 + 1]; ++ia)
    for (int ib = b.po[s2]; ib < b.po[s2 + 1]; ++ib) {
      const double a = b.exps[ia], bb = b.exps[ib];
      if (fabs(b.coefs[ia] * b.coefs[ib]) * exp(-a * bb / (a + bb) * AB2) >= thr) return false;
    }
  return true;
}

constexpr double kPairThr = 1e-17;

}  // namespace

// one thread per shell pair s1 >= s2: the block and its transpose
__global__ __launch_bounds__(256) void k_one_body(OneBodyView b, RysArgs ra, int op, double* __restrict__ out) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long npair = (long long)b.nsh * (b.nsh + 1) / 2;
  if (t >= npair) return;
  int s1 = (int)((sqrt(8.0 * (double)t + 1.0) - 1.0) / 2.0);
  while ((long long)s1 * (s1 + 1) / 2 > t) --s1;
  while ((long long)(s1 + 1) * (s1 + 2) / 2 <= t) ++s1;
  const int s2 = (int)(t - (long long)s1 * (s1 + 1) / 2);
  const int l1 = b.l[s1], l2 = b.l[s2];
  double cart[kMaxC * kMaxF], blk[kMaxF * kMaxF];
  ob_cart_block(b, ra, s1, s2, 0, 0, op, -1, cart);
  ob_to_out(b, l1, l2, cart, blk);
  const int n1 = ob_nfun(b, l1), n2 = ob_nfun(b, l2);
  const size_t n = (size_t)b.nbf;
  for (int i = 0; i < n1; ++i)
    for (int j = 0; j < n2; ++j) {
      out[(size_t)(b.bf0[s1] + i) * n + b.bf0[s2] + j] = blk[i * n2 + j];
      out[(size_t)(b.bf0[s2] + j) * n + b.bf0[s1] + i] = blk[i * n2 + j];
    }
}

// S and T derivatives: one thread per ordered shell pair (s1, s2); the
// derivative with respect to the centre of s1 goes to its atom
__global__ __launch_bounds__(256) void k_one_body_grad_st(OneBodyView b, RysArgs ra, int op,
                                                          const double* __restrict__ W, double* __restrict__ grad) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long long)b.nsh * b.nsh) return;
  const int s1 = (int)(t / b.nsh), s2 = (int)(t % b.nsh);
  double d[3 * kMaxC * kMaxF];
  ob_deriv_A(b, ra, s1, s2, op, -1, d);
  double g[3];
  ob_contract(b, s1, s2, d, W, g);
  const int at = b.atom[s1];
  for (int x = 0; x < 3; ++x)
    if (g[x] != 0.0) atomicAdd(&grad[3 * at + x], g[x]);
}

// V derivative: one thread per (ordered shell pair, nucleus); by translation
// invariance the nucleus takes minus the centre derivatives of both shells,
// and the pair (s2, s1) supplies the second one.  Pairs whose every
// primitive product prefactor is below 1e-17 are skipped (most pairs of a
// large molecule: their contributions are below 1e-14).
__global__ __launch_bounds__(256) void k_one_body_grad_v(OneBodyView b, RysArgs ra, const double* __restrict__ W,
                                                         double* __restrict__ grad) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long np = (long long)b.nsh * b.nsh;
  const bool valid = t < np * b.natoms;
  // the nucleus is the fast index: a wave works on one shell pair, so its
  // control flow (angular momenta, contraction lengths, the skip) is uniform
  // and its contributions to the pair's own atom are summed across the wave
  // before one atomic per component
  int cn = 0, s1 = 0;
  long long pq = -1;
  double g[3] = {0.0, 0.0, 0.0};
  if (valid) {
    cn = (int)(t % b.natoms);
    pq = t / b.natoms;
    s1 = (int)(pq / b.nsh);
    const int s2 = (int)(pq % b.nsh);
    if (!ob_pair_negligible(b, s1, s2, kPairThr)) {
      double d[3 * kMaxC * kMaxF];
      ob_deriv_A(b, ra, s1, s2, 2, cn, d);
      ob_contract(b, s1, s2, d, W, g);
      for (int x = 0; x < 3; ++x)
        if (g[x] != 0.0) atomicAdd(&grad[3 * cn + x], -g[x]);
    }
  }
  const long long pq0 = __shfl(pq, 0);
  if (__all(pq == pq0)) {
    if (pq0 < 0) return;  // the whole wave is past the end
    for (int x = 0; x < 3; ++x) {
      double v = g[x];
      for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
      if (threadIdx.x % 64 == 0 && v != 0.0) atomicAdd(&grad[3 * b.atom[s1] + x], v);
    }
  } else if (valid) {
    const int at = b.atom[s1];
    for (int x = 0; x < 3; ++x)
      if (g[x] != 0.0) atomicAdd(&grad[3 * at + x], g[x]);
  }
}

void launch_one_body(const OneBodyView& b, const RysArgs& ra, int op, double* out, hipStream_t s) {
  const long long npair = (long long)b.nsh * (b.nsh + 1) / 2;
  if (npair <= 0) return;
  hipLaunchKernelGGL(k_one_body, dim3((unsigned)((npair + 255) / 256)), dim3(256), 0, s, b, ra, op, out);
}

void launch_one_body_grad(const OneBodyView& b, const RysArgs& ra, int op, const double* W, double* grad,
                          hipStream_t s) {
  const long long np = (long long)b.nsh * b.nsh;
  if (np <= 0) return;
  if (op <= 1) {
    hipLaunchKernelGGL(k_one_body_grad_st, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, s, b, ra, op, W, grad);
  } else {
    const long long n = np * b.natoms;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_one_body_grad_v, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, b, ra, W, grad);
  }
}

}  // namespace qfe
