// One-electron integrals for the SCF harness (host, OpenMP): overlap S,
// kinetic T, nuclear attraction V, and the derivative contractions
// sum_mn W_mn dO_mn/dR used by the nuclear gradient.  These mirror what the
// reference computes with Libint2 in scf/src/util/int1e.cpp (S :518, T/V and
// derivatives :473-977); they are not on the Fock-build path.
//
// S and T: Obara-Saika 1-D overlap recursion, T from the second-derivative
// identity.  V: Rys quadrature with the engine's root table (roots of the
// Boys measure at T = p |PC|^2), 1-D recursion C00 = PA - u PC,
// B10 = (1-u)/2p, then horizontal transfer.
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/qfe.h"
#include "host_common.hpp"
#include "sph_coef.inc"

namespace qfe {

extern thread_local std::string g_last_error;

namespace {

constexpr double kPi = 3.14159265358979323846;
inline int ncart(int l) { return (l + 1) * (l + 2) / 2; }
inline void cart_xyz(int l, int c, int& x, int& y, int& z) {
  int k = 0;
  while ((k + 1) * (k + 2) / 2 <= c) ++k;
  x = l - k;
  z = c - k * (k + 1) / 2;
  y = l - x - z;
}

// 1-D overlap S[i][j] (without the sqrt(pi/p) exp prefactor), i <= im, j <= jm
void os1d(double PA, double PB, double p, int im, int jm, double* S, int ld) {
  const double o2p = 0.5 / p;
  for (int i = 0; i <= im; ++i)
    for (int j = 0; j <= jm; ++j) {
      double v;
      if (i == 0 && j == 0)
        v = 1.0;
      else if (j == 0) {
        v = PA * S[(i - 1) * ld];
        if (i >= 2) v += (i - 1) * o2p * S[(i - 2) * ld];
      } else {
        v = PB * S[i * ld + j - 1];
        if (i >= 1) v += i * o2p * S[(i - 1) * ld + j - 1];
        if (j >= 2) v += (j - 1) * o2p * S[i * ld + j - 2];
      }
      S[i * ld + j] = v;
    }
}

// Cartesian block of shell pair (s1 with l1 = l(s1) + dl, s2): primitive sum
// weighted by (2 alpha_a)^wpow.  op 0 = S, 1 = T, 2 = V (nucleus `nuc`, or all
// nuclei when nuc < 0).  out[n1c][n2c].
void cart_block(const BasisHost& b, int s1, int s2, int dl, int wpow, int op, int nuc, double* out) {
  const ShellHost& A = b.shells[s1];
  const ShellHost& B = b.shells[s2];
  const int la = A.l + dl, lb = B.l;
  const int na = ncart(la), nb = ncart(lb);
  std::memset(out, 0, sizeof(double) * na * nb);
  if (la < 0) return;
  double AB[3], AB2 = 0.0;
  for (int d = 0; d < 3; ++d) {
    AB[d] = A.O[d] - B.O[d];
    AB2 += AB[d] * AB[d];
  }
  const int LD = 12;
  for (size_t ia = 0; ia < A.a.size(); ++ia)
    for (size_t ib = 0; ib < B.a.size(); ++ib) {
      const double a = A.a[ia], bb = B.a[ib], p = a + bb;
      double cc = A.c[ia] * B.c[ib];
      for (int k = 0; k < wpow; ++k) cc *= 2.0 * a;
      const double Kab = std::exp(-a * bb / p * AB2) * cc;
      if (Kab == 0.0) continue;
      double P[3], PA[3], PB[3];
      for (int d = 0; d < 3; ++d) {
        P[d] = (a * A.O[d] + bb * B.O[d]) / p;
        PA[d] = P[d] - A.O[d];
        PB[d] = P[d] - B.O[d];
      }
      if (op <= 1) {
        double S[3][LD * LD];
        for (int d = 0; d < 3; ++d) os1d(PA[d], PB[d], p, la, lb + 2, S[d], LD);
        const double s00 = std::pow(kPi / p, 1.5) * Kab;
        for (int i = 0; i < na; ++i) {
          int ax, ay, az;
          cart_xyz(la, i, ax, ay, az);
          const int ea[3] = {ax, ay, az};
          for (int j = 0; j < nb; ++j) {
            int bx, by, bz;
            cart_xyz(lb, j, bx, by, bz);
            const int eb[3] = {bx, by, bz};
            double v;
            if (op == 0) {
              v = S[0][ax * LD + bx] * S[1][ay * LD + by] * S[2][az * LD + bz];
            } else {
              double Sd[3], Td[3];
              for (int d = 0; d < 3; ++d) {
                const int ii = ea[d], jj = eb[d];
                Sd[d] = S[d][ii * LD + jj];
                double t = -2.0 * bb * (2 * jj + 1) * S[d][ii * LD + jj] + 4.0 * bb * bb * S[d][ii * LD + jj + 2];
                if (jj >= 2) t += jj * (jj - 1) * S[d][ii * LD + jj - 2];
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
          const double* C = &b.R[3 * cn];
          double PC[3], PC2 = 0.0;
          for (int d = 0; d < 3; ++d) {
            PC[d] = P[d] - C[d];
            PC2 += PC[d] * PC[d];
          }
          double u[8], w[8];
          rys_host(nr, p * PC2, u, w);
          const double pref = -b.Z[cn] * 2.0 * kPi / p * Kab;
          for (int r = 0; r < nr; ++r) {
            double I[3][8][8];
            for (int d = 0; d < 3; ++d) {
              double G[16];
              const double C00 = PA[d] - u[r] * PC[d], B10 = (1.0 - u[r]) / (2.0 * p);
              G[0] = (d == 2) ? w[r] * pref : 1.0;
              if (L >= 1) G[1] = C00 * G[0];
              for (int i = 1; i < L; ++i) G[i + 1] = C00 * G[i] + i * B10 * G[i - 1];
              // horizontal transfer (a, b+1) = (a+1, b) + AB (a, b)
              double t[16][8];
              for (int i = 0; i <= L; ++i) t[i][0] = G[i];
              for (int jj = 1; jj <= lb; ++jj)
                for (int i = 0; i <= L - jj; ++i) t[i][jj] = t[i + 1][jj - 1] + AB[d] * t[i][jj - 1];
              for (int i = 0; i <= la; ++i)
                for (int jj = 0; jj <= lb; ++jj) I[d][i][jj] = t[i][jj];
            }
            for (int i = 0; i < na; ++i) {
              int ax, ay, az;
              cart_xyz(la, i, ax, ay, az);
              for (int j = 0; j < nb; ++j) {
                int bx, by, bz;
                cart_xyz(lb, j, bx, by, bz);
                out[i * nb + j] += I[0][ax][bx] * I[1][ay][by] * I[2][az][bz];
              }
            }
          }
        }
      }
    }
}

// transform a cartesian block [n1c][n2c] to output functions
void to_out(const BasisHost& b, int l1, int l2, const double* in, double* out) {
  const int n1c = ncart(l1), n2c = ncart(l2);
  const bool p1 = b.pure && l1 >= 2, p2 = b.pure && l2 >= 2;
  const int n1 = p1 ? 2 * l1 + 1 : n1c, n2 = p2 ? 2 * l2 + 1 : n2c;
  std::vector<double> t(n1 * n2c);
  for (int m = 0; m < n1; ++m)
    for (int j = 0; j < n2c; ++j) {
      double s = 0.0;
      if (p1)
        for (int c = 0; c < n1c; ++c) s += kSph[l1][m][c] * in[c * n2c + j];
      else
        s = in[m * n2c + j];
      t[m * n2c + j] = s;
    }
  for (int m = 0; m < n1; ++m)
    for (int n = 0; n < n2; ++n) {
      double s = 0.0;
      if (p2)
        for (int c = 0; c < n2c; ++c) s += kSph[l2][n][c] * t[m * n2c + c];
      else
        s = t[m * n2c + n];
      out[m * n2 + n] = s;
    }
}

// d/dA_i of the cartesian block (l1, l2) -> 3 blocks [3][n1c][n2c]
void deriv_A(const BasisHost& b, int s1, int s2, int op, int nuc, double* out) {
  const int l1 = b.shells[s1].l, l2 = b.shells[s2].l;
  const int n1 = ncart(l1), n2 = ncart(l2);
  std::vector<double> up(ncart(l1 + 1) * n2), dn(l1 > 0 ? ncart(l1 - 1) * n2 : 1);
  cart_block(b, s1, s2, +1, 1, op, nuc, up.data());
  if (l1 > 0) cart_block(b, s1, s2, -1, 0, op, nuc, dn.data());
  for (int i = 0; i < n1; ++i) {
    int x, y, z;
    cart_xyz(l1, i, x, y, z);
    const int e[3] = {x, y, z};
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

}  // namespace

double overlap_block_norm(const BasisHost& b, int s1, int s2) {
  const int l1 = b.shells[s1].l, l2 = b.shells[s2].l;
  std::vector<double> cart(ncart(l1) * ncart(l2)), out(ncart(l1) * ncart(l2));
  cart_block(b, s1, s2, 0, 0, 0, -1, cart.data());
  to_out(b, l1, l2, cart.data(), out.data());
  const int n = b.nfun(l1) * b.nfun(l2);
  double s = 0.0;
  for (int i = 0; i < n; ++i) s += out[i] * out[i];
  return std::sqrt(s);
}

}  // namespace qfe

using namespace qfe;

extern "C" int qfe_one_body(const qfe_basis* basis, int op, double* out) {
  try {
    if (op < 0 || op > 2) throw std::invalid_argument("op must be 0 (S), 1 (T) or 2 (V)");
    BasisHost b = BasisHost::from_c(basis);
    const int n = b.nbf;
#pragma omp parallel for schedule(dynamic)
    for (int s1 = 0; s1 < b.nsh; ++s1)
      for (int s2 = 0; s2 <= s1; ++s2) {
        const int l1 = b.shells[s1].l, l2 = b.shells[s2].l;
        std::vector<double> cart(ncart(l1) * ncart(l2)), blk(ncart(l1) * ncart(l2));
        cart_block(b, s1, s2, 0, 0, op, -1, cart.data());
        to_out(b, l1, l2, cart.data(), blk.data());
        const int n1 = b.nfun(l1), n2 = b.nfun(l2);
        for (int i = 0; i < n1; ++i)
          for (int j = 0; j < n2; ++j) {
            out[(size_t)(b.bf0[s1] + i) * n + b.bf0[s2] + j] = blk[i * n2 + j];
            out[(size_t)(b.bf0[s2] + j) * n + b.bf0[s1] + i] = blk[i * n2 + j];
          }
      }
    return QFE_OK;
  } catch (const std::invalid_argument& e) {
    g_last_error = e.what();
    return QFE_ERR_INVALID;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return QFE_ERR_RUNTIME;
  }
}

extern "C" int qfe_one_body_grad(const qfe_basis* basis, int op, const double* W, double* grad) {
  try {
    if (op < 0 || op > 2) throw std::invalid_argument("op must be 0 (S), 1 (T) or 2 (V)");
    BasisHost b = BasisHost::from_c(basis);
    const int n = b.nbf, na = b.natoms;
    std::vector<double> g(3 * (size_t)na, 0.0);
#pragma omp parallel
    {
      std::vector<double> gl(3 * (size_t)na, 0.0);
#pragma omp for schedule(dynamic)
      for (int s1 = 0; s1 < b.nsh; ++s1)
        for (int s2 = 0; s2 < b.nsh; ++s2) {
          if (overlap_block_norm(b, s1, s2) < 1e-17 && b.shells[s1].atom != b.shells[s2].atom) continue;
          const int l1 = b.shells[s1].l, l2 = b.shells[s2].l;
          const int n1c = ncart(l1), n2c = ncart(l2);
          const int n1 = b.nfun(l1), n2 = b.nfun(l2);
          std::vector<double> d(3 * n1c * n2c), o(n1 * n2);
          auto contract = [&](const double* blk3, double* acc3) {
            for (int x = 0; x < 3; ++x) {
              to_out(b, l1, l2, blk3 + x * n1c * n2c, o.data());
              double s = 0.0;
              for (int i = 0; i < n1; ++i)
                for (int j = 0; j < n2; ++j) s += W[(size_t)(b.bf0[s1] + i) * n + b.bf0[s2] + j] * o[i * n2 + j];
              acc3[x] += 2.0 * s;
            }
          };
          if (op <= 1) {
            deriv_A(b, s1, s2, op, -1, d.data());
            contract(d.data(), &gl[3 * b.shells[s1].atom]);
          } else {
            for (int cn = 0; cn < na; ++cn) {
              deriv_A(b, s1, s2, 2, cn, d.data());
              double t[3] = {0, 0, 0};
              contract(d.data(), t);
              for (int x = 0; x < 3; ++x) {
                gl[3 * b.shells[s1].atom + x] += t[x];
                gl[3 * cn + x] -= t[x];
              }
            }
          }
        }
#pragma omp critical(qfe_one_body_grad)
      for (size_t k = 0; k < g.size(); ++k) g[k] += gl[k];
    }
    std::memcpy(grad, g.data(), sizeof(double) * g.size());
    return QFE_OK;
  } catch (const std::invalid_argument& e) {
    g_last_error = e.what();
    return QFE_ERR_INVALID;
  } catch (const std::exception& e) {
    g_last_error = e.what();
    return QFE_ERR_RUNTIME;
  }
}
