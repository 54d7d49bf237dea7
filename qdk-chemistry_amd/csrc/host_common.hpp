// Host-side basis container and one-electron helpers shared by the engine and
// the one-electron C entry points.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/qfe.h"

namespace qfe {

struct ShellHost {
  int l = 0, atom = 0;
  double O[3] = {0, 0, 0};
  std::vector<double> a, c;
};

struct BasisHost {
  std::vector<ShellHost> shells;
  std::vector<int> bf0;  // nsh + 1
  int nsh = 0, nbf = 0, natoms = 0;
  bool pure = true;
  std::vector<double> Z, R;

  int nfun(int l) const { return (pure && l >= 2) ? 2 * l + 1 : (l + 1) * (l + 2) / 2; }

  static BasisHost from_c(const qfe_basis* b) {
    if (!b) throw std::invalid_argument("null basis");
    if (b->nshells < 0 || b->natoms < 0) throw std::invalid_argument("negative sizes");
    BasisHost h;
    h.nsh = b->nshells;
    h.natoms = b->natoms;
    h.pure = b->pure != 0;
    h.shells.resize(h.nsh);
    h.bf0.assign(h.nsh + 1, 0);
    for (int s = 0; s < h.nsh; ++s) {
      ShellHost& sh = h.shells[s];
      sh.l = b->l[s];
      sh.atom = b->atom[s];
      if (sh.l < 0 || sh.l > 6) throw std::invalid_argument("shell angular momentum out of range");
      if (sh.atom < 0 || (h.natoms > 0 && sh.atom >= h.natoms))
        throw std::invalid_argument("shell atom index out of range");
      for (int d = 0; d < 3; ++d) sh.O[d] = b->centers[3 * s + d];
      const int p0 = b->prim_offset[s], p1 = b->prim_offset[s + 1];
      if (p1 <= p0 || p1 - p0 > 64) throw std::invalid_argument("shell primitive count out of range");
      for (int k = p0; k < p1; ++k) {
        if (!(b->exps[k] > 0.0)) throw std::invalid_argument("Shell exponents must be positive");
        sh.a.push_back(b->exps[k]);
        sh.c.push_back(b->coefs[k]);
      }
      h.bf0[s + 1] = h.bf0[s] + h.nfun(sh.l);
    }
    h.nbf = h.bf0[h.nsh];
    h.Z.assign(h.natoms, 0.0);
    h.R.assign(3 * (size_t)h.natoms, 0.0);
    for (int a = 0; a < h.natoms; ++a) {
      h.Z[a] = b->charges ? b->charges[a] : 0.0;
      for (int d = 0; d < 3; ++d) h.R[3 * a + d] = b->coords ? b->coords[3 * a + d] : 0.0;
    }
    return h;
  }
};

// Frobenius norm of the overlap block (s1, s2) in output functions
double overlap_block_norm(const BasisHost& b, int s1, int s2);

// Rys roots and weights on the host (same table as the device)
void rys_host(int n, double T, double* u, double* w);

}  // namespace qfe
