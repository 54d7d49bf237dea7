// Density-fitted Coulomb: three-centre and metric kernels (dfj.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "eri_core.hpp"

namespace qfe {

// Device view of one basis: per shell l, primitive range [po[s], po[s+1])
// into exps/coefs (normalised), centre cen[3 s], first output function
// bf0[s]; pure (solid harmonics from d on) or cartesian.
struct DFBasisView {
  const int* l;
  const int* po;
  const int* bf0;
  const double* cen;
  const double* exps;
  const double* coefs;
  int nsh, nbf, pure;
};

// B[P][mu (mu + 1) / 2 + nu] = (mu nu | P) for mu >= nu, naux x nbf (nbf + 1) / 2
// (orbital shells up to d, auxiliary shells up to g); sph = the
// solid-harmonic table kSph (device)
void launch_df_three_centre(const DFBasisView& orb, const DFBasisView& aux, const RysArgs& ra, const double* sph,
                            double* B, hipStream_t s);
// the density D (n x n, device) in the packed layout of B, and the packed J
// back to the full symmetric matrix
void launch_df_pack(const double* D, double* d, int n, hipStream_t s);
void launch_df_unpack(const double* j, double* J, int n, hipStream_t s);
// V[P][Q] = (P | Q), naux x naux
void launch_df_metric(const DFBasisView& aux, const RysArgs& ra, const double* sph, double* V, hipStream_t s);

}  // namespace qfe
