// One-electron integrals on the GPU (kernels in onebody_gpu.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "eri_core.hpp"

namespace qfe {

// Device view of a basis: per shell l, atom, centre (3 doubles), primitive
// range [po[s], po[s+1]) into exps/coefs (normalised contraction
// coefficients), first output function bf0[s]; per atom the nuclear charge
// seen by the electrons Z and the position R (3 doubles); the solid-harmonic
// table sph[l][m][c] (5 x 9 x 15 doubles).  Shells up to l = 3.
struct OneBodyView {
  const int* l;
  const int* atom;
  const int* po;
  const int* bf0;
  const double* cen;
  const double* exps;
  const double* coefs;
  const double* Z;
  const double* R;
  const double* sph;
  int nsh, natoms, nbf, pure;
};

// op 0 = S, 1 = T, 2 = V into out (nbf x nbf, row-major, device)
void launch_one_body(const OneBodyView& b, const RysArgs& ra, int op, double* out, hipStream_t s);
// grad[natoms][3] (device, zeroed by the caller) += sum_mn W_mn dO_mn/dR
void launch_one_body_grad(const OneBodyView& b, const RysArgs& ra, int op, const double* W, double* grad,
                          hipStream_t s);

}  // namespace qfe
