// The dense AO two-electron tensor for the quarter and MO transforms
// (kernel in eri_dense.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "dfj.hpp"
#include "eri_core.hpp"

namespace qfe {

// eri[i][j][k][l] = (ij|kl), nbf^4 doubles (device, row-major), every shell
// quartet computed; shells up to d; sph = the solid-harmonic table kSph;
// pairs = every unordered shell pair once (nsh (nsh + 1) / 2 entries, any
// order and either orientation; the order only groups the work)
void launch_eri_dense(const DFBasisView& b, const RysArgs& ra, const double* sph, const int2* pairs, double* eri,
                      hipStream_t s);

}  // namespace qfe
