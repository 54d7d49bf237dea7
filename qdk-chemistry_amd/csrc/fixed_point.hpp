// Order-independent accumulation for deterministic builds (the reference's
// default use_atomics = false, core/scf.h:81-82, reduces thread-local buffers
// in a fixed order, libint2_direct.cpp:316-333,576-592).  Here every value
// added to J or K is split into two 64-bit integers,
//   v = hi 2^-30 + lo 2^-72 (to within 2^-73),
// and the integers are summed with integer atomics: integer addition is
// exact and commutative, so the sums -- and the doubles formed from them at
// the end -- do not depend on the order in which the GPU adds them.
//   range:      |v| < 2^20 per added value (checked; a larger one sets the
//               overflow flag and the build fails), |sum| < 2^33
//   resolution: 2^-72 = 2.1e-22 absolute per added value
// The split uses the 1.5 * 2^52 rounding constant: for |x| < 2^51,
// x + 1.5 * 2^52 rounds x to the nearest integer and leaves it in the low
// mantissa bits.  Before a sum leaves LDS, its lo part's carry above 2^42 is
// moved into hi, so each global add brings at most 2^42 in lo.
#pragma once

#include <hip/hip_runtime.h>

#include <cstring>

namespace qfe {

constexpr double kFxHi = 1073741824.0;             // 2^30
constexpr double kFxHiInv = 1.0 / 1073741824.0;    // 2^-30
constexpr double kFxLo = 4722366482869645213696.0;  // 2^72
constexpr double kFxLoInv = 1.0 / 4722366482869645213696.0;
constexpr double kFxMagic = 6755399441055744.0;    // 1.5 * 2^52
constexpr double kFxMax = 1048576.0;               // 2^20

__host__ __device__ __forceinline__ long long fx_bits(double x) {
#if defined(__HIP_DEVICE_COMPILE__)
  return __double_as_longlong(x);
#else
  long long i;
  std::memcpy(&i, &x, sizeof(i));
  return i;
#endif
}

__host__ __device__ __forceinline__ void fx_split(double v, long long& hi, long long& lo) {
  const double t = fma(v, kFxHi, kFxMagic);
  hi = fx_bits(t) - fx_bits(kFxMagic);
  const double r = fma(-(t - kFxMagic), kFxHiInv, v);  // exact remainder, |r| <= 2^-31
  const double u = fma(r, kFxLo, kFxMagic);
  lo = fx_bits(u) - fx_bits(kFxMagic);
}

__host__ __device__ __forceinline__ double fx_value(long long hi, long long lo) {
  return fma((double)hi, kFxHiInv, (double)lo * kFxLoInv);
}

// add v to the (hi, lo) pair at p in LDS; bad becomes nonzero when |v| is
// out of range (the caller reports it once, after its last add).  No
// branches: both words are added even when one of them is zero.
__device__ __forceinline__ void fx_lds_add(unsigned long long* p, double v, int& bad) {
  bad |= !(fabs(v) < kFxMax);
  long long hi, lo;
  fx_split(v, hi, lo);
  atomicAdd(p, (unsigned long long)hi);
  atomicAdd(p + 1, (unsigned long long)lo);
}

// add v to the (hi, lo) pair at g in global memory
__device__ __forceinline__ void fx_global_add(unsigned long long* g, double v, int& bad) {
  bad |= !(fabs(v) < kFxMax);
  long long hi, lo;
  fx_split(v, hi, lo);
  atomicAdd(g, (unsigned long long)hi);
  atomicAdd(g + 1, (unsigned long long)lo);
}

// move an LDS pair into a global pair: carry lo's bits above 2^42 into hi
__device__ __forceinline__ void fx_flush(const unsigned long long* p, unsigned long long* g) {
  long long hi = (long long)p[0], lo = (long long)p[1];
  if (hi == 0 && lo == 0) return;
  const long long carry = lo >> 42;
  hi += carry;
  lo -= carry * (1ll << 42);
  if (hi) atomicAdd(g, (unsigned long long)hi);
  if (lo) atomicAdd(g + 1, (unsigned long long)lo);
}

}  // namespace qfe
