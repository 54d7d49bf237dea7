// Host-side declarations of the kernel launchers (implemented in kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#ifndef QFE_MAX_L
#define QFE_MAX_L 2  // highest shell angular momentum compiled into the ERI kernels
#endif

namespace qfe {

struct PairHdr;
template <typename R>
struct RysArgsT;
using RysArgs = RysArgsT<double>;
using RysArgs32 = RysArgsT<float>;

struct JKLaunch {
  const int2* tasks;
  long long ntask;
  const PairHdr* pairs;
  const double4* ppos;
  const double* pk;
  const RysArgs* ra;  // host copy, passed by value to the kernels
  const double* P;
  double* J;
  double* K;
  int nbf;
  double kscale;
  bool fp32;          // evaluate in single precision (mixed-precision tasks)
  const float4* ppos32;
  const float* pk32;
  const RysArgs32* ra32;
};

struct GradLaunch {
  const int2* tasks;
  long long ntask;
  const PairHdr* pairs;
  const double4* ppos;
  const double* pk;
  const double* pa;       // exponent of each primitive pair's first shell
  const RysArgs* ra;
  const double* P;
  int nbf;
  const int* sh_atom;
  int natoms;
  double kscale;
  double* gJ;             // 3 * natoms, xyz-major
  double* gK;
};

struct SchwarzLaunch {
  const int* plist;
  int npair;
  const PairHdr* pairs;
  const double4* ppos;
  const double* pk;
  const RysArgs* ra;
  double* q2;
};

void launch_jk(int bra_cls, int ket_cls, bool pure, const JKLaunch& a, hipStream_t s);
void launch_schwarz(int pair_cls, bool pure, const SchwarzLaunch& a, hipStream_t s);
void launch_grad(int bra_cls, int ket_cls, bool pure, const GradLaunch& a, hipStream_t s);
void launch_shell_norm(const double* P, int nbf, const int* sh_bf0, const int* sh_nf, int nsh,
                       float* Pn, double* Pnd, int* bad, double* pmax, hipStream_t s);
void launch_task_count(const int2* items, long long nitems, const int2* buckets,
                       const int* pair_bucket, const int2* pair_shells, const double* Q,
                       const float* Pn, int nsh, const double* pmax, double thr, double mix_thr,
                       int* counts, hipStream_t s);
void launch_task_write(const int2* items, long long nitems, const int2* buckets,
                       const int* pair_bucket, const int2* pair_shells, const double* Q,
                       const float* Pn, int nsh, const double* pmax, double thr, double mix_thr,
                       const int* counts, const long long* offsets, int2* tasks, hipStream_t s);
size_t scan_temp_bytes(long long n);
void launch_exclusive_scan(void* temp, size_t temp_bytes, const int* in, long long* out,
                           long long n, hipStream_t s);
void launch_gather_offsets(const long long* offsets, const long long* idx, int n, long long* out,
                           hipStream_t s);
void launch_class_stats(const int* counts, const int2* items, long long nitems, const int* item_cp,
                        const int2* buckets, const PairHdr* pairs, int ncp, double* q_out, double* p_out,
                        hipStream_t s);
void launch_symmetrize(double* J, double* K, int n, hipStream_t s);

}  // namespace qfe
