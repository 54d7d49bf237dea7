// Launch descriptors and class dispatch of the ERI class kernels (implemented
// in the generated per-class translation units).
#pragma once

#include <hip/hip_runtime.h>

#ifndef QFE_MAX_L
#define QFE_MAX_L 3  // highest shell angular momentum compiled into the ERI kernels
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
  // segment view of the same tasks: segment s covers items [segs[s].x, segs[s].y)
  // whose tasks are [offs[x], offs[y]) of tasks_all; all share one bra pair
  const int2* segs;
  int nseg;
  const long long* offs;
  const int2* tasks_all;
  int lds_max;        // dynamic LDS bytes a block may request (negative: exchange rows only)
  // J_cd by ket pair (jbuf null: one global atomic per task function pair):
  // the task (p, q) of segment item range [x, y) writes its share to
  // jbuf + jslot[x + qpos[q].x] + qpos[q].y * NKETS (qpos = ket bucket of q
  // relative to the first bucket of its class, position of q in that
  // bucket); the first chunk of the class stores it, later chunks add to it
  double* jbuf;
  const long long* jslot;
  const int2* qpos;
  // keep the density rows in LDS when that still leaves at least this many
  // resident waves per CU (else the K-row-only variant, if it fits more)
  int seg_min_waves;
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
  double* blk;  // staging of the (ab|ab) blocks of f-shell classes: npair x schwarz_block(c) doubles
};

// doubles of the Schwarz staging block per pair of pair class c (0 when the
// class needs no staging); implemented with the class dispatch
int schwarz_block(int pair_cls);

void launch_jk(int bra_cls, int ket_cls, bool pure, const JKLaunch& a, hipStream_t s);
void launch_schwarz(int pair_cls, bool pure, const SchwarzLaunch& a, hipStream_t s);
void launch_grad(int bra_cls, int ket_cls, bool pure, const GradLaunch& a, hipStream_t s);
}  // namespace qfe
