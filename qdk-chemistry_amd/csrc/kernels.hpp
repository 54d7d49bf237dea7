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

// Per-launch extras of the segment J/K kernel (passed by value).
constexpr int kMaxColTiles = 8;
struct SegExtra {
  const double* pmax;       // max shell-block norm of the screening density (device scalar)
  int prim_screen;          // 1: skip primitive quartets below the engine precision
  int det;                  // 1: order-independent fixed-point accumulation (Jx, Kx)
  unsigned long long* Jx;   // (hi, lo) pairs per J element, deterministic builds
  unsigned long long* Kx;   // (hi, lo) pairs per K element, deterministic builds
  int* fx_over;             // set to 1 if a value leaves the fixed-point range
  const int4* kets;         // per pair {prim0, nprim, sa, sb}: the ket side of a task
  const double4* shells;    // per shell {x, y, z, first AO}
  int ntile;                // column tiles of the exchange rows (1: whole rows)
  int tile_lo[kMaxColTiles], tile_hi[kMaxColTiles];  // [lo, hi) columns of each tile (shell-aligned)
};

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
  // primitive screening, deterministic accumulation and column tiles
  const double* pmax;
  int prim_screen;
  int det;
  unsigned long long* Jx;
  unsigned long long* Kx;
  int* fx_over;
  const int* sh_bf0;        // host: first function of each shell, nsh + 1 entries (tile boundaries)
  int nsh;
  int kcol_tile;            // > 0: exchange-row tiles of at most this many columns (testing knob)
  const int4* kets;
  const double4* shells;
};

struct GradLaunch {
  const int2* tasks;
  long long ntask;
  const PairHdr* pairs;
  const double4* ppos;
  const double* pk;
  const double* pa;       // exponent of each primitive pair's first shell
  const RysArgs* ra;
  const double* P;        // Coulomb density (the total density of a spin pair)
  const double* Pk;       // exchange density: P itself (RHF) or the alpha density
  const double* Pk2;      // null (RHF) or the beta density
  int nbf;
  const int* sh_atom;
  int natoms;
  double kscale;
  double* gJ;             // 3 * natoms, xyz-major
  double* gK;
  const double* sph;      // device copy of the solid-harmonic table kSph (runtime-l kernel)
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
// highest shell angular momentum the gradient kernels of this build cover
int grad_max_l();
// the gradient of one class pair with runtime angular momenta (classes with
// f shells; eri_grad_any.hip)
void launch_grad_any(const GradLaunch& a, int la, int lb, int lc, int ld, hipStream_t s);
}  // namespace qfe
