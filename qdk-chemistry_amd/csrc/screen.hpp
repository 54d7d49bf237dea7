// Host-side declarations of the screening, task-list, digestion-support and
// copy kernels (implemented in kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace qfe {

struct PairHdr;

// Pn[s1][s2] = max |P_block| (double), pmax = max over all blocks, bad = 1 on NaN/Inf
void launch_shell_norm(const double* P, int nbf, const int* sh_bf0, const int* sh_nf, int nsh, double* Pn, int* bad,
                       double* pmax, hipStream_t s);
// Inputs of the task-list scan: work items (bra pair, ket bucket) and their
// segments (the items of one bra pair within one class pair), buckets, pair
// shells and Schwarz factors, the density norms Pn (nsh x nsh) and per pair
// (Pcd), both in double, the thresholds, and per item the class pair and per class pair the
// functions per ket pair.  With segments and 16 nsh bytes of LDS available
// (lds_max) the scan runs one block per segment, else one wave per item.
struct ScanArgs {
  const int2* items;
  long long nitems;
  const int2* segs;
  int nseg;
  const int2* buckets;
  const int* pair_bucket;
  const int2* psh;
  const double* Q;
  const double* Pn;
  const double* Pcd;
  int nsh;
  const double* pmax;
  double thr, mix_thr;
  const int* item_cp;
  const int* cp_nk;
  int lds_max;
};
void launch_pair_norm(const int2* psh, int npair, const double* Pn, int nsh, double* Pcd, hipStream_t s);
// counts [2][nitems] (FP64, FP32 survivors), nslot[nitems] (J_cd slots) and
// nmask[nitems] (64-bit survivor-mask words)
void launch_task_count(const ScanArgs& a, int* counts, int* nslot, int* nmask, hipStream_t s);
// tasks at the scanned offsets; with jmask, each item's survivor mask from
// its word base mbase[item]
void launch_task_write(const ScanArgs& a, const int* counts, const long long* offsets, int2* tasks,
                       const long long* mbase, unsigned long long* jmask, hipStream_t s);
// workspace of launch_exclusive_scan and launch_sum for up to n ints (both
// accumulate in 64 bits)
size_t scan_temp_bytes(long long n);
void launch_exclusive_scan(void* temp, size_t temp_bytes, const int* in, long long* out,
                           long long n, hipStream_t s);
// 64-bit sum of n ints (the check of a scan's total); same workspace
void launch_sum(void* temp, size_t temp_bytes, const int* in, long long* out, long long n, hipStream_t s);
void launch_gather_offsets(const long long* offsets, const long long* idx, int n, long long* out,
                           hipStream_t s);
void launch_class_stats(const int* counts, const int2* items, long long nitems, const int* item_cp,
                        const int2* buckets, const PairHdr* pairs, int ncp, double* q_out, double* p_out,
                        hipStream_t s);
void launch_symmetrize(double* J, double* K, int n, hipStream_t s);

// J_cd by ket pair.  Item i (bra pair p, ket bucket kb) owns nslot[i] / nk
// share slots of nk doubles from jbuf + jslot[i], one per position j of the
// Schwarz prefix of its bucket range (ket pair = first pair of kb + j), and
// survivor-mask words from jmask + mbase[i].  The class kernels store the
// shares of the surviving quartets.  launch_jcd_tiles sums them per ket pair
// into J for the tiles of one ket pair size nk (nsb = functions of the pair's
// second shell); a tile is int4 {first item, stride << 16 | bra pairs
// (<= 256), first pair of the ket bucket, first position (a multiple of 64)}.
// (Jx non-null: the sums go to the fixed-point pairs of a deterministic build)
void launch_jcd_tiles(const int4* tiles, int ntiles, int nk, int nsb, const int* nslot, const long long* jslot,
                      const long long* mbase, const unsigned long long* jmask, const double* jbuf,
                      const PairHdr* pairs, double* J, int nbf, unsigned long long* Jx, int* over, hipStream_t s);
// J, K (either may be null) from the fixed-point sums of a deterministic
// build, symmetrised: J = (Jx + Jx^T) / 4, K = kq (Kx + Kx^T) / 2
void launch_fixed_finish(const unsigned long long* Jx, const unsigned long long* Kx, double* J, double* K, int n,
                         double kq, hipStream_t s);
// Profiling: add to *count the primitive quartets that the class kernels
// evaluate for the tasks [0, ntask): the same prefix walk over primitives
// sorted by sig, in the same precision and with the same cut, without the
// integrals
void launch_count_prims(const int2* tasks, long long ntask, const PairHdr* pairs, const double* sig,
                        const float* sig32, const double* pmax, int prim_screen, int fp32,
                        unsigned long long* count, hipStream_t s);
constexpr int kJcdTileBra = 256;   // bra pairs per tile
constexpr int kJcdTileKet = 1024;  // ket positions per tile

// plain device-to-device copy of n 8-byte words (also the reference kernel for
// HBM traffic counters: it moves exactly 8n bytes each way)
void launch_copy_words(const unsigned long long* src, unsigned long long* dst, size_t n, hipStream_t s);

}  // namespace qfe
