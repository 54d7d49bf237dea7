// Host-side declarations of the screening, task-list, digestion-support and
// copy kernels (implemented in kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace qfe {

struct PairHdr;

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

// J_cd by ket pair: the class kernels write each task's J_cd share to jbuf
// (NKETS doubles at a per-task offset) with key = ket pair and value = that
// offset; the keys are sorted and every ket pair's run is summed into J with
// one atomic per J element.  jcd_sort_temp_bytes sizes the radix-sort
// workspace for n keys of `bits` bits; launch_jcd_reduce sorts (double
// buffers keys0/keys1, vals0/vals1), finds each pair's run (run_start,
// run_end: npair ints, cleared here) and adds the sums to J.
size_t jcd_sort_temp_bytes(long long n, int bits);
// cls_begin (host, ncls + 1 entries): first pair of each pair class (pairs
// are sorted by class)
void launch_jcd_reduce(void* temp, size_t temp_bytes, unsigned* keys0, unsigned* keys1, unsigned* vals0,
                       unsigned* vals1, long long n, int bits, const PairHdr* pairs, int npair,
                       const int* cls_begin, int ncls, int* run_start, int* run_end, const double* jbuf, double* J,
                       int nbf, hipStream_t s);

// plain device-to-device copy of n 8-byte words (also the reference kernel for
// HBM traffic counters: it moves exactly 8n bytes each way)
void launch_copy_words(const unsigned long long* src, unsigned long long* dst, size_t n, hipStream_t s);

}  // namespace qfe
