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

// plain device-to-device copy of n 8-byte words (also the reference kernel for
// HBM traffic counters: it moves exactly 8n bytes each way)
void launch_copy_words(const unsigned long long* src, unsigned long long* dst, size_t n, hipStream_t s);

}  // namespace qfe
