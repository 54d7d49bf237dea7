// Host side of the Fock-build engine: the C ABI of include/qfe.h.
//
// qfe_create   restates libint2::direct::ERI::ERI (libint2_direct.cpp:204-226):
//              significant shell pairs (same centre or ||S_ab||_F >=
//              shell_pair_threshold, :83-151), primitive-pair tables, Schwarz
//              factors on the GPU (schwarz.cpp:26-78); then buckets the pairs
//              by (class, primitive count), Schwarz-descending inside a bucket.
// qfe_build_jk restates ERI::build_JK (:256-626) as: shell-block density
//              norms -> two-pass screened task list (same test as :394-402,
//              same canonical quartet set and degeneracy as :357-411) ->
//              class kernels with J/K digestion -> symmetrisation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <limits>
#include <tuple>
#include <utility>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <dlfcn.h>
#include <fstream>
#include <map>
#include <mutex>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include "../../include/qfe.h"
#include "eri_core.hpp"
#include "host_common.hpp"
#include "kernels.hpp"
#include "dfj.hpp"
#include "eri_dense.hpp"
#include "onebody_gpu.hpp"
#include "screen.hpp"

namespace qfe {

thread_local std::string g_last_error;

struct DeviceError : std::runtime_error {
  using std::runtime_error::runtime_error;
};
// functions of a shell pair of pair class c = la (la + 1) / 2 + lb (pure)
static int nfun_pair_class(int c) {
  int la = 0;
  while ((la + 1) * (la + 2) / 2 <= c) ++la;
  const int lb = c - la * (la + 1) / 2;
  return nsph(la, true) * nsph(lb, true);
}

struct Unsupported : std::runtime_error {
  using std::runtime_error::runtime_error;
};

#define HIP_CHECK(x)                                                                       \
  do {                                                                                     \
    hipError_t e_ = (x);                                                                   \
    if (e_ != hipSuccess)                                                                  \
      throw ::qfe::DeviceError(std::string("HIP error: ") + hipGetErrorString(e_) + " at " #x);  \
  } while (0)

// Set while the no-alloc entry point runs on this thread: a DevBuf that would
// have to grow then throws instead of calling hipMalloc.
thread_local bool g_no_alloc = false;

template <typename T>
struct DevBuf {
  T* p = nullptr;
  size_t n = 0;
  DevBuf() = default;
  DevBuf(const DevBuf&) = delete;
  DevBuf& operator=(const DevBuf&) = delete;
  DevBuf(DevBuf&& o) noexcept : p(o.p), n(o.n) {
    o.p = nullptr;
    o.n = 0;
  }
  ~DevBuf() {
    if (p) (void)hipFree(p);
  }
  void alloc(size_t count) {
    if (count <= n && p) return;
    if (g_no_alloc) throw std::logic_error("a device allocation inside the no-alloc entry point");
    if (p) (void)hipFree(p);
    p = nullptr;
    n = 0;
    if (count == 0) count = 1;
    HIP_CHECK(hipMalloc((void**)&p, count * sizeof(T)));
    n = count;
  }
  void upload(const std::vector<T>& v, hipStream_t s) {
    alloc(v.size());
    if (!v.empty()) HIP_CHECK(hipMemcpyAsync(p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, s));
  }
};

// ------------------------------------------------------------ Rys table
static std::string library_dir() {
  Dl_info info;
  if (dladdr((void*)&library_dir, &info) && info.dli_fname) {
    std::string f = info.dli_fname;
    auto k = f.rfind('/');
    if (k != std::string::npos) return f.substr(0, k);
  }
  return ".";
}

struct RysTableHost {
  int nmax = 0, deg = 0;
  double h = 1.0;
  std::vector<std::vector<double>> coef;  // per n
  std::vector<int> nint;
  std::vector<double> tmax;
  std::vector<std::vector<double>> hx2, hw;
};

static RysTableHost load_rys_table() {
  RysTableHost t;
  const char* env = getenv("QFE_RYS_TABLE");
  std::string path = env ? env : library_dir() + "/rys_table.bin";
  std::ifstream f(path, std::ios::binary);
  if (!f) throw std::runtime_error("cannot open Rys table " + path);
  int32_t nmax, deg;
  double h;
  f.read((char*)&nmax, 4);
  f.read((char*)&deg, 4);
  f.read((char*)&h, 8);
  if (!f || deg != RYS_DEG || nmax < 1 || nmax > 8) throw std::runtime_error("bad Rys table header in " + path);
  t.nmax = nmax;
  t.deg = deg;
  t.h = h;
  t.coef.resize(nmax + 1);
  t.nint.assign(nmax + 1, 0);
  t.tmax.assign(nmax + 1, 0.0);
  t.hx2.resize(nmax + 1);
  t.hw.resize(nmax + 1);
  for (int n = 1; n <= nmax; ++n) {
    int32_t nint;
    double tmax;
    f.read((char*)&nint, 4);
    f.read((char*)&tmax, 8);
    if (!f || nint < 1 || nint > 4096) throw std::runtime_error("bad Rys table in " + path);
    t.nint[n] = nint;
    t.tmax[n] = tmax;
    t.coef[n].resize((size_t)nint * 2 * n * (deg + 1));
    f.read((char*)t.coef[n].data(), t.coef[n].size() * 8);
    t.hx2[n].resize(n);
    t.hw[n].resize(n);
    f.read((char*)t.hx2[n].data(), n * 8);
    f.read((char*)t.hw[n].data(), n * 8);
  }
  if (!f) throw std::runtime_error("truncated Rys table " + path);
  return t;
}

// loaded once, thread-safe (function-local static)
static const RysTableHost& rys_table() {
  static const RysTableHost t = load_rys_table();
  return t;
}

// host evaluation (used by the one-electron code)
void rys_host(int n, double T, double* u, double* w) {
  const RysTableHost& t = rys_table();
  if (T >= t.tmax[n]) {
    const double it = 1.0 / T, is = std::sqrt(it);
    for (int i = 0; i < n; ++i) {
      u[i] = t.hx2[n][i] * it;
      w[i] = t.hw[n][i] * is;
    }
    return;
  }
  const int k = (int)T;
  const double x = T - (k + 0.5);
  const double* c = t.coef[n].data() + (size_t)k * 2 * n * (t.deg + 1);
  for (int i = 0; i < 2 * n; ++i) {
    const double* ci = c + i * (t.deg + 1);
    double v = ci[t.deg];
    for (int d = t.deg - 1; d >= 0; --d) v = v * x + ci[d];
    if (i < n)
      u[i] = v;
    else
      w[i - n] = v;
  }
}

// ------------------------------------------------------------ the engine
// ------------------------------------------------------------ no-alloc scratch
// Per-build buffers of qfe_build_jk_device_noalloc (the task list, the J_cd
// share slots and their survivor masks): a file-scope device array of fixed
// size, so that the build itself allocates nothing.  Every GPU the library
// runs on holds its own copy, at its own address, until the process ends;
// host code reaches it only through the address hipGetSymbolAddress returns
// for the current device (the array's name is not a device address there).
constexpr size_t kNoAllocScratchBytes = (size_t)256 << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

// the device address of `scratch` on `device`, looked up once per device with
// that device current (the caller has made it current)
static unsigned char* noalloc_scratch(int device) {
  static std::mutex mu;
  static std::map<int, unsigned char*> addr;
  std::lock_guard<std::mutex> lock(mu);
  auto it = addr.find(device);
  if (it != addr.end()) return it->second;
  void* p = nullptr;
  HIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  if (!p) throw DeviceError("hipGetSymbolAddress returned no address for the no-alloc scratch");
  return addr.emplace(device, static_cast<unsigned char*>(p)).first->second;
}

struct Engine {
  // input
  BasisHost basis;
  double eri_thr = 1e-12, sp_thr = 1e-12;
  int device = 0;
  bool mix = false;
  double mix_thr = 1e-6;
  int verbose = 0;
  bool profile = false;
  int nbf = 0, nsh = 0, natoms = 0;
  std::vector<hipEvent_t> cls_ev;    // [2 precisions][ncp][2] when profiling
  std::vector<double> cls_ms;        // [2][ncp]

  hipStream_t stream = nullptr;
  hipEvent_t ev[4] = {};

  // pairs (sorted by class, nprim, Q desc)
  std::vector<PairHdr> pairs_h;
  std::vector<double> Q_h;
  std::vector<int2> buckets_h;       // [begin, end) in the sorted pair array
  std::vector<int> bucket_cls_h;
  std::vector<int> pair_bucket_h;
  std::vector<int2> items_h;         // (bra pair, ket bucket)
  std::vector<int> cp_item_begin;    // per class pair: first item (size ncp + 1)
  std::vector<int2> segs_h;          // item ranges [x, y) of one bra pair within a class pair
  std::vector<int> cp_seg_begin;     // per class pair: first segment (size ncp + 1)
  DevBuf<int2> segs_d;
  int lds_max = 65536;               // dynamic LDS bytes a block may request
  // J_cd by ket pair: share slots per item (counts nslot, bases jslot, both
  // nitems + 1 long), the slots themselves, per pair (ket bucket relative to
  // the first bucket of its class, position in its bucket), functions per
  // ket pair of each class pair, and the reduction tiles grouped by ket pair
  // size: tiles [begin, begin + count) of group g have nk, nsb = jt_groups[g]
  bool jcd_slots = true;
  int seg_min_waves = 8;
  // deterministic builds (use_atomics = false): J and K summed as
  // fixed-point integer pairs, so that repeated builds agree bitwise
  bool det = true;
  DevBuf<unsigned long long> Jx_d, Kx_d;
  DevBuf<int> fxo_d;                 // fixed-point range overflow flag
  // primitive-quartet screening at the engine precision (QFE_PRIM_SCREEN=0: off)
  bool prim_screen = true;
  int kcol_tile = 0;                 // QFE_KCOL_TILE: cap on the exchange-row tile width
  std::vector<int> sh_bf0_h;         // first function of each shell (+ nbf), host
  std::vector<int> cls_pair_begin;  // first pair of each pair class (pairs are class-sorted)
  DevBuf<int> nslot_d, nmask_d;
  DevBuf<long long> jslot_d, mbase_d;
  DevBuf<unsigned long long> jmask_d;  // survivor-mask words of every item's Schwarz prefix
  DevBuf<long long> sums_d;  // 64-bit sums of the task counts and of the slot counts
  DevBuf<double> jbuf_d;
  DevBuf<int2> qpos_d;
  DevBuf<int> cp_nk_d;
  DevBuf<int4> jtiles_d;
  struct JTileGroup {
    int nk, nsb, begin, count;
  };
  std::vector<JTileGroup> jt_groups;
  long long jslot_total = 0;
  std::vector<int> cp_bra, cp_ket;   // class pair -> (bra class, ket class)
  long long prim_pairs = 0;

  DevBuf<PairHdr> pairs_d;
  DevBuf<int2> psh_d;  // (sa, sb) per sorted pair, for the screening scan
  DevBuf<double> Pcd_d;  // per pair: density norm of its shell block, per build
  DevBuf<double4> ppos_d;
  DevBuf<double> pk_d;
  DevBuf<double> pa_d;      // exponent of each primitive pair's first shell (gradients)
  DevBuf<double> h2_d;      // 1/(2p) per primitive pair (p = exponent sum)
  DevBuf<float> h2_32_d;
  DevBuf<double> sig_d;     // primitive-quartet screening factor per primitive pair
  DevBuf<float> sig32_d;
  DevBuf<int> sh_atom_d;
  DevBuf<double> grad_d;    // [2][3 * natoms]
  DevBuf<float4> ppos32_d;  // single-precision copies for the FP32 tasks
  DevBuf<float> pk32_d;
  DevBuf<double> Q_d;
  DevBuf<int2> buckets_d;
  DevBuf<int> pair_bucket_d;
  DevBuf<int2> items_d;
  DevBuf<int> item_cp_d;
  DevBuf<double> cstat_d;           // [4][ncp]: quartets, prim quartets (FP64), then (FP32)
  std::vector<double> cstat_h;
  // executed primitive quartets per class pair [2][ncp] (profiled engines:
  // the census minus those the primitive screening skips), counted on demand
  // over the last J/K build's task list (last_bval: its class-pair bounds)
  DevBuf<unsigned long long> pqc_d;
  mutable std::vector<unsigned long long> pqc_h;
  mutable bool pqc_valid = false;
  std::vector<long long> last_bval;
  DevBuf<int> counts_d;
  DevBuf<long long> offsets_d;
  DevBuf<long long> bidx_d, bval_d;
  DevBuf<unsigned char> scan_tmp_d;
  size_t scan_tmp_bytes = 0;
  DevBuf<int2> tasks_d;
  // the task list, J_cd share slots and survivor masks of the current build
  // (the engine buffers above, or a caller's scratch)
  int2* tasks_p = nullptr;
  double* jbuf_p = nullptr;
  unsigned long long* jmask_p = nullptr;
  DevBuf<int> sh_bf0_d, sh_nf_d;
  DevBuf<int4> kets_d;        // per pair {prim0, nprim, sa, sb} (class kernels' ket side)
  DevBuf<double4> shells_d;   // per shell {x, y, z, first AO}
  DevBuf<double> Pn_d, pmax_d;  // shell-block density norms (nsh x nsh) and their maximum
  DevBuf<int> bad_d;
  DevBuf<double> P_d, J_d, K_d;      // device buffers handed out by qfe_device_buffers (SCF context too)
  DevBuf<double> hP_d, hJ_d, hK_d;   // staging of the host-pointer build (ndm x nbf x nbf)
  DevBuf<double> gP_d;               // gradient densities: [P or P_alpha][P_beta][P_alpha + P_beta]
  DevBuf<double> sph_d;              // solid-harmonic table kSph (runtime-l gradient kernel)
  std::vector<DevBuf<double>> rys_d;
  DevBuf<double> boys_d;
  std::vector<DevBuf<float>> rys32_d;
  DevBuf<float> boys32_d;
  RysArgs ra{};
  RysArgs32 ra32{};

  qfe_stats stats{};

  // extra streams for the class kernels of one build (default priority):
  // kernels of different classes run concurrently and fill each other's
  // tails; J and K take their contributions through atomics (integer ones in
  // deterministic builds, so the result does not depend on the order) and
  // every class writes its own J_cd share slots, so the classes are
  // independent.  The chunks of one class stay in order on one stream.
  // Profiled engines launch on one stream (clean per-class times).
  std::vector<hipStream_t> side;
  std::vector<hipEvent_t> side_ev;
  hipEvent_t fork_ev = nullptr;
  std::vector<hipEvent_t> done_ev;  // per (precision, class pair): its kernels are done

  ~Engine() {
    for (auto& e : done_ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : side_ev)
      if (e) (void)hipEventDestroy(e);
    if (fork_ev) (void)hipEventDestroy(fork_ev);
    for (auto& st : side)
      if (st) (void)hipStreamDestroy(st);
    for (auto& e : ev)
      if (e) (void)hipEventDestroy(e);
    for (auto& e : cls_ev)
      if (e) (void)hipEventDestroy(e);
    if (stream) (void)hipStreamDestroy(stream);
  }

  // F_m(k*BOYS_DT), m < BOYS_W: series for the top order in long double,
  // F_m(T) = e^-T sum_k (2T)^k / ((2m+1)(2m+3)...(2m+2k+1)), then downward
  static std::vector<double> boys_table() {
    std::vector<double> t((size_t)BOYS_NG * BOYS_W);
    for (int g = 0; g < BOYS_NG; ++g) {
      const long double T = g * (long double)BOYS_DT;
      const int mt = BOYS_W - 1;
      long double term = 1.0L / (2 * mt + 1), sum = term;
      for (int k = 1; k < 4000; ++k) {
        term *= 2.0L * T / (2 * mt + 2 * k + 1);
        sum += term;
        if (term < sum * 1e-22L) break;
      }
      const long double eT = expl(-T);
      long double F = eT * sum;
      t[(size_t)g * BOYS_W + mt] = (double)F;
      for (int m = mt - 1; m >= 0; --m) {
        F = (2.0L * T * F + eT) / (2 * m + 1);
        t[(size_t)g * BOYS_W + m] = (double)F;
      }
    }
    return t;
  }

  static std::vector<float> to_float(const std::vector<double>& v) {
    std::vector<float> f(v.size());
    for (size_t i = 0; i < v.size(); ++i) f[i] = (float)v[i];
    return f;
  }

  void setup_rys() {
    const RysTableHost& t = rys_table();
    const std::vector<double> bt = boys_table();
    boys_d.upload(bt, stream);
    boys32_d.upload(to_float(bt), stream);
    rys32_d.resize(t.nmax + 1);
    std::memset(&ra32, 0, sizeof(ra32));
    for (int n = 1; n <= t.nmax && n <= 8; ++n) {
      rys32_d[n].upload(to_float(t.coef[n]), stream);
      ra32.coef[n] = rys32_d[n].p;
      ra32.tmax[n] = (float)t.tmax[n];
      for (int i = 0; i < n; ++i) {
        ra32.hx2[n][i] = (float)t.hx2[n][i];
        ra32.hw[n][i] = (float)t.hw[n][i];
      }
    }
    ra32.boys = boys32_d.p;
    rys_d.resize(t.nmax + 1);
    std::memset(&ra, 0, sizeof(ra));
    for (int n = 1; n <= t.nmax && n <= 8; ++n) {
      rys_d[n].upload(t.coef[n], stream);
      ra.coef[n] = rys_d[n].p;
      ra.tmax[n] = t.tmax[n];
      for (int i = 0; i < n; ++i) {
        ra.hx2[n][i] = t.hx2[n][i];
        ra.hw[n][i] = t.hw[n][i];
      }
    }
    ra.boys = boys_d.p;
  }

  // primitive-quartet screening factor of one primitive pair with exponent
  // sum p and prefactor pk = c_a c_b exp(-ab/p AB^2) / p: ln(|pk| p^-1/4)
  // + ln(2^1/2 pi^5/2) / 2, so that 2 pi^5/2 |pk_ab pk_cd| / sqrt(p + q) <=
  // exp(sig_ab + sig_cd) (p + q >= 2 sqrt(pq))
  static double prim_sig(double p, double pk) {
    return std::log(std::fabs(pk)) - 0.25 * std::log(p) + 0.5 * std::log(std::sqrt(2.0) * 17.493418327624862);
  }

  // 1/(2p) and the screening factor of every primitive pair, in the order of
  // the ppos / pk tables they go with
  void upload_h2(const std::vector<double4>& pp, const std::vector<double>& pkv) {
    std::vector<double> h(pp.size()), sg(pp.size());
    for (size_t i = 0; i < pp.size(); ++i) {
      h[i] = 0.5 / pp[i].x;
      sg[i] = prim_sig(pp[i].x, pkv[i]);
    }
    h2_d.upload(h, stream);
    h2_32_d.upload(to_float(h), stream);
    sig_d.upload(sg, stream);
    sig32_d.upload(to_float(sg), stream);
    ra.h2 = h2_d.p;
    ra32.h2 = h2_32_d.p;
    ra.sig = sig_d.p;
    ra32.sig = sig32_d.p;
  }

  void build_pairs() {
    // significant pairs (libint2_direct.cpp:83-151)
    struct RawPair {
      int s1, s2;  // oriented: l(s1) >= l(s2)
    };
    std::vector<RawPair> raw;
    for (int s1 = 0; s1 < nsh; ++s1)
      for (int s2 = 0; s2 <= s1; ++s2) {
        const auto& A = basis.shells[s1];
        const auto& B = basis.shells[s2];
        bool same = A.O[0] == B.O[0] && A.O[1] == B.O[1] && A.O[2] == B.O[2];
        bool sig = same;
        if (!same) sig = overlap_block_norm(basis, s1, s2) >= sp_thr;
        if (!sig) continue;
        if (A.l >= B.l)
          raw.push_back({s1, s2});
        else
          raw.push_back({s2, s1});
      }
    // primitive pairs
    std::vector<PairHdr> hdr(raw.size());
    std::vector<double4> ppos;
    std::vector<double> pk, pa;
    for (size_t i = 0; i < raw.size(); ++i) {
      const auto& A = basis.shells[raw[i].s1];
      const auto& B = basis.shells[raw[i].s2];
      PairHdr& h = hdr[i];
      for (int d = 0; d < 3; ++d) {
        h.A[d] = A.O[d];
        h.B[d] = B.O[d];
      }
      h.sa = raw[i].s1;
      h.sb = raw[i].s2;
      h.bfa = basis.bf0[h.sa];
      h.bfb = basis.bf0[h.sb];
      h.cls = A.l * (A.l + 1) / 2 + B.l;
      h.prim0 = (int)ppos.size();
      const double AB2 = (A.O[0] - B.O[0]) * (A.O[0] - B.O[0]) + (A.O[1] - B.O[1]) * (A.O[1] - B.O[1]) +
                         (A.O[2] - B.O[2]) * (A.O[2] - B.O[2]);
      for (size_t ia = 0; ia < A.a.size(); ++ia)
        for (size_t ib = 0; ib < B.a.size(); ++ib) {
          const double a = A.a[ia], b = B.a[ib], p = a + b;
          const double K = A.c[ia] * B.c[ib] * std::exp(-a * b / p * AB2);
          // the primitive pairs the reference's shell-pair data keeps: built
          // at max_engine_precision = DBL_EPSILON / 1e10
          // (libint2_direct.cpp:29,87,146)
          if (std::fabs(K) < std::numeric_limits<double>::epsilon() / 1e10) continue;
          ppos.push_back(make_double4(p, (a * A.O[0] + b * B.O[0]) / p, (a * A.O[1] + b * B.O[1]) / p,
                                      (a * A.O[2] + b * B.O[2]) / p));
          pk.push_back(K / p);
          pa.push_back(a);
        }
      h.nprim = (int)ppos.size() - h.prim0;
      h.pad = 0;
    }
    prim_pairs = (long long)ppos.size();
    const int npair = (int)hdr.size();
    stats.significant_pairs = npair;

    // Schwarz on the GPU, per pair class
    pairs_d.upload(hdr, stream);
    ppos_d.upload(ppos, stream);
    pk_d.upload(pk, stream);
    upload_h2(ppos, pk);
    DevBuf<double> q2;
    q2.alloc(npair);
    HIP_CHECK(hipMemsetAsync(q2.p, 0, sizeof(double) * npair, stream));
    const int ncls = (QFE_MAX_L + 1) * (QFE_MAX_L + 2) / 2;
    std::vector<std::vector<int>> bycls(ncls);
    for (int i = 0; i < npair; ++i) bycls[hdr[i].cls].push_back(i);
    std::vector<DevBuf<int>> plist(ncls);
    // f-shell classes stage their (ab|ab) blocks: batches of pairs that fit
    // a staging buffer of at most kSchwarzStage doubles
    constexpr size_t kSchwarzStage = (size_t)1 << 24;
    DevBuf<double> stage;
    for (int c = 0; c < ncls; ++c) {
      if (bycls[c].empty()) continue;
      plist[c].upload(bycls[c], stream);
      const int n = (int)bycls[c].size();
      const size_t per = (size_t)schwarz_block(c);
      const int batch = per ? (int)std::max<size_t>(1, std::min<size_t>((size_t)n, kSchwarzStage / per)) : n;
      if (per) stage.alloc(per * (size_t)batch);
      for (int i0 = 0; i0 < n; i0 += batch) {
        SchwarzLaunch a{plist[c].p + i0, std::min(batch, n - i0), pairs_d.p, ppos_d.p, pk_d.p, &ra, q2.p,
                        per ? stage.p : nullptr};
        launch_schwarz(c, basis.pure, a, stream);
        HIP_CHECK(hipGetLastError());
      }
    }
    std::vector<double> q2h(npair);
    HIP_CHECK(hipMemcpyAsync(q2h.data(), q2.p, sizeof(double) * npair, hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    std::vector<double> Q(npair);
    for (int i = 0; i < npair; ++i) Q[i] = std::sqrt(std::sqrt(q2h[i]));

    // sort by (class, nprim, Q desc)
    std::vector<int> ord(npair);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(), [&](int x, int y) {
      if (hdr[x].cls != hdr[y].cls) return hdr[x].cls < hdr[y].cls;
      if (hdr[x].nprim != hdr[y].nprim) return hdr[x].nprim < hdr[y].nprim;
      return Q[x] > Q[y];
    });
    pairs_h.resize(npair);
    Q_h.resize(npair);
    std::vector<double4> ppos2;
    std::vector<double> pk2, pa2;
    ppos2.reserve(ppos.size());
    pk2.reserve(pk.size());
    for (int i = 0; i < npair; ++i) {
      PairHdr h = hdr[ord[i]];
      const int old0 = h.prim0;
      h.prim0 = (int)ppos2.size();
      // primitives by screening factor, largest first: the primitive
      // quartets a build keeps form a prefix of each pair's list
      std::vector<int> po(h.nprim);
      std::iota(po.begin(), po.end(), 0);
      std::stable_sort(po.begin(), po.end(), [&](int x, int y) {
        return prim_sig(ppos[old0 + x].x, pk[old0 + x]) > prim_sig(ppos[old0 + y].x, pk[old0 + y]);
      });
      for (int k : po) {
        ppos2.push_back(ppos[old0 + k]);
        pk2.push_back(pk[old0 + k]);
        pa2.push_back(pa[old0 + k]);
      }
      pairs_h[i] = h;
      Q_h[i] = Q[ord[i]];
    }
    // buckets
    buckets_h.clear();
    bucket_cls_h.clear();
    pair_bucket_h.assign(npair, 0);
    for (int i = 0; i < npair;) {
      int j = i;
      while (j < npair && pairs_h[j].cls == pairs_h[i].cls && pairs_h[j].nprim == pairs_h[i].nprim) ++j;
      for (int k = i; k < j; ++k) pair_bucket_h[k] = (int)buckets_h.size();
      buckets_h.push_back(make_int2(i, j));
      bucket_cls_h.push_back(pairs_h[i].cls);
      i = j;
    }
    // work items ordered by class pair, then (bra bucket, ket bucket, bra pair)
    const int nb = (int)buckets_h.size();
    const int ncp = ncls * (ncls + 1) / 2;
    cp_bra.assign(ncp, 0);
    cp_ket.assign(ncp, 0);
    for (int b = 0; b < ncls; ++b)
      for (int k = 0; k <= b; ++k) {
        cp_bra[b * (b + 1) / 2 + k] = b;
        cp_ket[b * (b + 1) / 2 + k] = k;
      }
    // items bra-pair-major inside each class pair: the items (and so the
    // tasks) of one bra pair are contiguous, one segment per (bra pair, class pair)
    std::vector<std::vector<int2>> per_cp(ncp);
    for (int bb = 0; bb < nb; ++bb)
      for (int p = buckets_h[bb].x; p < buckets_h[bb].y; ++p)
        for (int kb = 0; kb <= bb; ++kb) {
          const int cb = bucket_cls_h[bb], ck = bucket_cls_h[kb];
          const int cp = cb * (cb + 1) / 2 + ck;  // cb >= ck since buckets are class-sorted
          per_cp[cp].push_back(make_int2(p, kb));
        }
    items_h.clear();
    cp_item_begin.assign(ncp + 1, 0);
    segs_h.clear();
    cp_seg_begin.assign(ncp + 1, 0);
    for (int cp = 0; cp < ncp; ++cp) {
      cp_item_begin[cp] = (int)items_h.size();
      cp_seg_begin[cp] = (int)segs_h.size();
      const int base = (int)items_h.size();
      const std::vector<int2>& v = per_cp[cp];
      for (size_t i = 0; i < v.size();) {
        size_t e = i + 1;
        while (e < v.size() && v[e].x == v[i].x) ++e;
        segs_h.push_back(make_int2(base + (int)i, base + (int)e));
        i = e;
      }
      items_h.insert(items_h.end(), v.begin(), v.end());
    }
    cp_item_begin[ncp] = (int)items_h.size();
    cp_seg_begin[ncp] = (int)segs_h.size();

    {
      const int ncls_all = (QFE_MAX_L + 1) * (QFE_MAX_L + 2) / 2;
      cls_pair_begin.assign(ncls_all + 1, npair);
      for (int i = npair - 1; i >= 0; --i) cls_pair_begin[pairs_h[i].cls] = i;
      for (int c = ncls_all - 1; c >= 0; --c) cls_pair_begin[c] = std::min(cls_pair_begin[c], cls_pair_begin[c + 1]);
    }
    pairs_d.upload(pairs_h, stream);
    std::vector<int2> psh(npair);
    for (int i = 0; i < npair; ++i) psh[i] = make_int2(pairs_h[i].sa, pairs_h[i].sb);
    psh_d.upload(psh, stream);
    {
      std::vector<int4> kr(npair);
      for (int i = 0; i < npair; ++i)
        kr[i] = make_int4(pairs_h[i].prim0, pairs_h[i].nprim, pairs_h[i].sa, pairs_h[i].sb);
      kets_d.upload(kr, stream);
      std::vector<double4> sr(nsh);
      for (int s2 = 0; s2 < nsh; ++s2)
        sr[s2] = make_double4(basis.shells[s2].O[0], basis.shells[s2].O[1], basis.shells[s2].O[2],
                              (double)basis.bf0[s2]);
      shells_d.upload(sr, stream);
    }
    Pcd_d.alloc(npair);
    ppos_d.upload(ppos2, stream);
    pk_d.upload(pk2, stream);
    pa_d.upload(pa2, stream);
    upload_h2(ppos2, pk2);
    {
      std::vector<float4> p32(ppos2.size());
      for (size_t i = 0; i < ppos2.size(); ++i)
        p32[i] = make_float4((float)ppos2[i].x, (float)ppos2[i].y, (float)ppos2[i].z, (float)ppos2[i].w);
      ppos32_d.upload(p32, stream);
      pk32_d.upload(to_float(pk2), stream);
    }
    Q_d.upload(Q_h, stream);
    buckets_d.upload(buckets_h, stream);
    pair_bucket_d.upload(pair_bucket_h, stream);
    items_d.upload(items_h, stream);
    segs_d.upload(segs_h, stream);
    std::vector<int> item_cp(items_h.size());
    for (int cp = 0; cp < ncp; ++cp)
      for (int i = cp_item_begin[cp]; i < cp_item_begin[cp + 1]; ++i) item_cp[i] = cp;
    item_cp_d.upload(item_cp, stream);
    cstat_d.alloc(4 * (size_t)ncp);
    cstat_h.assign(4 * (size_t)ncp, 0.0);
    pqc_d.alloc(2 * (size_t)ncp);
    pqc_h.assign(2 * (size_t)ncp, 0ull);
    const long long nitems = (long long)items_h.size();
    counts_d.alloc(2 * nitems + 1);
    HIP_CHECK(hipMemsetAsync(counts_d.p, 0, sizeof(int) * (2 * nitems + 1), stream));
    offsets_d.alloc(2 * nitems + 1);
    scan_tmp_bytes = scan_temp_bytes(2 * nitems + 1);
    scan_tmp_d.alloc(scan_tmp_bytes);
    nslot_d.alloc(nitems + 1);
    HIP_CHECK(hipMemsetAsync(nslot_d.p, 0, sizeof(int) * (nitems + 1), stream));
    nmask_d.alloc(nitems + 1);
    HIP_CHECK(hipMemsetAsync(nmask_d.p, 0, sizeof(int) * (nitems + 1), stream));
    jslot_d.alloc(nitems + 1);
    mbase_d.alloc(nitems + 1);
    sums_d.alloc(2);
    HIP_CHECK(hipMemsetAsync(sums_d.p, 0, sizeof(long long) * 2, stream));
    build_jcd_tiles();
    // class-pair boundaries in the scanned counts: FP64 part, then FP32 part
    std::vector<long long> bidx(2 * (ncp + 1));
    for (int cp = 0; cp <= ncp; ++cp) {
      bidx[cp] = cp_item_begin[cp];
      bidx[ncp + 1 + cp] = nitems + cp_item_begin[cp];
    }
    bidx_d.upload(bidx, stream);
    bval_d.alloc(2 * (ncp + 1));
    HIP_CHECK(hipStreamSynchronize(stream));
    if (verbose)
      fprintf(stderr, "[qfe] nsh=%d nbf=%d pairs=%d prim_pairs=%lld buckets=%d items=%lld\n", nsh, nbf,
              npair, prim_pairs, nb, nitems);
  }

  // Share-slot geometry of the J_cd reduction.  Within class pair cp, the
  // segment of bra pair p (class cb) is cp_seg_begin[cp] + p - first pair of
  // cb, and its items are the ket buckets of class ck from the first one,
  // up to p's own bucket when cb == ck: every pair of one bra bucket has the
  // same number of items, so the items (p, kb) of a bra bucket are evenly
  // strided.  Tiles cover (bra bucket, ket bucket) in blocks of kJcdTileBra
  // bra pairs x kJcdTileKet ket positions.
  void build_jcd_tiles() {
    const int npair = (int)pairs_h.size();
    const int nb = (int)buckets_h.size();
    const int ncls_all = (int)cls_pair_begin.size() - 1;
    std::vector<int> cls_bucket0(ncls_all + 1, nb);
    for (int b = nb - 1; b >= 0; --b) cls_bucket0[bucket_cls_h[b]] = b;
    for (int c = ncls_all - 1; c >= 0; --c) cls_bucket0[c] = std::min(cls_bucket0[c], cls_bucket0[c + 1]);
    std::vector<int2> qpos(npair);
    for (int q = 0; q < npair; ++q) {
      const int kb = pair_bucket_h[q];
      qpos[q] = make_int2(kb - cls_bucket0[bucket_cls_h[kb]], q - buckets_h[kb].x);
    }
    qpos_d.upload(qpos, stream);
    const int ncp = (int)cp_bra.size();
    std::vector<int> cp_nk(ncp);
    for (int cp = 0; cp < ncp; ++cp) cp_nk[cp] = nfun_pair_class(cp_ket[cp]);
    cp_nk_d.upload(cp_nk, stream);
    std::map<int, std::vector<int4>> by_nk;
    std::map<int, int> nsb_of;
    for (int cp = 0; cp < ncp; ++cp) {
      const int cb = cp_bra[cp], ck = cp_ket[cp];
      const int kbf = cls_bucket0[ck], kbl = cls_bucket0[ck + 1];  // ket buckets [kbf, kbl)
      const int bbf = cls_bucket0[cb], bbl = cls_bucket0[cb + 1];
      if (kbf >= kbl || bbf >= bbl) continue;
      int la = 0;
      while ((la + 1) * (la + 2) / 2 <= ck) ++la;
      const int lb = ck - la * (la + 1) / 2;
      const int nk = cp_nk[cp];
      nsb_of[nk] = nsph(lb, true);
      std::vector<int4>& out = by_nk[nk];
      for (int kb = kbf; kb < kbl; ++kb) {
        const int2 kbk = buckets_h[kb];
        const int ksize = kbk.y - kbk.x;
        for (int bb = std::max(bbf, kb); bb < bbl; ++bb) {
          const int2 bbk = buckets_h[bb];
          const int stride = (cb == ck) ? bb - kbf + 1 : kbl - kbf;
          const int s0 = cp_seg_begin[cp] + (bbk.x - cls_pair_begin[cb]);
          const int item0 = segs_h[s0].x + (kb - kbf);
          if (stride >= (1 << 15))
            throw std::runtime_error("too many pair buckets in one class for the J_cd tiles");
          // the strided layout must name the items built above
          for (int p : {bbk.x, bbk.y - 1}) {
            const int2 it = items_h[item0 + (p - bbk.x) * stride];
            if (it.x != p || it.y != kb) throw std::runtime_error("J_cd tile layout does not match the work items");
          }
          for (int p0 = bbk.x; p0 < bbk.y; p0 += kJcdTileBra) {
            const int np = std::min(kJcdTileBra, bbk.y - p0);
            // same bucket: bra pair p holds positions j <= p - first pair only
            const int jmax = (bb == kb) ? std::min(ksize, p0 + np - kbk.x) : ksize;
            for (int j0 = 0; j0 < jmax; j0 += kJcdTileKet)
              out.push_back(make_int4(item0 + (p0 - bbk.x) * stride, (stride << 16) | np, kbk.x, j0));
          }
        }
      }
    }
    std::vector<int4> all;
    jt_groups.clear();
    for (auto& kv : by_nk) {
      jt_groups.push_back({kv.first, nsb_of[kv.first], (int)all.size(), (int)kv.second.size()});
      all.insert(all.end(), kv.second.begin(), kv.second.end());
    }
    jtiles_d.upload(all, stream);
  }

  void create(const qfe_basis* b, const qfe_config* cfg) {
    basis = BasisHost::from_c(b);
    nsh = basis.nsh;
    nbf = basis.nbf;
    natoms = basis.natoms;
    if (nsh <= 0) throw std::invalid_argument("empty basis");
    // the class kernels address the density with 32-bit byte offsets
    if (nbf >= 23170) throw Unsupported("nbf >= 23170: density offsets exceed 32 bits");
    if (!basis.pure) throw Unsupported("cartesian (non-pure) bases are not compiled into this engine");
    for (const auto& s : basis.shells)
      if (s.l > QFE_MAX_L)
        throw Unsupported("shell angular momentum " + std::to_string(s.l) + " exceeds the compiled maximum " +
                          std::to_string(QFE_MAX_L));
    if (cfg) {
      eri_thr = cfg->eri_threshold;
      sp_thr = cfg->shell_pair_threshold;
      device = cfg->device_id;
      mix = cfg->mix_precision != 0;
      mix_thr = cfg->mix_threshold;
      verbose = cfg->verbose;
      profile = cfg->profile != 0;
      det = cfg->use_atomics == 0;
    }
    int ndev = 0;
    HIP_CHECK(hipGetDeviceCount(&ndev));
    if (device < 0 || device >= ndev) throw std::invalid_argument("device_id out of range");
    HIP_CHECK(hipSetDevice(device));
    HIP_CHECK(hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock, device));
    // QFE_SEGMENT_DIGEST=0 keeps only the exchange rows in LDS (the density
    // rows are read from L2), the variant used when both do not fit
    if (const char* e = std::getenv("QFE_SEGMENT_DIGEST"))
      if (e[0] == '0') lds_max = -lds_max;
    // QFE_SEG_WAVES: resident waves per CU below which the density rows leave LDS
    if (const char* e = std::getenv("QFE_SEG_WAVES")) seg_min_waves = std::max(1, std::atoi(e));
    // QFE_JCD_SLOTS=0 adds every task's J_cd share to J with global atomics
    if (const char* e = std::getenv("QFE_JCD_SLOTS"))
      if (e[0] == '0') jcd_slots = false;
    // QFE_PRIM_SCREEN=0 evaluates every primitive quartet
    if (const char* e = std::getenv("QFE_PRIM_SCREEN"))
      if (e[0] == '0') prim_screen = false;
    // QFE_KCOL_TILE=<columns>: exchange-row tiles of at most that many
    // columns even where whole rows fit (exercises the tiled path)
    if (const char* e = std::getenv("QFE_KCOL_TILE")) kcol_tile = std::max(0, std::atoi(e));
    HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    for (auto& e : ev) HIP_CHECK(hipEventCreate(&e));
    // QFE_STREAMS=<n>: streams the class kernels of a build spread over (default 4)
    {
      int ns = 4;
      if (const char* e = std::getenv("QFE_STREAMS")) ns = std::max(1, std::min(8, std::atoi(e)));
      side.assign(ns - 1, nullptr);
      side_ev.assign(ns - 1, nullptr);
      for (auto& st : side) HIP_CHECK(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
      for (auto& e : side_ev) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      HIP_CHECK(hipEventCreateWithFlags(&fork_ev, hipEventDisableTiming));
    }
    setup_rys();
    std::vector<int> bf0(nsh), nf(nsh);
    for (int s = 0; s < nsh; ++s) {
      bf0[s] = basis.bf0[s];
      nf[s] = basis.bf0[s + 1] - basis.bf0[s];
    }
    sh_bf0_d.upload(bf0, stream);
    sh_nf_d.upload(nf, stream);
    sh_bf0_h = bf0;
    sh_bf0_h.push_back(nbf);
    fxo_d.alloc(1);
    if (det) {  // the fixed-point sums of J and K, (hi, lo) per element
      Jx_d.alloc(2 * (size_t)nbf * nbf);
      Kx_d.alloc(2 * (size_t)nbf * nbf);
    }
    std::vector<int> sha(nsh);
    for (int s2 = 0; s2 < nsh; ++s2) sha[s2] = basis.shells[s2].atom;
    sh_atom_d.upload(sha, stream);
    {
      const double* k0 = &::kSph[0][0][0];
      sph_d.upload(std::vector<double>(k0, k0 + sizeof(::kSph) / sizeof(double)), stream);
    }
    Pn_d.alloc((size_t)nsh * nsh);
    pmax_d.alloc(1);
    bad_d.alloc(1);
    scratch_ptr = noalloc_scratch(device);
    build_pairs();
    const int ncp = (int)cp_bra.size();
    cls_ms.assign(2 * (size_t)ncp, 0.0);
    if (profile) {
      cls_ev.assign(4 * (size_t)ncp, nullptr);
      for (auto& e : cls_ev) HIP_CHECK(hipEventCreate(&e));
    }
    if (!side.empty()) {
      done_ev.assign(2 * (size_t)ncp, nullptr);
      for (auto& e : done_ev) HIP_CHECK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
    }
  }

  unsigned char* scratch_ptr = nullptr;  // this device's no-alloc scratch (device address)

  // Device memory a build may take its per-build buffers from instead of the
  // engine's own (the no-alloc entry point): base address and capacity.
  struct Scratch {
    unsigned char* base;
    size_t bytes;
  };

  // Screened canonical task list for density dP (FP64 and FP32 parts);
  // returns the class-pair boundaries [2][ncp + 1] of the task array.  The
  // task list, the J_cd share slots and their survivor masks go to tasks_p,
  // jbuf_p and jmask_p: engine buffers grown as needed, or carved from ext,
  // which must hold them (checked before any kernel writes there).
  std::vector<long long> screen(const double* dP, hipStream_t s, bool allow_mix, bool slots = false,
                                const Scratch* ext = nullptr) {
    HIP_CHECK(hipMemsetAsync(bad_d.p, 0, sizeof(int), s));
    launch_shell_norm(dP, nbf, sh_bf0_d.p, sh_nf_d.p, nsh, Pn_d.p, bad_d.p, pmax_d.p, s);
    HIP_CHECK(hipGetLastError());
    const int npair = (int)pairs_h.size();
    launch_pair_norm(psh_d.p, npair, Pn_d.p, nsh, Pcd_d.p, s);
    HIP_CHECK(hipGetLastError());
    const long long nitems = (long long)items_h.size();
    const double mthr = (mix && allow_mix) ? mix_thr : -1.0;
    const ScanArgs sa{items_d.p, nitems, segs_d.p, (int)segs_h.size(), buckets_d.p, pair_bucket_d.p, psh_d.p,
                      Q_d.p, Pn_d.p, Pcd_d.p, nsh, pmax_d.p, eri_thr, mthr, item_cp_d.p, cp_nk_d.p,
                      std::abs(lds_max)};
    launch_task_count(sa, counts_d.p, nslot_d.p, nmask_d.p, s);
    HIP_CHECK(hipGetLastError());
    launch_exclusive_scan(scan_tmp_d.p, scan_tmp_bytes, counts_d.p, offsets_d.p, 2 * nitems + 1, s);
    HIP_CHECK(hipGetLastError());
    launch_sum(scan_tmp_d.p, scan_tmp_bytes, counts_d.p, sums_d.p, 2 * nitems, s);
    HIP_CHECK(hipGetLastError());
    if (slots) {
      launch_exclusive_scan(scan_tmp_d.p, scan_tmp_bytes, nslot_d.p, jslot_d.p, nitems + 1, s);
      HIP_CHECK(hipGetLastError());
      launch_exclusive_scan(scan_tmp_d.p, scan_tmp_bytes, nmask_d.p, mbase_d.p, nitems + 1, s);
      HIP_CHECK(hipGetLastError());
      launch_sum(scan_tmp_d.p, scan_tmp_bytes, nslot_d.p, sums_d.p + 1, nitems, s);
      HIP_CHECK(hipGetLastError());
    }
    const int ncp = (int)cp_bra.size();
    HIP_CHECK(hipMemsetAsync(cstat_d.p, 0, sizeof(double) * 4 * ncp, s));
    launch_class_stats(counts_d.p, items_d.p, nitems, item_cp_d.p, buckets_d.p, pairs_d.p, ncp, cstat_d.p,
                       cstat_d.p + ncp, s);
    launch_class_stats(counts_d.p + nitems, items_d.p, nitems, item_cp_d.p, buckets_d.p, pairs_d.p, ncp,
                       cstat_d.p + 2 * ncp, cstat_d.p + 3 * ncp, s);
    HIP_CHECK(hipGetLastError());
    launch_gather_offsets(offsets_d.p, bidx_d.p, 2 * (ncp + 1), bval_d.p, s);
    std::vector<long long> bval(2 * (ncp + 1));
    int bad = 0;
    long long total = 0;
    HIP_CHECK(hipMemcpyAsync(bval.data(), bval_d.p, sizeof(long long) * 2 * (ncp + 1), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(&total, offsets_d.p + 2 * nitems, sizeof(long long), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(&bad, bad_d.p, sizeof(int), hipMemcpyDeviceToHost, s));
    long long nslots = 0, nwords = 0, sums[2] = {0, 0};
    if (slots) {
      HIP_CHECK(hipMemcpyAsync(&nslots, jslot_d.p + nitems, sizeof(long long), hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipMemcpyAsync(&nwords, mbase_d.p + nitems, sizeof(long long), hipMemcpyDeviceToHost, s));
    }
    HIP_CHECK(hipMemcpyAsync(sums, sums_d.p, sizeof(sums), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipMemcpyAsync(cstat_h.data(), cstat_d.p, sizeof(double) * 4 * ncp, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (bad) throw std::runtime_error("Density matrix contains NaN/Inf values.");
    // the scanned totals size the task list and the share slots: check them
    // against an independent sum before any kernel writes there
    if (total != sums[0] || total < 0 || (slots && (nslots != sums[1] || nslots < 0 || nwords < 0)))
      throw std::runtime_error("task-list scan totals disagree with their sums");
    const bool use_slots = slots && nslots > 0;
    if (ext) {
      auto al = [](size_t b) { return (b + 255) & ~(size_t)255; };
      const size_t t_b = al(sizeof(int2) * (size_t)total);
      const size_t m_b = use_slots ? al(sizeof(unsigned long long) * (size_t)nwords) : 0;
      const size_t j_b = use_slots ? al(sizeof(double) * (size_t)nslots) : 0;
      if (t_b + m_b + j_b > ext->bytes)
        throw std::invalid_argument("the build needs " + std::to_string((t_b + m_b + j_b) >> 20) +
                                    " MiB of task list and J share slots, more than the " +
                                    std::to_string(ext->bytes >> 20) + " MiB scratch of the no-alloc entry point");
      tasks_p = reinterpret_cast<int2*>(ext->base);
      jmask_p = use_slots ? reinterpret_cast<unsigned long long*>(ext->base + t_b) : nullptr;
      jbuf_p = use_slots ? reinterpret_cast<double*>(ext->base + t_b + m_b) : nullptr;
    } else {
      if (total > 0) tasks_d.alloc((size_t)total);
      if (use_slots) {
        jbuf_d.alloc((size_t)nslots);
        jmask_d.alloc((size_t)nwords);
      }
      tasks_p = tasks_d.p;
      jmask_p = use_slots ? jmask_d.p : nullptr;
      jbuf_p = use_slots ? jbuf_d.p : nullptr;
    }
    jslot_total = nslots;
    launch_task_write(sa, counts_d.p, offsets_d.p, tasks_p, mbase_d.p, jmask_p, s);
    HIP_CHECK(hipGetLastError());
    return bval;
  }

  // ERI gradient contributions (libint2_direct.cpp:646-890) for ndm host
  // densities: ndm = 1 is the RHF total density; ndm = 2 the spin pair
  // [P_alpha; P_beta], whose Coulomb term takes P_alpha + P_beta and whose
  // exchange term sums the two spins with scale -1/2 (:683-692, :866-886).
  // Quartets are screened with the norms of the first density block, as the
  // reference's compute_shellblock_norm(obs_, P, nbf) reads (:693).  A null
  // dJ or dK leaves that output untouched.
  void gradients(const double* P, int ndm, double* dJ, double* dK, double kscale) {
    last_bval.clear();  // the gradient's screening replaces the J/K task list
    pqc_valid = false;
    if (natoms <= 0) throw std::invalid_argument("gradients need the atoms the shells sit on");
    if (ndm < 1 || ndm > 2) throw std::invalid_argument("ndm must be 1 or 2");
    for (const auto& sh : basis.shells)
      if (sh.l > grad_max_l())
        throw Unsupported("analytic ERI gradients are compiled for shells up to l = " +
                          std::to_string(grad_max_l()));
    HIP_CHECK(hipSetDevice(device));
    const size_t nn = (size_t)nbf * nbf;
    grad_d.alloc(6 * (size_t)natoms);
    hipStream_t s = stream;
    // device layout: [P_screen/P_alpha][P_beta][P_total]
    gP_d.alloc(ndm == 1 ? nn : 3 * nn);
    HIP_CHECK(hipMemcpyAsync(gP_d.p, P, ndm * nn * sizeof(double), hipMemcpyHostToDevice, s));
    const double* Pj = gP_d.p;
    const double* Pk = gP_d.p;
    const double* Pk2 = nullptr;
    if (ndm == 2) {
      std::vector<double> tot(nn);
      for (size_t i = 0; i < nn; ++i) tot[i] = P[i] + P[nn + i];
      HIP_CHECK(hipMemcpyAsync(gP_d.p + 2 * nn, tot.data(), nn * sizeof(double), hipMemcpyHostToDevice, s));
      HIP_CHECK(hipStreamSynchronize(s));
      Pj = gP_d.p + 2 * nn;
      Pk2 = gP_d.p + nn;
    }
    HIP_CHECK(hipMemsetAsync(grad_d.p, 0, sizeof(double) * 6 * natoms, s));
    HIP_CHECK(hipEventRecord(ev[0], s));
    std::vector<long long> bval = screen(gP_d.p, s, false);
    HIP_CHECK(hipEventRecord(ev[1], s));
    const int ncp = (int)cp_bra.size();
    long long nq = 0;
    for (int cp = 0; cp < ncp; ++cp) {
      const long long n = bval[cp + 1] - bval[cp];
      if (n <= 0) continue;
      nq += n;
      GradLaunch a{tasks_p + bval[cp], n, pairs_d.p, ppos_d.p, pk_d.p, pa_d.p, &ra, Pj, Pk, Pk2, nbf,
                   sh_atom_d.p, natoms, kscale, grad_d.p, grad_d.p + 3 * natoms, sph_d.p};
      launch_grad(cp_bra[cp], cp_ket[cp], basis.pure, a, s);
      HIP_CHECK(hipGetLastError());
    }
    HIP_CHECK(hipEventRecord(ev[2], s));
    std::vector<double> g(6 * (size_t)natoms);
    HIP_CHECK(hipMemcpyAsync(g.data(), grad_d.p, sizeof(double) * g.size(), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (dJ) std::memcpy(dJ, g.data(), sizeof(double) * 3 * natoms);
    if (dK) std::memcpy(dK, g.data() + 3 * natoms, sizeof(double) * 3 * natoms);
    float t01 = 0, t12 = 0;
    HIP_CHECK(hipEventElapsedTime(&t01, ev[0], ev[1]));
    HIP_CHECK(hipEventElapsedTime(&t12, ev[1], ev[2]));
    stats.quartets = nq;
    stats.quartets_fp32 = 0;
    double npq = 0.0;
    for (int cp = 0; cp < ncp; ++cp) npq += cstat_h[ncp + cp];
    stats.prim_quartets = (long long)npq;
    stats.ms_screen = t01;
    stats.ms_eri = t12;
    stats.ms_total = t01 + t12;
  }

  // Class kernels and J_cd reduction of density dP over the task list of the
  // last screen(): J += Coulomb (dJ null: skipped), K += exchange scaled by
  // kscale (dK null: skipped).  J and K must be zeroed by the caller.
  void digest(const std::vector<long long>& bval, const double* dP, double* dJ, double* dK, double kscale,
              hipStream_t s, long long& nq, long long& nq32) {
    const int ncp = (int)cp_bra.size();
    const bool jslots = dJ && jcd_slots && jslot_total > 0;
    if (!dK) kscale = 0.0;
    // deterministic builds accumulate into the fixed-point pairs (zeroed by
    // the caller) and leave dJ / dK to the finish kernel
    unsigned long long* Jx = det && dJ ? Jx_d.p : nullptr;
    unsigned long long* Kx = det && dK ? Kx_d.p : nullptr;
    double* Jd = det ? nullptr : dJ;
    nq = nq32 = 0;
    // the class kernels over 1 + side.size() streams: the classes by work
    // (quartets x functions per quartet), largest first, each to the stream
    // with the least work so far; profiled engines keep one stream
    const int nst = profile ? 1 : 1 + (int)side.size();
    struct Job {
      double cost;
      int prec, cp;
    };
    std::vector<Job> jobs;
    for (int prec = 0; prec < 2; ++prec) {
      const long long* b = bval.data() + prec * (ncp + 1);
      for (int cp = 0; cp < ncp; ++cp) {
        const long long n = b[cp + 1] - b[cp];
        if (n <= 0) continue;
        (prec ? nq32 : nq) += n;
        jobs.push_back({(double)n * nfun_pair_class(cp_bra[cp]) * nfun_pair_class(cp_ket[cp]), prec, cp});
      }
    }
    if (nst > 1)
      std::stable_sort(jobs.begin(), jobs.end(), [](const Job& x, const Job& y) { return x.cost > y.cost; });
    std::vector<double> load(nst, 0.0);
    if (nst > 1) {
      HIP_CHECK(hipEventRecord(fork_ev, s));
      for (int k = 1; k < nst; ++k) HIP_CHECK(hipStreamWaitEvent(side[k - 1], fork_ev, 0));
    }
    for (const Job& job : jobs) {
      {
        const int prec = job.prec, cp = job.cp;
        const long long* b = bval.data() + prec * (ncp + 1);
        const long long n = b[cp + 1] - b[cp];
        const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
        load[k] += job.cost;
        hipStream_t cs = k == 0 ? s : side[k - 1];
        JKLaunch a{tasks_p + b[cp], n, pairs_d.p, ppos_d.p, pk_d.p, &ra, dP, Jd, det ? nullptr : dK, nbf, kscale,
                   prec == 1, ppos32_d.p, pk32_d.p, &ra32,
                   segs_d.p + cp_seg_begin[cp], cp_seg_begin[cp + 1] - cp_seg_begin[cp],
                   offsets_d.p + (prec ? (long long)items_h.size() : 0), tasks_p, lds_max,
                   jslots ? jbuf_p : nullptr, jslot_d.p, qpos_d.p, seg_min_waves,
                   pmax_d.p, prim_screen ? 1 : 0, det ? 1 : 0, Jx, Kx, fxo_d.p, sh_bf0_h.data(), nsh,
                   kcol_tile, kets_d.p, shells_d.p};
        hipEvent_t* e = profile ? &cls_ev[4 * (size_t)cp + 2 * prec] : nullptr;
        if (e) HIP_CHECK(hipEventRecord(e[0], cs));
        launch_jk(cp_bra[cp], cp_ket[cp], basis.pure, a, cs);
        HIP_CHECK(hipGetLastError());
        if (e) HIP_CHECK(hipEventRecord(e[1], cs));
        if (nst > 1) HIP_CHECK(hipEventRecord(done_ev[(size_t)prec * ncp + cp], cs));
      }
    }
    // J_cd sums per ket size: each group starts once the classes whose kets
    // have that size are done (the other classes may still run)
    if (jslots) {
      for (size_t gi = 0; gi < jt_groups.size(); ++gi) {
        const JTileGroup& g = jt_groups[gi];
        hipStream_t gs = s;
        if (nst > 1) {
          const int k = (int)(gi % nst);
          gs = k == 0 ? s : side[k - 1];
          for (const Job& job : jobs)
            if (nfun_pair_class(cp_ket[job.cp]) == g.nk)
              HIP_CHECK(hipStreamWaitEvent(gs, done_ev[(size_t)job.prec * ncp + job.cp], 0));
        }
        launch_jcd_tiles(jtiles_d.p + g.begin, g.count, g.nk, g.nsb, nslot_d.p, jslot_d.p, mbase_d.p, jmask_p,
                         jbuf_p, pairs_d.p, Jd, nbf, Jx, fxo_d.p, gs);
        HIP_CHECK(hipGetLastError());
      }
    }
    for (int k = 1; k < nst; ++k) {  // join: everything after waits for every stream
      HIP_CHECK(hipEventRecord(side_ev[k - 1], side[k - 1]));
      HIP_CHECK(hipStreamWaitEvent(s, side_ev[k - 1], 0));
    }
  }

  // J/K on device buffers for ndm densities dP[idm] (nbf x nbf each), all
  // screened with the norms of dP[0] (libint2_direct.cpp:271 reads the first
  // block only); dJ or dK may be null (that matrix is not built).
  void build_device(const double* const* dP, double* const* dJ, double* const* dK, int ndm, double kscale,
                    hipStream_t s, const Scratch* ext = nullptr) {
    HIP_CHECK(hipSetDevice(device));
    const size_t nn = (size_t)nbf * nbf;
    const bool wantJ = dJ != nullptr && dJ[0] != nullptr;
    const bool wantK = dK != nullptr && dK[0] != nullptr;
    HIP_CHECK(hipEventRecord(ev[0], s));
    const int ncp = (int)cp_bra.size();
    std::vector<long long> bval = screen(dP[0], s, true, wantJ && jcd_slots, ext);
    HIP_CHECK(hipEventRecord(ev[1], s));
    long long nq = 0, nq32 = 0;
    last_bval = bval;
    pqc_valid = false;
    if (det) {
      if (wantJ) Jx_d.alloc(2 * nn);
      if (wantK) Kx_d.alloc(2 * nn);
      HIP_CHECK(hipMemsetAsync(fxo_d.p, 0, sizeof(int), s));
    }
    for (int idm = 0; idm < ndm; ++idm) {
      double* Jw = wantJ ? dJ[idm] : nullptr;
      double* Kw = wantK ? dK[idm] : nullptr;
      if (det) {
        if (Jw) HIP_CHECK(hipMemsetAsync(Jx_d.p, 0, 2 * nn * sizeof(unsigned long long), s));
        if (Kw) HIP_CHECK(hipMemsetAsync(Kx_d.p, 0, 2 * nn * sizeof(unsigned long long), s));
      } else {
        if (Jw) HIP_CHECK(hipMemsetAsync(Jw, 0, nn * sizeof(double), s));
        if (Kw) HIP_CHECK(hipMemsetAsync(Kw, 0, nn * sizeof(double), s));
      }
      digest(bval, dP[idm], Jw, Kw, kscale, s, nq, nq32);
      if (det) launch_fixed_finish(Jx_d.p, Kx_d.p, Jw, Kw, nbf, 0.25 * kscale, s);
    }
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(ev[2], s));
    if (!det)
      for (int idm = 0; idm < ndm; ++idm)
        launch_symmetrize(wantJ ? dJ[idm] : nullptr, wantK ? dK[idm] : nullptr, nbf, s);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipEventRecord(ev[3], s));
    HIP_CHECK(hipEventSynchronize(ev[3]));
    if (det) {
      int over = 0;
      HIP_CHECK(hipMemcpyAsync(&over, fxo_d.p, sizeof(int), hipMemcpyDeviceToHost, s));
      HIP_CHECK(hipStreamSynchronize(s));
      if (over)
        throw std::runtime_error("a J/K contribution left the range of the deterministic fixed-point sums "
                                 "(|value| >= 2^20); build with use_atomics = true");
    }
    float t01 = 0, t12 = 0, t03 = 0;
    HIP_CHECK(hipEventElapsedTime(&t01, ev[0], ev[1]));
    HIP_CHECK(hipEventElapsedTime(&t12, ev[1], ev[2]));
    HIP_CHECK(hipEventElapsedTime(&t03, ev[0], ev[3]));
    for (int prec = 0; prec < 2; ++prec) {
      const long long* b = bval.data() + prec * (ncp + 1);
      for (int cp = 0; cp < ncp; ++cp) {
        double& m = cls_ms[(size_t)prec * ncp + cp];
        m = 0.0;
        if (profile && b[cp + 1] > b[cp]) {
          float t = 0.f;
          HIP_CHECK(hipEventElapsedTime(&t, cls_ev[4 * (size_t)cp + 2 * prec], cls_ev[4 * (size_t)cp + 2 * prec + 1]));
          m = t;
        }
      }
    }
    stats.quartets = nq + nq32;
    stats.quartets_fp32 = nq32;
    double npq = 0.0;
    for (int cp = 0; cp < ncp; ++cp) npq += cstat_h[ncp + cp] + cstat_h[3 * ncp + cp];
    stats.prim_quartets = (long long)npq;
    stats.ms_screen = t01;
    stats.ms_eri = t12;
    stats.ms_total = t03;
  }

  // the primitive quartets the class kernels evaluated in the last J/K build
  // (profiled engines; zeros otherwise, or after a gradient replaced the
  // task list): the same prefix walk, without the integrals, outside any
  // timed build
  void count_executed() const {
    if (pqc_valid) return;
    std::fill(pqc_h.begin(), pqc_h.end(), 0ull);
    const int ncp = (int)cp_bra.size();
    if (profile && last_bval.size() == 2 * (size_t)(ncp + 1)) {
      HIP_CHECK(hipSetDevice(device));
      HIP_CHECK(hipMemsetAsync(pqc_d.p, 0, sizeof(unsigned long long) * 2 * ncp, stream));
      for (int prec = 0; prec < 2; ++prec) {
        const long long* b = last_bval.data() + prec * (ncp + 1);
        for (int cp = 0; cp < ncp; ++cp)
          launch_count_prims(tasks_p + b[cp], b[cp + 1] - b[cp], pairs_d.p, sig_d.p, sig32_d.p, pmax_d.p,
                             prim_screen ? 1 : 0, prec, pqc_d.p + (size_t)prec * ncp + cp, stream);
      }
      HIP_CHECK(hipGetLastError());
      HIP_CHECK(hipMemcpyAsync(pqc_h.data(), pqc_d.p, sizeof(unsigned long long) * 2 * ncp, hipMemcpyDeviceToHost,
                               stream));
      HIP_CHECK(hipStreamSynchronize(stream));
    }
    pqc_valid = true;
  }

  // one density on device buffers (the single-density form of build_device)
  void build_device(const double* dP, double* dJ, double* dK, double kscale, hipStream_t s,
                    const Scratch* ext = nullptr) {
    const double* p[1] = {dP};
    double* j[1] = {dJ};
    double* k[1] = {dK};
    build_device(p, j, k, 1, kscale, s, ext);
  }

  void build_host(const double* P, double* J, double* K, double alpha, double beta, double omega, int ndm) {
    if (std::fabs(omega) > 1e-12)
      throw Unsupported("range-separated exchange (omega != 0) is not supported by this engine");
    if (ndm < 1 || ndm > 2) throw std::invalid_argument("ndm must be 1 or 2");
    const size_t nn = (size_t)nbf * nbf;
    hP_d.alloc(ndm * nn);
    if (J) hJ_d.alloc(ndm * nn);
    if (K) hK_d.alloc(ndm * nn);
    HIP_CHECK(hipMemcpyAsync(hP_d.p, P, ndm * nn * sizeof(double), hipMemcpyHostToDevice, stream));
    const double* p[2] = {hP_d.p, hP_d.p + nn};
    double* j[2] = {J ? hJ_d.p : nullptr, J ? hJ_d.p + nn : nullptr};
    double* k[2] = {K ? hK_d.p : nullptr, K ? hK_d.p + nn : nullptr};
    build_device(p, J ? j : nullptr, K ? k : nullptr, ndm, alpha + beta, stream);
    if (J) HIP_CHECK(hipMemcpyAsync(J, hJ_d.p, ndm * nn * sizeof(double), hipMemcpyDeviceToHost, stream));
    if (K) HIP_CHECK(hipMemcpyAsync(K, hK_d.p, ndm * nn * sizeof(double), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
  }
};

// ------------------------------------------------------------ one-electron integrals on the GPU
// Device copies of the Rys and Boys tables and of the solid-harmonic table
// for the one-electron kernels: built once per device and kept until the
// process ends (about 0.5 MB; freeing them at exit would race the runtime's
// own teardown).
struct OneBodyTables {
  RysArgs ra;
  const double* sph;
};
static double* upload_forever(const std::vector<double>& v) {
  double* p = nullptr;
  HIP_CHECK(hipMalloc((void**)&p, std::max<size_t>(v.size(), 1) * sizeof(double)));
  if (!v.empty()) HIP_CHECK(hipMemcpy(p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
  return p;
}
static const OneBodyTables& one_body_tables(int device) {
  static std::mutex mu;
  static std::map<int, OneBodyTables> cache;
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(device);
  if (it != cache.end()) return it->second;
  OneBodyTables t;
  std::memset(&t.ra, 0, sizeof(t.ra));
  const RysTableHost& rt = rys_table();
  for (int n = 1; n <= rt.nmax && n <= 8; ++n) {
    t.ra.coef[n] = upload_forever(rt.coef[n]);
    t.ra.tmax[n] = rt.tmax[n];
    for (int i = 0; i < n; ++i) {
      t.ra.hx2[n][i] = rt.hx2[n][i];
      t.ra.hw[n][i] = rt.hw[n][i];
    }
  }
  t.ra.boys = upload_forever(Engine::boys_table());
  const double* k0 = &::kSph[0][0][0];
  t.sph = upload_forever(std::vector<double>(k0, k0 + sizeof(::kSph) / sizeof(double)));
  return cache.emplace(device, t).first->second;
}

// flat copy of a basis on the current device, for one call
struct OneBodyDevice {
  DevBuf<int> l, atom, po, bf0;
  DevBuf<double> cen, exps, coefs, Z, R;
  OneBodyView view{};
  OneBodyDevice(const BasisHost& b, const double* sph) {
    std::vector<int> hl(b.nsh), hat(b.nsh), hpo(b.nsh + 1, 0), hbf(b.nsh);
    std::vector<double> hc(3 * (size_t)b.nsh), he, hk;
    for (int s = 0; s < b.nsh; ++s) {
      const ShellHost& sh = b.shells[s];
      if (sh.l > 3) throw Unsupported("one-electron GPU integrals are compiled for shells up to f");
      hl[s] = sh.l;
      hat[s] = sh.atom;
      hbf[s] = b.bf0[s];
      for (int d = 0; d < 3; ++d) hc[3 * s + d] = sh.O[d];
      he.insert(he.end(), sh.a.begin(), sh.a.end());
      hk.insert(hk.end(), sh.c.begin(), sh.c.end());
      hpo[s + 1] = (int)he.size();
    }
    l.upload(hl, nullptr);
    atom.upload(hat, nullptr);
    po.upload(hpo, nullptr);
    bf0.upload(hbf, nullptr);
    cen.upload(hc, nullptr);
    exps.upload(he, nullptr);
    coefs.upload(hk, nullptr);
    Z.upload(b.Z, nullptr);
    R.upload(b.R, nullptr);
    view = OneBodyView{l.p, atom.p, po.p, bf0.p, cen.p, exps.p, coefs.p, Z.p, R.p, sph, b.nsh, b.natoms, b.nbf,
                       b.pure ? 1 : 0};
  }
};

// one rocBLAS handle per device for the eigensolver, kept until the process
// ends (creating one costs far more than a solve of SCF size)
static rocblas_handle solver_handle(int device) {
  static std::mutex mu;
  static std::map<int, rocblas_handle> cache;
  std::lock_guard<std::mutex> lock(mu);
  auto it = cache.find(device);
  if (it != cache.end()) return it->second;
  rocblas_handle h = nullptr;
  if (rocblas_create_handle(&h) != rocblas_status_success) throw DeviceError("rocblas_create_handle failed");
  return cache.emplace(device, h).first->second;
}

static void one_body_device(int device_id);

// Density-fitted J (the reference's INCORE DF engine, incore_impl_df.cpp):
// B[P][mu >= nu] = (mu nu|P) (packed lower triangle) and the Cholesky factor
// of V[P][Q] = (P|Q) kept in HBM; a build is X = B . d, V c = X, J = B^T c
// (two GEMVs over the packed B and a triangular solve pair), with d the
// density packed as D_mu,nu + D_nu,mu below the diagonal and D_mu,mu on it.
struct DFJ {
  int device = 0, nbf = 0, naux = 0;
  DevBuf<double> B, V, d, x, j, full;
  rocblas_handle blas = nullptr;

  static DFBasisView view(const BasisHost& b, std::vector<DevBuf<int>>& ib, std::vector<DevBuf<double>>& db,
                          int lmax, const char* what) {
    std::vector<int> l(b.nsh), po(b.nsh + 1, 0), bf0(b.nsh);
    std::vector<double> cen(3 * (size_t)b.nsh), ex, co;
    for (int s = 0; s < b.nsh; ++s) {
      const ShellHost& sh = b.shells[s];
      if (sh.l > lmax) throw Unsupported(std::string(what) + " shells above l = " + std::to_string(lmax));
      l[s] = sh.l;
      bf0[s] = b.bf0[s];
      for (int k = 0; k < 3; ++k) cen[3 * s + k] = sh.O[k];
      ex.insert(ex.end(), sh.a.begin(), sh.a.end());
      co.insert(co.end(), sh.c.begin(), sh.c.end());
      po[s + 1] = (int)ex.size();
    }
    ib.resize(3);
    db.resize(3);
    ib[0].upload(l, nullptr);
    ib[1].upload(po, nullptr);
    ib[2].upload(bf0, nullptr);
    db[0].upload(cen, nullptr);
    db[1].upload(ex, nullptr);
    db[2].upload(co, nullptr);
    return DFBasisView{ib[0].p, ib[1].p, ib[2].p, db[0].p, db[1].p, db[2].p, b.nsh, b.nbf, b.pure ? 1 : 0};
  }

  DFJ(const BasisHost& orb, const BasisHost& aux, int dev);

  void build_j(const double* P, double* J) {
    HIP_CHECK(hipSetDevice(device));
    const size_t n = (size_t)nbf, np = n * (n + 1) / 2;
    // D packed like B on the device: D_mu,nu + D_nu,mu below the diagonal,
    // D_mu,mu on it (full is the n x n staging square, for D and then J)
    HIP_CHECK(hipMemcpy(full.p, P, sizeof(double) * n * n, hipMemcpyHostToDevice));
    launch_df_pack(full.p, d.p, nbf, nullptr);
    HIP_CHECK(hipGetLastError());
    const double one = 1.0, zero = 0.0;
    // B is row-major [naux][np], i.e. column-major np x naux
    if (rocblas_dgemv(blas, rocblas_operation_transpose, (rocblas_int)np, naux, &one, B.p, (rocblas_int)np, d.p, 1,
                      &zero, x.p, 1) != rocblas_status_success)
      throw DeviceError("DF-J: rocblas_dgemv failed");
    if (rocsolver_dpotrs(blas, rocblas_fill_lower, naux, 1, V.p, naux, x.p, naux) != rocblas_status_success)
      throw DeviceError("DF-J: rocsolver_dpotrs failed");
    if (rocblas_dgemv(blas, rocblas_operation_none, (rocblas_int)np, naux, &one, B.p, (rocblas_int)np, x.p, 1, &zero,
                      j.p, 1) != rocblas_status_success)
      throw DeviceError("DF-J: rocblas_dgemv failed");
    launch_df_unpack(j.p, full.p, nbf, nullptr);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpy(J, full.p, sizeof(double) * n * n, hipMemcpyDeviceToHost));
  }
};

static void one_body_device(int device_id) {
  int ndev = 0;
  HIP_CHECK(hipGetDeviceCount(&ndev));
  if (device_id < 0 || device_id >= ndev) throw std::invalid_argument("device_id out of range");
  HIP_CHECK(hipSetDevice(device_id));
}

}  // namespace qfe

// ------------------------------------------------------------ C ABI
struct qfe_handle {
  qfe::Engine e;
};

using qfe::g_last_error;

template <typename F>
static int guarded(F&& f) {
  try {
    f();
    return QFE_OK;
  } catch (const std::invalid_argument& ex) {
    g_last_error = ex.what();
    return QFE_ERR_INVALID;
  } catch (const qfe::Unsupported& ex) {
    g_last_error = ex.what();
    return QFE_ERR_UNSUPPORTED;
  } catch (const qfe::DeviceError& ex) {
    g_last_error = ex.what();
    return QFE_ERR_DEVICE;
  } catch (const std::exception& ex) {
    g_last_error = ex.what();
    return QFE_ERR_RUNTIME;
  } catch (...) {
    g_last_error = "unknown error";
    return QFE_ERR_RUNTIME;
  }
}

extern "C" {

int qfe_create(const qfe_basis* basis, const qfe_config* cfg, qfe_handle** out) {
  return guarded([&] {
    if (!basis || !out) throw std::invalid_argument("null argument");
    *out = nullptr;
    auto h = std::make_unique<qfe_handle>();
    h->e.create(basis, cfg);
    *out = h.release();
  });
}

int qfe_build_jk(qfe_handle* h, const double* P, double* J, double* K, double alpha, double beta,
                 double omega, int ndm) {
  return guarded([&] {
    if (!h) throw std::runtime_error("ERI engine is not initialized");
    if (!P) throw std::invalid_argument("null density");
    h->e.build_host(P, J, K, alpha, beta, omega, ndm);
  });
}

int qfe_build_jk_device(qfe_handle* h, const double* dP, double* dJ, double* dK, double alpha, double beta,
                        double omega, void* stream) {
  return guarded([&] {
    if (!h) throw std::runtime_error("ERI engine is not initialized");
    if (!dP) throw std::invalid_argument("null density");
    if (std::fabs(omega) > 1e-12)
      throw qfe::Unsupported("range-separated exchange (omega != 0) is not supported by this engine");
    h->e.build_device(dP, dJ, dK, alpha + beta, stream ? (hipStream_t)stream : h->e.stream);
  });
}

int qfe_build_jk_device_noalloc(qfe_handle* h, const double* dP, double* dJ, double* dK, double alpha, double beta,
                                double omega, void* stream) {
  return guarded([&] {
    if (!h) throw std::runtime_error("ERI engine is not initialized");
    if (!dP) throw std::invalid_argument("null density");
    if (std::fabs(omega) > 1e-12)
      throw qfe::Unsupported("range-separated exchange (omega != 0) is not supported by this engine");
    auto& e = h->e;
    HIP_CHECK(hipSetDevice(e.device));
    if (!e.scratch_ptr) throw std::runtime_error("no-alloc scratch not set up on this device");
    const qfe::Engine::Scratch ext{e.scratch_ptr, qfe::kNoAllocScratchBytes};
    struct Guard {
      Guard() { qfe::g_no_alloc = true; }
      ~Guard() { qfe::g_no_alloc = false; }
    } guard;
    e.build_device(dP, dJ, dK, alpha + beta, stream ? (hipStream_t)stream : e.stream, &ext);
  });
}

unsigned long long qfe_noalloc_scratch_bytes(void) { return (unsigned long long)qfe::kNoAllocScratchBytes; }

int qfe_device_buffers(qfe_handle* h, double** dP, double** dJ, double** dK) {
  return guarded([&] {
    if (!h) throw std::runtime_error("ERI engine is not initialized");
    auto& e = h->e;
    HIP_CHECK(hipSetDevice(e.device));
    const size_t nn = (size_t)e.nbf * e.nbf;
    e.P_d.alloc(nn);
    e.J_d.alloc(nn);
    e.K_d.alloc(nn);
    if (dP) *dP = e.P_d.p;
    if (dJ) *dJ = e.J_d.p;
    if (dK) *dK = e.K_d.p;
  });
}

int qfe_copy(qfe_handle* h, void* dst, const void* src, unsigned long long bytes, int kind) {
  return guarded([&] {
    if (!h) throw std::runtime_error("ERI engine is not initialized");
    if (kind < 0 || kind > 2) throw std::invalid_argument("kind must be 0 (H2D), 1 (D2H) or 2 (D2D)");
    HIP_CHECK(hipSetDevice(h->e.device));
    if (kind == 2) {
      if (bytes % 8 != 0 || (reinterpret_cast<uintptr_t>(dst) | reinterpret_cast<uintptr_t>(src)) % 8 != 0)
        throw std::invalid_argument("device copies move whole, 8-byte aligned words");
      qfe::launch_copy_words(static_cast<const unsigned long long*>(src), static_cast<unsigned long long*>(dst),
                             bytes / 8, h->e.stream);
      HIP_CHECK(hipGetLastError());
    } else {
      HIP_CHECK(hipMemcpyAsync(dst, src, bytes, kind == 0 ? hipMemcpyHostToDevice : hipMemcpyDeviceToHost,
                               h->e.stream));
    }
    HIP_CHECK(hipStreamSynchronize(h->e.stream));
  });
}

int qfe_gradients(qfe_handle* h, const double* P, double* dJ, double* dK, double alpha, double beta,
                  double omega, int ndm) {
  return guarded([&] {
    if (!h) throw std::runtime_error("ERI engine is not initialized");
    if (!P) throw std::invalid_argument("null density");
    if (std::fabs(omega) > 1e-12)
      throw qfe::Unsupported("range-separated exchange (omega != 0) is not supported by this engine");
    if (dJ) std::fill_n(dJ, 3 * (size_t)std::max(h->e.natoms, 0), 0.0);
    if (dK) std::fill_n(dK, 3 * (size_t)std::max(h->e.natoms, 0), 0.0);
    // nothing requested, or an exchange term of zero weight alone
    // (libint2_direct.cpp:657-680)
    const bool need_k = dK != nullptr && std::fabs(alpha + beta) > 1e-12;
    if (!dJ && !need_k) return;
    h->e.gradients(P, ndm, dJ, need_k ? dK : nullptr, alpha + beta);
  });
}

int qfe_set_precision(qfe_handle* h, int mix, double threshold) {
  return guarded([&] {
    if (!h) throw std::runtime_error("ERI engine is not initialized");
    if (mix && !(threshold > 0.0)) throw std::invalid_argument("mixed-precision threshold must be positive");
    h->e.mix = mix != 0;
    if (mix) h->e.mix_thr = threshold;
  });
}

int qfe_get_stats(const qfe_handle* h, qfe_stats* out) {
  return guarded([&] {
    if (!h || !out) throw std::invalid_argument("null argument");
    *out = h->e.stats;
  });
}

int qfe_class_stats(const qfe_handle* h, int precision, int cap, int* cls, double* quartets,
                    double* prim_quartets, double* exec_prim_quartets, double* ms) {
  if (!h) return -1;
  const auto& e = h->e;
  if (exec_prim_quartets) {
    try {
      e.count_executed();
    } catch (const std::exception& ex) {
      qfe::g_last_error = ex.what();
      return -1;
    }
  }
  const int ncp = (int)e.cp_bra.size();
  for (int cp = 0; cp < ncp && cp < cap; ++cp) {
    const int b = e.cp_bra[cp], k = e.cp_ket[cp];
    auto la = [](int c) {
      int l = 0;
      while ((l + 1) * (l + 2) / 2 <= c) ++l;
      return l;
    };
    cls[4 * cp + 0] = la(b);
    cls[4 * cp + 1] = b - la(b) * (la(b) + 1) / 2;
    cls[4 * cp + 2] = la(k);
    cls[4 * cp + 3] = k - la(k) * (la(k) + 1) / 2;
    const size_t o = precision ? 2 * (size_t)ncp : 0;
    quartets[cp] = e.cstat_h.empty() ? 0.0 : e.cstat_h[o + cp];
    prim_quartets[cp] = e.cstat_h.empty() ? 0.0 : e.cstat_h[o + ncp + cp];
    if (ms) ms[cp] = e.cls_ms.empty() ? 0.0 : e.cls_ms[(precision ? ncp : 0) + cp];
    if (exec_prim_quartets)
      exec_prim_quartets[cp] = e.pqc_h.empty() ? 0.0 : (double)e.pqc_h[(precision ? ncp : 0) + cp];
  }
  return ncp;
}

int qfe_one_body_gpu(const qfe_basis* basis, int op, double* out, int device_id) {
  return guarded([&] {
    if (op < 0 || op > 2) throw std::invalid_argument("op must be 0 (S), 1 (T) or 2 (V)");
    if (!out) throw std::invalid_argument("null output");
    const qfe::BasisHost b = qfe::BasisHost::from_c(basis);
    qfe::one_body_device(device_id);
    const qfe::OneBodyTables& tb = qfe::one_body_tables(device_id);
    qfe::OneBodyDevice d(b, tb.sph);
    qfe::DevBuf<double> o;
    o.alloc((size_t)b.nbf * b.nbf);
    qfe::launch_one_body(d.view, tb.ra, op, o.p, nullptr);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpy(out, o.p, sizeof(double) * (size_t)b.nbf * b.nbf, hipMemcpyDeviceToHost));
  });
}

int qfe_one_body_grad_gpu(const qfe_basis* basis, int op, const double* W, double* grad, int device_id) {
  return guarded([&] {
    if (op < 0 || op > 2) throw std::invalid_argument("op must be 0 (S), 1 (T) or 2 (V)");
    if (!W || !grad) throw std::invalid_argument("null input or output");
    const qfe::BasisHost b = qfe::BasisHost::from_c(basis);
    qfe::one_body_device(device_id);
    const qfe::OneBodyTables& tb = qfe::one_body_tables(device_id);
    qfe::OneBodyDevice d(b, tb.sph);
    const size_t nn = (size_t)b.nbf * b.nbf;
    qfe::DevBuf<double> w, g;
    w.alloc(nn);
    g.alloc(3 * (size_t)b.natoms);
    HIP_CHECK(hipMemcpy(w.p, W, sizeof(double) * nn, hipMemcpyHostToDevice));
    HIP_CHECK(hipMemset(g.p, 0, sizeof(double) * 3 * (size_t)b.natoms));
    qfe::launch_one_body_grad(d.view, tb.ra, op, w.p, g.p, nullptr);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpy(grad, g.p, sizeof(double) * 3 * (size_t)b.natoms, hipMemcpyDeviceToHost));
  });
}

int qfe_eigh_gpu(int n, const double* A, double* w, double* V, int device_id) {
  return guarded([&] {
    if (n <= 0 || !A || !w || !V) throw std::invalid_argument("eigh: need n > 0 and non-null arrays");
    qfe::one_body_device(device_id);
    rocblas_handle h = qfe::solver_handle(device_id);
    const size_t nn = (size_t)n * n;
    qfe::DevBuf<double> a, d, e;
    qfe::DevBuf<rocblas_int> info;
    a.alloc(nn);
    d.alloc(n);
    e.alloc(n);
    info.alloc(1);
    // A is symmetric: its row-major and column-major layouts are the same
    HIP_CHECK(hipMemcpy(a.p, A, sizeof(double) * nn, hipMemcpyHostToDevice));
    if (rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_upper, n, a.p, n, d.p, e.p, info.p) !=
        rocblas_status_success)
      throw qfe::DeviceError("rocsolver_dsyevd failed");
    rocblas_int hinfo = 0;
    HIP_CHECK(hipMemcpy(&hinfo, info.p, sizeof(hinfo), hipMemcpyDeviceToHost));
    if (hinfo != 0) throw std::runtime_error("eigh: the divide-and-conquer solve did not converge");
    HIP_CHECK(hipMemcpy(w, d.p, sizeof(double) * n, hipMemcpyDeviceToHost));
    // eigenvector j is column j of the column-major result: transpose into
    // row-major V with the eigenvectors as columns
    std::vector<double> cm(nn);
    HIP_CHECK(hipMemcpy(cm.data(), a.p, sizeof(double) * nn, hipMemcpyDeviceToHost));
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < n; ++j) V[(size_t)i * n + j] = cm[(size_t)j * n + i];
  });
}

}  // extern "C"

namespace qfe {
DFJ::DFJ(const BasisHost& orb, const BasisHost& aux, int dev) : device(dev), nbf(orb.nbf), naux(aux.nbf) {
  one_body_device(dev);
  if (nbf <= 0 || naux <= 0) throw std::invalid_argument("DF-J: empty basis");
  const OneBodyTables& tb = one_body_tables(dev);
  std::vector<DevBuf<int>> oi, xi;
  std::vector<DevBuf<double>> od, xd;
  const DFBasisView ov = view(orb, oi, od, 2, "DF-J: orbital");
  const DFBasisView xv = view(aux, xi, xd, 4, "DF-J: auxiliary");
  const size_t np = (size_t)nbf * (nbf + 1) / 2;
  if (np > (size_t)std::numeric_limits<rocblas_int>::max())
    throw Unsupported("DF-J: nbf (nbf + 1) / 2 exceeds the rocBLAS integer range");
  B.alloc(np * naux);
  V.alloc((size_t)naux * naux);
  d.alloc(np);
  j.alloc(np);
  full.alloc((size_t)nbf * nbf);
  x.alloc(naux);
  launch_df_three_centre(ov, xv, tb.ra, tb.sph, B.p, nullptr);
  HIP_CHECK(hipGetLastError());
  launch_df_metric(xv, tb.ra, tb.sph, V.p, nullptr);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipDeviceSynchronize());
  blas = solver_handle(dev);
  DevBuf<rocblas_int> info;
  info.alloc(1);
  // V is symmetric: the column-major lower factor is the row-major upper one
  if (rocsolver_dpotrf(blas, rocblas_fill_lower, naux, V.p, naux, info.p) != rocblas_status_success)
    throw DeviceError("DF-J: rocsolver_dpotrf failed");
  rocblas_int hinfo = 0;
  HIP_CHECK(hipMemcpy(&hinfo, info.p, sizeof(hinfo), hipMemcpyDeviceToHost));
  if (hinfo != 0) throw std::runtime_error("DF-J: the auxiliary metric is not positive definite");
}

// ------------------------------------------------------------ MO transforms
// The dense AO tensor T[a][b][c][d] = (ab|cd) in HBM (nbf^4 doubles, shells
// up to d), for the quarter transform (ERI::quarter_trans, eri.h:116) and
// MOERI::compute (moeri.cpp:28-153).  Small systems only: nbf = 200 is
// 12.8 GB.
static void dense_tensor(const BasisHost& b, int dev, DevBuf<double>& T, size_t extra_doubles) {
  one_body_device(dev);
  const size_t n = (size_t)b.nbf, n4 = n * n * n * n;
  size_t free_b = 0, total_b = 0;
  HIP_CHECK(hipMemGetInfo(&free_b, &total_b));
  if ((double)(n4 + extra_doubles) * sizeof(double) > 0.9 * (double)free_b)
    throw Unsupported("MO transform: the dense AO tensor (" + std::to_string(n4 * sizeof(double) >> 20) +
                      " MiB) and its work buffer do not fit in free device memory");
  const OneBodyTables& tb = one_body_tables(dev);
  std::vector<DevBuf<int>> ib;
  std::vector<DevBuf<double>> db;
  const DFBasisView v = DFJ::view(b, ib, db, 2, "MO transform: orbital");
  // every unordered shell pair once, the higher l first, grouped by class
  // and contraction length so that a wave's quartets mostly share them
  std::vector<int2> pr;
  pr.reserve((size_t)b.nsh * (b.nsh + 1) / 2);
  for (int s1 = 0; s1 < b.nsh; ++s1)
    for (int s2 = 0; s2 <= s1; ++s2) {
      const bool swap = b.shells[s2].l > b.shells[s1].l;
      pr.push_back(int2{swap ? s2 : s1, swap ? s1 : s2});
    }
  auto key = [&](const int2& p) {
    const ShellHost &x = b.shells[p.x], &y = b.shells[p.y];
    return std::make_tuple(x.l, y.l, x.a.size() * y.a.size());
  };
  std::stable_sort(pr.begin(), pr.end(), [&](const int2& p, const int2& q) { return key(p) < key(q); });
  DevBuf<int2> pairs;
  pairs.upload(pr, nullptr);
  T.alloc(n4);
  launch_eri_dense(v, tb.ra, tb.sph, pairs.p, T.p, nullptr);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipDeviceSynchronize());
}

static rocblas_int blas_dim(size_t v) {
  if (v > (size_t)std::numeric_limits<rocblas_int>::max())
    throw Unsupported("MO transform: a GEMM dimension exceeds the rocBLAS integer range");
  return (rocblas_int)v;
}

// column-major C = op(A) op(B), alpha 1, beta 0
static void dgemm(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, size_t m, size_t n, size_t k,
                  const double* A, size_t lda, const double* B, size_t ldb, double* C, size_t ldc) {
  const double one = 1.0, zero = 0.0;
  if (rocblas_dgemm(h, ta, tb, blas_dim(m), blas_dim(n), blas_dim(k), &one, A, blas_dim(lda), B, blas_dim(ldb),
                    &zero, C, blas_dim(ldc)) != rocblas_status_success)
    throw DeviceError("MO transform: rocblas_dgemm failed");
}

// out[a][b][c][p] = sum_d T[a][b][c][d] C[d][p]  (C row-major nbf x nt):
// column-major out (nt x nbf^3) = C (nt x nbf) . T (nbf x nbf^3)
static void quarter_trans(const BasisHost& b, int dev, size_t nt, const double* C, double* out) {
  const size_t n = (size_t)b.nbf, n3 = n * n * n;
  if (nt == 0 || nt > n) throw std::invalid_argument("quarter_trans: need 0 < nt <= nbf");
  DevBuf<double> T, c, o;
  dense_tensor(b, dev, T, nt * n3 + nt * n);
  c.alloc(n * nt);
  o.alloc(nt * n3);
  HIP_CHECK(hipMemcpy(c.p, C, sizeof(double) * n * nt, hipMemcpyHostToDevice));
  rocblas_handle h = solver_handle(dev);
  dgemm(h, rocblas_operation_none, rocblas_operation_none, nt, n3, n, c.p, nt, T.p, n, o.p, nt);
  HIP_CHECK(hipMemcpy(out, o.p, sizeof(double) * nt * n3, hipMemcpyDeviceToHost));
}

// out[w][x][y][z] = sum (ab|cd) Cl[a][w] Ck[b][x] Cj[c][y] Ci[d][z], the
// result of MOERI::compute.  Each quarter contracts the last index of a
// row-major X[m][d] and moves the new index to the front,
// Y[t][m] = sum_d X[m][d] C[d][t], one GEMM: column-major Y (M x nt) =
// X^T (M x nbf) . C^T (nbf x nt).  The four steps take T[a][b][c][d] to
// [z][a][b][c], [y][z][a][b], [x][y][z][a] and [w][x][y][z].  Two HBM
// buffers: T (then the second and fourth results) and one of nt nbf^3.
static void mo_transform(const BasisHost& b, int dev, size_t nt, const double* const Cs[4], double* out) {
  const size_t n = (size_t)b.nbf;
  if (nt == 0 || nt > n) throw std::invalid_argument("MO transform: need 0 < nt <= nbf");
  const size_t n3 = n * n * n;
  DevBuf<double> T, W, c[4];
  dense_tensor(b, dev, T, nt * n3 + 4 * nt * n);
  W.alloc(nt * n3);
  for (int q = 0; q < 4; ++q) {
    c[q].alloc(n * nt);
    HIP_CHECK(hipMemcpy(c[q].p, Cs[q], sizeof(double) * n * nt, hipMemcpyHostToDevice));
  }
  rocblas_handle h = solver_handle(dev);
  const rocblas_operation t = rocblas_operation_transpose;
  size_t M = n3;  // rows of X: the product of the three leading extents
  double* src = T.p;
  double* dst = W.p;
  for (int q = 0; q < 4; ++q) {  // Ci, Cj, Ck, Cl
    dgemm(h, t, t, M, nt, n, src, n, c[q].p, nt, dst, M);
    M = M / n * nt;
    std::swap(src, dst);
  }
  HIP_CHECK(hipMemcpy(out, src, sizeof(double) * nt * nt * nt * nt, hipMemcpyDeviceToHost));
}

// ------------------------------------------------------------ device-resident RHF iterations
// The dense linear algebra of one RHF iteration kept in HBM (SCFImpl::
// iterate_, scf_impl.cpp:555-621, with the DIIS of scf_algorithm.cpp:
// 401-435): F = F + J - 1/2 K from a (difference) density, the energy
// E_nuc + 1/2 Tr(P H) + 1/2 Tr(P F), the orbital gradient FPS - SPF, the DIIS
// extrapolation, X^T F X, the eigensolve and P = 2 C_occ C_occ^T, all on the
// engine's stream with one rocBLAS handle; only scalars and the small DIIS
// system (at most diis_size + 1 square) cross to the host.  Matrices are
// held column-major; S, H, F and P are symmetric, X (nbf x norb) is
// transposed on upload.  The error matrix is stored as SPF - FPS, the
// transpose of the row-major FPS - SPF: DIIS uses only its inner products
// and its largest magnitude, which do not see the sign.
struct SCFDevice {
  Engine& e;
  int n = 0, m = 0, nocc = 0, diis_size = 10;
  double enuc = 0.0;
  rocblas_handle blas = nullptr;
  DevBuf<double> H, S, X, F, P, Plast, Pd, Pprev, T1, T2, Fo, C, w, ework, err;
  DevBuf<rocblas_int> info;
  std::vector<DevBuf<double>> Fh, Eh;  // DIIS ring of Fock and error matrices
  std::vector<int> ring;               // ring slots, oldest first
  std::vector<double> Bm;              // inner products of the stored errors [slot][slot]

  SCFDevice(Engine& eng, const double* Hh, const double* Sh, const double* Xh, int norb, int nocc_, double enuc_,
            int diis)
      : e(eng), n(eng.nbf), m(norb), nocc(nocc_), diis_size(diis), enuc(enuc_) {
    if (m <= 0 || m > n) throw std::invalid_argument("SCF: need 0 < norb <= nbf");
    if (nocc <= 0 || nocc > m) throw std::invalid_argument("SCF: need 0 < nocc <= norb");
    if (diis_size < 1) throw std::invalid_argument("SCF: diis_size must be positive");
    HIP_CHECK(hipSetDevice(e.device));
    if (rocblas_create_handle(&blas) != rocblas_status_success) throw DeviceError("rocblas_create_handle failed");
    rocblas_set_stream(blas, e.stream);
    const size_t nn = (size_t)n * n;
    for (DevBuf<double>* b : {&H, &S, &F, &P, &Plast, &Pd, &Pprev, &T2, &err}) b->alloc(nn);
    X.alloc((size_t)n * m);
    T1.alloc((size_t)n * m);
    C.alloc((size_t)n * m);
    Fo.alloc((size_t)m * m);
    w.alloc(m);
    ework.alloc(m);
    info.alloc(1);
    e.P_d.alloc(nn);
    e.J_d.alloc(nn);
    e.K_d.alloc(nn);
    std::vector<double> xc((size_t)n * m);
    for (int i = 0; i < n; ++i)
      for (int j = 0; j < m; ++j) xc[(size_t)j * n + i] = Xh[(size_t)i * m + j];
    HIP_CHECK(hipMemcpyAsync(H.p, Hh, nn * sizeof(double), hipMemcpyHostToDevice, e.stream));
    HIP_CHECK(hipMemcpyAsync(S.p, Sh, nn * sizeof(double), hipMemcpyHostToDevice, e.stream));
    HIP_CHECK(hipMemcpyAsync(X.p, xc.data(), xc.size() * sizeof(double), hipMemcpyHostToDevice, e.stream));
    HIP_CHECK(hipMemsetAsync(Pprev.p, 0, nn * sizeof(double), e.stream));
    HIP_CHECK(hipMemsetAsync(Plast.p, 0, nn * sizeof(double), e.stream));
    HIP_CHECK(hipStreamSynchronize(e.stream));
    Fh.resize(diis_size);
    Eh.resize(diis_size);
    Bm.assign((size_t)diis_size * diis_size, 0.0);
  }
  ~SCFDevice() {
    if (blas) rocblas_destroy_handle(blas);
  }

  static void ok(rocblas_status st, const char* what) {
    if (st != rocblas_status_success) throw DeviceError(std::string("SCF: ") + what + " failed");
  }
  // C = a A + b op(B), n x n
  void geam(double a, const double* A, double b, const double* B, bool tb, double* Cm) {
    ok(rocblas_dgeam(blas, rocblas_operation_none, tb ? rocblas_operation_transpose : rocblas_operation_none, n, n,
                     &a, A, n, &b, B, n, Cm, n),
       "dgeam");
  }
  double dot(const double* a, const double* b) {
    double r = 0.0;
    ok(rocblas_ddot(blas, n * n, a, 1, b, 1, &r), "ddot");
    return r;
  }
  void copy(const double* src, double* dst, size_t count) {
    HIP_CHECK(hipMemcpyAsync(dst, src, count * sizeof(double), hipMemcpyDeviceToDevice, e.stream));
  }

  void set_density(const double* Ph) {
    HIP_CHECK(hipMemcpyAsync(P.p, Ph, (size_t)n * n * sizeof(double), hipMemcpyHostToDevice, e.stream));
    HIP_CHECK(hipStreamSynchronize(e.stream));
  }

  // J/K of P (full) or of P - P_last (incremental), F, E, orbital gradient
  // and density change; out = {E, og, drms}
  void step(bool full, double out[3]) {
    const size_t nn = (size_t)n * n;
    const double one = 1.0, zero = 0.0;
    if (full) {
      copy(P.p, e.P_d.p, nn);
      copy(H.p, F.p, nn);
    } else {
      geam(1.0, P.p, -1.0, Plast.p, false, e.P_d.p);
    }
    copy(P.p, Plast.p, nn);
    e.build_device(e.P_d.p, e.J_d.p, e.K_d.p, 1.0, e.stream);
    geam(1.0, F.p, 1.0, e.J_d.p, false, T2.p);
    geam(1.0, T2.p, -0.5, e.K_d.p, false, F.p);
    const double E = enuc + 0.5 * dot(P.p, H.p) + 0.5 * dot(P.p, F.p);
    // err = (F P S) - (F P S)^T, column-major
    ok(rocblas_dgemm(blas, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, F.p, n, P.p, n, &zero,
                     T2.p, n),
       "dgemm");
    ok(rocblas_dgemm(blas, rocblas_operation_none, rocblas_operation_none, n, n, n, &one, T2.p, n, S.p, n, &zero,
                     Pd.p, n),
       "dgemm");
    geam(1.0, Pd.p, -1.0, Pd.p, true, err.p);
    rocblas_int imax = 0;
    ok(rocblas_idamax(blas, n * n, err.p, 1, &imax), "idamax");
    double emax = 0.0;
    HIP_CHECK(hipMemcpyAsync(&emax, err.p + (imax - 1), sizeof(double), hipMemcpyDeviceToHost, e.stream));
    geam(1.0, Pprev.p, -1.0, P.p, false, T2.p);
    double dn = 0.0;
    ok(rocblas_dnrm2(blas, n * n, T2.p, 1, &dn), "dnrm2");
    copy(P.p, Pprev.p, nn);
    HIP_CHECK(hipStreamSynchronize(e.stream));
    out[0] = E;
    out[1] = std::fabs(emax) / n;
    out[2] = dn / n;
  }

  // diagonalise X^T A X (A symmetric, n x n): C = X Cp, w = eigenvalues
  void diag(const double* A) {
    const double one = 1.0, zero = 0.0;
    ok(rocblas_dgemm(blas, rocblas_operation_none, rocblas_operation_none, n, m, n, &one, A, n, X.p, n, &zero, T1.p,
                     n),
       "dgemm");
    ok(rocblas_dgemm(blas, rocblas_operation_transpose, rocblas_operation_none, m, m, n, &one, X.p, n, T1.p, n, &zero,
                     T2.p, m),
       "dgemm");
    const double half = 0.5;
    ok(rocblas_dgeam(blas, rocblas_operation_none, rocblas_operation_transpose, m, m, &half, T2.p, m, &half, T2.p, m,
                     Fo.p, m),
       "dgeam");
    ok(rocsolver_dsyevd(blas, rocblas_evect_original, rocblas_fill_upper, m, Fo.p, m, w.p, ework.p, info.p),
       "dsyevd");
    rocblas_int hinfo = 0;
    HIP_CHECK(hipMemcpyAsync(&hinfo, info.p, sizeof(hinfo), hipMemcpyDeviceToHost, e.stream));
    HIP_CHECK(hipStreamSynchronize(e.stream));
    if (hinfo != 0) throw std::runtime_error("SCF: the divide-and-conquer eigensolve did not converge");
    ok(rocblas_dgemm(blas, rocblas_operation_none, rocblas_operation_none, n, m, m, &one, X.p, n, Fo.p, m, &zero,
                     C.p, n),
       "dgemm");
  }

  void density() {
    const double two = 2.0, zero = 0.0;
    ok(rocblas_dgemm(blas, rocblas_operation_none, rocblas_operation_transpose, n, n, nocc, &two, C.p, n, C.p, n,
                     &zero, P.p, n),
       "dgemm");
    HIP_CHECK(hipStreamSynchronize(e.stream));
  }

  // DIIS: store (F, err), extrapolate, diagonalise, new density
  void update() {
    const size_t nn = (size_t)n * n;
    int slot;
    if ((int)ring.size() < diis_size) {
      slot = (int)ring.size();
      ring.push_back(slot);
    } else {
      slot = ring.front();
      ring.erase(ring.begin());
      ring.push_back(slot);
    }
    Fh[slot].alloc(nn);
    Eh[slot].alloc(nn);
    copy(F.p, Fh[slot].p, nn);
    copy(err.p, Eh[slot].p, nn);
    for (int o : ring) {
      const double v = dot(Eh[slot].p, Eh[o].p);
      Bm[(size_t)slot * diis_size + o] = Bm[(size_t)o * diis_size + slot] = v;
    }
    const int k = (int)ring.size();
    // [B -1; -1 0] c = [0; -1], solved by Gaussian elimination with partial
    // pivoting (LU as in dgesv); an exactly singular system falls back to the
    // newest Fock matrix
    std::vector<double> A((size_t)(k + 1) * (k + 1), -1.0), rhs(k + 1, 0.0);
    for (int i = 0; i < k; ++i)
      for (int j = 0; j < k; ++j) A[(size_t)i * (k + 1) + j] = Bm[(size_t)ring[i] * diis_size + ring[j]];
    A[(size_t)k * (k + 1) + k] = 0.0;
    rhs[k] = -1.0;
    std::vector<double> c(k, 0.0);
    if (!solve_small(A, rhs, k + 1)) {
      c[k - 1] = 1.0;
    } else {
      for (int i = 0; i < k; ++i) c[i] = rhs[i];
    }
    HIP_CHECK(hipMemsetAsync(T2.p, 0, nn * sizeof(double), e.stream));
    for (int i = 0; i < k; ++i) ok(rocblas_daxpy(blas, n * n, &c[i], Fh[ring[i]].p, 1, T2.p, 1), "daxpy");
    diag(T2.p);
    density();
  }

  static bool solve_small(std::vector<double>& A, std::vector<double>& b, int N) {
    for (int col = 0; col < N; ++col) {
      int piv = col;
      for (int r = col + 1; r < N; ++r)
        if (std::fabs(A[(size_t)r * N + col]) > std::fabs(A[(size_t)piv * N + col])) piv = r;
      if (A[(size_t)piv * N + col] == 0.0) return false;  // exactly singular, as LAPACK dgesv reports
      if (piv != col) {
        for (int j = 0; j < N; ++j) std::swap(A[(size_t)piv * N + j], A[(size_t)col * N + j]);
        std::swap(b[piv], b[col]);
      }
      for (int r = col + 1; r < N; ++r) {
        const double f = A[(size_t)r * N + col] / A[(size_t)col * N + col];
        for (int j = col; j < N; ++j) A[(size_t)r * N + j] -= f * A[(size_t)col * N + j];
        b[r] -= f * b[col];
      }
    }
    for (int r = N - 1; r >= 0; --r) {
      double v = b[r];
      for (int j = r + 1; j < N; ++j) v -= A[(size_t)r * N + j] * b[j];
      b[r] = v / A[(size_t)r * N + r];
    }
    return true;
  }

  // final full rebuild and plain diagonalisation (scf_impl.cpp:612-621)
  double finish() {
    double out[3];
    step(true, out);
    diag(F.p);
    return out[0];
  }

  // 0: P, 1: F (n x n); 2: C (n x norb, row-major); 3: eigenvalues (norb)
  void get(int what, double* dst) {
    const size_t nn = (size_t)n * n;
    if (what == 0 || what == 1) {
      HIP_CHECK(hipMemcpyAsync(dst, what ? F.p : P.p, nn * sizeof(double), hipMemcpyDeviceToHost, e.stream));
      HIP_CHECK(hipStreamSynchronize(e.stream));
    } else if (what == 2) {
      std::vector<double> cm((size_t)n * m);
      HIP_CHECK(hipMemcpyAsync(cm.data(), C.p, cm.size() * sizeof(double), hipMemcpyDeviceToHost, e.stream));
      HIP_CHECK(hipStreamSynchronize(e.stream));
      for (int i = 0; i < n; ++i)
        for (int j = 0; j < m; ++j) dst[(size_t)i * m + j] = cm[(size_t)j * n + i];
    } else if (what == 3) {
      HIP_CHECK(hipMemcpyAsync(dst, w.p, (size_t)m * sizeof(double), hipMemcpyDeviceToHost, e.stream));
      HIP_CHECK(hipStreamSynchronize(e.stream));
    } else {
      throw std::invalid_argument("SCF get: what must be 0 (P), 1 (F), 2 (C) or 3 (eigenvalues)");
    }
  }
};
}  // namespace qfe

struct qfe_dfj {
  std::unique_ptr<qfe::DFJ> e;
};

struct qfe_scf {
  std::unique_ptr<qfe::SCFDevice> s;
};

extern "C" {

int qfe_dfj_create(const qfe_basis* basis, const qfe_basis* aux, int device_id, qfe_dfj** out) {
  return guarded([&] {
    if (!out) throw std::invalid_argument("null handle pointer");
    *out = nullptr;
    const qfe::BasisHost o = qfe::BasisHost::from_c(basis);
    const qfe::BasisHost a = qfe::BasisHost::from_c(aux);
    auto h = std::make_unique<qfe_dfj>();
    h->e = std::make_unique<qfe::DFJ>(o, a, device_id);
    *out = h.release();
  });
}

int qfe_dfj_build_j(qfe_dfj* h, const double* P, double* J) {
  return guarded([&] {
    if (!h || !h->e || !P || !J) throw std::invalid_argument("null argument");
    h->e->build_j(P, J);
  });
}

int qfe_dfj_naux(const qfe_dfj* h) { return (h && h->e) ? h->e->naux : -1; }

int qfe_quarter_trans(qfe_handle* h, unsigned long long nt, const double* C, double* out) {
  return guarded([&] {
    if (!h || !C || !out) throw std::invalid_argument("quarter_trans: null argument");
    qfe::quarter_trans(h->e.basis, h->e.device, (size_t)nt, C, out);
  });
}

int qfe_mo_eri(qfe_handle* h, unsigned long long nb, unsigned long long nt, const double* Ci, const double* Cj,
               const double* Ck, const double* Cl, double* out) {
  return guarded([&] {
    if (!h || !Ci || !Cj || !Ck || !Cl || !out) throw std::invalid_argument("mo_eri: null argument");
    if (nb != (unsigned long long)h->e.nbf) throw std::invalid_argument("mo_eri: nb differs from the basis size");
    if (nt > nb) throw std::runtime_error("MOERI does not support num_molecular_orbitals > num_atomic_orbitals");
    const double* Cs[4] = {Ci, Cj, Ck, Cl};
    qfe::mo_transform(h->e.basis, h->e.device, (size_t)nt, Cs, out);
  });
}

int qfe_scf_create(qfe_handle* h, const double* H, const double* S, const double* X, int norb, int nocc,
                   double enuc, int diis_size, qfe_scf** out) {
  return guarded([&] {
    if (!h || !H || !S || !X || !out) throw std::invalid_argument("scf_create: null argument");
    *out = nullptr;
    auto s = std::make_unique<qfe_scf>();
    s->s = std::make_unique<qfe::SCFDevice>(h->e, H, S, X, norb, nocc, enuc, diis_size);
    *out = s.release();
  });
}

int qfe_scf_set_density(qfe_scf* s, const double* P) {
  return guarded([&] {
    if (!s || !P) throw std::invalid_argument("scf_set_density: null argument");
    s->s->set_density(P);
  });
}

int qfe_scf_step(qfe_scf* s, int full_build, double* energy, double* orbital_gradient, double* density_rms) {
  return guarded([&] {
    if (!s || !energy || !orbital_gradient || !density_rms) throw std::invalid_argument("scf_step: null argument");
    double o[3];
    s->s->step(full_build != 0, o);
    if (!std::isfinite(o[0])) throw std::runtime_error("NaN or INF Encountered in SCF Energy");
    *energy = o[0];
    *orbital_gradient = o[1];
    *density_rms = o[2];
  });
}

int qfe_scf_update(qfe_scf* s) {
  return guarded([&] {
    if (!s) throw std::invalid_argument("scf_update: null handle");
    s->s->update();
  });
}

int qfe_scf_finish(qfe_scf* s, double* energy) {
  return guarded([&] {
    if (!s || !energy) throw std::invalid_argument("scf_finish: null argument");
    *energy = s->s->finish();
  });
}

int qfe_scf_get(qfe_scf* s, int what, double* out) {
  return guarded([&] {
    if (!s || !out) throw std::invalid_argument("scf_get: null argument");
    s->s->get(what, out);
  });
}

void qfe_scf_destroy(qfe_scf* s) { delete s; }

void qfe_dfj_destroy(qfe_dfj* h) { delete h; }

int qfe_max_l(void) { return QFE_MAX_L; }

int qfe_nbf(const qfe_handle* h) { return h ? h->e.nbf : -1; }

void qfe_destroy(qfe_handle* h) { delete h; }

const char* qfe_last_error(void) { return g_last_error.c_str(); }

}  // extern "C"
