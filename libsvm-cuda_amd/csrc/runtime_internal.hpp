// Internals shared by the runtime files (smo_runtime.cpp, serve.cpp): the
// solver handle with its error / allocation helpers, the kernel launchers
// and the SELL-64 packing of svm_node rows.
#pragma once
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <climits>
#include <map>
#include <memory>
#include <string>
#include <thread>
#include <unordered_map>
#include <utility>
#include <vector>

#include "../../include/svm_hip.h"
#include "smo_device.hpp"

extern "C" {
hipError_t smo_launch_self_dot(const smo::View *V, double *xsq, hipStream_t st);
hipError_t smo_launch_qd(const smo::View *V, hipStream_t st);
hipError_t smo_launch_ginit_row(const smo::View *V, int v, size_t shmem, hipStream_t st);
hipError_t smo_launch_ginit_multi(const smo::View *V, const int *list, int count, size_t shmem, hipStream_t st);
hipError_t smo_launch_select_i(const smo::View *V, size_t shmem, hipStream_t st);
hipError_t smo_launch_select_j(const smo::View *V, size_t shmem, hipStream_t st);
hipError_t smo_launch_alpha(const smo::View *V, hipStream_t st);
hipError_t smo_launch_gupdate(const smo::View *V, size_t shmem, hipStream_t st);
hipError_t smo_launch_status(const smo::View *V, hipStream_t st);
hipError_t smo_launch_persistent(const smo::View *V, int iter_limit, size_t shmem, hipStream_t st);
hipError_t smo_launch_lru_private_init(const smo::View *V, hipStream_t st);
hipError_t smo_persistent_capacity(int kind, size_t shmem, int *blocks_per_cu);
hipError_t smo_launch_finalize(const smo::View *V, hipStream_t st);
hipError_t smo_launch_finalize_views(const smo::View *views, const int *idx, int n, hipStream_t st);
hipError_t smo_launch_kernel_row(const smo::View *V, int q, float *out, size_t shmem, hipStream_t st);
hipError_t smo_launch_batch(const smo::View *views, const int *items, int nitems, int grid, const smo::BatchPool *pool,
                            size_t shmem, hipStream_t st);
hipError_t smo_batch_capacity(size_t shmem, int *blocks_per_cu);
hipError_t smo_launch_predict(const smo::PredictArgs *A, int nrows, int cus, hipStream_t st);
int smo_predict_tb(int ld, size_t lds_bytes, size_t *shmem);
hipError_t smo_launch_split_f16(const smo::View *V, _Float16 *xs, int *bad, hipStream_t st);
hipError_t smo_launch_gram_f16(const smo::View *V, const _Float16 *xs, hipStream_t st);
size_t smo_split_f16_bytes(int l, int d);
hipError_t smo_launch_predict_split(const smo::PRows *R, _Float16 *out, double *sq, int *bad, int d, hipStream_t st);
hipError_t smo_launch_predict_gram(const smo::PredictArgs *A, const _Float16 *xt, const _Float16 *xs,
                                   const double *sqt, const double *sqs, double *partial, hipStream_t st);
long long smo_predict_split_halves(int d);
int smo_predict_gram_tiles(int ns);
hipError_t smo_launch_dots_f64(const smo::View *V, double *D, long long ld, hipStream_t st);
hipError_t smo_launch_pack_rowset(const smo::View *src, const int *rows, int n, const long long *dst_slice_ptr,
                                  float *dst_vals, int *dst_cols, int *dst_len, double *dst_xsq, hipStream_t st);
}

namespace svmhip_detail {

constexpr size_t kLdsQueryLimit = 64 * 1024;  // query row staged in LDS up to this size
constexpr size_t kFlagLdsLimit = 64 * 1024;  // persistent kernel: query row + row-flag bitset in LDS
// Batched solves with neither cache_bytes nor hbm_fraction given: all the
// kernel-row arenas together at most this.  A device allocation of most of
// the free memory right after freeing as much takes seconds (the freed
// memory is cleared again: 190 GB in 5.7 s, 100 GB up to 3 s, 38 GB at once
// on one MI355X), so a handle made per call (svm_train, svm_cross_validation)
// paid it on every call after the first; rows gathered from the shared dot
// products make a larger cache worth little.
constexpr size_t kBatchArenaDefaultBytes = 32ull << 30;

struct Failure {
  svmhip_status code;
};

}  // namespace svmhip_detail
using namespace svmhip_detail;

struct svmhip_solver {
  int dev = 0;
  hipStream_t st = nullptr;
  hipEvent_t ev[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  svmhip_kind kind = SVMHIP_KIND_C;
  svmhip_options opt{};
  int svm_type = C_SVC;
  double param_cache_mb = 100;
  double eps = 1e-3;
  // problem
  int l = 0, d = 0, nslices = 0, dense = 0, ktype = RBF, degree = 3;
  double gamma = 0, coef0 = 0;
  long long total = 0;
  std::vector<int> host_row_len;  // SELL row lengths (kept for packing row sets)
  // solver
  int nvar = 0;
  double Cp = 0, Cn = 0;
  int S = 0, full = 0;
  long long row_stride = 0;
  int nblocks = 1;
  int xrep = 0;                   // exchange-record copies of the persistent kernel (0 = by grid size)
  size_t shmem = 0;
  int use_lds = 1;
  bool rbf_ready = false, setup_done = false, have_j = false;
  smo::Ctl *host_ctl = nullptr;  // pinned mirror
  // device memory
  long long *d_slice_ptr = nullptr;
  float *d_vals = nullptr;
  int *d_cols = nullptr, *d_row_len = nullptr, *d_serial = nullptr;
  double *d_xsq = nullptr, *d_G = nullptr, *d_alpha = nullptr, *d_p = nullptr, *d_QD = nullptr;
  int8_t *d_y = nullptr, *d_status = nullptr, *d_computed = nullptr;
  float *d_rows = nullptr, *d_xs = nullptr;
  int *d_slot_of_row = nullptr, *d_row_of_slot = nullptr, *d_prev = nullptr, *d_next = nullptr;
  smo::Ctl *d_ctl = nullptr;
  smo::Partial *d_part = nullptr;
  smo::XWord *d_xw = nullptr;
  int8_t *d_pm_flags = nullptr;
  int *d_pm_lru = nullptr;
  long long pm_stride = 0;
  bool shared_meta_dirty = false;  // the persistent kernel overwrote cache slots
  std::vector<void *> allocs;
  long long n_allocs = 0;         // hipMalloc calls made for this handle (all time)
  // capacities (elements) of the per-solve buffers svmhip_setup keeps from
  // one solve to the next on the same handle
  size_t cap_nvar = 0, cap_rows = 0, cap_flags = 0, cap_pm_flags = 0, cap_lru_map = 0, cap_lru_slots = 0,
         cap_pm_lru = 0, cap_list = 0;
  int *d_list = nullptr;          // G initialisation: variables with alpha != 0
  // Gram-tile mode (opt.mfma_rows): the FP16 hi/lo copy of X (made once per
  // handle), whether X fits FP16's range, and whether this solve's cache
  // already holds the whole matrix
  _Float16 *d_xsplit = nullptr;
  int *d_flag = nullptr;
  int split_state = 0;            // 0 = not made, 1 = ready, -1 = X out of FP16 range (exact rows)
  bool gram_done = false;
  long long gram_rows = 0;        // kernel rows the Gram-tile kernel computed in this solve
  svmhip_stats stats{};
  std::string err;

  void fail(svmhip_status code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    err = buf;
    throw Failure{code};
  }
  void ck(hipError_t e, const char *what) {
    if (e != hipSuccess) {
      (void)hipGetLastError();
      fail(e == hipErrorOutOfMemory ? SVMHIP_ERR_NOMEM : SVMHIP_ERR_DEVICE, "%s: %s", what,
           hipGetErrorString(e));
    }
  }
  // Freed blocks kept for reuse, by size class, up to pool_cap bytes (the
  // server turns it on: its requests come and go between two stages, where
  // hipMalloc / hipFree would leave the device idle).  A block is reused
  // only by a later dalloc, whose uploads and kernels follow on the same
  // stream every use of the block's previous owner.
  size_t pool_cap = 0, pooled = 0;
  std::multimap<size_t, void *> pool;
  std::unordered_map<void *, size_t> block_class;
  static size_t size_class(size_t b) {
    if (b <= 65536) return (b + 255) / 256 * 256;
    size_t p2 = 65536;
    while (p2 < b) p2 <<= 1;
    const size_t step = p2 / 16;
    return (b + step - 1) / step * step;
  }
  void release_pool() {
    for (auto &kv : pool) {
      block_class.erase(kv.second);
      auto it = std::find(allocs.begin(), allocs.end(), kv.second);
      if (it != allocs.end()) allocs.erase(it);
      (void)hipFree(kv.second);
    }
    pool.clear();
    pooled = 0;
  }
  template <class T>
  T *dalloc(size_t n, const char *what) {
    void *p = nullptr;
    size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    if (pool_cap > 0) {
      bytes = size_class(bytes);
      auto it = pool.find(bytes);
      if (it != pool.end()) {
        p = it->second;
        pool.erase(it);
        pooled -= bytes;
        return (T *)p;
      }
    }
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess && pooled > 0) {  // hand the pooled blocks back and try again
      (void)hipGetLastError();
      release_pool();
      e = hipMalloc(&p, bytes);
    }
    if (e != hipSuccess) {
      (void)hipGetLastError();
      fail(SVMHIP_ERR_NOMEM, "device allocation of %zu bytes for %s failed: %s", bytes, what,
           hipGetErrorString(e));
    }
    allocs.push_back(p);
    ++n_allocs;
    if (pool_cap > 0) block_class[p] = bytes;
    return (T *)p;
  }
  // Buffer p of at least n elements: kept when its capacity (cap) suffices,
  // else freed and allocated anew.
  template <class T>
  T *ensure(T *&p, size_t &cap, size_t n, const char *what) {
    n = std::max<size_t>(n, 1);
    if (p && cap >= n) return p;
    dfree(p);
    p = nullptr;
    cap = 0;
    p = dalloc<T>(n, what);
    cap = n;
    return p;
  }
  void dfree(void *p) {
    if (!p) return;
    auto bc = block_class.find(p);
    if (bc != block_class.end()) {
      if (pooled + bc->second <= pool_cap) {
        pool.emplace(bc->second, p);
        pooled += bc->second;
        return;
      }
      block_class.erase(bc);
    }
    auto it = std::find(allocs.begin(), allocs.end(), p);
    if (it != allocs.end()) allocs.erase(it);
    (void)hipFree(p);
  }
  template <class T>
  void upload(T *dst, const T *src, size_t n, const char *what) {
    if (n) ck(hipMemcpyAsync(dst, src, n * sizeof(T), hipMemcpyHostToDevice, st), what);
  }

  smo::View view() const {
    smo::View V;
    memset(&V, 0, sizeof V);
    V.l = l; V.d = d; V.nslices = nslices; V.dense = dense;
    V.ktype = ktype; V.degree = degree; V.gamma = gamma; V.coef0 = coef0;
    V.slice_ptr = d_slice_ptr; V.vals = d_vals; V.cols = d_cols; V.row_len = d_row_len;
    V.xsq = d_xsq; V.serial = d_serial;
    V.nvar = nvar; V.svr = (nvar == 2 * l && l > 0) ? 1 : 0; V.kind = (int)kind;
    V.max_iter = host_ctl ? host_ctl->max_iter : 0;
    V.Cp = Cp; V.Cn = Cn; V.eps = eps;
    V.G = d_G; V.alpha = d_alpha; V.p = d_p; V.QD = d_QD; V.y = d_y; V.status = d_status;
    V.rows = d_rows; V.row_stride = row_stride; V.S = S; V.full = full;
    V.slot_of_row = d_slot_of_row; V.row_of_slot = d_row_of_slot; V.lru_prev = d_prev; V.lru_next = d_next;
    V.computed = d_computed;
    V.ctl = d_ctl; V.part = d_part; V.xs_global = d_xs;
    V.use_lds = use_lds; V.nblocks = nblocks; V.rank = -1; V.xrep = xrep;
    V.xw = d_xw; V.pm_flags = d_pm_flags; V.pm_lru = d_pm_lru; V.pm_stride = pm_stride;
    return V;
  }

  // Ends the current solve before a new svmhip_setup on the same rows.  The
  // per-solve buffers (solver vectors, kernel-row cache, cache bookkeeping)
  // stay allocated and are reused when the next solve fits them: freeing
  // and re-allocating a cache of tens of GB per solve costs more than the
  // solve.
  void end_solve() {
    (void)hipStreamSynchronize(st);
    setup_done = false;
    have_j = false;
    stats = svmhip_stats{};
  }

  void read_ctl() {
    ck(hipMemcpyAsync(host_ctl, d_ctl, sizeof(smo::Ctl), hipMemcpyDeviceToHost, st), "read control record");
    ck(hipStreamSynchronize(st), "synchronise");
  }

  ~svmhip_solver() {
    if (dev >= 0) (void)hipSetDevice(dev);
    if (st) (void)hipStreamSynchronize(st);
    for (void *p : allocs) (void)hipFree(p);
    if (host_ctl) (void)hipHostFree(host_ctl);
    for (auto &e : ev)
      if (e) (void)hipEventDestroy(e);
    if (st) (void)hipStreamDestroy(st);
  }
};

#define GUARD_BEGIN try {
#define GUARD_END                          \
  }                                        \
  catch (const Failure &f) {               \
    return f.code;                         \
  }                                        \
  catch (const std::bad_alloc &) {         \
    h->err = "host allocation failed";     \
    return SVMHIP_ERR_NOMEM;               \
  }                                        \
  return SVMHIP_OK;


// Rows of a problem packed on the host in the SELL-64 layout
// (see smo_device.hpp), ready for upload.
// Allocator whose vectors leave new elements uninitialised (the packers
// zero and fill them in parallel, by slice ranges).
template <class T>
struct uninit_alloc : std::allocator<T> {
  template <class U>
  struct rebind {
    using other = uninit_alloc<U>;
  };
  uninit_alloc() = default;
  template <class U>
  uninit_alloc(const uninit_alloc<U> &) {}
  template <class U>
  void construct(U *p) noexcept {
    ::new ((void *)p) U;
  }
  template <class U, class... A>
  void construct(U *p, A &&...a) {
    ::new ((void *)p) U(std::forward<A>(a)...);
  }
};

// f(i0, i1) over [0, n) split among host threads (at most 16) when the work
// (n items of `cost` each) is large enough to pay for them.
template <class F>
void par_ranges(int n, double cost, F f) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  const int nt = (int)std::min<double>({(double)std::min(hw, 16u), (double)n, n * cost / 2e5});
  if (nt <= 1) {
    f(0, n);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nt; ++t) {
    const int i0 = (int)((long long)n * t / nt), i1 = (int)((long long)n * (t + 1) / nt);
    th.emplace_back([=, &f] { f(i0, i1); });
  }
  for (auto &x : th) x.join();
}

struct HostSell {
  int l = 0, d = 1, nslices = 0, dense = 0;
  long long total = 0;
  std::vector<long long> sp;   // nslices + 1 slice offsets
  std::vector<float, uninit_alloc<float>> vals;
  std::vector<int, uninit_alloc<int>> cols;  // empty when dense
  std::vector<int> rl;         // stored entries per row
  std::vector<int> serial;     // PRECOMPUTED only
};

// Packs prob's rows (features truncated to float32, as the reference does on
// upload, libsvm/cuda_solver.cpp:366); errors go through h->fail.
void make_host_sell(svmhip_solver *h, const svm_problem *prob, int ktype, HostSell &out);

// Dynamic LDS of a batched launch whose rows have up to maxd features: room
// for 8 staged query rows (G initialisation, speculative rows), at least one
// row, at most 64 KiB.
inline size_t batch_lds_bytes(int maxd) {
  const size_t one = (size_t)std::max(maxd, 1) * sizeof(float);
  return std::max(one, std::min<size_t>(8 * one, 64 * 1024));
}

// Words of the LDS row-flag bitset a batched launch adds past
// batch_lds_bytes(maxd) for subproblems of up to maxn rows (0 = none: the
// launch's dynamic LDS stays within 64 KiB).
inline int batch_flag_words(int maxd, long long maxn) {
  const long long w = (std::max<long long>(maxn, 1) + 31) / 32;
  return batch_lds_bytes(maxd) + (size_t)w * sizeof(unsigned) <= 64 * 1024 ? (int)w : 0;
}

// Device wall-clock ticks per second (wall_clock64 in kernels).
inline double device_clock_hz(int dev) {
  int khz = 0;
  if (hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, dev) != hipSuccess || khz <= 0) khz = 100000;
  return 1e3 * khz;
}

// LDS bytes of the persistent kernel's resident vectors for nv variables per
// row and m rows: G, alpha, QD (FP64), three kernel-row segments (float),
// labels and statuses (int8).
inline size_t smo_res_bytes(int nv, size_t m) { return (24 * (size_t)nv + 12 + 2 * (size_t)nv) * m; }

// Dynamic LDS one workgroup may use (bytes).
inline int device_max_lds(int dev) {
  int v = 0;
  if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess || v <= 0) {
    (void)hipGetLastError();
    v = 64 * 1024;
  }
  return v;
}

inline int device_cus(int dev) {
  int cus = 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) cus = 0;
  return cus > 0 ? cus : 1;
}
