// Host runtime of the device SMO solver: the svmhip_* C-ABI (include/svm_hip.h).
//
// Owns every device buffer of one solve in a handle (the reference keeps them
// in a process-global CudaSolver plus __constant__/__device__ symbols,
// libsvm/cuda_solver.cpp:5, libsvm/svm_device.cu:32-53), converts the
// svm_node rows to the SELL-64 layout on upload, and drives either the
// step-wise kernels (mirroring CudaSolver's method-per-phase interface) or
// the persistent cooperative kernel (svmhip_solve).
#include <dlfcn.h>

#include <mutex>

#include "runtime_internal.hpp"

// ------------------------------------------------------- launch table
//
// Launch geometry chosen by the autotuner (libsvm-cuda_amd/tune/autotune.py,
// which times whole solves over a sweep of grid sizes and exchange-record
// copies on the device and writes the fastest per problem shape).  The
// table is a text file, tune/launch_table.txt beside the library's lib/
// directory (or the file SVMHIP_TUNE_TABLE names), one row per shape range:
//     mode full dense l_min l_max d_min d_max grid xrep
// mode "exact" (kernel rows on demand) or "gram" (mfma_rows = 1); full = 1
// when the whole kernel matrix fits the row cache.  Without the file, or
// for a shape no row covers, the fixed size rule in svmhip_create applies.
struct TuneRow {
  int gram, full, dense;
  long long lmin, lmax;
  int dmin, dmax, grid, xrep;
};

static const std::vector<TuneRow> &tune_table() {
  static const std::vector<TuneRow> rows = [] {
    std::vector<TuneRow> out;
    std::string path;
    const char *env = getenv("SVMHIP_TUNE_TABLE");
    if (env && *env) {
      path = env;
    } else {
      Dl_info info;
      if (dladdr((void *)&tune_table, &info) && info.dli_fname) {
        std::string so = info.dli_fname;
        const size_t slash = so.rfind('/');
        path = (slash == std::string::npos ? std::string(".") : so.substr(0, slash)) + "/../tune/launch_table.txt";
      }
    }
    FILE *f = path.empty() ? nullptr : fopen(path.c_str(), "r");
    if (!f) return out;
    char line[256];
    while (fgets(line, sizeof line, f)) {
      if (line[0] == '#') continue;
      char mode[16];
      TuneRow r;
      if (sscanf(line, "%15s %d %d %lld %lld %d %d %d %d", mode, &r.full, &r.dense, &r.lmin, &r.lmax, &r.dmin,
                 &r.dmax, &r.grid, &r.xrep) == 9 &&
          r.grid > 0) {
        r.gram = strcmp(mode, "gram") == 0;
        out.push_back(r);
      }
    }
    fclose(f);
    return out;
  }();
  return rows;
}

// The table's row for a problem shape, or null.
static const TuneRow *tuned_launch(int gram, int full, int dense, int l, int d) {
  for (const TuneRow &r : tune_table())
    if (r.gram == gram && r.full == full && r.dense == dense && r.lmin <= l && l <= r.lmax && r.dmin <= d &&
        d <= r.dmax)
      return &r;
  return nullptr;
}

// ------------------------------------------------------------- upload

// Rows of prob -> SELL-64 (see smo_device.hpp).  Features are truncated to
// float32 as the reference does on upload (libsvm/cuda_solver.cpp:366).
void make_host_sell(svmhip_solver *h, const svm_problem *prob, int ktype, HostSell &o) {
  const int l = prob->l;
  const bool pre = ktype == PRECOMPUTED;
  std::vector<int> len(l);
  long long nnz = 0;
  int maxidx = -1, maxlen = 0;
  // row lengths and checks, in parallel over rows; the first bad row is the
  // one reported (kind 1: null, 2: negative index, 3: not ascending, 4:
  // precomputed row without its serial)
  int bad_row = INT_MAX, bad_kind = 0, bad_index = 0;
  std::mutex mu;
  par_ranges(l, 64.0, [&](int r0, int r1) {
    long long z = 0;
    int mx = -1, ml = 0, br = INT_MAX, bk = 0, bi = 0;
    for (int r = r0; r < r1 && br == INT_MAX; ++r) {
      const svm_node *x = prob->x[r];
      if (!x) {
        br = r, bk = 1;
        break;
      }
      int n = 0, prev = INT_MIN;
      for (; x[n].index != -1; ++n) {
        if (!pre) {
          if (x[n].index < 0) {
            br = r, bk = 2, bi = x[n].index;
            break;
          }
          if (x[n].index <= prev) {
            br = r, bk = 3;
            break;
          }
          prev = x[n].index;
          mx = std::max(mx, x[n].index);
        }
      }
      if (br != INT_MAX) break;
      if (pre && n == 0) {
        br = r, bk = 4;
        break;
      }
      len[r] = n;
      z += n;
      ml = std::max(ml, n);
    }
    std::lock_guard<std::mutex> g(mu);
    nnz += z;
    maxidx = std::max(maxidx, mx);
    maxlen = std::max(maxlen, ml);
    if (br < bad_row) bad_row = br, bad_kind = bk, bad_index = bi;
  });
  if (bad_kind == 1) h->fail(SVMHIP_ERR_ARG, "row %d is null", bad_row);
  if (bad_kind == 2) h->fail(SVMHIP_ERR_ARG, "row %d: negative feature index %d", bad_row, bad_index);
  if (bad_kind == 3) h->fail(SVMHIP_ERR_ARG, "row %d: feature indices not ascending", bad_row);
  if (bad_kind == 4) h->fail(SVMHIP_ERR_ARG, "row %d: precomputed row needs 0:serial", bad_row);
  o.l = l;
  o.d = pre ? std::max(maxlen, 1) : std::max(maxidx + 1, 1);
  o.dense = (!pre && l > 0 && (double)nnz >= 0.5 * (double)l * (double)o.d) ? 1 : 0;
  o.nslices = (l + smo::SLICE - 1) / smo::SLICE;
  o.sp.assign(o.nslices + 1, 0);
  o.rl.assign(l, 0);
  for (int s = 0; s < o.nslices; ++s) {
    int w = 0;
    for (int r = s * smo::SLICE; r < std::min(l, (s + 1) * smo::SLICE); ++r) {
      o.rl[r] = o.dense ? o.d : len[r];
      w = std::max(w, o.rl[r]);
    }
    o.sp[s + 1] = o.sp[s] + (long long)w * smo::SLICE;
  }
  o.total = o.sp[o.nslices];
  o.vals.resize(std::max<long long>(o.total, 1));  // (not initialised: zeroed below)
  o.cols.clear();
  if (!o.dense) o.cols.resize(std::max<long long>(o.total, 1));
  if (o.total == 0) {
    o.vals[0] = 0.0f;
    if (!o.dense) o.cols[0] = 0;
  }
  o.serial.clear();
  if (pre) o.serial.resize(l);
  // each thread zeroes and fills whole slices
  const double per_slice = o.nslices ? (double)o.total / o.nslices : 0.0;
  par_ranges(o.nslices, per_slice, [&](int s0, int s1) {
    std::fill(o.vals.data() + o.sp[s0], o.vals.data() + o.sp[s1], 0.0f);
    if (!o.dense) std::fill(o.cols.data() + o.sp[s0], o.cols.data() + o.sp[s1], 0);
    for (int r = s0 * smo::SLICE; r < std::min(l, s1 * smo::SLICE); ++r) {
      const svm_node *x = prob->x[r];
      const long long base = o.sp[r / smo::SLICE] + (r % smo::SLICE);
      if (pre) o.serial[r] = (int)x[0].value;
      for (int p = 0; p < len[r]; ++p) {
        if (o.dense) {
          o.vals[base + (long long)x[p].index * smo::SLICE] = (float)x[p].value;
        } else {
          o.vals[base + (long long)p * smo::SLICE] = (float)x[p].value;
          o.cols[base + (long long)p * smo::SLICE] = pre ? p : x[p].index;
        }
      }
    }
  });
}

static void build_sell(svmhip_solver *h, const svm_problem *prob) {
  HostSell o;
  make_host_sell(h, prob, h->ktype, o);
  const int l = o.l;
  h->d = o.d;
  h->dense = o.dense;
  h->nslices = o.nslices;
  h->total = o.total;
  h->d_slice_ptr = h->dalloc<long long>(o.sp.size(), "slice pointers");
  h->d_vals = h->dalloc<float>(o.vals.size(), "feature values");
  h->d_row_len = h->dalloc<int>(l, "row lengths");
  h->upload(h->d_slice_ptr, o.sp.data(), o.sp.size(), "upload slice pointers");
  h->upload(h->d_vals, o.vals.data(), o.vals.size(), "upload feature values");
  h->upload(h->d_row_len, o.rl.data(), (size_t)l, "upload row lengths");
  if (!o.dense) {
    h->d_cols = h->dalloc<int>(o.cols.size(), "feature columns");
    h->upload(h->d_cols, o.cols.data(), o.cols.size(), "upload feature columns");
  }
  if (!o.serial.empty()) {
    h->d_serial = h->dalloc<int>(l, "serials");
    h->upload(h->d_serial, o.serial.data(), (size_t)l, "upload serials");
  }
  h->d_xsq = h->dalloc<double>(l, "squared norms");
  h->host_row_len = o.rl;
  // the host vectors die at return: wait for the copies
  h->ck(hipStreamSynchronize(h->st), "upload");
}

// Empties the cache bookkeeping the step kernels share (all rows invalid).
static void reset_shared_cache(svmhip_solver *h) {
  if (h->full) {
    h->ck(hipMemsetAsync(h->d_computed, 0, std::max(h->l, 1), h->st), "clear cache flags");
  } else {
    std::vector<int> init(std::max(h->l, h->S), -1);
    h->upload(h->d_slot_of_row, init.data(), h->l, "init cache map");
    h->upload(h->d_row_of_slot, init.data(), h->S, "init cache map");
    std::vector<int> prv(h->S), nxt(h->S);
    for (int s = 0; s < h->S; ++s) {
      prv[s] = s - 1;
      nxt[s] = s + 1 < h->S ? s + 1 : -1;
    }
    h->upload(h->d_prev, prv.data(), h->S, "init LRU");
    h->upload(h->d_next, nxt.data(), h->S, "init LRU");
    h->ck(hipStreamSynchronize(h->st), "init LRU");  // host vectors die here
  }
  h->host_ctl->lru_head = 0;
  h->host_ctl->lru_tail = h->S - 1;
  h->ck(hipMemcpyAsync(&h->d_ctl->lru_head, &h->host_ctl->lru_head, 2 * sizeof(int), hipMemcpyHostToDevice, h->st),
        "reset LRU ends");
  h->ck(hipStreamSynchronize(h->st), "reset cache");
  h->shared_meta_dirty = false;
}


// Gram-tile mode: makes the FP16 hi/lo copy of X the first time (false if
// X leaves FP16's range, in which case the solve computes exact rows).
static bool gram_prepare(svmhip_solver *h) {
  if (h->split_state == 0) {
    h->d_xsplit = h->dalloc<_Float16>(smo_split_f16_bytes(h->l, h->d) / sizeof(_Float16), "FP16 copy of the rows");
    h->d_flag = h->dalloc<int>(1, "range flag");
    h->ck(hipMemsetAsync(h->d_flag, 0, sizeof(int), h->st), "clear range flag");
    smo::View V = h->view();
    h->ck(smo_launch_split_f16(&V, h->d_xsplit, h->d_flag, h->st), "FP16 split kernel");
    int bad = 0;
    h->ck(hipMemcpyAsync(&bad, h->d_flag, sizeof(int), hipMemcpyDeviceToHost, h->st), "read range flag");
    h->ck(hipStreamSynchronize(h->st), "FP16 split");
    h->split_state = bad ? -1 : 1;
    if (bad && h->opt.verbose) fprintf(stderr, "svmhip: features beyond FP16 range, exact rows instead of Gram tiles\n");
  }
  return h->split_state == 1;
}

extern "C" {

int svmhip_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) {
    (void)hipGetLastError();
    return 0;
  }
  return n;
}

const char *svmhip_last_error(const svmhip_solver *h) { return h ? h->err.c_str() : "null handle"; }

svmhip_status svmhip_create(const svm_problem *prob, const svm_parameter *param, svmhip_kind kind,
                            const svmhip_options *opt, svmhip_solver **out) {
  if (!out) return SVMHIP_ERR_ARG;
  *out = nullptr;
  if (!prob || !param || prob->l < 0 || (prob->l > 0 && !prob->x)) return SVMHIP_ERR_ARG;
  if (kind != SVMHIP_KIND_C && kind != SVMHIP_KIND_NU) return SVMHIP_ERR_ARG;
  svmhip_solver *h = new (std::nothrow) svmhip_solver;
  if (!h) return SVMHIP_ERR_NOMEM;
  h->dev = -1;
  svmhip_status rc = [&]() -> svmhip_status {
    GUARD_BEGIN
    if (opt) {
      size_t n = std::min(opt->struct_size ? opt->struct_size : sizeof(svmhip_options), sizeof(svmhip_options));
      memcpy(&h->opt, opt, n);
    }
    if (h->opt.block_threads != 0 && h->opt.block_threads != 512)
      h->fail(SVMHIP_ERR_ARG, "block_threads must be 512 (got %d)", h->opt.block_threads);
    int ndev = svmhip_device_count();
    if (ndev <= 0) h->fail(SVMHIP_ERR_DEVICE, "no HIP device available");
    if (h->opt.device < 0 || h->opt.device >= ndev)
      h->fail(SVMHIP_ERR_ARG, "device %d out of range (%d devices)", h->opt.device, ndev);
    h->dev = h->opt.device;
    h->ck(hipSetDevice(h->dev), "hipSetDevice");
    h->ck(hipStreamCreateWithFlags(&h->st, hipStreamNonBlocking), "create stream");
    for (auto &e : h->ev) h->ck(hipEventCreate(&e), "create event");
    h->ck(hipHostMalloc((void **)&h->host_ctl, sizeof(smo::Ctl), hipHostMallocDefault), "pinned control record");
    memset(h->host_ctl, 0, sizeof(smo::Ctl));
    h->kind = kind;
    h->svm_type = param->svm_type;
    h->ktype = param->kernel_type;
    h->degree = param->degree;
    h->gamma = param->gamma;
    h->coef0 = param->coef0;
    h->eps = param->eps;
    h->param_cache_mb = param->cache_size;
    if (h->ktype < LINEAR || h->ktype > PRECOMPUTED) h->fail(SVMHIP_ERR_ARG, "unknown kernel type %d", h->ktype);
    if (!(h->eps > 0)) h->fail(SVMHIP_ERR_ARG, "eps <= 0");
    h->l = prob->l;
    build_sell(h, prob);
    // launch geometry: at most one workgroup per slice, one per CU by default
    int cus = device_cus(h->dev);
    // default: one workgroup per CU once there are two 64-row slices per CU
    // (row passes then use the whole device); below that the largest power
    // of two (at least 32) that leaves each workgroup about 2.4 slices:
    // fewer workgroups make each exchange cheaper, and powers of two spread
    // evenly over the 8 XCDs.  Measured on dense 256-feature problems: 10k
    // rows 64 workgroups (79.7 ms; 32: 89.9), 20k and 30k rows 128 (106.9,
    // 129.2 ms; 256: 111.5, 131.9), 100k rows 256 (6% ahead of 128).
    int want = h->opt.grid_blocks;
    // the cache regime the solves will likely run in (the cache is sized at
    // setup; this uses the same byte budget)
    const double cache_b = h->opt.cache_bytes > 0 ? h->opt.cache_bytes : h->param_cache_mb * 1048576.0;
    const int full_cache = cache_b >= 4.0 * (double)h->l * (double)(((long long)h->l + 63) / 64 * 64);
    const int gram_mode = h->opt.mfma_rows == 1 && h->dense && h->ktype != PRECOMPUTED && full_cache;
    const TuneRow *tr = tuned_launch(gram_mode, full_cache, h->dense, h->l, h->d);
    if (want <= 0 && tr) want = tr->grid;
    h->xrep = h->opt.exchange_copies > 0 ? std::min(h->opt.exchange_copies, smo::XREP) : (tr ? tr->xrep : 0);
    if (want <= 0) {
      if (h->nslices >= 2 * cus) {
        want = cus;
      } else {
        want = 32;
        while (2 * want <= cus && 2 * want <= h->nslices * 5 / 12) want *= 2;
      }
    }
    h->nblocks = std::max(1, std::min(want, std::max(h->nslices, 1)));
    size_t qbytes = (size_t)h->d * sizeof(float);
    h->use_lds = qbytes <= kLdsQueryLimit;
    h->shmem = h->use_lds ? qbytes : 0;
    int per_cu = 0;
    h->ck(smo_persistent_capacity((int)kind, h->shmem, &per_cu), "occupancy query");
    int cap = std::max(1, per_cu) * cus;
    h->nblocks = std::min(h->nblocks, cap);
    h->nblocks = std::min(h->nblocks, 512);  // the exchange reads one record per thread
    if (!h->use_lds) h->d_xs = h->dalloc<float>((size_t)h->nblocks * h->d, "query-row scratch");
    h->d_ctl = h->dalloc<smo::Ctl>(1, "control record");
    h->d_part = h->dalloc<smo::Partial>(2 * (size_t)h->nblocks, "partials");
    h->d_xw = h->dalloc<smo::XWord>(smo::xrec_words(h->nblocks), "exchange records");
    h->ck(hipMemsetAsync(h->d_ctl, 0, sizeof(smo::Ctl), h->st), "clear control record");
    h->ck(hipStreamSynchronize(h->st), "create");
    GUARD_END
  }();
  if (rc != SVMHIP_OK) {
    if (h->err.empty()) h->err = "create failed";
    fprintf(stderr, "svmhip_create: %s\n", h->err.c_str());
    delete h;
    return rc;
  }
  *out = h;
  return SVMHIP_OK;
}

svmhip_status svmhip_setup_rbf(svmhip_solver *h, int l) {
  if (!h) return SVMHIP_ERR_ARG;
  GUARD_BEGIN
  if (l != h->l) h->fail(SVMHIP_ERR_ARG, "setup_rbf: l=%d but the problem has %d rows", l, h->l);
  h->ck(hipSetDevice(h->dev), "hipSetDevice");
  smo::View V = h->view();
  if (h->l > 0) h->ck(smo_launch_self_dot(&V, h->d_xsq, h->st), "x^2 kernel");
  h->rbf_ready = true;
  GUARD_END
}

svmhip_status svmhip_setup(svmhip_solver *h, const signed char *y, const double *G, const double *alpha,
                           const char *alpha_status, double Cp, double Cn, int n_var) {
  if (!h) return SVMHIP_ERR_ARG;
  GUARD_BEGIN
  if (!y || !G || !alpha || !alpha_status) h->fail(SVMHIP_ERR_ARG, "setup: null array");
  if (n_var != h->l && n_var != 2 * h->l) h->fail(SVMHIP_ERR_ARG, "setup: n_var=%d must be l or 2l (l=%d)", n_var, h->l);
  if (n_var == 2 * h->l && h->l > 0 && h->svm_type != EPSILON_SVR && h->svm_type != NU_SVR)
    h->fail(SVMHIP_ERR_ARG, "setup: n_var=2l is only valid for regression");
  if (h->setup_done) h->end_solve();  // a new solve over the uploaded rows, reusing the buffers
  if (!(Cp > 0) || !(Cn > 0)) h->fail(SVMHIP_ERR_ARG, "setup: Cp, Cn must be > 0");
  h->ck(hipSetDevice(h->dev), "hipSetDevice");
  if (!h->rbf_ready) {
    svmhip_status s = svmhip_setup_rbf(h, h->l);
    if (s != SVMHIP_OK) throw Failure{s};
  }
  h->nvar = n_var;
  h->Cp = Cp;
  h->Cn = Cn;
  for (int v = 0; v < n_var; ++v)
    if (y[v] != 1 && y[v] != -1) h->fail(SVMHIP_ERR_ARG, "setup: y[%d]=%d is not +1/-1", v, (int)y[v]);
  const int nv = n_var;
  if (h->cap_nvar < (size_t)std::max(nv, 1)) {  // the six solver vectors share one capacity
    for (void **p : {(void **)&h->d_G, (void **)&h->d_alpha, (void **)&h->d_p, (void **)&h->d_QD, (void **)&h->d_y,
                     (void **)&h->d_status}) {
      h->dfree(*p);
      *p = nullptr;
    }
    h->cap_nvar = 0;
    h->d_G = h->dalloc<double>(nv, "G");
    h->d_alpha = h->dalloc<double>(nv, "alpha");
    h->d_p = h->dalloc<double>(nv, "p");
    h->d_QD = h->dalloc<double>(nv, "QD");
    h->d_y = h->dalloc<int8_t>(nv, "y");
    h->d_status = h->dalloc<int8_t>(nv, "status");
    h->cap_nvar = (size_t)std::max(nv, 1);
  }
  h->upload(h->d_G, G, nv, "upload G");
  h->upload(h->d_p, G, nv, "upload p");  // Solver::Solve passes G = p here
  h->upload(h->d_alpha, alpha, nv, "upload alpha");
  h->upload(h->d_y, (const int8_t *)y, nv, "upload y");
  h->upload(h->d_status, (const int8_t *)alpha_status, nv, "upload status");

  // kernel-row cache: S rows of l floats (64-bit sized; the reference's
  // int byte count overflows at 2 GB, libsvm/cuda_solver.cpp:278).  The
  // cache a previous solve on this handle left allocated counts as free.
  double bytes = h->opt.cache_bytes > 0 ? h->opt.cache_bytes : h->param_cache_mb * 1048576.0;
  h->row_stride = ((long long)h->l + 63) / 64 * 64;
  long long rowb = std::max<long long>(h->row_stride, 1) * (long long)sizeof(float);
  long long want = (long long)std::floor(bytes / (double)rowb);
  size_t freeb = 0, totb = 0;
  if (hipMemGetInfo(&freeb, &totb) == hipSuccess) {
    const double avail = (double)freeb + (double)h->cap_rows * sizeof(float);
    long long cap = (long long)(avail * 0.85) / rowb;
    want = std::min(want, cap);
  }
  want = std::max<long long>(want, 2);
  if (want >= h->l) {
    h->full = 1;
    h->S = std::max(h->l, 1);
  } else {
    h->full = 0;
    h->S = (int)want;
  }
  h->ensure(h->d_rows, h->cap_rows, (size_t)h->S * (size_t)h->row_stride, "kernel-row cache");
  if (h->full) {
    h->ensure(h->d_computed, h->cap_flags, (size_t)std::max(h->l, 1), "cache flags");
    h->ensure(h->d_pm_flags, h->cap_pm_flags, (size_t)h->nblocks * std::max(h->l, 1), "private cache flags");
  } else {
    h->ensure(h->d_slot_of_row, h->cap_lru_map, (size_t)h->l, "cache map");
    if (h->cap_lru_slots < (size_t)h->S) {  // row_of_slot, prev, next share one capacity
      for (void **p : {(void **)&h->d_row_of_slot, (void **)&h->d_prev, (void **)&h->d_next}) {
        h->dfree(*p);
        *p = nullptr;
      }
      h->cap_lru_slots = 0;
      h->d_row_of_slot = h->dalloc<int>(h->S, "cache map");
      h->d_prev = h->dalloc<int>(h->S, "LRU links");
      h->d_next = h->dalloc<int>(h->S, "LRU links");
      h->cap_lru_slots = (size_t)h->S;
    }
    h->pm_stride = ((long long)h->l + 3LL * h->S + 2 + 31) / 32 * 32;
    h->ensure(h->d_pm_lru, h->cap_pm_lru, (size_t)h->nblocks * (size_t)h->pm_stride, "private LRU bookkeeping");
  }
  memset(h->host_ctl, 0, sizeof(smo::Ctl));
  reset_shared_cache(h);
  int mi = h->opt.max_iter;
  // Solver::Solve's cap over its own l, the number of variables (2l for
  // regression), libsvm/svm.cpp:577
  if (mi <= 0) mi = std::max(10000000, nv > INT_MAX / 100 ? INT_MAX : 100 * nv);
  h->host_ctl->max_iter = mi;
  h->ck(hipMemcpyAsync(h->d_ctl, h->host_ctl, sizeof(smo::Ctl), hipMemcpyHostToDevice, h->st), "init control");

  smo::View V = h->view();
  h->ck(hipEventRecord(h->ev[2], h->st), "event");
  if (nv > 0) h->ck(smo_launch_qd(&V, h->st), "QD kernel");
  // G = p + sum_{alpha_v > 0} alpha_v Q_v, in increasing v (svm.cpp:555-569):
  // one launch over the list of such v, several rows staged per pass over X
  // (each G_k still accumulates in increasing v); rows too wide for LDS go
  // one launch per v
  std::vector<int> list;
  for (int v = 0; v < nv; ++v)
    if (alpha_status[v] != smo::AT_LOWER) list.push_back(v);
  if (!list.empty()) {
    if (h->use_lds) {
      h->ensure(h->d_list, h->cap_list, list.size(), "G initialisation list");
      h->upload(h->d_list, list.data(), list.size(), "upload G initialisation list");
      h->ck(smo_launch_ginit_multi(&V, h->d_list, (int)list.size(), 0, h->st), "G init kernel");
    } else {
      for (int v : list) h->ck(smo_launch_ginit_row(&V, v, h->shmem, h->st), "G init kernel");
    }
  }
  h->ck(hipEventRecord(h->ev[3], h->st), "event");
  h->ck(hipStreamSynchronize(h->st), "setup");
  float ms = 0;
  (void)hipEventElapsedTime(&ms, h->ev[2], h->ev[3]);
  h->stats.setup_ms = ms;
  h->stats.cache_slots = h->S;
  h->stats.grid_blocks = h->nblocks;
  h->setup_done = true;
  h->have_j = false;
  h->gram_done = false;
  h->gram_rows = 0;
  if (h->opt.verbose)
    fprintf(stderr, "svmhip: l=%d d=%d %s, n_var=%d, cache %d rows%s, grid %d x 512, exchange copies %d\n", h->l,
            h->d, h->dense ? "dense" : "sparse", nv, h->S, h->full ? " (all)" : "", h->nblocks, h->xrep);
  GUARD_END
}

svmhip_status svmhip_select_working_set(svmhip_solver *h, int n_var, int *optimal, int *out_i, int *out_j) {
  if (!h) return SVMHIP_ERR_ARG;
  GUARD_BEGIN
  if (!h->setup_done) h->fail(SVMHIP_ERR_STATE, "select_working_set before setup");
  if (n_var != h->nvar) h->fail(SVMHIP_ERR_ARG, "n_var mismatch");
  h->ck(hipSetDevice(h->dev), "hipSetDevice");
  if (h->shared_meta_dirty) reset_shared_cache(h);
  smo::View V = h->view();
  h->ck(smo_launch_select_i(&V, h->shmem, h->st), "select i");
  h->read_ctl();  // the Gmax / Gmax2 copy the reference makes (cuda_solver.cpp:83-86)
  if (h->host_ctl->err) h->fail(SVMHIP_ERR_TIMEOUT, "device wait timed out");
  if (h->host_ctl->stop) {
    if (optimal) *optimal = 1;
    h->have_j = false;
    return SVMHIP_OK;
  }
  h->ck(smo_launch_select_j(&V, h->shmem, h->st), "select j");
  if (optimal) *optimal = 0;
  if (out_i || out_j) {
    h->read_ctl();
    if (h->host_ctl->stop) {
      if (optimal) *optimal = 1;
      return SVMHIP_OK;
    }
    if (out_i) *out_i = h->host_ctl->i;
    if (out_j) *out_j = h->host_ctl->j;
  }
  h->have_j = true;
  GUARD_END
}

svmhip_status svmhip_compute_alpha(svmhip_solver *h) {
  if (!h) return SVMHIP_ERR_ARG;
  GUARD_BEGIN
  if (!h->have_j) h->fail(SVMHIP_ERR_STATE, "compute_alpha without a selected pair");
  smo::View V = h->view();
  h->ck(smo_launch_alpha(&V, h->st), "alpha kernel");
  GUARD_END
}

svmhip_status svmhip_update_gradient(svmhip_solver *h, int n_var) {
  if (!h) return SVMHIP_ERR_ARG;
  GUARD_BEGIN
  if (!h->have_j) h->fail(SVMHIP_ERR_STATE, "update_gradient without a selected pair");
  if (n_var != h->nvar) h->fail(SVMHIP_ERR_ARG, "n_var mismatch");
  smo::View V = h->view();
  h->ck(smo_launch_gupdate(&V, h->shmem, h->st), "gradient kernel");
  GUARD_END
}

svmhip_status svmhip_update_alpha_status(svmhip_solver *h) {
  if (!h) return SVMHIP_ERR_ARG;
  GUARD_BEGIN
  if (!h->have_j) h->fail(SVMHIP_ERR_STATE, "update_alpha_status without a selected pair");
  smo::View V = h->view();
  h->ck(smo_launch_status(&V, h->st), "status kernel");
  h->have_j = false;
  GUARD_END
}

svmhip_status svmhip_fetch(svmhip_solver *h, double *G, double *alpha, char *alpha_status, int n_var) {
  if (!h) return SVMHIP_ERR_ARG;
  GUARD_BEGIN
  if (!h->setup_done) h->fail(SVMHIP_ERR_STATE, "fetch before setup");
  if (n_var != h->nvar) h->fail(SVMHIP_ERR_ARG, "n_var mismatch");
  h->ck(hipSetDevice(h->dev), "hipSetDevice");
  if (G && n_var) h->ck(hipMemcpyAsync(G, h->d_G, sizeof(double) * n_var, hipMemcpyDeviceToHost, h->st), "fetch G");
  if (alpha && n_var)
    h->ck(hipMemcpyAsync(alpha, h->d_alpha, sizeof(double) * n_var, hipMemcpyDeviceToHost, h->st), "fetch alpha");
  if (alpha_status && n_var)
    h->ck(hipMemcpyAsync(alpha_status, h->d_status, n_var, hipMemcpyDeviceToHost, h->st), "fetch status");
  h->ck(hipStreamSynchronize(h->st), "fetch");
  if (h->opt.verbose) {
    h->read_ctl();
    fprintf(stderr, "svmhip: cache hits %llu, rows computed %llu\n", h->host_ctl->hits, h->host_ctl->misses);
  }
  GUARD_END
}

svmhip_status svmhip_finalize(svmhip_solver *h, svmhip_stats *stats) {
  if (!h) return SVMHIP_ERR_ARG;
  GUARD_BEGIN
  if (!h->setup_done) h->fail(SVMHIP_ERR_STATE, "finalize before setup");
  h->ck(hipSetDevice(h->dev), "hipSetDevice");
  smo::View V = h->view();
  if (h->nvar > 0) h->ck(smo_launch_finalize(&V, h->st), "finalize kernel");
  h->read_ctl();
  h->stats.obj = h->host_ctl->obj;
  h->stats.rho = h->host_ctl->rho;
  h->stats.r = h->host_ctl->r;
  h->stats.iters = h->host_ctl->iter;
  h->stats.row_hits = (long long)h->host_ctl->hits;
  h->stats.row_misses = (long long)h->host_ctl->misses + h->gram_rows;
  h->stats.device_allocs = h->n_allocs;
  if (stats) *stats = h->stats;
  GUARD_END
}

svmhip_status svmhip_solve(svmhip_solver *h, int max_iter, svmhip_stats *stats) {
  if (!h) return SVMHIP_ERR_ARG;
  svmhip_status rc = [&]() -> svmhip_status {
    GUARD_BEGIN
    if (!h->setup_done) h->fail(SVMHIP_ERR_STATE, "solve before setup");
    h->ck(hipSetDevice(h->dev), "hipSetDevice");
    if (max_iter > 0) {
      h->host_ctl->max_iter = max_iter;
      h->ck(hipMemcpyAsync(&h->d_ctl->max_iter, &h->host_ctl->max_iter, sizeof(int), hipMemcpyHostToDevice, h->st),
            "set max_iter");
    }
    h->read_ctl();
    // Gram-tile mode: the whole kernel matrix on the matrix cores, once per
    // setup, when the rows are dense and every row fits the cache
    const bool use_gram = h->opt.mfma_rows == 1 && h->full && h->dense && h->ktype != PRECOMPUTED && h->l > 0 &&
                          h->nvar > 0 && gram_prepare(h);
    // persistent-kernel state: start from the current alpha / G / status with
    // every workgroup's private cache bookkeeping empty (Gram-tile mode: every
    // row present)
    {
      smo::Dec &dc = smo::ctl_dec(h->host_ctl);
      memset(&dc, 0, sizeof dc);
      dc.iter = h->host_ctl->iter;
      h->host_ctl->stop = 0;
      h->host_ctl->err = 0;
      h->ck(hipMemcpyAsync(h->d_ctl, h->host_ctl, sizeof(smo::Ctl), hipMemcpyHostToDevice, h->st), "seed control");
      h->ck(hipMemsetAsync(h->d_xw, 0, sizeof(smo::XWord) * smo::xrec_words(h->nblocks), h->st), "clear exchange records");
      smo::View V0 = h->view();
      if (h->full)
        h->ck(hipMemsetAsync(h->d_pm_flags, use_gram ? 1 : 0, (size_t)h->nblocks * std::max(h->l, 1), h->st),
              use_gram ? "mark every row cached" : "clear cache flags");
      else
        h->ck(smo_launch_lru_private_init(&V0, h->st), "init private LRU");
      h->shared_meta_dirty = true;
    }
    smo::View V = h->view();
    V.max_iter = h->host_ctl->max_iter;
    // speculative rows (C rule, full cache, 8 staged rows in LDS)
    size_t shmem = h->shmem;
    if (h->opt.speculative_rows > 0 && h->full && h->use_lds && h->kind == SVMHIP_KIND_C &&
        h->ktype != PRECOMPUTED && 8 * (size_t)h->d * sizeof(float) <= kLdsQueryLimit) {
      V.spec = 1;
      shmem = 8 * (size_t)h->d * sizeof(float);
    }
    // full cache: each workgroup's row flags as an LDS bitset past the query
    // row for the launch (the lookup after every exchange is on the
    // iteration's critical path), while the grid stays co-resident
    if (h->full && !V.spec) {
      const size_t off = std::max<size_t>(4, (shmem / sizeof(float) + 3) / 4 * 4);
      const size_t need = off * sizeof(float) + (size_t)((h->l + 31) / 32) * sizeof(unsigned);
      int per_cu = 0;
      if (need <= kFlagLdsLimit && smo_persistent_capacity((int)h->kind, need, &per_cu) == hipSuccess &&
          per_cu * device_cus(h->dev) >= h->nblocks) {
        V.flag_lds = (int)off;
        shmem = need;
      }
    }
    V.qpre = h->full && !use_gram && h->dense && h->d <= smo::BLOCK && V.flag_lds > 0 ? 1 : 0;
    // the owned variables' vectors and row segments resident in LDS past
    // the query row and the flag bits (SVMHIP_RESIDENT=0 keeps them in
    // device memory; results are identical either way)
    {
      const char *e = getenv("SVMHIP_RESIDENT");
      const int nv = h->nvar == 2 * h->l ? 2 : 1;
      const size_t m = (size_t)((h->nslices + h->nblocks - 1) / h->nblocks) * smo::SLICE;
      const size_t off = std::max<size_t>(4, (shmem + 15) / 16 * 4);
      const size_t need = off * sizeof(float) + smo_res_bytes(nv, m);
      int per_cu = 0;
      if (!(e && e[0] == '0') && h->nvar > 0 && need <= (size_t)device_max_lds(h->dev) &&
          smo_persistent_capacity((int)h->kind, need, &per_cu) == hipSuccess &&
          per_cu * device_cus(h->dev) >= h->nblocks) {
        V.res_off = (int)off;
        V.res_m = (int)m;
        shmem = need;
      }
      h->stats.resident_vectors = V.res_off > 0 ? 1 : 0;
    }
    int chunk = h->opt.iters_per_launch > 0 ? h->opt.iters_per_launch : INT_MAX;
    h->ck(hipEventRecord(h->ev[0], h->st), "event");
    h->stats.gram_ms = 0;
    if (use_gram && !h->gram_done) {
      h->ck(hipEventRecord(h->ev[4], h->st), "event");
      smo::View Vg = h->view();
      h->ck(smo_launch_gram_f16(&Vg, h->d_xsplit, h->st), "Gram-tile kernel");
      h->ck(hipEventRecord(h->ev[5], h->st), "event");
      h->ck(hipEventSynchronize(h->ev[5]), "Gram-tile kernel");
      float gms = 0;
      (void)hipEventElapsedTime(&gms, h->ev[4], h->ev[5]);
      h->stats.gram_ms = gms;
      h->gram_done = true;
      h->gram_rows = h->l;
    }
    if (h->nvar > 0) {
      h->stats.kernel_ms = 0;
      h->stats.launches = 0;
      for (;;) {
        h->ck(hipEventRecord(h->ev[2], h->st), "event");
        h->ck(smo_launch_persistent(&V, chunk, shmem, h->st), "persistent SMO kernel");
        h->ck(hipEventRecord(h->ev[3], h->st), "event");
        h->read_ctl();
        float kms = 0;
        (void)hipEventElapsedTime(&kms, h->ev[2], h->ev[3]);
        h->stats.kernel_ms += kms;
        h->stats.launches += 1;
        if (h->host_ctl->err) h->fail(SVMHIP_ERR_TIMEOUT, "device wait timed out at iteration %d", h->host_ctl->iter);
        if (h->host_ctl->stop) break;
      }
    }
    h->ck(hipEventRecord(h->ev[1], h->st), "event");
    h->ck(hipStreamSynchronize(h->st), "solve");
    float ms = 0;
    (void)hipEventElapsedTime(&ms, h->ev[0], h->ev[1]);
    h->stats.solve_ms = ms;
    h->have_j = false;
    svmhip_status s = svmhip_finalize(h, nullptr);
    if (s != SVMHIP_OK) throw Failure{s};
    h->stats.optimal = h->host_ctl->stop == smo::STOP_OPTIMAL || h->nvar == 0;
    GUARD_END
  }();
  h->stats.device_allocs = h->n_allocs;
  if (stats) *stats = h->stats;
  return rc;
}

svmhip_status svmhip_get_stats(svmhip_solver *h, svmhip_stats *stats) {
  if (!h || !stats) return SVMHIP_ERR_ARG;
  h->stats.device_allocs = h->n_allocs;
  *stats = h->stats;
  return SVMHIP_OK;
}

svmhip_status svmhip_kernel_row(svmhip_solver *h, int i, float *out) {
  if (!h) return SVMHIP_ERR_ARG;
  GUARD_BEGIN
  if (i < 0 || i >= h->l || !out) h->fail(SVMHIP_ERR_ARG, "kernel_row: bad row %d", i);
  h->ck(hipSetDevice(h->dev), "hipSetDevice");
  if (!h->rbf_ready) {
    svmhip_status s = svmhip_setup_rbf(h, h->l);
    if (s != SVMHIP_OK) throw Failure{s};
  }
  if (h->gram_done) {  // Gram-tile mode: the row as the cache holds it
    h->ck(hipMemcpyAsync(out, h->d_rows + (long long)i * h->row_stride, sizeof(float) * h->l, hipMemcpyDeviceToHost,
                         h->st),
          "copy cached row");
    h->ck(hipStreamSynchronize(h->st), "kernel row");
    return SVMHIP_OK;
  }
  float *buf = h->dalloc<float>(h->l, "row buffer");
  smo::View V = h->view();
  h->ck(smo_launch_kernel_row(&V, i, buf, h->shmem, h->st), "kernel row");
  h->ck(hipMemcpyAsync(out, buf, sizeof(float) * h->l, hipMemcpyDeviceToHost, h->st), "copy row");
  h->ck(hipStreamSynchronize(h->st), "kernel row");
  h->dfree(buf);
  GUARD_END
}

void svmhip_destroy(svmhip_solver *h) { delete h; }

}  // extern "C"

// ------------------------------------------------------------ batched solves

struct svmhip_batch {
  svmhip_solver *g = nullptr;  // owns the uploaded problem and every allocation
  struct Rowset {
    int n, nslices;
    double width;            // mean stored entries per row (SELL slice widths)
    long long *slice_ptr;
    float *vals;
    int *cols, *len;
    double *xsq;
    int *grow;               // the base problem's row of each row of the set
  };
  std::vector<Rowset> rowsets;
  double param_cache_mb = 100;
  std::string err;
  // exact dot products of every pair of base rows (dense problems, when
  // they fit): a kernel row of any subproblem is then a gather of one row of
  // them instead of a pass over the subproblem's X
  double *dots = nullptr;
  long long dots_ld = 0;
  int dots_state = 0;        // 0 = not decided, 1 = built, -1 = not used
  float dots_ms = 0;
  // device buffers of the previous solve, kept for the next one (a batch is
  // usually solved again at the same sizes, and allocating and freeing the
  // row-cache arenas, most of the device memory, costs more than a solve);
  // they are released with the handle
  struct Kept {
    char *p = nullptr;
    size_t cap = 0;
  };
  enum { K_G, K_ALPHA, K_P, K_QD, K_Y, K_STATUS, K_CTL, K_VIEWS, K_ROWS, K_FLAGS, K_LRU, K_WORDS, K_X, K_ITEMS, K_N };
  Kept kept[K_N];
};

extern "C" {

const char *svmhip_batch_last_error(const svmhip_batch *b) {
  if (!b) return "null batch";
  return !b->err.empty() ? b->err.c_str() : (b->g ? b->g->err.c_str() : "");
}

svmhip_status svmhip_batch_create(const svm_problem *prob, const svm_parameter *param, const svmhip_options *opt,
                                  svmhip_batch **out) {
  if (!out) return SVMHIP_ERR_ARG;
  *out = nullptr;
  if (param && param->kernel_type == PRECOMPUTED) return SVMHIP_ERR_ARG;  // row sets keep no serials
  svmhip_solver *g = nullptr;
  svmhip_status st = svmhip_create(prob, param, SVMHIP_KIND_C, opt, &g);
  if (st != SVMHIP_OK) return st;
  st = svmhip_setup_rbf(g, prob->l);
  if (st != SVMHIP_OK) {
    svmhip_destroy(g);
    return st;
  }
  if (!g->use_lds) {
    fprintf(stderr, "svmhip_batch_create: %d features exceed the batched solver's LDS query row\n", g->d);
    svmhip_destroy(g);
    return SVMHIP_ERR_ARG;
  }
  svmhip_batch *b = new (std::nothrow) svmhip_batch;
  if (!b) {
    svmhip_destroy(g);
    return SVMHIP_ERR_NOMEM;
  }
  b->g = g;
  b->param_cache_mb = param->cache_size;
  *out = b;
  return SVMHIP_OK;
}

int svmhip_batch_add_rowset(svmhip_batch *b, const int *rows, int n) {
  if (!b || !rows || n <= 0) return -(int)SVMHIP_ERR_ARG;
  svmhip_solver *h = b->g;
  svmhip_status rc = [&]() -> svmhip_status {
    GUARD_BEGIN
    std::vector<char> seen(h->l, 0);
    for (int r = 0; r < n; ++r) {
      if (rows[r] < 0 || rows[r] >= h->l) h->fail(SVMHIP_ERR_ARG, "row set: index %d out of range", rows[r]);
      if (seen[rows[r]]) h->fail(SVMHIP_ERR_ARG, "row set: row %d repeated", rows[r]);
      seen[rows[r]] = 1;
    }
    svmhip_batch::Rowset rs;
    rs.n = n;
    rs.nslices = (n + smo::SLICE - 1) / smo::SLICE;
    std::vector<long long> sp(rs.nslices + 1, 0);
    for (int s = 0; s < rs.nslices; ++s) {
      int w = 0;
      for (int r = s * smo::SLICE; r < std::min(n, (s + 1) * smo::SLICE); ++r) w = std::max(w, h->host_row_len[rows[r]]);
      sp[s + 1] = sp[s] + (long long)w * smo::SLICE;
    }
    rs.width = (double)sp.back() / ((double)rs.nslices * smo::SLICE);
    rs.slice_ptr = h->dalloc<long long>(sp.size(), "row-set slice pointers");
    rs.vals = h->dalloc<float>(std::max<long long>(sp.back(), 1), "row-set values");
    rs.cols = h->dense ? nullptr : h->dalloc<int>(std::max<long long>(sp.back(), 1), "row-set columns");
    rs.len = h->dalloc<int>(n, "row-set lengths");
    rs.xsq = h->dalloc<double>(n, "row-set norms");
    int *d_rows = h->dalloc<int>(n, "row-set indices");
    h->upload(rs.slice_ptr, sp.data(), sp.size(), "upload row-set slices");
    h->upload(d_rows, rows, (size_t)n, "upload row-set indices");
    if (sp.back() > 0)
      h->ck(hipMemsetAsync(rs.vals, 0, sizeof(float) * sp.back(), h->st), "clear row-set values");
    smo::View V = h->view();
    h->ck(smo_launch_pack_rowset(&V, d_rows, n, rs.slice_ptr, rs.vals, rs.cols, rs.len, rs.xsq, h->st), "pack row set");
    h->ck(hipStreamSynchronize(h->st), "pack row set");
    rs.grow = d_rows;
    b->rowsets.push_back(rs);
    GUARD_END
  }();
  if (rc != SVMHIP_OK) return -(int)rc;
  return (int)b->rowsets.size() - 1;
}

svmhip_status svmhip_batch_solve(svmhip_batch *b, svmhip_subproblem *subs, int nsub, svmhip_stats *stats) {
  if (!b || (nsub > 0 && !subs)) return SVMHIP_ERR_ARG;
  svmhip_solver *h = b->g;
  svmhip_stats tot{};
  svmhip_status rc = [&]() -> svmhip_status {
    GUARD_BEGIN
    h->ck(hipSetDevice(h->dev), "hipSetDevice");
    long long nv_total = 0;
    int maxn = 1;
    for (int k = 0; k < nsub; ++k) {
      const svmhip_subproblem &sp = subs[k];
      if (sp.rowset < 0 || sp.rowset >= (int)b->rowsets.size()) h->fail(SVMHIP_ERR_ARG, "subproblem %d: bad row set", k);
      const int n = b->rowsets[sp.rowset].n;
      if (sp.n_var != n && sp.n_var != 2 * n) h->fail(SVMHIP_ERR_ARG, "subproblem %d: n_var must be n or 2n", k);
      if (sp.kind != SVMHIP_KIND_C && sp.kind != SVMHIP_KIND_NU) h->fail(SVMHIP_ERR_ARG, "subproblem %d: bad kind", k);
      if (!sp.y || !sp.p || !sp.alpha) h->fail(SVMHIP_ERR_ARG, "subproblem %d: null array", k);
      if (!(sp.Cp > 0) || !(sp.Cn > 0) || !(sp.eps > 0)) h->fail(SVMHIP_ERR_ARG, "subproblem %d: Cp, Cn, eps must be > 0", k);
      for (int v = 0; v < sp.n_var; ++v)
        if (sp.y[v] != 1 && sp.y[v] != -1) h->fail(SVMHIP_ERR_ARG, "subproblem %d: y must be +1/-1", k);
      nv_total += sp.n_var;
      maxn = std::max(maxn, n);
    }
    if (nsub == 0) return SVMHIP_OK;
    // the shared dot products, once per handle: dense rows, not disabled by
    // SVMHIP_SHARED_DOTS=0, and at most a quarter of the free device memory
    if (b->dots_state == 0) {
      b->dots_state = -1;
      const char *e = getenv("SVMHIP_SHARED_DOTS");
      const long long ld = ((long long)h->l + 7) / 8 * 8;
      const double bytes = (double)h->l * (double)ld * sizeof(double);
      size_t fb = 0, tb = 0;
      h->ck(hipMemGetInfo(&fb, &tb), "hipMemGetInfo");
      if (!(e && e[0] == '0') && h->dense && h->ktype != PRECOMPUTED && h->l >= 2 && bytes <= 0.25 * (double)fb) {
        b->dots = h->dalloc<double>((size_t)h->l * (size_t)ld, "shared dot products");
        b->dots_ld = ld;
        smo::View V = h->view();
        h->ck(hipEventRecord(h->ev[4], h->st), "event");
        h->ck(smo_launch_dots_f64(&V, b->dots, ld, h->st), "shared dot products");
        h->ck(hipEventRecord(h->ev[5], h->st), "event");
        h->ck(hipEventSynchronize(h->ev[5]), "shared dot products");
        (void)hipEventElapsedTime(&b->dots_ms, h->ev[4], h->ev[5]);
        b->dots_state = 1;
        tot.gram_ms = b->dots_ms;
        if (h->opt.verbose)
          fprintf(stderr, "[svmhip] batch: shared dot products of %d rows (%.1f GB) in %.2f ms\n", h->l, bytes / 1e9,
                  b->dots_ms);
      }
    }
    // pooled per-subproblem state, resident across the stages of the solve
    auto talloc = [&](int slot, size_t bytes, const char *what) {
      svmhip_batch::Kept &k = b->kept[slot];
      return (void *)h->ensure(k.p, k.cap, bytes, what);
    };
    double *dG = (double *)talloc(svmhip_batch::K_G, sizeof(double) * nv_total, "batch G");
    double *dA = (double *)talloc(svmhip_batch::K_ALPHA, sizeof(double) * nv_total, "batch alpha");
    double *dP = (double *)talloc(svmhip_batch::K_P, sizeof(double) * nv_total, "batch p");
    double *dQ = (double *)talloc(svmhip_batch::K_QD, sizeof(double) * nv_total, "batch QD");
    int8_t *dY = (int8_t *)talloc(svmhip_batch::K_Y, nv_total, "batch y");
    int8_t *dS = (int8_t *)talloc(svmhip_batch::K_STATUS, nv_total, "batch status");
    smo::Ctl *dC = (smo::Ctl *)talloc(svmhip_batch::K_CTL, sizeof(smo::Ctl) * nsub, "batch control");
    smo::View *dV = (smo::View *)talloc(svmhip_batch::K_VIEWS, sizeof(smo::View) * nsub, "batch views");
    std::vector<double> hA(nv_total), hP(nv_total);
    std::vector<int8_t> hY(nv_total), hS(nv_total);
    std::vector<smo::View> views(nsub);
    std::vector<int> true_max(nsub);
    long long off = 0;
    const smo::View base = h->view();
    for (int k = 0; k < nsub; ++k) {
      const svmhip_subproblem &sp = subs[k];
      const svmhip_batch::Rowset &rs = b->rowsets[sp.rowset];
      for (int v = 0; v < sp.n_var; ++v) {
        hA[off + v] = sp.alpha[v];
        hP[off + v] = sp.p[v];
        hY[off + v] = sp.y[v];
        double C = sp.y[v] > 0 ? sp.Cp : sp.Cn;
        hS[off + v] = sp.alpha[v] >= C ? smo::AT_UPPER : (sp.alpha[v] <= 0 ? smo::AT_LOWER : smo::IS_FREE);
      }
      smo::View V = base;
      V.l = rs.n; V.nslices = rs.nslices;
      V.slice_ptr = rs.slice_ptr; V.vals = rs.vals; V.cols = rs.cols; V.row_len = rs.len; V.xsq = rs.xsq;
      V.serial = nullptr;
      V.gamma = sp.gamma;
      V.nvar = sp.n_var; V.svr = sp.n_var == 2 * rs.n ? 1 : 0; V.kind = sp.kind;
      // over the subproblem's variables, as Solver::Solve (libsvm/svm.cpp:577)
      true_max[k] = h->opt.max_iter > 0 ? h->opt.max_iter
                                        : std::max(10000000, sp.n_var > INT_MAX / 100 ? INT_MAX : 100 * sp.n_var);
      V.max_iter = true_max[k];
      V.Cp = sp.Cp; V.Cn = sp.Cn; V.eps = sp.eps;
      V.G = dG + off; V.alpha = dA + off; V.p = dP + off; V.QD = dQ + off; V.y = dY + off; V.status = dS + off;
      V.rows = nullptr; V.S = 0; V.full = 0;
      V.slot_of_row = V.row_of_slot = V.lru_prev = V.lru_next = nullptr;
      V.computed = nullptr;
      V.ctl = dC + k; V.part = nullptr; V.xw = nullptr; V.xs_global = nullptr;
      V.use_lds = 1; V.nblocks = 1; V.rank = 0; V.resume = 0; V.iter0 = 0; V.xrep = 0;
      V.ginit_from = 0; V.ginit_to = sp.n_var;
      if (b->dots_state == 1) {
        V.gdot = b->dots;
        V.gdot_ld = b->dots_ld;
        V.grow = rs.grow;
      }
      views[k] = V;
      off += sp.n_var;
    }
    h->upload(dA, hA.data(), nv_total, "upload batch alpha");
    h->upload(dP, hP.data(), nv_total, "upload batch p");
    h->upload(dG, hP.data(), nv_total, "upload batch G");
    h->upload(dY, hY.data(), nv_total, "upload batch y");
    h->upload(dS, hS.data(), nv_total, "upload batch status");
    // one cache arena per resident workgroup
    const size_t lds0 = batch_lds_bytes(h->d);
    const int flag_words = batch_flag_words(h->d, maxn);
    const size_t lds = lds0 + (size_t)flag_words * sizeof(unsigned);
    int per_cu = 0;
    h->ck(smo_batch_capacity(lds, &per_cu), "occupancy query");
    const int G = std::max(1, per_cu) * device_cus(h->dev);
    const int K = G;
    // arena size: opt.cache_bytes if set, else an equal share of 70% of the
    // free device memory (kernel-row caching never changes a result, so the
    // LIBSVM cache_size, sized for host memory, is not applied here)
    long long maxstride = ((long long)maxn + 63) / 64 * 64;
    size_t freeb = 0, totb = 0;
    h->ck(hipMemGetInfo(&freeb, &totb), "hipMemGetInfo");
    svmhip_batch::Kept &kept_rows = b->kept[svmhip_batch::K_ROWS];
    freeb += kept_rows.cap;  // the previous solve's arenas count as free
    const double frac = h->opt.hbm_fraction > 0 ? std::min(h->opt.hbm_fraction, 0.95) : 0.7;
    const long long share = (long long)(frac * (double)freeb / sizeof(float)) / K;
    long long arena = h->opt.cache_bytes > 0 ? (long long)(h->opt.cache_bytes / sizeof(float)) : share;
    arena = std::max<long long>(std::min(arena, std::max<long long>(share, 2 * maxstride)), 2 * maxstride);
    // the default budget (kBatchArenaDefaultBytes in all)
    if (h->opt.cache_bytes <= 0 && h->opt.hbm_fraction <= 0)
      arena = std::max<long long>(std::min<long long>(arena, (long long)(kBatchArenaDefaultBytes / sizeof(float)) / K),
                                  2 * maxstride);
    // no subproblem needs more than its whole matrix
    arena = std::min<long long>(arena, std::max<long long>((long long)maxn * maxstride, 2 * maxstride));
    // within 5% of the kept arenas: keep them rather than allocate anew
    // (free memory moves a little between calls)
    const long long kept_arena = (long long)(kept_rows.cap / sizeof(float)) / K;
    if (kept_arena >= 2 * maxstride && kept_arena < arena && kept_arena >= arena - arena / 20) arena = kept_arena;
    if (h->opt.verbose)
      fprintf(stderr, "[svmhip] batch: %d kernel-row arenas of %.1f MB (%.1f GB of %.1f GB free, fraction %.2f)\n", K,
              arena * 4.0 / 1048576.0, arena * 4.0 * K / 1073741824.0, freeb / 1073741824.0, frac);
    smo::BatchPool pool;
    memset(&pool, 0, sizeof pool);
    pool.K = K;
    pool.arena_floats = arena;
    pool.rows = (float *)talloc(svmhip_batch::K_ROWS, sizeof(float) * (size_t)arena * K, "batch row-cache arenas");
    pool.flag_stride = maxn;
    pool.flags = (int8_t *)talloc(svmhip_batch::K_FLAGS, (size_t)maxn * K, "batch cache flags");
    pool.lru_stride = 4LL * maxn + 2;
    pool.lru = (int *)talloc(svmhip_batch::K_LRU, sizeof(int) * (size_t)pool.lru_stride * K, "batch LRU bookkeeping");
    // queue, arena counter, launch start and idle count (each cleared before
    // every launch)
    char *words = (char *)talloc(svmhip_batch::K_WORDS, 32, "batch counters");
    pool.queue = (int *)words;
    pool.arena_next = (int *)(words + 4);
    pool.t0 = (unsigned long long *)(words + 8);
    unsigned *dIdle = (unsigned *)(words + 16);
    pool.lds_floats = (int)(lds0 / sizeof(float));
    pool.flag_words = flag_words;
    pool.spec = h->opt.speculative_rows >= 0 ? 1 : 0;
    smo::XWord *dX =
        (smo::XWord *)talloc(svmhip_batch::K_X, sizeof(smo::XWord) * smo::xrec_words(G), "team exchange records");
    // items: one per subproblem plus the extra members of teams (at most G)
    int *dItems = (int *)talloc(svmhip_batch::K_ITEMS, sizeof(int) * 2 * ((size_t)nsub + (size_t)G), "batch items");

    // Stages.  Every unfinished subproblem runs until it converges or, after
    // about tau of device time, pauses at an iteration boundary (team members
    // agree through their exchange); tau starts at 0.2 s and grows 3x each
    // stage.  A stage also ends once three quarters of its workgroups have
    // found the queue empty (the drain rule).  While more subproblems remain
    // than the grid has workgroups, each is solved by one workgroup (largest
    // first) except the few whose estimated work would outlast a
    // workgroup's share, which get small teams on up to half the grid; once
    // fewer remain, the grid is shared out as teams in proportion to their
    // work so far.  Where a solve pauses changes
    // nothing: a resumed subproblem continues the same sequence of working
    // sets, but its kernel-row cache starts empty, so stages are long; the
    // first stage is the one that learns which subproblems run long
    // (measured on 900 solves of 3k-14k rows, one box, all 900 in 1.3-1.4 s:
    // first stages of 0.1 / 0.15 / 0.2 / 0.3 / 0.4 / 0.6 s, 4x growth, took
    // 1.42 / 1.30 / 1.31 / 1.36 / 1.37 / 1.46 s; 0.2 s with 3x growth
    // 1.30 s, with 6x 1.31 s).
    const int team_max = h->opt.grid_blocks > 0 ? std::min(h->opt.grid_blocks, 64) : 32;
    const int fixed_chunk = h->opt.iters_per_launch;
    std::vector<int> rem(nsub), iters(nsub, 0);
    std::vector<char> started(nsub, 0);
    bool all_optimal = true;
    for (int k = 0; k < nsub; ++k) rem[k] = k;
    std::vector<smo::Ctl> hC(nsub);
    std::vector<int> T(nsub, 1), items;
    double tau = 0.2;
    // team weights: with fewer subproblems than workgroups, n x (iterations
    // so far + 1); with more, an estimate of a subproblem's whole work,
    // n^1.5 x (iterations so far + n) (a one-workgroup iteration costs about
    // n, and larger problems take more iterations)
    auto work = [&](int k, bool fewer) {
      const double n = (double)std::max(subs[k].n_var, 1);
      return fewer ? n * (iters[k] + 1.0) : n * std::sqrt(n) * ((double)iters[k] + n);
    };
    const double clock_hz = device_clock_hz(h->dev);
    int stages = 0;
    float total_ms = 0;
    while (!rem.empty()) {
      // team sizes
      // team shares in proportion to n_var x (iterations so far + 1): the
      // subproblems that have run longest are the likeliest to run longest
      const bool fewer = (int)rem.size() < G;
      double wsum = 0;
      for (int k : rem) wsum += work(k, fewer);
      int tsum = 0;
      for (int k : rem) {
        const svmhip_batch::Rowset &rs = b->rowsets[subs[k].rowset];
        int t = 1;
        if (team_max > 1) {
          // fewer subproblems than workgroups: the grid is shared out in
          // proportion to the weights; more: a subproblem whose estimated
          // work is more than about 1.5 workgroups' share of the whole gets
          // a small team (up to 8), the rest run one workgroup each
          const double share = (double)G * work(k, fewer) / wsum;
          t = fewer ? (int)share : (int)(share + 0.5);
          t = std::max(1, std::min({t, fewer ? team_max : std::min(team_max, 8), rs.nslices / 2}));
        }
        T[k] = t;
        tsum += t;
      }
      // teams may hold the whole grid when every subproblem fits in it, and
      // half the grid otherwise (the rest drains the single-workgroup items):
      // take members back from the largest teams
      auto team_members = [&]() {
        int m = 0;
        for (int k : rem)
          if (T[k] > 1) m += T[k];
        return m;
      };
      const int team_cap = fewer ? G : G / 2;
      while ((fewer ? tsum : team_members()) > team_cap) {
        int big = -1;
        for (int k : rem)
          if (T[k] > 1 && (big < 0 || T[k] > T[big])) big = k;
        if (big < 0) break;
        --T[big];
        --tsum;
      }
      const int team_items = team_members();
      if (team_items > G) h->fail(SVMHIP_ERR_STATE, "team plan exceeds the grid");
      // queue order: teams (largest first), then single workgroups (largest first)
      std::vector<int> ord = rem;
      std::stable_sort(ord.begin(), ord.end(), [&](int a, int c) {
        if ((T[a] > 1) != (T[c] > 1)) return T[a] > 1;
        return subs[a].n_var > subs[c].n_var;
      });
      items.clear();
      int xoff = 0;
      for (int k : ord) {
        smo::View &V = views[k];
        V.nblocks = T[k];
        V.xw = T[k] > 1 ? dX + xoff : nullptr;
        if (T[k] > 1) xoff += (int)smo::xrec_words(T[k]);
        V.resume = started[k];
        V.iter0 = iters[k];
        if (fixed_chunk > 0) {  // pause after a fixed number of iterations
          V.max_iter = (int)std::min<long long>((long long)iters[k] + fixed_chunk, true_max[k]);
          V.pause_ticks = 0;
        } else {  // pause on the device clock about tau after the launch starts
          V.max_iter = true_max[k];
          V.pause_ticks = (long long)(tau * clock_hz);
        }
        for (int r = 0; r < T[k]; ++r) {
          items.push_back(k);
          items.push_back(r);
        }
      }
      const int nitems = (int)items.size() / 2;
      pool.team_items = team_items;
      h->ck(hipMemsetAsync(pool.arena_next, 0, sizeof(int), h->st), "clear batch arena counter");
      h->upload(dV, views.data(), nsub, "upload batch views");
      if (items.size() > 2 * ((size_t)nsub + (size_t)G)) h->fail(SVMHIP_ERR_STATE, "batch item list overflows");
      h->upload(dItems, items.data(), items.size(), "upload batch items");
      h->ck(hipMemsetAsync(pool.queue, 0, sizeof(int), h->st), "clear batch queue");
      h->ck(hipMemsetAsync(pool.t0, 0, sizeof(unsigned long long), h->st), "clear launch start");
      pool.deadline_ticks = fixed_chunk > 0 ? 0 : (long long)(tau * clock_hz);
      // drain rule (not with fixed chunks): once three quarters of the
      // launch's workgroups find the queue empty, the solves still running
      // pause and the next stage gives them the idle workgroups as teams
      pool.idle = fixed_chunk > 0 ? nullptr : dIdle;
      pool.drain_at = std::max(1, std::min(G, nitems) * 3 / 4);
      if (pool.idle) h->ck(hipMemsetAsync(pool.idle, 0, sizeof(unsigned), h->st), "clear batch idle count");
      if (xoff > 0) h->ck(hipMemsetAsync(dX, 0, sizeof(smo::XWord) * xoff, h->st), "clear team records");
      h->ck(hipMemsetAsync(dC, 0, sizeof(smo::Ctl) * nsub, h->st), "clear batch control");
      h->ck(hipEventRecord(h->ev[0], h->st), "event");
      h->ck(smo_launch_batch(dV, dItems, nitems, std::min(G, nitems), &pool, lds, h->st), "batched SMO kernel");
      h->ck(hipEventRecord(h->ev[1], h->st), "event");
      h->ck(hipMemcpyAsync(hC.data(), dC, sizeof(smo::Ctl) * nsub, hipMemcpyDeviceToHost, h->st), "fetch batch control");
      h->ck(hipStreamSynchronize(h->st), "batched solve");
      float ms = 0;
      (void)hipEventElapsedTime(&ms, h->ev[0], h->ev[1]);
      total_ms += ms;
      ++stages;
      std::vector<int> next;
      for (int k : rem) {
        const smo::Ctl &c = hC[k];
        if (c.stop == smo::STOP_NONE && !c.err) {  // left for the next launch
          next.push_back(k);
          continue;
        }
        started[k] = 1;
        if (c.err) h->fail(SVMHIP_ERR_TIMEOUT, "subproblem %d: a team member did not respond", k);
        iters[k] = c.iter;
        tot.row_hits += (long long)c.hits;
        tot.row_misses += (long long)c.misses;
        svmhip_subproblem &sp = subs[k];
        sp.rho = c.rho;
        sp.obj = c.obj;
        sp.r = c.r;
        sp.iters = c.iter;
        if (c.stop == smo::STOP_OPTIMAL) {
          sp.status = SVMHIP_OK;
        } else if (c.iter >= true_max[k]) {
          sp.status = SVMHIP_ERR_MAXITER;
          all_optimal = false;
        } else {
          next.push_back(k);
        }
      }
      if (h->opt.verbose) {
        int nteam = 0, tmax = 1;
        for (int k : rem)
          if (T[k] > 1) { ++nteam; tmax = std::max(tmax, T[k]); }
        fprintf(stderr, "[svmhip] batch stage %d: %zu subproblems (%d teams, largest %d), %d items, %.1f ms, %zu left\n",
                stages, rem.size(), nteam, tmax, nitems, ms, next.size());
      }
      rem.swap(next);
      tau *= 3;
    }
    h->ck(hipMemcpyAsync(hA.data(), dA, sizeof(double) * nv_total, hipMemcpyDeviceToHost, h->st), "fetch batch alpha");
    h->ck(hipStreamSynchronize(h->st), "fetch batch alpha");
    tot.solve_ms = total_ms;
    tot.kernel_ms = total_ms;
    tot.launches = stages;
    tot.optimal = all_optimal ? 1 : 0;
    off = 0;
    for (int k = 0; k < nsub; ++k) {
      svmhip_subproblem &sp = subs[k];
      for (int v = 0; v < sp.n_var; ++v) sp.alpha[v] = hA[off + v];
      off += sp.n_var;
      tot.iters += sp.iters;
    }
    tot.cache_slots = (int)std::min<long long>(INT_MAX, arena / maxstride);
    tot.grid_blocks = G;
    tot.shared_dots = b->dots_state == 1 ? 1 : 0;
    tot.device_allocs = h->n_allocs;
    GUARD_END
  }();
  if (stats) *stats = tot;
  return rc;
}

void svmhip_batch_destroy(svmhip_batch *b) {
  if (!b) return;
  svmhip_destroy(b->g);
  delete b;
}

}  // extern "C"
