// Batched prediction on the device: svmhip_predict_batch, the batched form of
// svm_predict_values (libsvm/svm.cpp:2569-2643, which the reference runs on
// the CPU only), and svmhip_predict_batch_noalloc, the same computation
// staged in a library-owned device scratch area instead of allocations.
//
// Both pack the model's support vectors and the test rows on the host in the
// SELL-64 layout with the caller's FP64 feature values (svm_predict_values
// works on the svm_node doubles; only training truncates features to float,
// as the reference uploads them), copy them to device memory, run k_predict
// and copy the decision values back.  The two differ only in where the
// device memory comes from, so their results are identical.
#include <memory>
#include <mutex>
#include <thread>

#include "runtime_internal.hpp"

namespace {

// Rows in SELL-64 with FP64 values (host side), as smo_device.hpp lays them
// out: element p of row 64s+lane at vals[sp[s] + 64p + lane].
struct Sell64 {
  int l = 0, d = 1, nslices = 0, dense = 0;
  std::vector<long long> sp;     // nslices + 1
  std::unique_ptr<double[]> vals;
  std::unique_ptr<int[]> cols;   // null when dense
  std::vector<int> rl;           // stored entries per row
};

bool pack_sell64(const svm_node *const *x, int n, Sell64 &o, std::string &err) {
  std::vector<int> len(std::max(n, 0));
  for (int r = 0; r < n; ++r)
    if (!x[r]) {
      err = "row " + std::to_string(r) + " is null";
      return false;
    }
  // row lengths and index checks, in parallel over rows
  std::vector<int> bad(n > 0 ? 1 : 0, -1);
  std::mutex mu;
  long long nnz = 0;
  int maxidx = -1;
  par_ranges(n, 64.0, [&](int r0, int r1) {
    long long z = 0;
    int mx = -1, b = -1;
    for (int r = r0; r < r1 && b < 0; ++r) {
      const svm_node *p = x[r];
      int k = 0, prev = INT_MIN;
      for (; p[k].index != -1; ++k) {
        if (p[k].index < 0 || p[k].index <= prev) {
          b = r;
          break;
        }
        prev = p[k].index;
      }
      mx = std::max(mx, prev);
      len[r] = k;
      z += k;
    }
    std::lock_guard<std::mutex> g(mu);
    nnz += z;
    maxidx = std::max(maxidx, mx);
    if (b >= 0 && (bad[0] < 0 || b < bad[0])) bad[0] = b;
  });
  if (n > 0 && bad[0] >= 0) {
    err = "row " + std::to_string(bad[0]) + ": feature indices must be >= 0 and ascending";
    return false;
  }
  o.l = n;
  o.d = std::max(maxidx + 1, 1);
  o.dense = (n > 0 && (double)nnz >= 0.5 * (double)n * (double)o.d) ? 1 : 0;
  o.nslices = (n + smo::SLICE - 1) / smo::SLICE;
  o.sp.assign(o.nslices + 1, 0);
  o.rl.assign(n, 0);
  for (int s = 0; s < o.nslices; ++s) {
    int w = 0;
    for (int r = s * smo::SLICE; r < std::min(n, (s + 1) * smo::SLICE); ++r) {
      o.rl[r] = o.dense ? o.d : len[r];
      w = std::max(w, o.rl[r]);
    }
    o.sp[s + 1] = o.sp[s] + (long long)w * smo::SLICE;
  }
  const long long total = std::max<long long>(o.sp[o.nslices], 1);
  o.vals.reset(new double[total]);
  o.cols.reset(o.dense ? nullptr : new int[total]);
  if (o.sp[o.nslices] == 0) o.vals[0] = 0.0;
  // each thread zeroes and fills whole slices
  const double per_slice = o.nslices ? (double)o.sp[o.nslices] / o.nslices : 0.0;
  par_ranges(o.nslices, per_slice, [&](int s0, int s1) {
    std::fill(o.vals.get() + o.sp[s0], o.vals.get() + o.sp[s1], 0.0);
    if (!o.dense) std::fill(o.cols.get() + o.sp[s0], o.cols.get() + o.sp[s1], 0);
    for (int r = s0 * smo::SLICE; r < std::min(n, s1 * smo::SLICE); ++r) {
      const long long base = o.sp[r / smo::SLICE] + (r % smo::SLICE);
      for (int k = 0; k < len[r]; ++k) {
        if (o.dense) {
          o.vals[base + (long long)x[r][k].index * smo::SLICE] = x[r][k].value;
        } else {
          o.vals[base + (long long)k * smo::SLICE] = x[r][k].value;
          o.cols[base + (long long)k * smo::SLICE] = x[r][k].index;
        }
      }
    }
  });
  return true;
}

// Device memory of one prediction: hipMalloc'd blocks freed at the end, or
// consecutive 256-byte aligned pieces of the no-alloc scratch area.
struct Arena {
  bool scratch = false;
  char *base = nullptr;
  size_t cap = 0, used = 0;
  std::vector<void *> owned;
  void *take(size_t bytes) {
    bytes = std::max<size_t>((bytes + 255) / 256 * 256, 256);
    if (scratch) {
      if (used + bytes > cap) return nullptr;
      void *p = base + used;
      used += bytes;
      return p;
    }
    void *p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
    owned.push_back(p);
    return p;
  }
  static size_t rounded(size_t bytes) { return std::max<size_t>((bytes + 255) / 256 * 256, 256); }
  ~Arena() {
    for (void *p : owned) (void)hipFree(p);
  }
};

struct PredictFail {
  svmhip_status code;
};

[[noreturn]] void pfail(svmhip_status code, const char *fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  fprintf(stderr, "svmhip_predict: %s\n", buf);
  throw PredictFail{code};
}

void pck(hipError_t e, const char *what) {
  if (e != hipSuccess) {
    (void)hipGetLastError();
    pfail(e == hipErrorOutOfMemory ? SVMHIP_ERR_NOMEM : SVMHIP_ERR_DEVICE, "%s: %s", what, hipGetErrorString(e));
  }
}

template <class T>
T *put(Arena &ar, const T *src, size_t n, const char *what) {
  T *p = (T *)ar.take(std::max<size_t>(n, 1) * sizeof(T));
  if (!p)
    pfail(SVMHIP_ERR_NOMEM, "%s: %zu bytes do not fit the %s", what, std::max<size_t>(n, 1) * sizeof(T),
          ar.scratch ? "no-alloc scratch area" : "device");
  if (n) pck(hipMemcpy(p, src, n * sizeof(T), hipMemcpyHostToDevice), what);
  return p;
}

// Device bytes of rows [r0, r1) of a packed operand (slice-aligned r0).
size_t rows_bytes(const Sell64 &R, int s0, int s1) {
  const size_t vals = (size_t)(R.sp[s1] - R.sp[s0]);
  const int r0 = s0 * smo::SLICE, r1 = std::min(R.l, s1 * smo::SLICE);
  return Arena::rounded(8 * (size_t)(s1 - s0 + 1)) + Arena::rounded(8 * vals) +
         (R.dense ? 0 : Arena::rounded(4 * vals)) + Arena::rounded(4 * (size_t)(r1 - r0));
}

// Uploads slices [s0, s1) of R and describes them as one PRows operand.
smo::PRows upload_rows(Arena &ar, const Sell64 &R, int s0, int s1) {
  const int r0 = s0 * smo::SLICE, r1 = std::min(R.l, s1 * smo::SLICE);
  std::vector<long long> sp(s1 - s0 + 1);
  for (int s = s0; s <= s1; ++s) sp[s - s0] = R.sp[s] - R.sp[s0];
  const size_t nv = (size_t)(R.sp[s1] - R.sp[s0]);
  smo::PRows P;
  memset(&P, 0, sizeof P);
  P.l = r1 - r0;
  P.d = R.d;
  P.nslices = s1 - s0;
  P.dense = R.dense;
  P.f64 = 1;
  P.slice_ptr = put(ar, sp.data(), sp.size(), "slice pointers");
  P.vals = put(ar, R.vals.get() + R.sp[s0], nv, "feature values");
  P.cols = R.dense ? nullptr : put(ar, R.cols.get() + R.sp[s0], nv, "feature columns");
  P.row_len = put(ar, R.rl.data() + r0, (size_t)(r1 - r0), "row lengths");
  return P;
}

// The prediction both entry points run; ar provides the device memory.
svmhip_status predict_run(const svm_model *model, const svm_node *const *x, int n, const svmhip_options *opt,
                          double *dec_values, double *labels, Arena &ar) {
  if (!model || (n > 0 && (!x || !dec_values))) return SVMHIP_ERR_ARG;
  if (model->param.kernel_type == PRECOMPUTED) return SVMHIP_ERR_ARG;
  const int st_type = model->param.svm_type;
  const bool reg = st_type == ONE_CLASS || st_type == EPSILON_SVR || st_type == NU_SVR;
  const int k = model->nr_class;
  const int npair = reg ? 1 : k * (k - 1) / 2;
  if (n == 0) return SVMHIP_OK;
  const int nsv = model->l;
  try {
    if (nsv == 0) {  // no support vectors: every decision value is -rho
      for (int t = 0; t < n; ++t)
        for (int q = 0; q < npair; ++q) dec_values[(long long)t * npair + q] = -model->rho[q];
    } else {
      if (svmhip_device_count() <= 0) pfail(SVMHIP_ERR_DEVICE, "no HIP device available");
      const int dev = opt ? opt->device : 0;
      pck(hipSetDevice(dev), "hipSetDevice");
      Sell64 S, T;
      std::string err;
      if (!pack_sell64(model->SV, nsv, S, err)) pfail(SVMHIP_ERR_ARG, "support vectors: %s", err.c_str());
      if (!pack_sell64(x, n, T, err)) pfail(SVMHIP_ERR_ARG, "test rows: %s", err.c_str());
      const int ld = std::max(S.d, T.d);
      // dynamic LDS: the device's per-workgroup limit less the kernel's own
      const int tb = smo_predict_tb(ld, (size_t)device_max_lds(dev) - 1024, nullptr);
      if (tb < 1) pfail(SVMHIP_ERR_ARG, "rows of %d features are too wide for the prediction kernel", ld);
      const int ncoef = reg ? 1 : k - 1;
      std::vector<double> coef((size_t)ncoef * nsv);
      for (int c = 0; c < ncoef; ++c)
        for (int s = 0; s < nsv; ++s) coef[(size_t)c * nsv + s] = model->sv_coef[c][s];
      std::vector<int> cstart(std::max(k, 1), 0), ccount(std::max(k, 1), 0);
      if (!reg) {
        for (int c = 0; c < k; ++c) ccount[c] = model->nSV[c];
        for (int c = 1; c < k; ++c) cstart[c] = cstart[c - 1] + ccount[c - 1];
      }
      smo::PredictArgs A;
      memset(&A, 0, sizeof A);
      A.sv = upload_rows(ar, S, 0, S.nslices);
      A.coef = put(ar, coef.data(), coef.size(), "coefficients");
      A.rho = put(ar, model->rho, (size_t)npair, "rho");
      A.cls_start = put(ar, cstart.data(), cstart.size(), "class starts");
      A.cls_count = put(ar, ccount.data(), ccount.size(), "class counts");
      A.ktype = model->param.kernel_type;
      A.degree = model->param.degree;
      A.gamma = model->param.gamma;
      A.coef0 = model->param.coef0;
      A.k = k;
      A.npair = npair;
      A.tb = tb;
      A.regression_like = reg ? 1 : 0;
      A.ld = ld;
      // mfma_rows = 1: models with one coefficient per support vector on
      // Gram tiles (FP16 hi/lo split, FP32 accumulation), when memory is
      // allocated (the no-alloc entry point always runs the FP64 kernel)
      const bool mfma = opt && opt->mfma_rows == 1 && !ar.scratch && (reg || k == 2);
      _Float16 *sv_split = nullptr;
      double *sv_sq = nullptr;
      int *dbad = nullptr;
      if (mfma) {
        dbad = (int *)ar.take(sizeof(int));
        if (!dbad) pfail(SVMHIP_ERR_NOMEM, "range flag does not fit");
        pck(hipMemset(dbad, 0, sizeof(int)), "clear range flag");
      }
      // test rows in chunks of whole slices: all of them when memory is
      // allocated, as many as the scratch area has room for otherwise
      const size_t mark = ar.used;
      for (int s0 = 0; s0 < T.nslices;) {
        int s1 = T.nslices;
        if (ar.scratch) {
          auto need = [&](int e) {
            const int rows = std::min(T.l, e * smo::SLICE) - s0 * smo::SLICE;
            return rows_bytes(T, s0, e) + Arena::rounded(8 * (size_t)rows * npair);
          };
          const size_t room = ar.cap - mark;
          if (need(s0 + 1) > room)
            pfail(SVMHIP_ERR_NOMEM,
                  "the model (%d support vectors of %d features) leaves no room for one slice of test rows in "
                  "the %zu-byte no-alloc scratch area",
                  nsv, S.d, ar.cap);
          int lo = s0 + 1, hi = T.nslices;  // largest e with need(e) <= room
          while (lo < hi) {
            const int mid = (lo + hi + 1) / 2;
            if (need(mid) <= room) lo = mid;
            else hi = mid - 1;
          }
          s1 = lo;
          ar.used = mark;
        }
        const int r0 = s0 * smo::SLICE, r1 = std::min(T.l, s1 * smo::SLICE);
        A.tv = upload_rows(ar, T, s0, s1);
        A.dec = (double *)ar.take(sizeof(double) * (size_t)(r1 - r0) * npair);
        if (!A.dec) pfail(SVMHIP_ERR_NOMEM, "decision values do not fit");
        bool done = false;
        if (mfma) {  // Gram tiles on the matrix cores, unless a value is out of FP16's range
          const long long rh = smo_predict_split_halves(ld);
          auto split = [&](const smo::PRows &R, const char *what, double *&sq) -> _Float16 * {
            const size_t bytes = sizeof(_Float16) * (size_t)R.l * (size_t)rh;
            _Float16 *p = (_Float16 *)ar.take(bytes);
            sq = (double *)ar.take(sizeof(double) * (size_t)std::max(R.l, 1));
            if (!p || !sq) pfail(SVMHIP_ERR_NOMEM, "%s: FP16 split copy does not fit", what);
            pck(hipMemset(p, 0, bytes), "clear split copy");
            pck(smo_launch_predict_split(&R, p, sq, dbad, ld, 0), "split rows");
            return p;
          };
          if (!sv_split) sv_split = split(A.sv, "support vectors", sv_sq);
          double *tv_sq = nullptr;
          const _Float16 *tv_split = split(A.tv, "test rows", tv_sq);
          int bad = 0;
          pck(hipMemcpy(&bad, dbad, sizeof bad, hipMemcpyDeviceToHost), "range check");
          if (!bad) {
            double *partial = (double *)ar.take(sizeof(double) * (size_t)smo_predict_gram_tiles(nsv) * (r1 - r0));
            if (!partial) pfail(SVMHIP_ERR_NOMEM, "partial sums do not fit");
            pck(smo_launch_predict_gram(&A, tv_split, sv_split, tv_sq, sv_sq, partial, 0), "prediction Gram tiles");
            done = true;
          }
        }
        if (!done) pck(smo_launch_predict(&A, r1 - r0, device_cus(dev), 0), "prediction kernel");
        pck(hipMemcpy(dec_values + (long long)r0 * npair, A.dec, sizeof(double) * (size_t)(r1 - r0) * npair,
                      hipMemcpyDeviceToHost),
            "fetch decision values");
        s0 = s1;
      }
    }
  } catch (const PredictFail &f) {
    return f.code;
  } catch (const std::bad_alloc &) {
    fprintf(stderr, "svmhip_predict: host allocation failed\n");
    return SVMHIP_ERR_NOMEM;
  }
  if (labels) {
    std::vector<int> vote(std::max(k, 1));
    for (int t = 0; t < n; ++t) {
      const double *dv = dec_values + (long long)t * npair;
      if (reg) {
        labels[t] = st_type == ONE_CLASS ? (dv[0] > 0 ? 1 : -1) : dv[0];
        continue;
      }
      std::fill(vote.begin(), vote.end(), 0);
      int q = 0;
      for (int a = 0; a < k; ++a)
        for (int b = a + 1; b < k; ++b, ++q) {
          if (dv[q] > 0) ++vote[a];
          else ++vote[b];
        }
      int best = 0;
      for (int c = 1; c < k; ++c)
        if (vote[c] > vote[best]) best = c;
      labels[t] = model->label[best];
    }
  }
  return SVMHIP_OK;
}

// The no-alloc variant's device memory: one array per device the library
// runs on, held until the process ends.  Host code must not pass `scratch`
// itself to a kernel or copy (in host code it names a host-side stand-in):
// its device address, per device, comes from hipGetSymbolAddress.
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[SVMHIP_NOALLOC_SCRATCH_BYTES];
std::mutex g_scratch_mu;            // serialises the no-alloc calls: they share each device's array
void *g_scratch_ptr[kMaxDevices];   // device address of scratch on each device (null: not looked up yet)

}  // namespace

extern "C" {

svmhip_status svmhip_predict_batch(const svm_model *model, const svm_node *const *x, int n, const svmhip_options *opt,
                                   double *dec_values, double *labels) {
  Arena ar;
  return predict_run(model, x, n, opt, dec_values, labels, ar);
}

svmhip_status svmhip_predict_batch_noalloc(const svm_model *model, const svm_node *const *x, int n,
                                           const svmhip_options *opt, double *dec_values, double *labels) {
  const int dev = opt ? opt->device : 0;
  if (dev < 0 || dev >= kMaxDevices) return SVMHIP_ERR_ARG;
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev] && model && model->l > 0 && n > 0) {
    // once per device, with that device current
    if (svmhip_device_count() <= dev) return SVMHIP_ERR_DEVICE;
    if (hipSetDevice(dev) != hipSuccess) {
      (void)hipGetLastError();
      return SVMHIP_ERR_DEVICE;
    }
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      fprintf(stderr, "svmhip_predict_batch_noalloc: no device address for the scratch area: %s\n",
              hipGetErrorString(e));
      return SVMHIP_ERR_DEVICE;
    }
    g_scratch_ptr[dev] = p;
  }
  Arena ar;
  ar.scratch = true;
  ar.base = (char *)g_scratch_ptr[dev];
  ar.cap = SVMHIP_NOALLOC_SCRATCH_BYTES;
  return predict_run(model, x, n, opt, dec_values, labels, ar);
}

}  // extern "C"
