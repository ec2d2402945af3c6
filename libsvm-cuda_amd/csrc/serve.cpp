// Serving runtime: the svmhip_server_* C-ABI (include/svm_hip.h).
//
// A server holds requests on the device between rounds.  Train requests are
// binary svm_train_one solves (solve_c_svc / solve_nu_svc,
// libsvm/svm.cpp:1500-1584) over their rows in the order given; each owns its
// SELL rows, its FP64 solver state and a kernel-row cache arena that stay on
// the device until the solve ends, so the stages of the batched kernel
// (k_smo_batch) resume it with a warm cache.  A finished solve leaves a
// resident model: its support vectors packed into their own SELL arrays
// (k_pack_rowset) plus the signed coefficients, which predict requests
// evaluate with k_predict.  Admission is first come first served against an
// HBM byte budget.
#include <deque>

#include "runtime_internal.hpp"

namespace {

constexpr double kDefaultStageMs = 10.0;
constexpr double kDefaultCacheCap = 2.0 * 1073741824.0;  // bytes of kernel-row cache per solve

enum { REQ_QUEUED = 0, REQ_ACTIVE = 1, REQ_DONE = 2 };

// SELL rows of one request (or of a packed model) on the device.
struct DevRows {
  long long *sp = nullptr;
  float *vals = nullptr;
  int *cols = nullptr, *len = nullptr;
  double *xsq = nullptr;
  int l = 0, d = 1, nslices = 0, dense = 0;
  long long total = 0;
  double bytes = 0;
};

double sell_bytes(long long total, int l, int nslices, int dense) {
  return 8.0 * (nslices + 1) + 4.0 * (double)std::max<long long>(total, 1) * (dense ? 1 : 2) + 12.0 * std::max(l, 1);
}

struct Request {
  int id = 0, type = 0, state = REQ_QUEUED, status = SVMHIP_OK;
  double bytes = 0;       // device bytes held now
  double device_ms = 0;
  HostSell hs;            // packed rows until admission
  std::vector<int> rl;    // stored entries per row (for packing the model)
  DevRows rows;
  // train
  int svm_type = C_SVC, kind = SVMHIP_KIND_C, ktype = RBF, degree = 3;
  double gamma = 0, coef0 = 0, eps = 1e-3, Cp = 1, Cn = 1;
  double labels[2] = {0, 0};
  std::vector<signed char> y;
  std::vector<double> p, a0;
  int max_iter = 0, T = 1, Tcap = 1, iters = 0;
  int ginit_next = 0;     // G = p + Q alpha0 is done for variables below this
  bool started = false, cache_valid = false;
  long long hits = 0, misses = 0;
  double *G = nullptr, *alpha = nullptr, *pv = nullptr, *QD = nullptr;
  int8_t *yd = nullptr, *st = nullptr;
  float *arena = nullptr;
  long long row_stride = 0, pm_stride = 0;
  int S = 0, full = 0;
  int8_t *flags = nullptr;
  int *lru = nullptr;
  smo::XWord *xw = nullptr;
  double rho = 0, obj = 0, r = 0;
  std::vector<double> coef;  // signed, all rows
  int nSV = 0;
  // resident model (train, after the solve)
  DevRows model;
  double *mcoef = nullptr, *mrho = nullptr;
  int *mcs = nullptr, *mcc = nullptr;
  int mcount[2] = {0, 0};
  // predict
  int model_id = -1;
  double mlabels[2] = {0, 0};  // the model's two labels, copied at submission (the model may be released first)
  std::vector<double> dec;
  double *ddec = nullptr;
  int tb = 1;  // most test rows per workgroup of the prediction kernel
};

}  // namespace

struct svmhip_server {
  svmhip_solver *g = nullptr;  // device context: stream, events, allocation helpers
  double budget = 0, used = 0;
  int team_max = 64;
  double cache_cap = kDefaultCacheCap;
  std::vector<Request *> reqs;
  std::deque<int> queue;
  std::vector<int> active, predicts;
  std::vector<svmhip_completion> done;
  smo::View *dV = nullptr;
  int *dItems = nullptr, *dQueue = nullptr, *dFin = nullptr;
  unsigned long long *dT0 = nullptr;
  smo::Ctl *dCtl = nullptr;
  size_t capV = 0, capI = 0, capC = 0, capF = 0;
  // the round in flight: its stage's solves and its predictions
  bool inflight = false, stage_launched = false;
  std::vector<int> stage, stage_preds;
  std::vector<smo::View> h_views;  // host sources of the round's uploads, kept until it ends
  std::vector<int> h_items;
  svmhip_stats tot{};
  std::string err;
};

namespace {

void free_rows(svmhip_solver *h, DevRows &R) {
  for (void *p : {(void *)R.sp, (void *)R.vals, (void *)R.cols, (void *)R.len, (void *)R.xsq}) h->dfree(p);
  R = DevRows{};
}

// A resident model's or a predict request's rows as a prediction operand
// (float values: the rows as training uploads them).
smo::PRows prows(const DevRows &R) {
  smo::PRows P;
  memset(&P, 0, sizeof P);
  P.l = R.l; P.d = R.d; P.nslices = R.nslices; P.dense = R.dense; P.f64 = 0;
  P.slice_ptr = R.sp; P.vals = R.vals; P.cols = R.cols; P.row_len = R.len;
  return P;
}

smo::View rows_view(const svmhip_solver *h, const DevRows &R, const Request &q) {
  smo::View V;
  memset(&V, 0, sizeof V);
  V.l = R.l; V.d = R.d; V.nslices = R.nslices; V.dense = R.dense;
  V.ktype = q.ktype; V.degree = q.degree; V.gamma = q.gamma; V.coef0 = q.coef0;
  V.slice_ptr = R.sp; V.vals = R.vals; V.cols = R.cols; V.row_len = R.len; V.xsq = R.xsq;
  V.use_lds = 1; V.nblocks = 1; V.rank = -1;
  (void)h;
  return V;
}

void upload_rows(svmhip_solver *h, const HostSell &o, DevRows &R) {
  R.l = o.l; R.d = o.d; R.nslices = o.nslices; R.dense = o.dense; R.total = o.total;
  R.sp = h->dalloc<long long>(o.sp.size(), "request slice pointers");
  R.vals = h->dalloc<float>(o.vals.size(), "request values");
  R.len = h->dalloc<int>(std::max(o.l, 1), "request row lengths");
  R.xsq = h->dalloc<double>(std::max(o.l, 1), "request norms");
  h->upload(R.sp, o.sp.data(), o.sp.size(), "upload request slices");
  h->upload(R.vals, o.vals.data(), o.vals.size(), "upload request values");
  h->upload(R.len, o.rl.data(), o.rl.size(), "upload request row lengths");
  if (!o.dense) {
    R.cols = h->dalloc<int>(o.cols.size(), "request columns");
    h->upload(R.cols, o.cols.data(), o.cols.size(), "upload request columns");
  }
  R.bytes = sell_bytes(o.total, o.l, o.nslices, o.dense);
}

double train_footprint(const Request &q, double cache_cap, long long *arena_floats) {
  const HostSell &o = q.hs;
  const int l = o.l, nvar = (int)q.y.size();
  const long long stride = ((long long)l + 63) / 64 * 64;
  long long want = (long long)l * stride;
  long long cap = (long long)(cache_cap / sizeof(float));
  long long af = std::max<long long>(std::min(want, cap), 2 * stride);
  *arena_floats = af;
  const long long S = std::min<long long>(af / stride, l);
  const double meta = S >= l ? (double)q.Tcap * l : 4.0 * q.Tcap * (l + 3 * S + 2);
  return sell_bytes(o.total, l, o.nslices, o.dense) + 34.0 * nvar + 4.0 * af + meta +
         (double)smo::xrec_words(q.Tcap) * sizeof(smo::XWord) + sizeof(smo::Ctl);
}

// Estimated device seconds of one SMO iteration of q on one workgroup: two
// passes over the variables (about 44 bytes each) plus, per kernel-row miss,
// every stored entry of the rows (4 bytes dense, 8 sparse), at the ~40-60
// GB/s one workgroup streams; the miss rate is the one measured so far.
double est_solo_s(const Request &q) {
  const double width = q.rows.nslices > 0 ? (double)q.rows.total / ((double)q.rows.nslices * smo::SLICE) : 1.0;
  const double miss = q.iters > 0 ? std::max(0.05, std::min(2.0, (double)q.misses / q.iters)) : 1.0;
  const double row_bytes = width * (q.rows.dense ? 4.0 : 8.0);
  return (double)q.y.size() * (0.7e-9 + miss * row_bytes / 40e9) + 2e-6;
}

// Two exchanges per iteration make a team worth forming only above this.
constexpr double kTeamWorthS = 18e-6;


void complete(svmhip_server *s, Request &q, int status) {
  q.state = REQ_DONE;
  q.status = status;
  svmhip_completion c;
  memset(&c, 0, sizeof c);
  c.id = q.id;
  c.type = q.type;
  c.status = status;
  c.n = q.type == 0 ? (int)q.y.size() : (int)q.dec.size();
  c.iters = q.iters;
  c.nSV = q.nSV;
  c.rho = q.rho;
  c.obj = q.obj;
  c.r = q.r;
  c.device_ms = q.device_ms;
  s->done.push_back(c);
}

// Frees the solver state of a train request (keeps rows and model).
void free_train_state(svmhip_solver *h, Request &q) {
  for (void *p : {(void *)q.G, (void *)q.alpha, (void *)q.pv, (void *)q.QD, (void *)q.yd, (void *)q.st,
                  (void *)q.arena, (void *)q.flags, (void *)q.lru, (void *)q.xw})
    h->dfree(p);
  q.G = q.alpha = q.pv = q.QD = nullptr;
  q.yd = q.st = nullptr;
  q.arena = nullptr;
  q.flags = nullptr;
  q.lru = nullptr;
  q.xw = nullptr;
}

void admit_train(svmhip_server *s, Request &q) {
  svmhip_solver *h = s->g;
  long long af = 0;
  q.bytes = train_footprint(q, s->cache_cap, &af);
  upload_rows(h, q.hs, q.rows);
  const int l = q.hs.l, nvar = (int)q.y.size();
  q.rl = q.hs.rl;
  q.G = h->dalloc<double>(nvar, "request G");
  q.alpha = h->dalloc<double>(nvar, "request alpha");
  q.pv = h->dalloc<double>(nvar, "request p");
  q.QD = h->dalloc<double>(nvar, "request QD");
  q.yd = h->dalloc<int8_t>(nvar, "request y");
  q.st = h->dalloc<int8_t>(nvar, "request status");
  std::vector<int8_t> st0(nvar);
  for (int v = 0; v < nvar; ++v) {
    const double C = q.y[v] > 0 ? q.Cp : q.Cn;
    st0[v] = q.a0[v] >= C ? smo::AT_UPPER : (q.a0[v] <= 0 ? smo::AT_LOWER : smo::IS_FREE);
  }
  h->upload(q.G, q.p.data(), nvar, "upload request G");
  h->upload(q.alpha, q.a0.data(), nvar, "upload request alpha");
  h->upload(q.pv, q.p.data(), nvar, "upload request p");
  h->upload(q.yd, (const int8_t *)q.y.data(), nvar, "upload request y");
  h->upload(q.st, st0.data(), nvar, "upload request status");
  q.row_stride = ((long long)l + 63) / 64 * 64;
  q.arena = h->dalloc<float>(af, "request row cache");
  q.S = (int)std::min<long long>(af / q.row_stride, l);
  q.S = std::max(q.S, 2);
  q.full = q.S >= l ? 1 : 0;
  if (q.full) {
    q.S = std::max(l, 1);
    q.flags = h->dalloc<int8_t>((size_t)q.Tcap * std::max(l, 1), "request cache flags");
    q.pm_stride = 0;
  } else {
    q.pm_stride = (long long)l + 3LL * q.S + 2;
    q.lru = h->dalloc<int>((size_t)q.Tcap * q.pm_stride, "request LRU bookkeeping");
  }
  if (q.Tcap > 1) q.xw = h->dalloc<smo::XWord>(smo::xrec_words(q.Tcap), "request exchange records");
  smo::View V = rows_view(h, q.rows, q);
  if (l > 0) h->ck(smo_launch_self_dot(&V, q.rows.xsq, h->st), "request x^2");
  // the host copies above read q.hs / st0: wait before they go
  h->ck(hipStreamSynchronize(h->st), "admit");
  q.hs = HostSell{};
  s->used += q.bytes;
  q.state = REQ_ACTIVE;
  s->active.push_back(q.id);
}

void admit_predict(svmhip_server *s, Request &q) {
  svmhip_solver *h = s->g;
  Request &m = *s->reqs[q.model_id];
  upload_rows(h, q.hs, q.rows);
  const int n = q.hs.l;
  const int d = std::max(q.rows.d, m.model.d);
  q.tb = smo_predict_tb(d, device_max_lds(h->dev) - 1024, nullptr);
  if (q.tb < 1) h->fail(SVMHIP_ERR_ARG, "server: prediction rows of %d features are too wide", d);
  q.ddec = h->dalloc<double>(std::max(n, 1), "prediction output");
  q.bytes = q.rows.bytes + 8.0 * n;
  h->ck(hipStreamSynchronize(h->st), "admit");
  q.hs = HostSell{};
  s->used += q.bytes;
  q.state = REQ_ACTIVE;
  s->predicts.push_back(q.id);
}

double predict_footprint(const svmhip_server *s, const Request &q) {
  const Request &m = *s->reqs[q.model_id];
  (void)m;
  return sell_bytes(q.hs.total, q.hs.l, q.hs.nslices, q.hs.dense) + 8.0 * q.hs.l;
}

// Packs the support vectors (class +1 first, then -1, each in row order, as
// svm_train lays out a model, libsvm/svm.cpp:2359-2386) into the model.
void make_model(svmhip_server *s, Request &q) {
  svmhip_solver *h = s->g;
  const int l = q.rows.l;
  std::vector<int> sv;
  for (int pass = 0; pass < 2; ++pass) {
    const signed char want = pass == 0 ? +1 : -1;
    int c = 0;
    for (int i = 0; i < l; ++i)
      if (q.y[i] == want && q.coef[i] != 0) {
        sv.push_back(i);
        ++c;
      }
    q.mcount[pass] = c;
  }
  q.nSV = (int)sv.size();
  DevRows &M = q.model;
  M.l = q.nSV; M.d = q.rows.d; M.dense = q.rows.dense;
  M.nslices = (M.l + smo::SLICE - 1) / smo::SLICE;
  std::vector<long long> sp(M.nslices + 1, 0);
  for (int sl = 0; sl < M.nslices; ++sl) {
    int w = 0;
    for (int r = sl * smo::SLICE; r < std::min(M.l, (sl + 1) * smo::SLICE); ++r) w = std::max(w, q.rl[sv[r]]);
    sp[sl + 1] = sp[sl] + (long long)w * smo::SLICE;
  }
  M.total = sp[M.nslices];
  M.sp = h->dalloc<long long>(sp.size(), "model slices");
  M.vals = h->dalloc<float>(std::max<long long>(M.total, 1), "model values");
  M.cols = M.dense ? nullptr : h->dalloc<int>(std::max<long long>(M.total, 1), "model columns");
  M.len = h->dalloc<int>(std::max(M.l, 1), "model row lengths");
  M.xsq = h->dalloc<double>(std::max(M.l, 1), "model norms");
  M.bytes = sell_bytes(M.total, M.l, M.nslices, M.dense);
  std::vector<double> mc(std::max(q.nSV, 1), 0.0);
  for (int k = 0; k < q.nSV; ++k) mc[k] = q.coef[sv[k]];
  int *dsv = h->dalloc<int>(std::max(q.nSV, 1), "support vector indices");
  q.mcoef = h->dalloc<double>(mc.size(), "model coefficients");
  q.mrho = h->dalloc<double>(1, "model rho");
  q.mcs = h->dalloc<int>(2, "model class starts");
  q.mcc = h->dalloc<int>(2, "model class counts");
  int cs[2] = {0, q.mcount[0]};
  h->upload(M.sp, sp.data(), sp.size(), "upload model slices");
  h->upload(dsv, sv.data(), sv.size(), "upload support vector indices");
  h->upload(q.mcoef, mc.data(), mc.size(), "upload model coefficients");
  h->upload(q.mrho, &q.rho, 1, "upload model rho");
  h->upload(q.mcs, cs, 2, "upload model class starts");
  h->upload(q.mcc, q.mcount, 2, "upload model class counts");
  if (M.total > 0) h->ck(hipMemsetAsync(M.vals, 0, sizeof(float) * M.total, h->st), "clear model values");
  smo::View src = rows_view(h, q.rows, q);
  if (q.nSV > 0)
    h->ck(smo_launch_pack_rowset(&src, dsv, q.nSV, M.sp, M.vals, M.cols, M.len, M.xsq, h->st), "pack model");
  h->ck(hipStreamSynchronize(h->st), "pack model");  // host vectors die here
  h->dfree(dsv);
  free_rows(h, q.rows);
  s->used -= q.bytes;
  q.bytes = M.bytes + 8.0 * mc.size() + 8.0 + 16.0;
  s->used += q.bytes;
}

// A finished solve: rho, obj and r as k_finalize_views left them in c (FP64
// on the device, Solver::calculate_rho, libsvm/svm.cpp:1013-1049,
// Solver_NU::calculate_rho, :1260-1308, and the objective of :794-803), the
// signed coefficients from the final alpha, and the resident model.
void finish_train(svmhip_server *s, Request &q, const smo::Ctl &c) {
  svmhip_solver *h = s->g;
  const int nvar = (int)q.y.size();
  std::vector<double> a(nvar);
  h->ck(hipMemcpyAsync(a.data(), q.alpha, sizeof(double) * nvar, hipMemcpyDeviceToHost, h->st), "fetch alpha");
  h->ck(hipStreamSynchronize(h->st), "fetch alpha");
  q.rho = c.rho;
  q.obj = c.obj;
  q.r = c.r;
  q.coef.assign(nvar, 0.0);
  if (q.svm_type == NU_SVC) {  // libsvm/svm.cpp:1574-1580
    for (int i = 0; i < nvar; ++i) q.coef[i] = a[i] * q.y[i] / q.r;
    q.rho /= q.r;
    q.obj /= (q.r * q.r);
  } else {
    for (int i = 0; i < nvar; ++i) q.coef[i] = a[i] * q.y[i];
  }
  free_train_state(h, q);
  make_model(s, q);
  complete(s, q, c.stop == smo::STOP_OPTIMAL ? SVMHIP_OK : SVMHIP_ERR_MAXITER);
}

void run_predict(svmhip_server *s, Request &q) {
  svmhip_solver *h = s->g;
  Request &m = *s->reqs[q.model_id];
  const int n = q.rows.l;
  q.dec.assign(n, 0.0);
  if (m.model.l == 0) {  // no support vectors: every decision value is -rho
    for (int t = 0; t < n; ++t) q.dec[t] = -m.rho;
    return;
  }
  smo::PredictArgs A;
  memset(&A, 0, sizeof A);
  A.sv = prows(m.model);
  A.tv = prows(q.rows);
  A.ktype = m.ktype;
  A.degree = m.degree;
  A.gamma = m.gamma;
  A.coef0 = m.coef0;
  A.coef = m.mcoef;
  A.rho = m.mrho;
  A.cls_start = m.mcs;
  A.cls_count = m.mcc;
  A.k = 2;
  A.npair = 1;
  A.tb = q.tb;
  A.regression_like = 0;
  A.ld = std::max(q.rows.d, m.model.d);
  A.dec = q.ddec;
  h->ck(smo_launch_predict(&A, n, device_cus(h->dev), h->st), "prediction kernel");
}

// After the round's synchronisation: decision values to the host.
void finish_predict(svmhip_server *s, Request &q) {
  svmhip_solver *h = s->g;
  if (s->reqs[q.model_id]->model.l > 0 && !q.dec.empty())
    h->ck(hipMemcpy(q.dec.data(), q.ddec, sizeof(double) * q.dec.size(), hipMemcpyDeviceToHost), "fetch decisions");
  free_rows(h, q.rows);
  h->dfree(q.ddec);
  q.ddec = nullptr;
  s->used -= q.bytes;
  q.bytes = 0;
  complete(s, q, SVMHIP_OK);
}

template <class T>
T *grow(svmhip_solver *h, T *p, size_t &cap, size_t need, const char *what) {
  if (need <= cap && p) return p;
  h->dfree(p);
  cap = std::max(need, 2 * cap);
  return h->dalloc<T>(cap, what);
}

}  // namespace

extern "C" {

const char *svmhip_server_last_error(const svmhip_server *s) {
  if (!s) return "null server";
  return !s->err.empty() ? s->err.c_str() : (s->g ? s->g->err.c_str() : "");
}

svmhip_status svmhip_server_create(const svmhip_options *opt, double hbm_budget_bytes, svmhip_server **out) {
  if (!out) return SVMHIP_ERR_ARG;
  *out = nullptr;
  svm_problem p0{0, nullptr, nullptr};
  svm_parameter prm;
  memset(&prm, 0, sizeof prm);
  prm.cuda_flag = 1;
  prm.kernel_type = RBF;
  prm.eps = 1e-3;
  prm.cache_size = 100;
  svmhip_solver *g = nullptr;
  svmhip_status rc = svmhip_create(&p0, &prm, SVMHIP_KIND_C, opt, &g);
  if (rc != SVMHIP_OK) return rc;
  svmhip_server *s = new (std::nothrow) svmhip_server;
  if (!s) {
    svmhip_destroy(g);
    return SVMHIP_ERR_NOMEM;
  }
  s->g = g;
  svmhip_solver *h = g;
  rc = [&]() -> svmhip_status {
    GUARD_BEGIN
    size_t freeb = 0, totb = 0;
    h->ck(hipMemGetInfo(&freeb, &totb), "hipMemGetInfo");
    s->budget = hbm_budget_bytes > 0 ? hbm_budget_bytes : 0.8 * (double)freeb;
    // freed request buffers are kept for reuse, up to 2 GiB or a tenth of the budget
    h->pool_cap = (size_t)std::min(2.0 * 1024 * 1024 * 1024, 0.1 * s->budget);
    if (h->opt.grid_blocks > 0) s->team_max = std::min(h->opt.grid_blocks, 64);
    if (h->opt.cache_bytes > 0) s->cache_cap = h->opt.cache_bytes;
    s->dQueue = h->dalloc<int>(1, "server queue counter");
    s->dT0 = h->dalloc<unsigned long long>(1, "server launch start");
    GUARD_END
  }();
  if (rc != SVMHIP_OK) {
    svmhip_destroy(g);
    delete s;
    return rc;
  }
  *out = s;
  return SVMHIP_OK;
}

int svmhip_server_submit_train(svmhip_server *s, const svm_problem *prob, const svm_parameter *param) {
  if (!s || !prob || !param || prob->l < 2 || !prob->x || !prob->y) return -(int)SVMHIP_ERR_ARG;
  svmhip_solver *h = s->g;
  Request *q = new (std::nothrow) Request;
  if (!q) return -(int)SVMHIP_ERR_NOMEM;
  svmhip_status rc = [&]() -> svmhip_status {
    GUARD_BEGIN
    if (param->svm_type != C_SVC && param->svm_type != NU_SVC)
      h->fail(SVMHIP_ERR_ARG, "server: train requests are C_SVC or NU_SVC");
    if (param->kernel_type < LINEAR || param->kernel_type > SIGMOID)
      h->fail(SVMHIP_ERR_ARG, "server: kernel must be linear, poly, rbf or sigmoid");
    if (!(param->eps > 0)) h->fail(SVMHIP_ERR_ARG, "server: eps <= 0");
    const int l = prob->l;
    q->type = 0;
    q->svm_type = param->svm_type;
    q->kind = param->svm_type == NU_SVC ? SVMHIP_KIND_NU : SVMHIP_KIND_C;
    q->ktype = param->kernel_type;
    q->degree = param->degree;
    q->gamma = param->gamma;
    q->coef0 = param->coef0;
    q->eps = param->eps;
    // two labels, the first one seen is +1
    q->labels[0] = prob->y[0];
    int have2 = 0;
    for (int i = 1; i < l; ++i)
      if (prob->y[i] != q->labels[0]) {
        if (!have2) {
          q->labels[1] = prob->y[i];
          have2 = 1;
        } else if (prob->y[i] != q->labels[1]) {
          h->fail(SVMHIP_ERR_ARG, "server: train requests are binary (a third label %g)", prob->y[i]);
        }
      }
    if (!have2) h->fail(SVMHIP_ERR_ARG, "server: train requests need two labels");
    q->y.resize(l);
    int npos = 0;
    for (int i = 0; i < l; ++i) {
      q->y[i] = prob->y[i] == q->labels[0] ? +1 : -1;
      npos += q->y[i] > 0;
    }
    q->p.assign(l, 0.0);
    q->a0.assign(l, 0.0);
    if (q->svm_type == C_SVC) {  // libsvm/svm.cpp:1500-1533
      if (!(param->C > 0)) h->fail(SVMHIP_ERR_ARG, "server: C <= 0");
      for (int i = 0; i < l; ++i) q->p[i] = -1;
      q->Cp = q->Cn = param->C;
    } else {  // libsvm/svm.cpp:1535-1584
      const double nu = param->nu;
      if (!(nu > 0) || nu > 1) h->fail(SVMHIP_ERR_ARG, "nu <= 0 or nu > 1");
      if (nu * l / 2 > std::min(npos, l - npos)) h->fail(SVMHIP_ERR_ARG, "specified nu is infeasible");
      double sp = nu * l / 2, sn = nu * l / 2;
      for (int i = 0; i < l; ++i) {
        if (q->y[i] == +1) {
          q->a0[i] = std::min(1.0, sp);
          sp -= q->a0[i];
        } else {
          q->a0[i] = std::min(1.0, sn);
          sn -= q->a0[i];
        }
      }
      q->Cp = q->Cn = 1.0;
    }
    q->max_iter = h->opt.max_iter > 0 ? h->opt.max_iter : std::max(10000000, l > INT_MAX / 100 ? INT_MAX : 100 * l);
    q->ginit_next = 0;
    make_host_sell(h, prob, q->ktype, q->hs);
    if ((size_t)q->hs.d * sizeof(float) > kLdsQueryLimit)
      h->fail(SVMHIP_ERR_ARG, "server: %d features exceed the batched solver's LDS query row", q->hs.d);
    // largest team: a power of two, at least four 64-row slices per member
    q->Tcap = 1;
    while (2 * q->Tcap <= s->team_max && 2 * q->Tcap * 4 <= q->hs.nslices) q->Tcap *= 2;
    q->T = 1;
    GUARD_END
  }();
  if (rc != SVMHIP_OK) {
    delete q;
    return -(int)rc;
  }
  q->id = (int)s->reqs.size();
  s->reqs.push_back(q);
  s->queue.push_back(q->id);
  return q->id;
}

int svmhip_server_submit_predict(svmhip_server *s, int model_id, const svm_node *const *x, int n) {
  if (!s || !x || n <= 0) return -(int)SVMHIP_ERR_ARG;
  if (model_id < 0 || model_id >= (int)s->reqs.size() || !s->reqs[model_id]) return -(int)SVMHIP_ERR_ARG;
  Request &m = *s->reqs[model_id];
  if (m.type != 0 || m.state != REQ_DONE || (m.status != SVMHIP_OK && m.status != SVMHIP_ERR_MAXITER))
    return -(int)SVMHIP_ERR_STATE;
  svmhip_solver *h = s->g;
  Request *q = new (std::nothrow) Request;
  if (!q) return -(int)SVMHIP_ERR_NOMEM;
  svmhip_status rc = [&]() -> svmhip_status {
    GUARD_BEGIN
    std::vector<double> dummy(n, 0.0);
    svm_problem pt{n, dummy.data(), const_cast<svm_node **>(x)};
    q->type = 1;
    q->model_id = model_id;
    q->mlabels[0] = m.labels[0];
    q->mlabels[1] = m.labels[1];
    make_host_sell(h, &pt, m.ktype, q->hs);
    if (smo_predict_tb(std::max(q->hs.d, m.model.d), device_max_lds(h->dev) - 1024, nullptr) < 1)
      h->fail(SVMHIP_ERR_ARG, "server: prediction rows of %d features are too wide", std::max(q->hs.d, m.model.d));
    GUARD_END
  }();
  if (rc != SVMHIP_OK) {
    delete q;
    return -(int)rc;
  }
  q->id = (int)s->reqs.size();
  s->reqs.push_back(q);
  s->queue.push_back(q->id);
  return q->id;
}

svmhip_status svmhip_server_step(svmhip_server *s, double stage_ms, int *n_active, int *n_queued) {
  if (!s) return SVMHIP_ERR_ARG;
  svmhip_solver *h = s->g;
  return [&]() -> svmhip_status {
    GUARD_BEGIN
    h->ck(hipSetDevice(h->dev), "hipSetDevice");
    const double tau = (stage_ms > 0 ? stage_ms : kDefaultStageMs) * 1e-3;
    const double clock_hz = device_clock_hz(h->dev);
    // 1. results of the round in flight
    if (s->inflight) {
      h->ck(hipStreamSynchronize(h->st), "server round");
      s->inflight = false;
      if (s->stage_launched) {
        float ms = 0;
        (void)hipEventElapsedTime(&ms, h->ev[0], h->ev[1]);
        s->tot.solve_ms += ms;
        s->tot.launches += 1;
        std::vector<smo::Ctl> hc(s->stage.size());
        h->ck(hipMemcpy(hc.data(), s->dCtl, sizeof(smo::Ctl) * s->stage.size(), hipMemcpyDeviceToHost),
              "fetch server control");
        std::vector<int> still, fin;  // fin: stage positions of the solves that finished
        std::vector<char> ran(s->reqs.size(), 0);
        for (size_t k = 0; k < s->stage.size(); ++k) {
          Request &q = *s->reqs[s->stage[k]];
          ran[q.id] = 1;
          const smo::Ctl &c = hc[k];
          if (c.stop == smo::STOP_NONE && !c.err) {  // left for the next stage
            still.push_back(q.id);
            continue;
          }
          q.device_ms += ms;
          q.cache_valid = true;
          if (!q.started && c.stop == smo::STOP_PAUSE && !c.err && c.iter == q.iters &&
              c.ginit_next < (int)q.y.size() && c.ginit_next > q.ginit_next) {
            q.ginit_next = c.ginit_next;  // still initialising G
            still.push_back(q.id);
            continue;
          }
          q.started = true;
          s->tot.iters += c.iter - q.iters;
          q.iters = c.iter;
          q.hits += (long long)c.hits;
          q.misses += (long long)c.misses;
          s->tot.row_hits += (long long)c.hits;
          s->tot.row_misses += (long long)c.misses;
          if (c.err) {
            free_train_state(h, q);
            free_rows(h, q.rows);
            s->used -= q.bytes;
            q.bytes = 0;
            complete(s, q, SVMHIP_ERR_TIMEOUT);
          } else if (c.stop == smo::STOP_OPTIMAL || c.iter >= q.max_iter) {
            fin.push_back((int)k);
          } else {
            still.push_back(q.id);
          }
        }
        if (!fin.empty()) {  // rho / obj / r of the finished solves on the device
          s->dFin = grow(h, s->dFin, s->capF, fin.size(), "server finalize list");
          h->upload(s->dFin, fin.data(), fin.size(), "upload finalize list");
          h->ck(smo_launch_finalize_views(s->dV, s->dFin, (int)fin.size(), h->st), "finalize kernel");
          h->ck(hipMemcpyAsync(hc.data(), s->dCtl, sizeof(smo::Ctl) * s->stage.size(), hipMemcpyDeviceToHost, h->st),
                "fetch server control");
          h->ck(hipStreamSynchronize(h->st), "finalize");
          for (int k : fin) finish_train(s, *s->reqs[s->stage[k]], hc[k]);
        }
        for (int id : s->active)
          if (!ran[id]) still.push_back(id);
        s->active.swap(still);
      }
      for (int id : s->stage_preds) finish_predict(s, *s->reqs[id]);
      s->stage.clear();
      s->stage_preds.clear();
      s->stage_launched = false;
    }
    // 2. admission, first come first served
    while (!s->queue.empty()) {
      Request &q = *s->reqs[s->queue.front()];
      long long af = 0;
      const double need = q.type == 0 ? train_footprint(q, s->cache_cap, &af) : predict_footprint(s, q);
      if (need > s->budget) {
        s->queue.pop_front();
        q.hs = HostSell{};
        complete(s, q, SVMHIP_ERR_NOMEM);
        continue;
      }
      if (s->used + need > s->budget) break;
      s->queue.pop_front();
      if (q.type == 0) admit_train(s, q);
      else admit_predict(s, q);
    }
    // 3. predictions
    s->stage_preds.swap(s->predicts);
    for (int id : s->stage_preds) {
      Request &q = *s->reqs[id];
      q.dec.assign(q.rows.l, 0.0);
      if (s->reqs[q.model_id]->model.l == 0)  // no support vectors: every decision value is -rho
        for (double &v : q.dec) v = -s->reqs[q.model_id]->rho;
      else
        run_predict(s, q);
    }
    // 4. one stage of every admitted solve, launched without waiting: teams
    // first (as many as fit in the grid at once), then single workgroups,
    // oldest first
    if (!s->active.empty()) {
      int maxd = 1, maxl = 1;
      for (int id : s->active) {
        maxd = std::max(maxd, s->reqs[id]->rows.d);
        maxl = std::max(maxl, s->reqs[id]->rows.l);
      }
      const size_t shmem0 = batch_lds_bytes(maxd);
      const int flag_words = batch_flag_words(maxd, maxl);
      const size_t shmem = shmem0 + (size_t)flag_words * sizeof(unsigned);
      int per_cu = 0;
      h->ck(smo_batch_capacity(shmem, &per_cu), "occupancy query");
      const int G = std::max(1, per_cu) * device_cus(h->dev);
      // team sizes follow the load: each solve's share of the grid in
      // proportion to its estimated iteration cost, as a power of two up to
      // its cap; a size changes only on a 2x gain or a 4x excess (each
      // change restarts that solve's cache)
      double wsum = 0;
      for (int id : s->active) wsum += est_solo_s(*s->reqs[id]);
      for (int id : s->active) {
        Request &q = *s->reqs[id];
        const double share = (double)G * est_solo_s(q) / wsum;
        int t = 1;
        if (est_solo_s(q) > kTeamWorthS)
          while (2 * t <= q.Tcap && 2 * t <= share) t *= 2;
        int nt = q.T;
        if (!q.started || t >= 2 * q.T || 4 * t <= q.T) nt = t;
        nt = std::min(nt, std::max(1, G));
        if (nt != q.T) {
          q.T = nt;
          q.cache_valid = false;
        }
      }
      std::vector<int> &stage = s->stage;
      stage.clear();
      std::vector<int> teams, solos;
      for (int id : s->active) (s->reqs[id]->T > 1 ? teams : solos).push_back(id);
      std::stable_sort(teams.begin(), teams.end(), [&](int a, int b) { return a < b; });
      int tsum = 0;
      for (int id : teams) {
        Request &q = *s->reqs[id];
        if (tsum + q.T > G) continue;  // waits for a later stage
        tsum += q.T;
        stage.push_back(id);
      }
      // oldest first: a stage's deadline leaves the queue's tail for the next
      // stage, and that tail must not always be the same requests
      std::stable_sort(solos.begin(), solos.end(), [&](int a, int b) { return a < b; });
      stage.insert(stage.end(), solos.begin(), solos.end());
      std::vector<smo::View> &views = s->h_views;
      std::vector<int> &items = s->h_items;
      views.assign(stage.size(), smo::View{});
      items.clear();
      for (size_t k = 0; k < stage.size(); ++k) {
        Request &q = *s->reqs[stage[k]];
        smo::View V = rows_view(h, q.rows, q);
        V.nvar = (int)q.y.size();
        V.svr = 0;
        V.kind = q.kind;
        V.Cp = q.Cp; V.Cn = q.Cn; V.eps = q.eps;
        V.G = q.G; V.alpha = q.alpha; V.p = q.pv; V.QD = q.QD; V.y = q.yd; V.status = q.st;
        V.rows = q.arena; V.row_stride = q.row_stride; V.S = q.S; V.full = q.full;
        V.pm_flags = q.flags; V.pm_lru = q.lru; V.pm_stride = q.pm_stride;
        V.own_cache = 1;
        V.keep_cache = q.cache_valid ? 1 : 0;
        V.resume = q.started ? 1 : 0;
        V.iter0 = q.iters;
        if (!q.started) {
          // this stage's share of G = p + Q alpha0: about tau of staged
          // passes (each query row costs one multiply-add per stored entry
          // of the owned rows, ~7e10 per second per workgroup)
          const double width = (double)q.rows.total / std::max(1.0, (double)q.rows.nslices * smo::SLICE);
          const double row_s = (double)q.rows.l * width / q.T / 7e10 + 1e-7;
          long long want = std::max<long long>(16, (long long)(tau / row_s));
          int to = q.ginit_next;
          for (long long c = 0; to < V.nvar && c < want; ++to)
            if (q.a0[to] > 0) ++c;
          // the range ends just before the next nonzero alpha (or at nvar)
          while (to < V.nvar && !(q.a0[to] > 0)) ++to;
          V.ginit_from = q.ginit_next;
          V.ginit_to = to;
        }
        V.max_iter = q.max_iter;
        V.pause_ticks = std::max<long long>(1, (long long)(tau * clock_hz));
        V.nblocks = q.T;
        V.xw = q.xw;
        views[k] = V;
        for (int r = 0; r < q.T; ++r) {
          items.push_back((int)k);
          items.push_back(r);
        }
      }
      s->dV = grow(h, s->dV, s->capV, views.size(), "server views");
      s->dItems = grow(h, s->dItems, s->capI, items.size(), "server items");
      s->dCtl = grow(h, s->dCtl, s->capC, stage.size(), "server control");
      for (size_t k = 0; k < stage.size(); ++k) {
        views[k].ctl = s->dCtl + k;
        Request &q = *s->reqs[stage[k]];
        if (q.T > 1) h->ck(hipMemsetAsync(q.xw, 0, sizeof(smo::XWord) * smo::xrec_words(q.T), h->st), "clear team records");
      }
      h->upload(s->dV, views.data(), views.size(), "upload server views");
      h->upload(s->dItems, items.data(), items.size(), "upload server items");
      h->ck(hipMemsetAsync(s->dQueue, 0, sizeof(int), h->st), "clear server queue");
      h->ck(hipMemsetAsync(s->dT0, 0, sizeof(unsigned long long), h->st), "clear launch start");
      h->ck(hipMemsetAsync(s->dCtl, 0, sizeof(smo::Ctl) * stage.size(), h->st), "clear server control");
      smo::BatchPool pool;
      memset(&pool, 0, sizeof pool);
      pool.queue = s->dQueue;
      pool.t0 = s->dT0;
      pool.deadline_ticks = std::max<long long>(1, (long long)(tau * clock_hz));
      pool.K = 0;
      const int nitems = (int)items.size() / 2;
      h->ck(hipEventRecord(h->ev[0], h->st), "event");
      pool.lds_floats = (int)(shmem0 / sizeof(float));
      pool.flag_words = flag_words;
      pool.spec = h->opt.speculative_rows >= 0 ? 1 : 0;
      h->ck(smo_launch_batch(s->dV, s->dItems, nitems, std::min(G, nitems), &pool, shmem, h->st),
            "batched SMO kernel");
      h->ck(hipEventRecord(h->ev[1], h->st), "event");
      s->stage_launched = true;
    }
    s->inflight = s->stage_launched || !s->stage_preds.empty();
    if (n_active) *n_active = (int)s->active.size() + (int)s->stage_preds.size();
    if (n_queued) *n_queued = (int)s->queue.size();
    GUARD_END
  }();
}

int svmhip_server_poll(svmhip_server *s, svmhip_completion *out, int max) {
  if (!s || max <= 0 || !out) return 0;
  const int n = std::min<int>(max, (int)s->done.size());
  for (int k = 0; k < n; ++k) out[k] = s->done[k];
  s->done.erase(s->done.begin(), s->done.begin() + n);
  return n;
}

svmhip_status svmhip_server_train_result(svmhip_server *s, int id, double *coef, int n, double *labels2) {
  if (!s || id < 0 || id >= (int)s->reqs.size() || !s->reqs[id]) return SVMHIP_ERR_ARG;
  const Request &q = *s->reqs[id];
  if (q.type != 0) return SVMHIP_ERR_ARG;
  if (q.state != REQ_DONE || q.coef.empty()) return SVMHIP_ERR_STATE;
  if (coef) {
    if (n != (int)q.coef.size()) return SVMHIP_ERR_ARG;
    memcpy(coef, q.coef.data(), sizeof(double) * n);
  }
  if (labels2) {
    labels2[0] = q.labels[0];
    labels2[1] = q.labels[1];
  }
  return SVMHIP_OK;
}

svmhip_status svmhip_server_predict_result(svmhip_server *s, int id, double *dec_values, double *labels, int n) {
  if (!s || id < 0 || id >= (int)s->reqs.size() || !s->reqs[id]) return SVMHIP_ERR_ARG;
  const Request &q = *s->reqs[id];
  if (q.type != 1) return SVMHIP_ERR_ARG;
  if (q.state != REQ_DONE) return SVMHIP_ERR_STATE;
  if (n != (int)q.dec.size()) return SVMHIP_ERR_ARG;
  for (int t = 0; t < n; ++t) {
    if (dec_values) dec_values[t] = q.dec[t];
    if (labels) labels[t] = q.dec[t] > 0 ? q.mlabels[0] : q.mlabels[1];
  }
  return SVMHIP_OK;
}

svmhip_status svmhip_server_release(svmhip_server *s, int id) {
  if (!s || id < 0 || id >= (int)s->reqs.size() || !s->reqs[id]) return SVMHIP_ERR_ARG;
  Request *q = s->reqs[id];
  if (q->state != REQ_DONE) return SVMHIP_ERR_STATE;
  svmhip_solver *h = s->g;
  for (Request *o : s->reqs)
    if (o && o->type == 1 && o->model_id == id && o->state != REQ_DONE) return SVMHIP_ERR_STATE;
  free_rows(h, q->model);
  for (void *p : {(void *)q->mcoef, (void *)q->mrho, (void *)q->mcs, (void *)q->mcc}) h->dfree(p);
  s->used -= q->bytes;
  delete q;
  s->reqs[id] = nullptr;
  return SVMHIP_OK;
}

double svmhip_server_bytes_in_use(const svmhip_server *s) { return s ? s->used : 0.0; }

svmhip_status svmhip_server_stats(const svmhip_server *s, svmhip_stats *stats) {
  if (!s || !stats) return SVMHIP_ERR_ARG;
  *stats = s->tot;
  return SVMHIP_OK;
}

void svmhip_server_destroy(svmhip_server *s) {
  if (!s) return;
  for (Request *q : s->reqs) delete q;
  svmhip_destroy(s->g);  // frees every device allocation made through it
  delete s;
}

}  // extern "C"
