// The kept LIBSVM C API (include/svm.h) on top of the device solver.
//
// Training (svm_train, svm_cross_validation) follows the reference's
// orchestration (libsvm/svm.cpp:1497-1769 formulations and svm_train_one,
// :2079-2404 one-vs-one, :2407-2525 cross-validation) and hands every binary,
// one-class or regression subproblem to the device solver through the
// svmhip_* C-ABI, which runs the SMO loop on the GPU (the role CudaSolver /
// CudaSolverNU play in the reference, svm.cpp:1712-1718).  There is no CPU
// solver in this library: cuda_flag must be 1, and a missing or failing
// device is reported, never replaced.
//
// Prediction, model I/O, parameter checks and accessors are host code, as in
// the reference (svm.cpp:2527-3232).
#include <ctype.h>
#include <float.h>
#include <limits.h>
#include <locale.h>
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <map>
#include <string>
#include <vector>

#include "../../include/svm.h"
#include "../../include/svm_hip.h"

int libsvm_version = LIBSVM_VERSION;

namespace {

typedef signed char schar;

void print_stdout(const char *s) {
  fputs(s, stdout);
  fflush(stdout);
}
void (*g_print)(const char *) = &print_stdout;

void info(const char *fmt, ...) {
  char buf[4096];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof buf, fmt, ap);
  va_end(ap);
  g_print(buf);
}

template <class T>
T *xmalloc(size_t n) {
  return (T *)malloc(std::max<size_t>(n, 1) * sizeof(T));
}

// A device-solver failure; svm_train / svm_cross_validation report it and
// return without a model (the reference calls exit(1), cuda_solver.cpp:439-446).
struct DeviceFailure {
  svmhip_status code;
};

// Solution of one subproblem (Solver::SolutionInfo + alpha).
struct Solution {
  std::vector<double> alpha;  // signed coefficients, length l
  double rho = 0, obj = 0, r = 0, ub_p = 0, ub_n = 0;
  int iters = 0;
};

// One subproblem: rows of the caller's problem (in solver order), their
// labels / targets, and the formulation Solver::Solve receives.
struct Job {
  std::vector<int> rows;
  std::vector<double> ylab;
  double Cp = 0, Cn = 0;
  svmhip_kind kind = SVMHIP_KIND_C;
  int n_var = 0;
  std::vector<schar> y;
  std::vector<double> p, alpha;
  double fCp = 0, fCn = 0;
  Solution sol;
};

// Formulations, libsvm/svm.cpp:1500-1696: y, p, the initial alpha and the
// bounds of each svm_type.
void formulate(const svm_parameter *param, Job &j) {
  const int l = (int)j.rows.size();
  switch (param->svm_type) {
    case C_SVC:
      j.kind = SVMHIP_KIND_C;
      j.n_var = l;
      j.y.resize(l);
      for (int i = 0; i < l; ++i) j.y[i] = j.ylab[i] > 0 ? +1 : -1;
      j.p.assign(l, -1.0);
      j.alpha.assign(l, 0.0);
      j.fCp = j.Cp;
      j.fCn = j.Cn;
      break;
    case NU_SVC: {
      j.kind = SVMHIP_KIND_NU;
      j.n_var = l;
      j.y.resize(l);
      for (int i = 0; i < l; ++i) j.y[i] = j.ylab[i] > 0 ? +1 : -1;
      double pos = param->nu * l / 2, neg = param->nu * l / 2;
      j.alpha.resize(l);
      for (int i = 0; i < l; ++i) {
        if (j.y[i] == +1) { j.alpha[i] = std::min(1.0, pos); pos -= j.alpha[i]; }
        else { j.alpha[i] = std::min(1.0, neg); neg -= j.alpha[i]; }
      }
      j.p.assign(l, 0.0);
      j.fCp = j.fCn = 1.0;
      break;
    }
    case ONE_CLASS: {
      j.kind = SVMHIP_KIND_C;
      j.n_var = l;
      int n = (int)(param->nu * l);
      j.alpha.assign(l, 0.0);
      for (int i = 0; i < n; ++i) j.alpha[i] = 1;
      if (n < l) j.alpha[n] = param->nu * l - n;
      j.y.assign(l, 1);
      j.p.assign(l, 0.0);
      j.fCp = j.fCn = 1.0;
      break;
    }
    case EPSILON_SVR:
      j.kind = SVMHIP_KIND_C;
      j.n_var = 2 * l;
      j.alpha.assign(2 * l, 0.0);
      j.p.resize(2 * l);
      j.y.resize(2 * l);
      for (int i = 0; i < l; ++i) {
        j.p[i] = param->p - j.ylab[i];
        j.y[i] = 1;
        j.p[i + l] = param->p + j.ylab[i];
        j.y[i + l] = -1;
      }
      j.fCp = j.fCn = param->C;
      break;
    case NU_SVR: {
      j.kind = SVMHIP_KIND_NU;
      j.n_var = 2 * l;
      const double C = param->C;
      double sum = C * param->nu * l / 2;
      j.alpha.resize(2 * l);
      j.p.resize(2 * l);
      j.y.resize(2 * l);
      for (int i = 0; i < l; ++i) {
        j.alpha[i] = j.alpha[i + l] = std::min(sum, C);
        sum -= j.alpha[i];
        j.p[i] = -j.ylab[i];
        j.y[i] = 1;
        j.p[i + l] = j.ylab[i];
        j.y[i + l] = -1;
      }
      j.fCp = j.fCn = C;
      break;
    }
  }
}

// Post-processing of the solver output per svm_type (libsvm/svm.cpp:1520-1696)
// and the summary lines of svm_train_one (:1739-1766).
void finish(const svm_parameter *param, Job &j, double rho, double obj, double r, int iters, bool optimal) {
  const int l = (int)j.rows.size();
  Solution &out = j.sol;
  out.alpha.assign(l, 0.0);
  out.rho = rho;
  out.obj = obj;
  out.r = r;
  out.iters = iters;
  out.ub_p = j.fCp;
  out.ub_n = j.fCn;
  if (!optimal) fprintf(stderr, "\nWARNING: reaching max number of iterations\n");
  info("\noptimization finished, #iter = %d\n", iters);
  switch (param->svm_type) {
    case C_SVC: {
      double sum = 0;
      for (int i = 0; i < l; ++i) sum += j.alpha[i];
      if (j.Cp == j.Cn) info("nu = %f\n", sum / (j.Cp * l));
      for (int i = 0; i < l; ++i) out.alpha[i] = j.alpha[i] * j.y[i];
      break;
    }
    case NU_SVC:
      info("C = %f\n", 1 / r);
      for (int i = 0; i < l; ++i) out.alpha[i] = j.alpha[i] * j.y[i] / r;
      out.rho /= r;
      out.obj /= (r * r);
      out.ub_p = out.ub_n = 1 / r;
      break;
    case ONE_CLASS:
      for (int i = 0; i < l; ++i) out.alpha[i] = j.alpha[i];
      break;
    case EPSILON_SVR: {
      double sum = 0;
      for (int i = 0; i < l; ++i) {
        out.alpha[i] = j.alpha[i] - j.alpha[i + l];
        sum += fabs(out.alpha[i]);
      }
      info("nu = %f\n", sum / (param->C * l));
      break;
    }
    case NU_SVR:
      info("epsilon = %f\n", -r);
      for (int i = 0; i < l; ++i) out.alpha[i] = j.alpha[i] - j.alpha[i + l];
      break;
  }
  info("obj = %f, rho = %f\n", out.obj, out.rho);
  int nSV = 0, nBSV = 0;
  for (int i = 0; i < l; ++i) {
    if (fabs(out.alpha[i]) > 0) {
      ++nSV;
      double ub = j.ylab[i] > 0 ? out.ub_p : out.ub_n;
      if (fabs(out.alpha[i]) >= ub) ++nBSV;
    }
  }
  info("nSV = %d, nBSV = %d\n", nSV, nBSV);
}

svmhip_options device_options() {
  svmhip_options opt;
  memset(&opt, 0, sizeof opt);
  opt.struct_size = sizeof opt;
  const char *e = getenv("SVMHIP_VERBOSE");
  opt.verbose = e && *e == '1';
  const char *f = getenv("SVMHIP_HBM_FRACTION");
  if (f && *f) opt.hbm_fraction = atof(f);
  return opt;
}

// Problems up to this many rows go to the batched (one workgroup) solver
// even when alone; larger single problems use the persistent grid.
constexpr int kSoloMaxRows = 16384;

// One large subproblem on the persistent multi-workgroup solver.
void run_single(const svm_problem *prob, const svm_parameter *param, Job &j) {
  const int l = (int)j.rows.size();
  std::vector<svm_node *> sx(std::max(l, 1));
  for (int i = 0; i < l; ++i) sx[i] = prob->x[j.rows[i]];
  svm_problem sub;
  sub.l = l;
  sub.x = sx.data();
  sub.y = j.ylab.data();
  svmhip_options opt = device_options();
  svmhip_solver *h = nullptr;
  svmhip_status s = svmhip_create(&sub, param, j.kind, &opt, &h);
  if (s != SVMHIP_OK) {
    fprintf(stderr, "ERROR: device solver unavailable (status %d)\n", (int)s);
    throw DeviceFailure{s};
  }
  std::vector<char> status(j.n_var);
  for (int v = 0; v < j.n_var; ++v) {
    double C = j.y[v] > 0 ? j.fCp : j.fCn;
    status[v] = j.alpha[v] >= C ? 1 : (j.alpha[v] <= 0 ? 0 : 2);
  }
  svmhip_stats st;
  memset(&st, 0, sizeof st);
  if ((s = svmhip_setup_rbf(h, l)) != SVMHIP_OK ||
      (s = svmhip_setup(h, j.y.data(), j.p.data(), j.alpha.data(), status.data(), j.fCp, j.fCn, j.n_var)) != SVMHIP_OK ||
      (s = svmhip_solve(h, 0, &st)) != SVMHIP_OK ||
      (s = svmhip_fetch(h, nullptr, j.alpha.data(), nullptr, j.n_var)) != SVMHIP_OK) {
    fprintf(stderr, "ERROR: device solver failed (status %d): %s\n", (int)s, svmhip_last_error(h));
    svmhip_destroy(h);
    throw DeviceFailure{s};
  }
  svmhip_destroy(h);
  finish(param, j, st.rho, st.obj, st.r, st.iters, st.optimal != 0);
}

// Every subproblem on the device: one large one on the persistent grid,
// otherwise all of them in one batched launch (one workgroup each).  This is
// where the reference would create and run one CudaSolver per subproblem
// (libsvm/svm.cpp:1712-1738).
void run_jobs(const svm_problem *prob, const svm_parameter *param, std::vector<Job> &jobs) {
  for (Job &j : jobs) formulate(param, j);
  if (jobs.empty()) return;
  if (jobs.size() == 1 && (int)jobs[0].rows.size() > kSoloMaxRows) {
    run_single(prob, param, jobs[0]);
    return;
  }
  svmhip_options opt = device_options();
  svmhip_batch *b = nullptr;
  svmhip_status s = svmhip_batch_create(prob, param, &opt, &b);
  if (s != SVMHIP_OK) {
    if (s == SVMHIP_ERR_ARG) {  // e.g. precomputed kernels or very wide rows: one solve at a time
      for (Job &j : jobs) run_single(prob, param, j);
      return;
    }
    fprintf(stderr, "ERROR: device solver unavailable (status %d)\n", (int)s);
    throw DeviceFailure{s};
  }
  std::map<std::vector<int>, int> ids;
  std::vector<svmhip_subproblem> subs(jobs.size());
  for (size_t k = 0; k < jobs.size(); ++k) {
    Job &j = jobs[k];
    auto it = ids.find(j.rows);
    int id;
    if (it != ids.end()) {
      id = it->second;
    } else {
      id = svmhip_batch_add_rowset(b, j.rows.data(), (int)j.rows.size());
      if (id < 0) {
        fprintf(stderr, "ERROR: device row set failed: %s\n", svmhip_batch_last_error(b));
        svmhip_batch_destroy(b);
        throw DeviceFailure{(svmhip_status)(-id)};
      }
      ids.emplace(j.rows, id);
    }
    svmhip_subproblem &sp = subs[k];
    memset(&sp, 0, sizeof sp);
    sp.rowset = id;
    sp.kind = j.kind;
    sp.n_var = j.n_var;
    sp.y = j.y.data();
    sp.p = j.p.data();
    sp.alpha = j.alpha.data();
    sp.Cp = j.fCp;
    sp.Cn = j.fCn;
    sp.gamma = param->gamma;
    sp.eps = param->eps;
  }
  svmhip_stats st;
  s = svmhip_batch_solve(b, subs.data(), (int)subs.size(), &st);
  if (s != SVMHIP_OK) {
    fprintf(stderr, "ERROR: batched device solve failed (status %d): %s\n", (int)s, svmhip_batch_last_error(b));
    svmhip_batch_destroy(b);
    throw DeviceFailure{s};
  }
  svmhip_batch_destroy(b);
  for (size_t k = 0; k < jobs.size(); ++k)
    finish(param, jobs[k], subs[k].rho, subs[k].obj, subs[k].r, subs[k].iters, subs[k].status == SVMHIP_OK);
}

// svm_group_classes, libsvm/svm.cpp:2081-2157: labels in order of first
// appearance, except that a -1/+1 problem puts +1 first.
void group_classes(const svm_problem *prob, std::vector<int> &label, std::vector<int> &start,
                   std::vector<int> &count, std::vector<int> &perm) {
  const int l = prob->l;
  label.clear();
  count.clear();
  std::vector<int> cls(l);
  for (int i = 0; i < l; ++i) {
    int lab = (int)prob->y[i];
    size_t c = 0;
    while (c < label.size() && label[c] != lab) ++c;
    if (c == label.size()) {
      label.push_back(lab);
      count.push_back(0);
    }
    ++count[c];
    cls[i] = (int)c;
  }
  const int k = (int)label.size();
  if (k == 2 && label[0] == -1 && label[1] == 1) {
    std::swap(label[0], label[1]);
    std::swap(count[0], count[1]);
    for (int i = 0; i < l; ++i) cls[i] = 1 - cls[i];
  }
  start.assign(k, 0);
  for (int c = 1; c < k; ++c) start[c] = start[c - 1] + count[c - 1];
  perm.assign(l, 0);
  std::vector<int> fill(start);
  for (int i = 0; i < l; ++i) perm[fill[cls[i]]++] = i;
}

// Kernel::k_function, libsvm/svm.cpp:327-384 (prediction form).
double dot(const svm_node *a, const svm_node *b) {
  double s = 0;
  while (a->index != -1 && b->index != -1) {
    if (a->index == b->index) {
      s += a->value * b->value;
      ++a;
      ++b;
    } else if (a->index > b->index) {
      ++b;
    } else {
      ++a;
    }
  }
  return s;
}

double powi(double base, int times) {
  double acc = 1.0, sq = base;
  for (int t = times; t > 0; t /= 2) {
    if (t % 2 == 1) acc *= sq;
    sq = sq * sq;
  }
  return acc;
}

double k_function(const svm_node *x, const svm_node *y, const svm_parameter &p) {
  switch (p.kernel_type) {
    case LINEAR: return dot(x, y);
    case POLY: return powi(p.gamma * dot(x, y) + p.coef0, p.degree);
    case RBF: {
      double s = 0;
      while (x->index != -1 && y->index != -1) {
        if (x->index == y->index) {
          double d = x->value - y->value;
          s += d * d;
          ++x;
          ++y;
        } else if (x->index > y->index) {
          s += y->value * y->value;
          ++y;
        } else {
          s += x->value * x->value;
          ++x;
        }
      }
      for (; x->index != -1; ++x) s += x->value * x->value;
      for (; y->index != -1; ++y) s += y->value * y->value;
      return exp(-p.gamma * s);
    }
    case SIGMOID: return tanh(p.gamma * dot(x, y) + p.coef0);
    case PRECOMPUTED: return x[(int)(y->value)].value;
    default: return 0;
  }
}

bool is_regression_like(int t) { return t == ONE_CLASS || t == EPSILON_SVR || t == NU_SVR; }

const char *svm_type_table[] = {"c_svc", "nu_svc", "one_class", "epsilon_svr", "nu_svr", NULL};
const char *kernel_type_table[] = {"linear", "polynomial", "rbf", "sigmoid", "precomputed", NULL};

}  // namespace

extern "C" {

// Trains one model per training subset (rows of prob, in order) with every
// subproblem of every model in one device batch.  Model layout as in
// svm_train (libsvm/svm.cpp:2159-2404): one-vs-one pairs, weighted C,
// support vectors grouped by class, sv_coef[class][...] packing.
static std::vector<svm_model *> train_models(const svm_problem *prob, const svm_parameter *param,
                                             const std::vector<std::vector<int>> &subsets) {
  struct Plan {
    std::vector<int> rows;            // subset, grouped by class for classification
    std::vector<int> label, start, count;
    std::vector<double> wC;
    size_t first_job = 0;
  };
  const bool regression_like = is_regression_like(param->svm_type);
  std::vector<Plan> plans(subsets.size());
  std::vector<Job> jobs;
  for (size_t m = 0; m < subsets.size(); ++m) {
    Plan &pl = plans[m];
    const std::vector<int> &sub = subsets[m];
    pl.first_job = jobs.size();
    if (regression_like) {
      Job j;
      j.rows = sub;
      for (int r : sub) j.ylab.push_back(prob->y[r]);
      j.Cp = j.Cn = 0;
      jobs.push_back(std::move(j));
      continue;
    }
    // svm_group_classes over the subset
    std::vector<svm_node *> sx(sub.size());
    std::vector<double> sy(sub.size());
    for (size_t i = 0; i < sub.size(); ++i) { sx[i] = prob->x[sub[i]]; sy[i] = prob->y[sub[i]]; }
    svm_problem sp;
    sp.l = (int)sub.size();
    sp.x = sx.data();
    sp.y = sy.data();
    std::vector<int> perm;
    group_classes(&sp, pl.label, pl.start, pl.count, perm);
    const int k = (int)pl.label.size();
    if (k == 1) info("WARNING: training data in only one class. See README for details.\n");
    pl.rows.resize(sub.size());
    for (size_t i = 0; i < sub.size(); ++i) pl.rows[i] = sub[perm[i]];
    pl.wC.assign(k, param->C);
    for (int w = 0; w < param->nr_weight; ++w) {
      int c = 0;
      while (c < k && param->weight_label[w] != pl.label[c]) ++c;
      if (c == k) fprintf(stderr, "WARNING: class label %d specified in weight is not found\n", param->weight_label[w]);
      else pl.wC[c] *= param->weight[w];
    }
    for (int a = 0; a < k; ++a)
      for (int b = a + 1; b < k; ++b) {
        Job j;
        for (int t = 0; t < pl.count[a]; ++t) { j.rows.push_back(pl.rows[pl.start[a] + t]); j.ylab.push_back(+1); }
        for (int t = 0; t < pl.count[b]; ++t) { j.rows.push_back(pl.rows[pl.start[b] + t]); j.ylab.push_back(-1); }
        j.Cp = pl.wC[a];
        j.Cn = pl.wC[b];
        jobs.push_back(std::move(j));
      }
  }
  run_jobs(prob, param, jobs);

  std::vector<svm_model *> models(subsets.size());
  for (size_t m = 0; m < subsets.size(); ++m) {
    const Plan &pl = plans[m];
    const std::vector<int> &sub = subsets[m];
    svm_model *model = (svm_model *)calloc(1, sizeof(svm_model));
    model->param = *param;
    model->free_sv = 0;
    models[m] = model;
    if (regression_like) {
      const Solution &f = jobs[pl.first_job].sol;
      model->nr_class = 2;
      model->sv_coef = xmalloc<double *>(1);
      model->rho = xmalloc<double>(1);
      model->rho[0] = f.rho;
      int nSV = 0;
      for (size_t i = 0; i < sub.size(); ++i)
        if (fabs(f.alpha[i]) > 0) ++nSV;
      model->l = nSV;
      model->SV = xmalloc<svm_node *>(nSV);
      model->sv_coef[0] = xmalloc<double>(nSV);
      model->sv_indices = xmalloc<int>(nSV);
      int j = 0;
      for (size_t i = 0; i < sub.size(); ++i)
        if (fabs(f.alpha[i]) > 0) {
          model->SV[j] = prob->x[sub[i]];
          model->sv_coef[0][j] = f.alpha[i];
          model->sv_indices[j] = (int)i + 1;
          ++j;
        }
      continue;
    }
    const int k = (int)pl.label.size();
    const int l = (int)sub.size();
    const int npair = k * (k - 1) / 2;
    // position in the subset of each grouped row (for sv_indices)
    std::vector<int> pos_of(prob->l, -1);
    for (int i = 0; i < l; ++i) pos_of[sub[i]] = i;
    std::vector<char> nonzero(l, 0);
    for (int a = 0, q = 0; a < k; ++a)
      for (int b = a + 1; b < k; ++b, ++q) {
        const Solution &f = jobs[pl.first_job + q].sol;
        for (int t = 0; t < pl.count[a]; ++t)
          if (fabs(f.alpha[t]) > 0) nonzero[pl.start[a] + t] = 1;
        for (int t = 0; t < pl.count[b]; ++t)
          if (fabs(f.alpha[pl.count[a] + t]) > 0) nonzero[pl.start[b] + t] = 1;
      }
    model->nr_class = k;
    model->label = xmalloc<int>(k);
    for (int c = 0; c < k; ++c) model->label[c] = pl.label[c];
    model->rho = xmalloc<double>(npair);
    for (int q = 0; q < npair; ++q) model->rho[q] = jobs[pl.first_job + q].sol.rho;
    model->probA = model->probB = NULL;
    int total = 0;
    std::vector<int> nz_count(k);
    model->nSV = xmalloc<int>(k);
    for (int c = 0; c < k; ++c) {
      int n = 0;
      for (int t = 0; t < pl.count[c]; ++t)
        if (nonzero[pl.start[c] + t]) { ++n; ++total; }
      model->nSV[c] = n;
      nz_count[c] = n;
    }
    info("Total nSV = %d\n", total);
    model->l = total;
    model->SV = xmalloc<svm_node *>(total);
    model->sv_indices = xmalloc<int>(total);
    int p = 0;
    for (int i = 0; i < l; ++i)
      if (nonzero[i]) {
        model->SV[p] = prob->x[pl.rows[i]];
        model->sv_indices[p++] = pos_of[pl.rows[i]] + 1;
      }
    std::vector<int> nz_start(k, 0);
    for (int c = 1; c < k; ++c) nz_start[c] = nz_start[c - 1] + nz_count[c - 1];
    model->sv_coef = xmalloc<double *>(std::max(k - 1, 1));
    for (int c = 0; c < k - 1; ++c) model->sv_coef[c] = xmalloc<double>(total);
    for (int a = 0, q = 0; a < k; ++a)
      for (int b = a + 1; b < k; ++b, ++q) {
        // classifier (a,b): class-a coefficients in sv_coef[b-1], class-b in sv_coef[a]
        const Solution &f = jobs[pl.first_job + q].sol;
        int t2 = nz_start[a];
        for (int t = 0; t < pl.count[a]; ++t)
          if (nonzero[pl.start[a] + t]) model->sv_coef[b - 1][t2++] = f.alpha[t];
        t2 = nz_start[b];
        for (int t = 0; t < pl.count[b]; ++t)
          if (nonzero[pl.start[b] + t]) model->sv_coef[a][t2++] = f.alpha[pl.count[a] + t];
      }
  }
  return models;
}

// Returns NULL (after printing the reason) if the device solver fails.
svm_model *svm_train(const svm_problem *prob, const svm_parameter *param) {
  try {
    std::vector<int> all(prob->l);
    for (int i = 0; i < prob->l; ++i) all[i] = i;
    return train_models(prob, param, {all})[0];
  } catch (const DeviceFailure &) {
    return NULL;
  }
}

// Stratified folds drawn with rand() as the reference does
// (libsvm/svm.cpp:2407-2525); all folds' subproblems train in one batch.
void svm_cross_validation(const svm_problem *prob, const svm_parameter *param, int nr_fold, double *target) {
  const int l = prob->l;
  if (nr_fold > l) {
    nr_fold = l;
    fprintf(stderr, "WARNING: # folds > # data. Will use # folds = # data instead (i.e., leave-one-out cross validation)\n");
  }
  std::vector<int> perm(l), fold_start(nr_fold + 1);
  if ((param->svm_type == C_SVC || param->svm_type == NU_SVC) && nr_fold < l) {
    std::vector<int> label, start, count;
    group_classes(prob, label, start, count, perm);
    const int k = (int)label.size();
    std::vector<int> index(perm), fold_count(nr_fold, 0);
    for (int c = 0; c < k; ++c)
      for (int i = 0; i < count[c]; ++i) {
        int j = i + rand() % (count[c] - i);
        std::swap(index[start[c] + j], index[start[c] + i]);
      }
    for (int i = 0; i < nr_fold; ++i)
      for (int c = 0; c < k; ++c) fold_count[i] += (i + 1) * count[c] / nr_fold - i * count[c] / nr_fold;
    fold_start[0] = 0;
    for (int i = 1; i <= nr_fold; ++i) fold_start[i] = fold_start[i - 1] + fold_count[i - 1];
    for (int c = 0; c < k; ++c)
      for (int i = 0; i < nr_fold; ++i) {
        int b = start[c] + i * count[c] / nr_fold, e = start[c] + (i + 1) * count[c] / nr_fold;
        for (int j = b; j < e; ++j) perm[fold_start[i]++] = index[j];
      }
    fold_start[0] = 0;
    for (int i = 1; i <= nr_fold; ++i) fold_start[i] = fold_start[i - 1] + fold_count[i - 1];
  } else {
    for (int i = 0; i < l; ++i) perm[i] = i;
    for (int i = 0; i < l; ++i) std::swap(perm[i], perm[i + rand() % (l - i)]);
    for (int i = 0; i <= nr_fold; ++i) fold_start[i] = i * l / nr_fold;
  }
  std::vector<std::vector<int>> subsets(nr_fold);
  for (int f = 0; f < nr_fold; ++f) {
    for (int j = 0; j < fold_start[f]; ++j) subsets[f].push_back(perm[j]);
    for (int j = fold_start[f + 1]; j < l; ++j) subsets[f].push_back(perm[j]);
  }
  std::vector<svm_model *> models;
  try {
    models = train_models(prob, param, subsets);
  } catch (const DeviceFailure &) {  // device failure: no predictions
    for (int i = 0; i < l; ++i) target[i] = NAN;
    return;
  }
  for (int f = 0; f < nr_fold; ++f) {
    const int b = fold_start[f], e = fold_start[f + 1];
    std::vector<const svm_node *> rows;
    for (int j = b; j < e; ++j) rows.push_back(prob->x[perm[j]]);
    std::vector<double> lab(rows.size()), dec;
    const int k = models[f]->nr_class;
    dec.resize(rows.size() * (size_t)(is_regression_like(param->svm_type) ? 1 : std::max(1, k * (k - 1) / 2)));
    svmhip_options opt = device_options();
    svmhip_status s = rows.empty() ? SVMHIP_OK
                                   : svmhip_predict_batch(models[f], rows.data(), (int)rows.size(), &opt, dec.data(),
                                                          lab.data());
    if (s == SVMHIP_OK) {
      for (int j = b; j < e; ++j) target[perm[j]] = lab[j - b];
    } else if (s == SVMHIP_ERR_ARG) {  // precomputed kernels / very wide rows: per-row host prediction
      for (int j = b; j < e; ++j) target[perm[j]] = svm_predict(models[f], prob->x[perm[j]]);
    } else {
      fprintf(stderr, "ERROR: device prediction failed (status %d)\n", (int)s);
      for (int j = b; j < e; ++j) target[perm[j]] = NAN;
    }
    svm_free_and_destroy_model(&models[f]);
  }
}

int svm_get_svm_type(const svm_model *model) { return model->param.svm_type; }
int svm_get_nr_class(const svm_model *model) { return model->nr_class; }
void svm_get_labels(const svm_model *model, int *label) {
  if (model->label)
    for (int i = 0; i < model->nr_class; ++i) label[i] = model->label[i];
}
void svm_get_sv_indices(const svm_model *model, int *idx) {
  if (model->sv_indices)
    for (int i = 0; i < model->l; ++i) idx[i] = model->sv_indices[i];
}
int svm_get_nr_sv(const svm_model *model) { return model->l; }
double svm_get_svr_probability(const svm_model *model) {
  if ((model->param.svm_type == EPSILON_SVR || model->param.svm_type == NU_SVR) && model->probA)
    return model->probA[0];
  fprintf(stderr, "Model doesn't contain information for SVR probability inference\n");
  return 0;
}

double svm_predict_values(const svm_model *model, const svm_node *x, double *dec) {
  if (is_regression_like(model->param.svm_type)) {
    const double *c = model->sv_coef[0];
    double s = 0;
    for (int i = 0; i < model->l; ++i) s += c[i] * k_function(x, model->SV[i], model->param);
    s -= model->rho[0];
    *dec = s;
    if (model->param.svm_type == ONE_CLASS) return s > 0 ? 1 : -1;
    return s;
  }
  const int k = model->nr_class, l = model->l;
  std::vector<double> kv(std::max(l, 1));
  for (int i = 0; i < l; ++i) kv[i] = k_function(x, model->SV[i], model->param);
  std::vector<int> start(k, 0), vote(k, 0);
  for (int c = 1; c < k; ++c) start[c] = start[c - 1] + model->nSV[c - 1];
  int p = 0;
  for (int a = 0; a < k; ++a)
    for (int b = a + 1; b < k; ++b) {
      double s = 0;
      const double *ca = model->sv_coef[b - 1], *cb = model->sv_coef[a];
      for (int t = 0; t < model->nSV[a]; ++t) s += ca[start[a] + t] * kv[start[a] + t];
      for (int t = 0; t < model->nSV[b]; ++t) s += cb[start[b] + t] * kv[start[b] + t];
      s -= model->rho[p];
      dec[p] = s;
      if (dec[p] > 0) ++vote[a];
      else ++vote[b];
      ++p;
    }
  int best = 0;
  for (int c = 1; c < k; ++c)
    if (vote[c] > vote[best]) best = c;
  return model->label[best];
}

double svm_predict(const svm_model *model, const svm_node *x) {
  const int k = model->nr_class;
  std::vector<double> dec(is_regression_like(model->param.svm_type) ? 1 : std::max(1, k * (k - 1) / 2));
  return svm_predict_values(model, x, dec.data());
}

// Platt's sigmoid 1 / (1 + exp(A f + B)), written so that 1 - p keeps its
// precision (libsvm/svm.cpp:1885-1893).
static double platt_probability(double f, double A, double B) {
  const double t = f * A + B;
  if (t >= 0) return exp(-t) / (1.0 + exp(-t));
  return 1.0 / (1 + exp(t));
}

// Pairwise coupling: class probabilities p from the pairwise ones r
// (r[i*k+j] = P(i | i or j)), method 2 of Wu, Lin & Weng, "Probability
// estimates for multi-class classification by pairwise coupling", JMLR 5
// (2004) 975-1005: p minimises sum_t sum_{j != t} (r_jt p_t - r_tj p_j)^2 on
// the simplex, by the paper's coordinate fixed-point iteration with
// LIBSVM's stopping rule (libsvm/svm.cpp:1896-1957).
static void couple_pairwise(int k, const std::vector<double> &r, double *p) {
  std::vector<double> Q((size_t)k * k), Qp(k);
  for (int t = 0; t < k; ++t) {
    p[t] = 1.0 / k;
    double qtt = 0;
    for (int j = 0; j < t; ++j) {
      qtt += r[(size_t)j * k + t] * r[(size_t)j * k + t];
      Q[(size_t)t * k + j] = Q[(size_t)j * k + t];
    }
    for (int j = t + 1; j < k; ++j) {
      qtt += r[(size_t)j * k + t] * r[(size_t)j * k + t];
      Q[(size_t)t * k + j] = -r[(size_t)j * k + t] * r[(size_t)t * k + j];
    }
    Q[(size_t)t * k + t] = qtt;
  }
  const int max_iter = std::max(100, k);
  const double eps = 0.005 / k;
  int it = 0;
  for (; it < max_iter; ++it) {
    double pQp = 0;  // recomputed each round for accuracy
    for (int t = 0; t < k; ++t) {
      double v = 0;
      for (int j = 0; j < k; ++j) v += Q[(size_t)t * k + j] * p[j];
      Qp[t] = v;
      pQp += p[t] * Qp[t];
    }
    double worst = 0;
    for (int t = 0; t < k; ++t) worst = std::max(worst, fabs(Qp[t] - pQp));
    if (worst < eps) break;
    for (int t = 0; t < k; ++t) {
      const double qtt = Q[(size_t)t * k + t];
      const double diff = (-Qp[t] + pQp) / qtt;
      p[t] += diff;
      pQp = (pQp + diff * (diff * qtt + 2 * Qp[t])) / (1 + diff) / (1 + diff);
      for (int j = 0; j < k; ++j) {
        Qp[j] = (Qp[j] + diff * Q[(size_t)t * k + j]) / (1 + diff);
        p[j] /= (1 + diff);
      }
    }
  }
  if (it >= max_iter) info("Exceeds max_iter in multiclass_prob\n");
}

// Class probabilities of x from a model that carries Platt parameters
// (probA/probB, e.g. a model file written by LIBSVM with -b 1), as
// libsvm/svm.cpp:2660-2697 computes them; other models predict as
// svm_predict and leave prob_estimates untouched.  Host code: one row.
double svm_predict_probability(const svm_model *model, const svm_node *x, double *prob_estimates) {
  const int st = model->param.svm_type;
  if (!((st == C_SVC || st == NU_SVC) && model->probA && model->probB)) return svm_predict(model, x);
  const int k = model->nr_class;
  std::vector<double> dec(std::max(1, k * (k - 1) / 2)), r((size_t)k * k, 0.0);
  svm_predict_values(model, x, dec.data());
  const double min_prob = 1e-7;
  int q = 0;
  for (int i = 0; i < k; ++i)
    for (int j = i + 1; j < k; ++j, ++q) {
      const double pij = std::min(std::max(platt_probability(dec[q], model->probA[q], model->probB[q]), min_prob),
                                  1 - min_prob);
      r[(size_t)i * k + j] = pij;
      r[(size_t)j * k + i] = 1 - pij;
    }
  couple_pairwise(k, r, prob_estimates);
  int best = 0;
  for (int i = 1; i < k; ++i)
    if (prob_estimates[i] > prob_estimates[best]) best = i;
  return model->label[best];
}

// Model file writer, libsvm/svm.cpp:2709-2801 (same keys and formats).
int svm_save_model(const char *name, const svm_model *model) {
  FILE *fp = fopen(name, "w");
  if (!fp) return -1;
  char *old = strdup(setlocale(LC_ALL, NULL));
  setlocale(LC_ALL, "C");
  const svm_parameter &pr = model->param;
  fprintf(fp, "svm_type %s\n", svm_type_table[pr.svm_type]);
  fprintf(fp, "kernel_type %s\n", kernel_type_table[pr.kernel_type]);
  if (pr.kernel_type == POLY) fprintf(fp, "degree %d\n", pr.degree);
  if (pr.kernel_type == POLY || pr.kernel_type == RBF || pr.kernel_type == SIGMOID) fprintf(fp, "gamma %g\n", pr.gamma);
  if (pr.kernel_type == POLY || pr.kernel_type == SIGMOID) fprintf(fp, "coef0 %g\n", pr.coef0);
  const int k = model->nr_class, l = model->l;
  fprintf(fp, "nr_class %d\n", k);
  fprintf(fp, "total_sv %d\n", l);
  fprintf(fp, "rho");
  for (int i = 0; i < k * (k - 1) / 2; ++i) fprintf(fp, " %g", model->rho[i]);
  fprintf(fp, "\n");
  if (model->label) {
    fprintf(fp, "label");
    for (int i = 0; i < k; ++i) fprintf(fp, " %d", model->label[i]);
    fprintf(fp, "\n");
  }
  if (model->probA) {
    fprintf(fp, "probA");
    for (int i = 0; i < k * (k - 1) / 2; ++i) fprintf(fp, " %g", model->probA[i]);
    fprintf(fp, "\n");
  }
  if (model->probB) {
    fprintf(fp, "probB");
    for (int i = 0; i < k * (k - 1) / 2; ++i) fprintf(fp, " %g", model->probB[i]);
    fprintf(fp, "\n");
  }
  if (model->nSV) {
    fprintf(fp, "nr_sv");
    for (int i = 0; i < k; ++i) fprintf(fp, " %d", model->nSV[i]);
    fprintf(fp, "\n");
  }
  fprintf(fp, "SV\n");
  for (int i = 0; i < l; ++i) {
    for (int j = 0; j < k - 1; ++j) fprintf(fp, "%.16g ", model->sv_coef[j][i]);
    const svm_node *p = model->SV[i];
    if (pr.kernel_type == PRECOMPUTED) fprintf(fp, "0:%d ", (int)(p->value));
    else
      for (; p->index != -1; ++p) fprintf(fp, "%d:%.8g ", p->index, p->value);
    fprintf(fp, "\n");
  }
  setlocale(LC_ALL, old);
  free(old);
  if (ferror(fp) != 0 || fclose(fp) != 0) return -1;
  return 0;
}

// Model file reader, libsvm/svm.cpp:2803-3044.
static bool read_header(FILE *fp, svm_model *m) {
  svm_parameter &pr = m->param;
  char cmd[81];
  for (;;) {
    if (fscanf(fp, "%80s", cmd) != 1) return false;
    if (!strcmp(cmd, "svm_type")) {
      if (fscanf(fp, "%80s", cmd) != 1) return false;
      int i = 0;
      for (; svm_type_table[i]; ++i)
        if (!strcmp(svm_type_table[i], cmd)) { pr.svm_type = i; break; }
      if (!svm_type_table[i]) { fprintf(stderr, "unknown svm type.\n"); return false; }
    } else if (!strcmp(cmd, "kernel_type")) {
      if (fscanf(fp, "%80s", cmd) != 1) return false;
      int i = 0;
      for (; kernel_type_table[i]; ++i)
        if (!strcmp(kernel_type_table[i], cmd)) { pr.kernel_type = i; break; }
      if (!kernel_type_table[i]) { fprintf(stderr, "unknown kernel function.\n"); return false; }
    } else if (!strcmp(cmd, "degree")) {
      if (fscanf(fp, "%d", &pr.degree) != 1) return false;
    } else if (!strcmp(cmd, "gamma")) {
      if (fscanf(fp, "%lf", &pr.gamma) != 1) return false;
    } else if (!strcmp(cmd, "coef0")) {
      if (fscanf(fp, "%lf", &pr.coef0) != 1) return false;
    } else if (!strcmp(cmd, "nr_class")) {
      if (fscanf(fp, "%d", &m->nr_class) != 1) return false;
    } else if (!strcmp(cmd, "total_sv")) {
      if (fscanf(fp, "%d", &m->l) != 1) return false;
    } else if (!strcmp(cmd, "rho") || !strcmp(cmd, "probA") || !strcmp(cmd, "probB")) {
      int n = m->nr_class * (m->nr_class - 1) / 2;
      double *v = xmalloc<double>(n);
      for (int i = 0; i < n; ++i)
        if (fscanf(fp, "%lf", &v[i]) != 1) { free(v); return false; }
      if (!strcmp(cmd, "rho")) m->rho = v;
      else if (!strcmp(cmd, "probA")) m->probA = v;
      else m->probB = v;
    } else if (!strcmp(cmd, "label") || !strcmp(cmd, "nr_sv")) {
      int n = m->nr_class;
      int *v = xmalloc<int>(n);
      for (int i = 0; i < n; ++i)
        if (fscanf(fp, "%d", &v[i]) != 1) { free(v); return false; }
      if (!strcmp(cmd, "label")) m->label = v;
      else m->nSV = v;
    } else if (!strcmp(cmd, "SV")) {
      for (;;) {
        int c = getc(fp);
        if (c == EOF || c == '\n') break;
      }
      break;
    } else {
      fprintf(stderr, "unknown text in model file: [%s]\n", cmd);
      return false;
    }
  }
  return true;
}

svm_model *svm_load_model(const char *name) {
  FILE *fp = fopen(name, "rb");
  if (!fp) return NULL;
  char *old = strdup(setlocale(LC_ALL, NULL));
  setlocale(LC_ALL, "C");
  svm_model *m = (svm_model *)calloc(1, sizeof(svm_model));
  if (!read_header(fp, m)) {
    fprintf(stderr, "ERROR: fscanf failed to read model\n");
    setlocale(LC_ALL, old);
    free(old);
    free(m->rho);
    free(m->label);
    free(m->nSV);
    free(m->probA);
    free(m->probB);
    free(m);
    fclose(fp);
    return NULL;
  }
  // read the SV lines whole
  std::vector<std::string> lines;
  {
    std::string cur;
    int c;
    while ((c = getc(fp)) != EOF) {
      if (c == '\n') { lines.push_back(cur); cur.clear(); }
      else cur.push_back((char)c);
    }
    if (!cur.empty()) lines.push_back(cur);
  }
  const int mcls = m->nr_class - 1, l = m->l;
  long elements = 0;
  for (int i = 0; i < l && i < (int)lines.size(); ++i)
    for (char ch : lines[i])
      if (ch == ':') ++elements;
  elements += l;
  m->sv_coef = xmalloc<double *>(std::max(mcls, 1));
  for (int i = 0; i < mcls; ++i) m->sv_coef[i] = xmalloc<double>(l);
  m->SV = xmalloc<svm_node *>(l);
  svm_node *xs = l > 0 ? xmalloc<svm_node>(elements) : NULL;
  long j = 0;
  for (int i = 0; i < l; ++i) {
    std::string line = i < (int)lines.size() ? lines[i] : std::string();
    std::vector<char> buf(line.begin(), line.end());
    buf.push_back('\0');
    m->SV[i] = &xs[j];
    char *save = nullptr, *tok = strtok_r(buf.data(), " \t", &save);
    for (int c = 0; c < mcls; ++c) {
      m->sv_coef[c][i] = tok ? strtod(tok, NULL) : 0.0;
      tok = c + 1 < mcls ? strtok_r(NULL, " \t", &save) : tok;
    }
    for (;;) {
      char *idx = strtok_r(NULL, ":", &save);
      char *val = strtok_r(NULL, " \t", &save);
      if (!val) break;
      xs[j].index = (int)strtol(idx, NULL, 10);
      xs[j].value = strtod(val, NULL);
      ++j;
    }
    xs[j++].index = -1;
  }
  setlocale(LC_ALL, old);
  free(old);
  if (ferror(fp) != 0 || fclose(fp) != 0) return NULL;
  m->free_sv = 1;
  return m;
}

void svm_free_model_content(svm_model *m) {
  if (m->free_sv && m->l > 0 && m->SV != NULL) free((void *)(m->SV[0]));
  if (m->sv_coef)
    for (int i = 0; i < m->nr_class - 1; ++i) free(m->sv_coef[i]);
  free(m->SV); m->SV = NULL;
  free(m->sv_coef); m->sv_coef = NULL;
  free(m->rho); m->rho = NULL;
  free(m->label); m->label = NULL;
  free(m->probA); m->probA = NULL;
  free(m->probB); m->probB = NULL;
  free(m->sv_indices); m->sv_indices = NULL;
  free(m->nSV); m->nSV = NULL;
}

void svm_free_and_destroy_model(svm_model **pp) {
  if (pp && *pp) {
    svm_free_model_content(*pp);
    free(*pp);
    *pp = NULL;
  }
}

void svm_destroy_param(svm_parameter *param) {
  free(param->weight_label);
  free(param->weight);
}

// libsvm/svm.cpp:3097-3216, plus the two restrictions of this library.
const char *svm_check_parameter(const svm_problem *prob, const svm_parameter *param) {
  int t = param->svm_type;
  if (t != C_SVC && t != NU_SVC && t != ONE_CLASS && t != EPSILON_SVR && t != NU_SVR) return "unknown svm type";
  int kt = param->kernel_type;
  if (kt != LINEAR && kt != POLY && kt != RBF && kt != SIGMOID && kt != PRECOMPUTED) return "unknown kernel type";
  if (param->gamma < 0) return "gamma < 0";
  if (param->degree < 0) return "degree of polynomial kernel < 0";
  if (param->cache_size <= 0) return "cache_size <= 0";
  if (param->eps <= 0) return "eps <= 0";
  if ((t == C_SVC || t == EPSILON_SVR || t == NU_SVR) && param->C <= 0) return "C <= 0";
  if ((t == NU_SVC || t == ONE_CLASS || t == NU_SVR) && (param->nu <= 0 || param->nu > 1)) return "nu <= 0 or nu > 1";
  if (t == EPSILON_SVR && param->p < 0) return "p < 0";
  if (param->shrinking != 0 && param->shrinking != 1) return "shrinking != 0 and shrinking != 1";
  if (param->probability != 0 && param->probability != 1) return "probability != 0 and probability != 1";
  if (param->probability == 1 && t == ONE_CLASS) return "one-class SVM probability output not supported yet";
  if (param->cuda_flag != 1) return "this library has only the GPU solver: set cuda_flag = 1";
  if (param->probability == 1) return "probability estimates are not supported by the GPU solver";
  if (t == NU_SVC) {
    std::vector<int> label, count;
    for (int i = 0; i < prob->l; ++i) {
      int lab = (int)prob->y[i];
      size_t c = 0;
      while (c < label.size() && label[c] != lab) ++c;
      if (c == label.size()) { label.push_back(lab); count.push_back(0); }
      ++count[c];
    }
    for (size_t a = 0; a < label.size(); ++a)
      for (size_t b = a + 1; b < label.size(); ++b)
        if (param->nu * (count[a] + count[b]) / 2 > std::min(count[a], count[b])) return "specified nu is infeasible";
  }
  return NULL;
}

int svm_check_probability_model(const svm_model *model) {
  return ((model->param.svm_type == C_SVC || model->param.svm_type == NU_SVC) && model->probA && model->probB) ||
         ((model->param.svm_type == EPSILON_SVR || model->param.svm_type == NU_SVR) && model->probA);
}

void svm_set_print_string_function(void (*print_func)(const char *)) {
  g_print = print_func ? print_func : &print_stdout;
}

}  // extern "C"
