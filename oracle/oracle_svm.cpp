/*
 * ORACLE — test infrastructure only.  Not part of the product.
 *
 * A from-scratch CPU restatement, in FP64, of the LIBSVM 3.18 training and
 * prediction path that libsvm-cuda wraps (libsvm/svm.cpp).  Only tests/,
 * __graft_entry__.smoke() and bench.py's CPU-baseline leg load it, and only
 * as the checker or the CPU baseline; the product library (libsvm_hip.so)
 * never links or calls it.
 *
 * The arithmetic follows the reference expression by expression, so that
 * with Qfloat = float ("reference-exact" mode) the iterate sequence matches
 * LIBSVM 3.18 bit for bit when compiled with -ffp-contract=off.  The code
 * structure is our own; each routine cites the reference lines it restates.
 *
 * Exported API: the LIBSVM C API under an "ora_" prefix, plus
 * ora_train_one() which exposes one binary/one-class/regression solve with
 * its raw outputs (alpha, G, rho, obj, iterations, working-set trace).
 *
 * Pinning status (see tests/test_oracle_*.py):
 *   Solver / Solver_NU / SVC_Q / SVR_Q / ONE_CLASS_Q : pinned against
 *     scikit-learn's bundled LIBSVM (same iteration counts and dual
 *     coefficients), brute-force QP on tiny problems, KKT certificates,
 *     and the two-point closed form.
 *   kernel functions : pinned against closed forms (numpy).
 *   one-vs-one training (class order, SV packing, sv_coef layout, rho
 *     order) : pinned against sklearn SVC support_, dual_coef_ and
 *     intercept_ on a 4-class set.
 *   predict_values : pinned against sklearn decision_function.
 *   model I/O : the library's reader and writer are pinned by a hand-written
 *     model file in LIBSVM 3.18's format (tests/golden/poly_3class.model).
 *   full-size runs of both Q variants : committed as fixtures by
 *     tests/golden/make_fullsize_fixtures.py and checked for feasibility and
 *     consistency.
 *   cross-validation fold assignment (glibc rand()) : parity unpinned.
 */
#include <math.h>
#include <float.h>
#include <limits.h>
#include <locale.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "../include/svm.h"

typedef signed char schar;

/* ------------------------------------------------------------------------ */
/* options and logging                                                      */
/* ------------------------------------------------------------------------ */

extern "C" {
struct ora_options {
	int max_iter;        /* 0 = LIBSVM default max(10M, 100*l) */
	int q_double;        /* 0 = cache Q in float (reference-exact), 1 = double */
	int trace_cap;       /* capacity of trace_ij (pairs) */
	int *trace_ij;       /* if non-null, records (i, j) of each iteration */
};

struct ora_result {
	int l;               /* number of training rows */
	int n_var;           /* solver variables (l, or 2l for SVR) */
	double *alpha;       /* [l] signed coefficients, as svm_train_one returns */
	double *alpha_raw;   /* [n_var] solver alpha (unsigned, before y / r scaling) */
	double *G;           /* [n_var] final gradient */
	double rho;
	double obj;
	double r;            /* Solver_NU only */
	double upper_bound_p;
	double upper_bound_n;
	int iters;
	int trace_len;
	int nSV;
	int nBSV;
};
}

static ora_options g_opt = {0, 0, 0, nullptr};

/* Threads used to fill a missing kernel row (1 = LIBSVM's single thread).
 * Each element of a row is computed independently and identically, so the
 * thread count never changes a result; it only shortens full-size runs. */
static int g_threads = 1;

static void print_to_stdout(const char *s) { fputs(s, stdout); fflush(stdout); }
static void print_nothing(const char *) {}
static void (*g_print)(const char *) = &print_nothing;

static void info(const char *fmt, ...)
{
	char buf[4096];
	va_list ap;
	va_start(ap, fmt);
	vsnprintf(buf, sizeof buf, fmt, ap);
	va_end(ap);
	g_print(buf);
}

namespace ora {

static const double INF = HUGE_VAL;
static const double TAU = 1e-12;   /* libsvm/svm.cpp:42 */

template <class T> static inline void swap_(T &a, T &b) { T t = a; a = b; b = t; }

/* powi: repeated squaring, libsvm/svm.cpp:30-40. */
static inline double powi(double base, int times)
{
	double acc = 1.0, sq = base;
	for (int t = times; t > 0; t /= 2) {
		if (t % 2 == 1) acc *= sq;
		sq = sq * sq;
	}
	return acc;
}

/* ------------------------------------------------------------------------ */
/* Row cache: LRU over variable-length column prefixes.                     */
/* Restates Cache (libsvm/svm.cpp:65-201): a budget in Qf elements, one     */
/* entry per column, entries may hold a prefix [0,len); the least recently  */
/* used prefixes are released until a request fits.                        */
/* ------------------------------------------------------------------------ */
template <class Qf>
class RowCache {
public:
	RowCache(int l, long bytes) : l_(l), ent_(l)
	{
		budget_ = bytes / (long)sizeof(Qf);
		budget_ -= (long)l * (long)sizeof(Entry) / (long)sizeof(Qf);
		budget_ = std::max(budget_, 2L * (long)l);
		ring_.prev = ring_.next = &ring_;
	}
	~RowCache()
	{
		for (Entry *e = ring_.next; e != &ring_; e = e->next) free(e->data);
	}
	/* Returns the first position that still needs filling. */
	int fetch(int idx, Qf **out, int len)
	{
		Entry *e = &ent_[idx];
		if (e->len) unlink(e);
		int need = len - e->len;
		if (need > 0) {
			while (budget_ < need) {
				Entry *old = ring_.next;
				unlink(old);
				free(old->data);
				budget_ += old->len;
				old->data = nullptr;
				old->len = 0;
			}
			e->data = (Qf *)realloc(e->data, sizeof(Qf) * (size_t)len);
			budget_ -= need;
			swap_(e->len, len);
		}
		append(e);
		*out = e->data;
		return len;
	}
	void swap_index(int i, int j)
	{
		if (i == j) return;
		if (ent_[i].len) unlink(&ent_[i]);
		if (ent_[j].len) unlink(&ent_[j]);
		swap_(ent_[i].data, ent_[j].data);
		swap_(ent_[i].len, ent_[j].len);
		if (ent_[i].len) append(&ent_[i]);
		if (ent_[j].len) append(&ent_[j]);
		if (i > j) swap_(i, j);
		for (Entry *e = ring_.next; e != &ring_; e = e->next) {
			if (e->len > i) {
				if (e->len > j) {
					swap_(e->data[i], e->data[j]);
				} else {
					/* the prefix covers i but not j: drop it */
					Entry *keep = e->prev;
					unlink(e);
					free(e->data);
					budget_ += e->len;
					e->data = nullptr;
					e->len = 0;
					e = keep;
				}
			}
		}
	}

private:
	struct Entry {
		Entry *prev = nullptr, *next = nullptr;
		Qf *data = nullptr;
		int len = 0;
	};
	void unlink(Entry *e) { e->prev->next = e->next; e->next->prev = e->prev; }
	void append(Entry *e)
	{
		e->next = &ring_;
		e->prev = ring_.prev;
		e->prev->next = e;
		e->next->prev = e;
	}
	int l_;
	long budget_;
	std::vector<Entry> ent_;
	Entry ring_;
};

/* ------------------------------------------------------------------------ */
/* Kernel evaluation, libsvm/svm.cpp:203-384.                              */
/* ------------------------------------------------------------------------ */

/* Sparse merge dot product in increasing feature order (svm.cpp:305-325). */
static double sparse_dot(const svm_node *a, const svm_node *b)
{
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

/* Prediction-time kernel, svm.cpp:327-384; RBF uses sum (x-y)^2 directly. */
static double k_function(const svm_node *x, const svm_node *y, const svm_parameter &p)
{
	switch (p.kernel_type) {
	case LINEAR:
		return sparse_dot(x, y);
	case POLY:
		return powi(p.gamma * sparse_dot(x, y) + p.coef0, p.degree);
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
	case SIGMOID:
		return tanh(p.gamma * sparse_dot(x, y) + p.coef0);
	case PRECOMPUTED:
		return x[(int)(y->value)].value;
	default:
		return 0;
	}
}

/* Training-time kernel over the (swappable) row table, svm.cpp:203-303. */
class KernelRows {
public:
	KernelRows(int l, svm_node *const *x, const svm_parameter &p)
		: type_(p.kernel_type), degree_(p.degree), gamma_(p.gamma), coef0_(p.coef0),
		  x_(x, x + l)
	{
		if (type_ == RBF) {
			xsq_.resize(l);
			for (int i = 0; i < l; i++) xsq_[i] = sparse_dot(x_[i], x_[i]);
		}
	}
	double eval(int i, int j) const
	{
		switch (type_) {
		case LINEAR:
			return sparse_dot(x_[i], x_[j]);
		case POLY:
			return powi(gamma_ * sparse_dot(x_[i], x_[j]) + coef0_, degree_);
		case RBF:
			return exp(-gamma_ * (xsq_[i] + xsq_[j] - 2 * sparse_dot(x_[i], x_[j])));
		case SIGMOID:
			return tanh(gamma_ * sparse_dot(x_[i], x_[j]) + coef0_);
		case PRECOMPUTED:
			return x_[i][(int)(x_[j][0].value)].value;
		default:
			return 0;
		}
	}
	void swap_rows(int i, int j)
	{
		swap_(x_[i], x_[j]);
		if (!xsq_.empty()) swap_(xsq_[i], xsq_[j]);
	}

private:
	int type_, degree_;
	double gamma_, coef0_;
	std::vector<const svm_node *> x_;
	std::vector<double> xsq_;
};

/* ------------------------------------------------------------------------ */
/* Q matrices, svm.cpp:1310-1495.                                           */
/* ------------------------------------------------------------------------ */
template <class Qf>
class QMat {
public:
	virtual ~QMat() {}
	virtual Qf *column(int i, int len) = 0;
	virtual const double *diag() const = 0;
	virtual void swap_index(int i, int j) = 0;
};

/* SVC_Q: Q_ij = y_i y_j K(i,j) (svm.cpp:1310-1358).  ONE_CLASS_Q is the
 * same with y == 1 (svm.cpp:1360-1403). */
template <class Qf>
class ClassQ : public QMat<Qf> {
public:
	ClassQ(const svm_problem &prob, const svm_parameter &p, const schar *y)
		: k_(prob.l, prob.x, p), cache_(prob.l, (long)(p.cache_size * (1 << 20))),
		  y_(prob.l), qd_(prob.l)
	{
		for (int i = 0; i < prob.l; i++) {
			y_[i] = y ? y[i] : 1;
			qd_[i] = k_.eval(i, i);
		}
	}
	Qf *column(int i, int len) override
	{
		Qf *d;
		int from = cache_.fetch(i, &d, len);
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1 && len - from > 2048)
		for (int j = from; j < len; j++)
			d[j] = (Qf)(y_[i] * y_[j] * k_.eval(i, j));
		return d;
	}
	const double *diag() const override { return qd_.data(); }
	void swap_index(int i, int j) override
	{
		cache_.swap_index(i, j);
		k_.swap_rows(i, j);
		swap_(y_[i], y_[j]);
		swap_(qd_[i], qd_[j]);
	}

private:
	KernelRows k_;
	RowCache<Qf> cache_;
	std::vector<schar> y_;
	std::vector<double> qd_;
};

/* SVR_Q: 2l variables over l kernel rows, Q_ij = s_i s_j K(r_i, r_j)
 * (svm.cpp:1405-1495). */
template <class Qf>
class RegressionQ : public QMat<Qf> {
public:
	RegressionQ(const svm_problem &prob, const svm_parameter &p)
		: l_(prob.l), k_(prob.l, prob.x, p), cache_(prob.l, (long)(p.cache_size * (1 << 20))),
		  sign_(2 * prob.l), real_(2 * prob.l), qd_(2 * prob.l)
	{
		for (int k = 0; k < l_; k++) {
			sign_[k] = 1;
			sign_[k + l_] = -1;
			real_[k] = k;
			real_[k + l_] = k;
			qd_[k] = k_.eval(k, k);
			qd_[k + l_] = qd_[k];
		}
		buf_[0].resize(2 * l_);
		buf_[1].resize(2 * l_);
	}
	Qf *column(int i, int len) override
	{
		Qf *d;
		int ri = real_[i];
		if (cache_.fetch(ri, &d, l_) < l_) {
#pragma omp parallel for schedule(static) num_threads(g_threads) if (g_threads > 1 && l_ > 2048)
			for (int j = 0; j < l_; j++) d[j] = (Qf)k_.eval(ri, j);
		}
		Qf *out = buf_[next_].data();
		next_ = 1 - next_;
		schar si = sign_[i];
		for (int j = 0; j < len; j++) out[j] = (Qf)si * (Qf)sign_[j] * d[real_[j]];
		return out;
	}
	const double *diag() const override { return qd_.data(); }
	void swap_index(int i, int j) override
	{
		swap_(sign_[i], sign_[j]);
		swap_(real_[i], real_[j]);
		swap_(qd_[i], qd_[j]);
	}

private:
	int l_;
	KernelRows k_;
	RowCache<Qf> cache_;
	std::vector<schar> sign_;
	std::vector<int> real_;
	std::vector<double> qd_;
	std::vector<Qf> buf_[2];
	int next_ = 0;
};

/* ------------------------------------------------------------------------ */
/* SMO solver (WSS3 of Fan, Chen & Lin, JMLR 6 (2005)), svm.cpp:386-1054.   */
/* ------------------------------------------------------------------------ */
struct SolveInfo {
	double obj, rho, upper_bound_p, upper_bound_n, r;
	int iters;
};

enum { AT_LOWER = 0, AT_UPPER = 1, IS_FREE = 2 };

template <class Qf>
class SMO {
public:
	virtual ~SMO() {}

	/* Solver::Solve, svm.cpp:515-830. */
	void run(int l, QMat<Qf> &Q, const double *p_in, const schar *y_in, double *alpha_io,
		 double Cp, double Cn, double eps, SolveInfo *si, int shrinking, std::vector<double> *G_out)
	{
		l_ = l;
		Q_ = &Q;
		QD_ = Q.diag();
		p_.assign(p_in, p_in + l);
		y_.assign(y_in, y_in + l);
		a_.assign(alpha_io, alpha_io + l);
		Cp_ = Cp;
		Cn_ = Cn;
		eps_ = eps;
		unshrink_ = false;

		status_.resize(l);
		for (int i = 0; i < l; i++) set_status(i);

		active_.resize(l);
		for (int i = 0; i < l; i++) active_[i] = i;
		n_active_ = l;

		/* gradient G = p + Q alpha and G_bar = sum_{alpha_i = C_i} C_i Q_i */
		G_.assign(p_.begin(), p_.end());
		Gbar_.assign(l, 0.0);
		for (int i = 0; i < l; i++) {
			if (at_lower(i)) continue;
			const Qf *Qi = Q.column(i, l);
			double ai = a_[i];
			for (int j = 0; j < l; j++) G_[j] += ai * Qi[j];
			if (at_upper(i))
				for (int j = 0; j < l; j++) Gbar_[j] += C_of(i) * Qi[j];
		}

		int iter = 0;
		int max_iter = std::max(10000000, l > INT_MAX / 100 ? INT_MAX : 100 * l);
		if (g_opt.max_iter > 0) max_iter = g_opt.max_iter;
		int countdown = std::min(l, 1000) + 1;
		trace_len_ = 0;

		while (iter < max_iter) {
			if (--countdown == 0) {
				countdown = std::min(l, 1000);
				if (shrinking) shrink();
				info(".");
			}
			int i, j;
			if (pick_pair(i, j) != 0) {
				rebuild_gradient();
				n_active_ = l;
				info("*");
				if (pick_pair(i, j) != 0) break;
				countdown = 1; /* shrink at the next iteration */
			}
			++iter;
			if (g_opt.trace_ij && trace_len_ < g_opt.trace_cap) {
				g_opt.trace_ij[2 * trace_len_] = active_[i];
				g_opt.trace_ij[2 * trace_len_ + 1] = active_[j];
				++trace_len_;
			}
			step(i, j);
		}

		if (iter >= max_iter) {
			if (n_active_ < l) {
				rebuild_gradient();
				n_active_ = l;
				info("*");
			}
			if (g_opt.max_iter <= 0) fprintf(stderr, "\nWARNING: reaching max number of iterations\n");
		}

		si->rho = rho();
		double v = 0;
		for (int i = 0; i < l; i++) v += a_[i] * (G_[i] + p_[i]);
		si->obj = v / 2;
		si->iters = iter;

		for (int i = 0; i < l; i++) alpha_io[active_[i]] = a_[i];
		if (G_out) {
			G_out->assign(l, 0.0);
			for (int i = 0; i < l; i++) (*G_out)[active_[i]] = G_[i];
		}
		si->upper_bound_p = Cp;
		si->upper_bound_n = Cn;
		info("\noptimization finished, #iter = %d\n", iter);
	}

	int trace_len() const { return trace_len_; }

protected:
	double C_of(int i) const { return (y_[i] > 0) ? Cp_ : Cn_; }
	void set_status(int i)
	{
		if (a_[i] >= C_of(i)) status_[i] = AT_UPPER;
		else if (a_[i] <= 0) status_[i] = AT_LOWER;
		else status_[i] = IS_FREE;
	}
	bool at_upper(int i) const { return status_[i] == AT_UPPER; }
	bool at_lower(int i) const { return status_[i] == AT_LOWER; }
	bool is_free(int i) const { return status_[i] == IS_FREE; }

	/* One SMO step on the pair (i, j): analytic two-variable update with
	 * clipping (svm.cpp:626-721), G update (:724-736), G_bar (:739-770). */
	void step(int i, int j)
	{
		Qf *Qi = Q_->column(i, n_active_);
		Qf *Qj = Q_->column(j, n_active_);
		double Ci = C_of(i), Cj = C_of(j);
		double ai_old = a_[i], aj_old = a_[j];

		if (y_[i] != y_[j]) {
			double quad = QD_[i] + QD_[j] + 2 * Qi[j];
			if (quad <= 0) quad = TAU;
			double delta = (-G_[i] - G_[j]) / quad;
			double diff = a_[i] - a_[j];
			a_[i] += delta;
			a_[j] += delta;
			if (diff > 0) {
				if (a_[j] < 0) { a_[j] = 0; a_[i] = diff; }
			} else {
				if (a_[i] < 0) { a_[i] = 0; a_[j] = -diff; }
			}
			if (diff > Ci - Cj) {
				if (a_[i] > Ci) { a_[i] = Ci; a_[j] = Ci - diff; }
			} else {
				if (a_[j] > Cj) { a_[j] = Cj; a_[i] = Cj + diff; }
			}
		} else {
			double quad = QD_[i] + QD_[j] - 2 * Qi[j];
			if (quad <= 0) quad = TAU;
			double delta = (G_[i] - G_[j]) / quad;
			double sum = a_[i] + a_[j];
			a_[i] -= delta;
			a_[j] += delta;
			if (sum > Ci) {
				if (a_[i] > Ci) { a_[i] = Ci; a_[j] = sum - Ci; }
			} else {
				if (a_[j] < 0) { a_[j] = 0; a_[i] = sum; }
			}
			if (sum > Cj) {
				if (a_[j] > Cj) { a_[j] = Cj; a_[i] = sum - Cj; }
			} else {
				if (a_[i] < 0) { a_[i] = 0; a_[j] = sum; }
			}
		}

		double dai = a_[i] - ai_old;
		double daj = a_[j] - aj_old;
		for (int k = 0; k < n_active_; k++) G_[k] += Qi[k] * dai + Qj[k] * daj;

		bool ui = at_upper(i), uj = at_upper(j);
		set_status(i);
		set_status(j);
		if (ui != at_upper(i)) {
			Qi = Q_->column(i, l_);
			if (ui) for (int k = 0; k < l_; k++) Gbar_[k] -= Ci * Qi[k];
			else    for (int k = 0; k < l_; k++) Gbar_[k] += Ci * Qi[k];
		}
		if (uj != at_upper(j)) {
			Qj = Q_->column(j, l_);
			if (uj) for (int k = 0; k < l_; k++) Gbar_[k] -= Cj * Qj[k];
			else    for (int k = 0; k < l_; k++) Gbar_[k] += Cj * Qj[k];
		}
	}

	void swap_vars(int i, int j)
	{
		Q_->swap_index(i, j);
		swap_(y_[i], y_[j]);
		swap_(G_[i], G_[j]);
		swap_(status_[i], status_[j]);
		swap_(a_[i], a_[j]);
		swap_(p_[i], p_[j]);
		swap_(active_[i], active_[j]);
		swap_(Gbar_[i], Gbar_[j]);
	}

	/* reconstruct_gradient, svm.cpp:473-513. */
	void rebuild_gradient()
	{
		if (n_active_ == l_) return;
		int nfree = 0;
		for (int j = n_active_; j < l_; j++) G_[j] = Gbar_[j] + p_[j];
		for (int j = 0; j < n_active_; j++)
			if (is_free(j)) nfree++;
		if (2 * nfree < n_active_) info("\nWARNING: using -h 0 may be faster\n");
		if ((long)nfree * l_ > 2L * n_active_ * (l_ - n_active_)) {
			for (int i = n_active_; i < l_; i++) {
				const Qf *Qi = Q_->column(i, n_active_);
				for (int j = 0; j < n_active_; j++)
					if (is_free(j)) G_[i] += a_[j] * Qi[j];
			}
		} else {
			for (int i = 0; i < n_active_; i++) {
				if (!is_free(i)) continue;
				const Qf *Qi = Q_->column(i, l_);
				double ai = a_[i];
				for (int j = n_active_; j < l_; j++) G_[j] += ai * Qi[j];
			}
		}
	}

	/* select_working_set, svm.cpp:833-930.  Ties go to the larger index
	 * (>= / <= comparisons in increasing t). */
	virtual int pick_pair(int &out_i, int &out_j)
	{
		double Gmax = -INF, Gmax2 = -INF, best = INF;
		int imax = -1, jmin = -1;
		for (int t = 0; t < n_active_; t++) {
			if (y_[t] == +1) {
				if (!at_upper(t) && -G_[t] >= Gmax) { Gmax = -G_[t]; imax = t; }
			} else {
				if (!at_lower(t) && G_[t] >= Gmax) { Gmax = G_[t]; imax = t; }
			}
		}
		int i = imax;
		const Qf *Qi = (i != -1) ? Q_->column(i, n_active_) : nullptr;
		for (int j = 0; j < n_active_; j++) {
			if (y_[j] == +1) {
				if (at_lower(j)) continue;
				double gd = Gmax + G_[j];
				if (G_[j] >= Gmax2) Gmax2 = G_[j];
				if (gd > 0) {
					double quad = QD_[i] + QD_[j] - 2.0 * y_[i] * Qi[j];
					double od = (quad > 0) ? -(gd * gd) / quad : -(gd * gd) / TAU;
					if (od <= best) { jmin = j; best = od; }
				}
			} else {
				if (at_upper(j)) continue;
				double gd = Gmax - G_[j];
				if (-G_[j] >= Gmax2) Gmax2 = -G_[j];
				if (gd > 0) {
					double quad = QD_[i] + QD_[j] + 2.0 * y_[i] * Qi[j];
					double od = (quad > 0) ? -(gd * gd) / quad : -(gd * gd) / TAU;
					if (od <= best) { jmin = j; best = od; }
				}
			}
		}
		if (Gmax + Gmax2 < eps_) return 1;
		out_i = imax;
		out_j = jmin;
		return 0;
	}

	/* be_shrunk, svm.cpp:932-950. */
	bool prunable(int i, double g1, double g2) const
	{
		if (at_upper(i)) return (y_[i] == +1) ? (-G_[i] > g1) : (-G_[i] > g2);
		if (at_lower(i)) return (y_[i] == +1) ? (G_[i] > g2) : (G_[i] > g1);
		return false;
	}

	/* do_shrinking, svm.cpp:952-1011. */
	virtual void shrink()
	{
		double g1 = -INF, g2 = -INF;
		for (int i = 0; i < n_active_; i++) {
			if (y_[i] == +1) {
				if (!at_upper(i) && -G_[i] >= g1) g1 = -G_[i];
				if (!at_lower(i) && G_[i] >= g2) g2 = G_[i];
			} else {
				if (!at_upper(i) && -G_[i] >= g2) g2 = -G_[i];
				if (!at_lower(i) && G_[i] >= g1) g1 = G_[i];
			}
		}
		if (!unshrink_ && g1 + g2 <= eps_ * 10) {
			unshrink_ = true;
			rebuild_gradient();
			n_active_ = l_;
			info("*");
		}
		compact([&](int i) { return prunable(i, g1, g2); });
	}

	/* Moves every prunable variable behind n_active_ (svm.cpp:999-1010). */
	template <class F>
	void compact(F prune)
	{
		for (int i = 0; i < n_active_; i++) {
			if (!prune(i)) continue;
			n_active_--;
			while (n_active_ > i) {
				if (!prune(n_active_)) {
					swap_vars(i, n_active_);
					break;
				}
				n_active_--;
			}
		}
	}

	/* calculate_rho, svm.cpp:1013-1049. */
	virtual double rho()
	{
		int nfree = 0;
		double ub = INF, lb = -INF, sum = 0;
		for (int i = 0; i < n_active_; i++) {
			double yG = y_[i] * G_[i];
			if (at_upper(i)) {
				if (y_[i] == -1) ub = std::min(ub, yG);
				else lb = std::max(lb, yG);
			} else if (at_lower(i)) {
				if (y_[i] == +1) ub = std::min(ub, yG);
				else lb = std::max(lb, yG);
			} else {
				++nfree;
				sum += yG;
			}
		}
		return nfree > 0 ? sum / nfree : (ub + lb) / 2;
	}

	int l_ = 0, n_active_ = 0, trace_len_ = 0;
	QMat<Qf> *Q_ = nullptr;
	const double *QD_ = nullptr;
	std::vector<double> p_, a_, G_, Gbar_;
	std::vector<schar> y_;
	std::vector<char> status_;
	std::vector<int> active_;
	double Cp_ = 0, Cn_ = 0, eps_ = 0;
	bool unshrink_ = false;
};

/* Solver_NU: same step, working set restricted to y_i == y_j, two rho
 * halves (svm.cpp:1056-1308). */
template <class Qf>
class SMO_NU : public SMO<Qf> {
	using B = SMO<Qf>;

public:
	double r_ = 0;

protected:
	int pick_pair(int &out_i, int &out_j) override
	{
		double Gp = -INF, Gp2 = -INF, Gn = -INF, Gn2 = -INF, best = INF;
		int ip = -1, in = -1, jmin = -1;
		for (int t = 0; t < B::n_active_; t++) {
			if (B::y_[t] == +1) {
				if (!B::at_upper(t) && -B::G_[t] >= Gp) { Gp = -B::G_[t]; ip = t; }
			} else {
				if (!B::at_lower(t) && B::G_[t] >= Gn) { Gn = B::G_[t]; in = t; }
			}
		}
		const Qf *Qp = (ip != -1) ? B::Q_->column(ip, B::n_active_) : nullptr;
		const Qf *Qn = (in != -1) ? B::Q_->column(in, B::n_active_) : nullptr;
		for (int j = 0; j < B::n_active_; j++) {
			if (B::y_[j] == +1) {
				if (B::at_lower(j)) continue;
				double gd = Gp + B::G_[j];
				if (B::G_[j] >= Gp2) Gp2 = B::G_[j];
				if (gd > 0) {
					double quad = B::QD_[ip] + B::QD_[j] - 2 * Qp[j];
					double od = (quad > 0) ? -(gd * gd) / quad : -(gd * gd) / TAU;
					if (od <= best) { jmin = j; best = od; }
				}
			} else {
				if (B::at_upper(j)) continue;
				double gd = Gn - B::G_[j];
				if (-B::G_[j] >= Gn2) Gn2 = -B::G_[j];
				if (gd > 0) {
					double quad = B::QD_[in] + B::QD_[j] - 2 * Qn[j];
					double od = (quad > 0) ? -(gd * gd) / quad : -(gd * gd) / TAU;
					if (od <= best) { jmin = j; best = od; }
				}
			}
		}
		if (std::max(Gp + Gp2, Gn + Gn2) < B::eps_) return 1;
		out_i = (B::y_[jmin] == +1) ? ip : in;
		out_j = jmin;
		return 0;
	}

	bool prunable4(int i, double g1, double g2, double g3, double g4) const
	{
		if (B::at_upper(i)) return (B::y_[i] == +1) ? (-B::G_[i] > g1) : (-B::G_[i] > g4);
		if (B::at_lower(i)) return (B::y_[i] == +1) ? (B::G_[i] > g2) : (B::G_[i] > g3);
		return false;
	}

	void shrink() override
	{
		double g1 = -INF, g2 = -INF, g3 = -INF, g4 = -INF;
		for (int i = 0; i < B::n_active_; i++) {
			if (!B::at_upper(i)) {
				if (B::y_[i] == +1) { if (-B::G_[i] > g1) g1 = -B::G_[i]; }
				else if (-B::G_[i] > g4) g4 = -B::G_[i];
			}
			if (!B::at_lower(i)) {
				if (B::y_[i] == +1) { if (B::G_[i] > g2) g2 = B::G_[i]; }
				else if (B::G_[i] > g3) g3 = B::G_[i];
			}
		}
		if (!B::unshrink_ && std::max(g1 + g2, g3 + g4) <= B::eps_ * 10) {
			B::unshrink_ = true;
			B::rebuild_gradient();
			B::n_active_ = B::l_;
		}
		B::compact([&](int i) { return prunable4(i, g1, g2, g3, g4); });
	}

	double rho() override
	{
		int nf1 = 0, nf2 = 0;
		double ub1 = INF, ub2 = INF, lb1 = -INF, lb2 = -INF, s1 = 0, s2 = 0;
		for (int i = 0; i < B::n_active_; i++) {
			double g = B::G_[i];
			if (B::y_[i] == +1) {
				if (B::at_upper(i)) lb1 = std::max(lb1, g);
				else if (B::at_lower(i)) ub1 = std::min(ub1, g);
				else { ++nf1; s1 += g; }
			} else {
				if (B::at_upper(i)) lb2 = std::max(lb2, g);
				else if (B::at_lower(i)) ub2 = std::min(ub2, g);
				else { ++nf2; s2 += g; }
			}
		}
		double r1 = nf1 > 0 ? s1 / nf1 : (ub1 + lb1) / 2;
		double r2 = nf2 > 0 ? s2 / nf2 : (ub2 + lb2) / 2;
		r_ = (r1 + r2) / 2;
		return (r1 - r2) / 2;
	}
};

/* ------------------------------------------------------------------------ */
/* Formulations, svm.cpp:1497-1705.  Each returns the solver's alpha in     */
/* alpha_raw and fills the signed coefficients in alpha[l].                 */
/* ------------------------------------------------------------------------ */
struct OneSolve {
	std::vector<double> alpha;      /* [l] coefficients */
	std::vector<double> alpha_raw;  /* [n_var] */
	std::vector<double> G;          /* [n_var] */
	SolveInfo si;
	int trace_len = 0;
};

template <class Qf>
static void formulate_and_solve(const svm_problem *prob, const svm_parameter *param, double Cp, double Cn,
				OneSolve &out)
{
	int l = prob->l;
	out.alpha.assign(l, 0.0);
	out.si = SolveInfo{0, 0, 0, 0, 0, 0};
	switch (param->svm_type) {
	case C_SVC: { /* solve_c_svc, svm.cpp:1500-1533 */
		std::vector<double> a(l, 0.0), p(l, -1.0);
		std::vector<schar> y(l);
		for (int i = 0; i < l; i++) y[i] = prob->y[i] > 0 ? +1 : -1;
		ClassQ<Qf> Q(*prob, *param, y.data());
		SMO<Qf> s;
		s.run(l, Q, p.data(), y.data(), a.data(), Cp, Cn, param->eps, &out.si, param->shrinking, &out.G);
		out.trace_len = s.trace_len();
		out.alpha_raw = a;
		double sum = 0;
		for (int i = 0; i < l; i++) sum += a[i];
		if (Cp == Cn) info("nu = %f\n", sum / (Cp * prob->l));
		for (int i = 0; i < l; i++) out.alpha[i] = a[i] * y[i];
		break;
	}
	case NU_SVC: { /* solve_nu_svc, svm.cpp:1535-1588 */
		double nu = param->nu;
		std::vector<schar> y(l);
		for (int i = 0; i < l; i++) y[i] = prob->y[i] > 0 ? +1 : -1;
		double pos = nu * l / 2, neg = nu * l / 2;
		std::vector<double> a(l), zeros(l, 0.0);
		for (int i = 0; i < l; i++) {
			if (y[i] == +1) { a[i] = std::min(1.0, pos); pos -= a[i]; }
			else            { a[i] = std::min(1.0, neg); neg -= a[i]; }
		}
		ClassQ<Qf> Q(*prob, *param, y.data());
		SMO_NU<Qf> s;
		s.run(l, Q, zeros.data(), y.data(), a.data(), 1.0, 1.0, param->eps, &out.si, param->shrinking, &out.G);
		out.trace_len = s.trace_len();
		out.alpha_raw = a;
		double r = s.r_;
		out.si.r = r;
		info("C = %f\n", 1 / r);
		for (int i = 0; i < l; i++) out.alpha[i] = a[i] * y[i] / r;
		out.si.rho /= r;
		out.si.obj /= (r * r);
		out.si.upper_bound_p = 1 / r;
		out.si.upper_bound_n = 1 / r;
		break;
	}
	case ONE_CLASS: { /* solve_one_class, svm.cpp:1590-1621 */
		int n = (int)(param->nu * prob->l);
		std::vector<double> a(l, 0.0), zeros(l, 0.0);
		std::vector<schar> ones(l, 1);
		for (int i = 0; i < n; i++) a[i] = 1;
		if (n < prob->l) a[n] = param->nu * prob->l - n;
		ClassQ<Qf> Q(*prob, *param, nullptr);
		SMO<Qf> s;
		s.run(l, Q, zeros.data(), ones.data(), a.data(), 1.0, 1.0, param->eps, &out.si, param->shrinking, &out.G);
		out.trace_len = s.trace_len();
		out.alpha_raw = a;
		out.alpha = a;
		break;
	}
	case EPSILON_SVR: { /* solve_epsilon_svr, svm.cpp:1623-1659 */
		std::vector<double> a2(2 * l, 0.0), lin(2 * l);
		std::vector<schar> y(2 * l);
		for (int i = 0; i < l; i++) {
			lin[i] = param->p - prob->y[i];
			y[i] = 1;
			lin[i + l] = param->p + prob->y[i];
			y[i + l] = -1;
		}
		RegressionQ<Qf> Q(*prob, *param);
		SMO<Qf> s;
		s.run(2 * l, Q, lin.data(), y.data(), a2.data(), param->C, param->C, param->eps, &out.si,
		      param->shrinking, &out.G);
		out.trace_len = s.trace_len();
		out.alpha_raw = a2;
		double sum = 0;
		for (int i = 0; i < l; i++) {
			out.alpha[i] = a2[i] - a2[i + l];
			sum += fabs(out.alpha[i]);
		}
		info("nu = %f\n", sum / (param->C * l));
		break;
	}
	case NU_SVR: { /* solve_nu_svr, svm.cpp:1661-1696 */
		double C = param->C;
		std::vector<double> a2(2 * l), lin(2 * l);
		std::vector<schar> y(2 * l);
		double sum = C * param->nu * l / 2;
		for (int i = 0; i < l; i++) {
			a2[i] = a2[i + l] = std::min(sum, C);
			sum -= a2[i];
			lin[i] = -prob->y[i];
			y[i] = 1;
			lin[i + l] = prob->y[i];
			y[i + l] = -1;
		}
		RegressionQ<Qf> Q(*prob, *param);
		SMO_NU<Qf> s;
		s.run(2 * l, Q, lin.data(), y.data(), a2.data(), C, C, param->eps, &out.si, param->shrinking, &out.G);
		out.trace_len = s.trace_len();
		out.si.r = s.r_;
		out.alpha_raw = a2;
		info("epsilon = %f\n", -s.r_);
		for (int i = 0; i < l; i++) out.alpha[i] = a2[i] - a2[i + l];
		break;
	}
	}
}

/* svm_train_one, svm.cpp:1707-1769 (minus the device-solver hooks). */
static void train_one(const svm_problem *prob, const svm_parameter *param, double Cp, double Cn, OneSolve &out)
{
	if (g_opt.q_double) formulate_and_solve<double>(prob, param, Cp, Cn, out);
	else formulate_and_solve<float>(prob, param, Cp, Cn, out);
	info("obj = %f, rho = %f\n", out.si.obj, out.si.rho);
	int nSV = 0, nBSV = 0;
	for (int i = 0; i < prob->l; i++) {
		if (fabs(out.alpha[i]) > 0) {
			++nSV;
			double ub = prob->y[i] > 0 ? out.si.upper_bound_p : out.si.upper_bound_n;
			if (fabs(out.alpha[i]) >= ub) ++nBSV;
		}
	}
	info("nSV = %d, nBSV = %d\n", nSV, nBSV);
}

/* svm_group_classes, svm.cpp:2081-2157. */
static void group_classes(const svm_problem *prob, std::vector<int> &label, std::vector<int> &start,
			  std::vector<int> &count, std::vector<int> &perm)
{
	int l = prob->l;
	label.clear();
	count.clear();
	std::vector<int> cls(l);
	for (int i = 0; i < l; i++) {
		int lab = (int)prob->y[i];
		size_t c = 0;
		while (c < label.size() && label[c] != lab) c++;
		if (c == label.size()) { label.push_back(lab); count.push_back(0); }
		count[c]++;
		cls[i] = (int)c;
	}
	int k = (int)label.size();
	if (k == 2 && label[0] == -1 && label[1] == 1) {
		swap_(label[0], label[1]);
		swap_(count[0], count[1]);
		for (int i = 0; i < l; i++) cls[i] = 1 - cls[i];
	}
	start.assign(k, 0);
	for (int c = 1; c < k; c++) start[c] = start[c - 1] + count[c - 1];
	perm.assign(l, 0);
	std::vector<int> fill(start);
	for (int i = 0; i < l; i++) perm[fill[cls[i]]++] = i;
}

static svm_model *train(const svm_problem *prob, const svm_parameter *param)
{
	svm_model *model = (svm_model *)calloc(1, sizeof(svm_model));
	model->param = *param;
	model->free_sv = 0;

	if (param->svm_type == ONE_CLASS || param->svm_type == EPSILON_SVR || param->svm_type == NU_SVR) {
		/* svm.cpp:2166-2206 */
		model->nr_class = 2;
		model->sv_coef = (double **)malloc(sizeof(double *));
		OneSolve f;
		train_one(prob, param, 0, 0, f);
		model->rho = (double *)malloc(sizeof(double));
		model->rho[0] = f.si.rho;
		int nSV = 0;
		for (int i = 0; i < prob->l; i++)
			if (fabs(f.alpha[i]) > 0) ++nSV;
		model->l = nSV;
		model->SV = (svm_node **)malloc(sizeof(svm_node *) * std::max(nSV, 1));
		model->sv_coef[0] = (double *)malloc(sizeof(double) * std::max(nSV, 1));
		model->sv_indices = (int *)malloc(sizeof(int) * std::max(nSV, 1));
		int j = 0;
		for (int i = 0; i < prob->l; i++) {
			if (fabs(f.alpha[i]) > 0) {
				model->SV[j] = prob->x[i];
				model->sv_coef[0][j] = f.alpha[i];
				model->sv_indices[j] = i + 1;
				++j;
			}
		}
		return model;
	}

	/* classification, svm.cpp:2207-2404 */
	int l = prob->l;
	std::vector<int> label, start, count, perm;
	group_classes(prob, label, start, count, perm);
	int k = (int)label.size();
	if (k == 1) info("WARNING: training data in only one class. See README for details.\n");
	std::vector<svm_node *> x(l);
	for (int i = 0; i < l; i++) x[i] = prob->x[perm[i]];

	std::vector<double> wC(k, param->C);
	for (int w = 0; w < param->nr_weight; w++) {
		int c = 0;
		while (c < k && param->weight_label[w] != label[c]) c++;
		if (c == k) fprintf(stderr, "WARNING: class label %d specified in weight is not found\n", param->weight_label[w]);
		else wC[c] *= param->weight[w];
	}

	std::vector<char> nonzero(l, 0);
	int npair = k * (k - 1) / 2;
	std::vector<OneSolve> f(npair);
	int p = 0;
	for (int a = 0; a < k; a++) {
		for (int b = a + 1; b < k; b++) {
			int sa = start[a], sb = start[b], ca = count[a], cb = count[b];
			svm_problem sub;
			sub.l = ca + cb;
			std::vector<svm_node *> sx(sub.l);
			std::vector<double> sy(sub.l);
			for (int t = 0; t < ca; t++) { sx[t] = x[sa + t]; sy[t] = +1; }
			for (int t = 0; t < cb; t++) { sx[ca + t] = x[sb + t]; sy[ca + t] = -1; }
			sub.x = sx.data();
			sub.y = sy.data();
			train_one(&sub, param, wC[a], wC[b], f[p]);
			for (int t = 0; t < ca; t++)
				if (fabs(f[p].alpha[t]) > 0) nonzero[sa + t] = 1;
			for (int t = 0; t < cb; t++)
				if (fabs(f[p].alpha[ca + t]) > 0) nonzero[sb + t] = 1;
			++p;
		}
	}

	model->nr_class = k;
	model->label = (int *)malloc(sizeof(int) * k);
	for (int c = 0; c < k; c++) model->label[c] = label[c];
	model->rho = (double *)malloc(sizeof(double) * std::max(npair, 1));
	for (int q = 0; q < npair; q++) model->rho[q] = f[q].si.rho;
	model->probA = nullptr;
	model->probB = nullptr;

	int total = 0;
	std::vector<int> nz_count(k);
	model->nSV = (int *)malloc(sizeof(int) * k);
	for (int c = 0; c < k; c++) {
		int n = 0;
		for (int t = 0; t < count[c]; t++)
			if (nonzero[start[c] + t]) { ++n; ++total; }
		model->nSV[c] = n;
		nz_count[c] = n;
	}
	info("Total nSV = %d\n", total);
	model->l = total;
	model->SV = (svm_node **)malloc(sizeof(svm_node *) * std::max(total, 1));
	model->sv_indices = (int *)malloc(sizeof(int) * std::max(total, 1));
	p = 0;
	for (int i = 0; i < l; i++) {
		if (nonzero[i]) {
			model->SV[p] = x[i];
			model->sv_indices[p++] = perm[i] + 1;
		}
	}
	std::vector<int> nz_start(k, 0);
	for (int c = 1; c < k; c++) nz_start[c] = nz_start[c - 1] + nz_count[c - 1];
	model->sv_coef = (double **)malloc(sizeof(double *) * std::max(k - 1, 1));
	for (int c = 0; c < k - 1; c++) model->sv_coef[c] = (double *)malloc(sizeof(double) * std::max(total, 1));
	p = 0;
	for (int a = 0; a < k; a++) {
		for (int b = a + 1; b < k; b++) {
			/* classifier (a,b): coefficients of class a go to sv_coef[b-1],
			 * those of class b to sv_coef[a] (svm.cpp:2364-2384) */
			int q = nz_start[a];
			for (int t = 0; t < count[a]; t++)
				if (nonzero[start[a] + t]) model->sv_coef[b - 1][q++] = f[p].alpha[t];
			q = nz_start[b];
			for (int t = 0; t < count[b]; t++)
				if (nonzero[start[b] + t]) model->sv_coef[a][q++] = f[p].alpha[count[a] + t];
			++p;
		}
	}
	return model;
}

/* svm_predict_values, svm.cpp:2569-2643. */
static double predict_values(const svm_model *m, const svm_node *x, double *dec)
{
	if (m->param.svm_type == ONE_CLASS || m->param.svm_type == EPSILON_SVR || m->param.svm_type == NU_SVR) {
		const double *c = m->sv_coef[0];
		double s = 0;
		for (int i = 0; i < m->l; i++) s += c[i] * k_function(x, m->SV[i], m->param);
		s -= m->rho[0];
		*dec = s;
		if (m->param.svm_type == ONE_CLASS) return (s > 0) ? 1 : -1;
		return s;
	}
	int k = m->nr_class, l = m->l;
	std::vector<double> kv(l);
	for (int i = 0; i < l; i++) kv[i] = k_function(x, m->SV[i], m->param);
	std::vector<int> start(k, 0), vote(k, 0);
	for (int c = 1; c < k; c++) start[c] = start[c - 1] + m->nSV[c - 1];
	int p = 0;
	for (int a = 0; a < k; a++) {
		for (int b = a + 1; b < k; b++) {
			double s = 0;
			const double *ca = m->sv_coef[b - 1];
			const double *cb = m->sv_coef[a];
			for (int t = 0; t < m->nSV[a]; t++) s += ca[start[a] + t] * kv[start[a] + t];
			for (int t = 0; t < m->nSV[b]; t++) s += cb[start[b] + t] * kv[start[b] + t];
			s -= m->rho[p];
			dec[p] = s;
			if (dec[p] > 0) ++vote[a];
			else ++vote[b];
			p++;
		}
	}
	int best = 0;
	for (int c = 1; c < k; c++)
		if (vote[c] > vote[best]) best = c;
	return m->label[best];
}

} /* namespace ora */

/* ------------------------------------------------------------------------ */
/* exported C API                                                           */
/* ------------------------------------------------------------------------ */
extern "C" {

void ora_free_model(svm_model *m);

void ora_set_options(const ora_options *opt)
{
	if (opt) g_opt = *opt;
	else g_opt = ora_options{0, 0, 0, nullptr};
}

void ora_set_verbose(int verbose) { g_print = verbose ? &print_to_stdout : &print_nothing; }

void ora_set_threads(int threads) { g_threads = threads > 1 ? threads : 1; }

/* One svm_train_one solve with raw outputs.  Caller owns out->alpha
 * (length prob->l) and may pass alpha_raw / G of length n_var (l or 2l)
 * or null.  Returns 0 on success. */
int ora_train_one(const svm_problem *prob, const svm_parameter *param, double Cp, double Cn, ora_result *out)
{
	ora::OneSolve f;
	ora::train_one(prob, param, Cp, Cn, f);
	int l = prob->l;
	int nvar = (int)f.alpha_raw.size();
	out->l = l;
	out->n_var = nvar;
	if (out->alpha) memcpy(out->alpha, f.alpha.data(), sizeof(double) * l);
	if (out->alpha_raw) memcpy(out->alpha_raw, f.alpha_raw.data(), sizeof(double) * nvar);
	if (out->G) memcpy(out->G, f.G.data(), sizeof(double) * nvar);
	out->rho = f.si.rho;
	out->obj = f.si.obj;
	out->r = f.si.r;
	out->upper_bound_p = f.si.upper_bound_p;
	out->upper_bound_n = f.si.upper_bound_n;
	out->iters = f.si.iters;
	out->trace_len = f.trace_len;
	int nSV = 0, nBSV = 0;
	for (int i = 0; i < l; i++) {
		if (fabs(f.alpha[i]) > 0) {
			++nSV;
			double ub = prob->y[i] > 0 ? f.si.upper_bound_p : f.si.upper_bound_n;
			if (fabs(f.alpha[i]) >= ub) ++nBSV;
		}
	}
	out->nSV = nSV;
	out->nBSV = nBSV;
	return 0;
}

/* Row i of the kernel matrix K(i, 0..l-1) exactly as the solver caches it
 * (Qfloat-rounded, no label signs); used to pin device row evaluation. */
void ora_kernel_row(const svm_problem *prob, const svm_parameter *param, int i, float *out)
{
	ora::KernelRows k(prob->l, prob->x, *param);
	for (int j = 0; j < prob->l; j++) out[j] = (float)k.eval(i, j);
}

/* Same row in double before rounding. */
void ora_kernel_row_f64(const svm_problem *prob, const svm_parameter *param, int i, double *out)
{
	ora::KernelRows k(prob->l, prob->x, *param);
	for (int j = 0; j < prob->l; j++) out[j] = k.eval(i, j);
}

double ora_k_function(const svm_node *x, const svm_node *y, const svm_parameter *param)
{
	return ora::k_function(x, y, *param);
}

svm_model *ora_svm_train(const svm_problem *prob, const svm_parameter *param) { return ora::train(prob, param); }

double ora_svm_predict_values(const svm_model *m, const svm_node *x, double *dec) { return ora::predict_values(m, x, dec); }

double ora_svm_predict(const svm_model *m, const svm_node *x)
{
	int k = m->nr_class;
	bool reg = m->param.svm_type == ONE_CLASS || m->param.svm_type == EPSILON_SVR || m->param.svm_type == NU_SVR;
	std::vector<double> dec(reg ? 1 : std::max(1, k * (k - 1) / 2));
	return ora::predict_values(m, x, dec.data());
}

/* svm_cross_validation with caller-supplied fold ids (fold[i] in
 * [0, nr_fold)); replaces the glibc rand() shuffle of svm.cpp:2407-2525
 * so that device and oracle see the same folds. */
void ora_cross_validation_folds(const svm_problem *prob, const svm_parameter *param, int nr_fold, const int *fold,
				double *target)
{
	int l = prob->l;
	for (int f = 0; f < nr_fold; f++) {
		std::vector<svm_node *> sx;
		std::vector<double> sy;
		for (int i = 0; i < l; i++)
			if (fold[i] != f) { sx.push_back(prob->x[i]); sy.push_back(prob->y[i]); }
		svm_problem sub;
		sub.l = (int)sx.size();
		sub.x = sx.data();
		sub.y = sy.data();
		svm_model *m = ora::train(&sub, param);
		for (int i = 0; i < l; i++)
			if (fold[i] == f) target[i] = ora_svm_predict(m, prob->x[i]);
		/* free the model (SVs point into prob) */
		for (int c = 0; c < m->nr_class - 1; c++) free(m->sv_coef[c]);
		free(m->sv_coef);
		free(m->SV);
		free(m->rho);
		free(m->label);
		free(m->nSV);
		free(m->sv_indices);
		free(m);
	}
}

/* svm_cross_validation with the reference's rand()-drawn stratified folds
 * (libsvm/svm.cpp:2407-2525). */
void ora_svm_cross_validation(const svm_problem *prob, const svm_parameter *param, int nr_fold, double *target)
{
	int l = prob->l;
	if (nr_fold > l) nr_fold = l;
	std::vector<int> perm(l), fold_start(nr_fold + 1);
	if ((param->svm_type == C_SVC || param->svm_type == NU_SVC) && nr_fold < l) {
		std::vector<int> label, start, count;
		ora::group_classes(prob, label, start, count, perm);
		int k = (int)label.size();
		std::vector<int> index(perm), fold_count(nr_fold, 0);
		for (int c = 0; c < k; c++)
			for (int i = 0; i < count[c]; i++) {
				int j = i + rand() % (count[c] - i);
				std::swap(index[start[c] + j], index[start[c] + i]);
			}
		for (int i = 0; i < nr_fold; i++)
			for (int c = 0; c < k; c++) fold_count[i] += (i + 1) * count[c] / nr_fold - i * count[c] / nr_fold;
		fold_start[0] = 0;
		for (int i = 1; i <= nr_fold; i++) fold_start[i] = fold_start[i - 1] + fold_count[i - 1];
		for (int c = 0; c < k; c++)
			for (int i = 0; i < nr_fold; i++) {
				int b = start[c] + i * count[c] / nr_fold, e = start[c] + (i + 1) * count[c] / nr_fold;
				for (int j = b; j < e; j++) perm[fold_start[i]++] = index[j];
			}
		fold_start[0] = 0;
		for (int i = 1; i <= nr_fold; i++) fold_start[i] = fold_start[i - 1] + fold_count[i - 1];
	} else {
		for (int i = 0; i < l; i++) perm[i] = i;
		for (int i = 0; i < l; i++) std::swap(perm[i], perm[i + rand() % (l - i)]);
		for (int i = 0; i <= nr_fold; i++) fold_start[i] = i * l / nr_fold;
	}
	for (int f = 0; f < nr_fold; f++) {
		std::vector<svm_node *> sx;
		std::vector<double> sy;
		for (int j = 0; j < fold_start[f]; j++) { sx.push_back(prob->x[perm[j]]); sy.push_back(prob->y[perm[j]]); }
		for (int j = fold_start[f + 1]; j < l; j++) { sx.push_back(prob->x[perm[j]]); sy.push_back(prob->y[perm[j]]); }
		svm_problem sub;
		sub.l = (int)sx.size();
		sub.x = sx.data();
		sub.y = sy.data();
		svm_model *m = ora::train(&sub, param);
		for (int j = fold_start[f]; j < fold_start[f + 1]; j++) target[perm[j]] = ora_svm_predict(m, prob->x[perm[j]]);
		ora_free_model(m);
	}
}

void ora_free_model(svm_model *m)
{
	if (!m) return;
	if (m->sv_coef)
		for (int c = 0; c < m->nr_class - 1; c++) free(m->sv_coef[c]);
	free(m->sv_coef);
	free(m->SV);
	free(m->rho);
	free(m->label);
	free(m->nSV);
	free(m->sv_indices);
	free(m->probA);
	free(m->probB);
	free(m);
}

} /* extern "C" */
