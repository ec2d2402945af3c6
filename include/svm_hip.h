/*
 * libsvm-mi355x device-solver C-ABI.
 *
 * Replaces the reference's process-global C++ object
 *     CudaSolver *cudaSolver        (libsvm/cuda_solver.h:224-246)
 *     class CudaSolverNU            (libsvm/cuda_solverNU.h:24-57)
 * with an opaque, handle-scoped solver: every piece of device state lives in
 * the handle (no __constant__ / __device__ globals), so several solves can
 * coexist in one process and none of them resets the device.
 *
 * Errors: every entry point returns an svmhip_status; nothing throws across
 * the ABI, nothing calls exit() or hipDeviceReset() (the reference does both,
 * libsvm/svm_defs.h:108-117, libsvm/cuda_solver.cpp:439-454).  The message of
 * the last failure on a handle is available from svmhip_last_error().
 *
 * Threading: a handle may be used by one host thread at a time; distinct
 * handles are independent.  Each handle owns one HIP stream.
 *
 * Shrinking: the device solver always works on the full variable set (the
 * reference path requires shrinking=0, svm-train/svm-train.c:260-262);
 * svm_parameter.shrinking is accepted and ignored.
 */
#ifndef LIBSVM_MI355X_SVM_HIP_H
#define LIBSVM_MI355X_SVM_HIP_H

#include <stddef.h>

#include "svm.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
	SVMHIP_OK = 0,
	SVMHIP_ERR_ARG = 1,       /* invalid argument / parameter */
	SVMHIP_ERR_DEVICE = 2,    /* HIP runtime error or no usable device */
	SVMHIP_ERR_NOMEM = 3,     /* device allocation failed */
	SVMHIP_ERR_STATE = 4,     /* call out of order (e.g. solve before setup) */
	SVMHIP_ERR_TIMEOUT = 5,   /* a device-side wait exceeded its spin limit */
	SVMHIP_ERR_MAXITER = 6    /* informational: iteration cap reached */
} svmhip_status;

/* Which working-set rule the solver uses: C_SVC / ONE_CLASS / EPSILON_SVR use
 * the standard WSS3 (libsvm/cuda_solver.cpp:493-518); NU_SVC / NU_SVR use the
 * nu rule (libsvm/cuda_solverNU.cpp:152-178). */
typedef enum { SVMHIP_KIND_C = 0, SVMHIP_KIND_NU = 1 } svmhip_kind;

/* Device options kept out of svm_parameter so the kept struct stays
 * byte-identical.  Zero-initialise and set struct_size; zero means "default". */
typedef struct {
	size_t struct_size;       /* sizeof(svmhip_options) */
	int device;               /* HIP device ordinal (default 0) */
	int grid_blocks;          /* persistent-grid workgroups (0 = the launch table written by the
				     autotuner, libsvm-cuda_amd/tune/launch_table.txt, else the size
				     rule: one per CU on large problems, fewer on mid-size ones) */
	int block_threads;        /* threads per workgroup (0 = 512, the only size built) */
	int max_iter;             /* 0 = LIBSVM rule max(10^7, 100*n_var), n_var = 2l for regression */
	double cache_bytes;       /* kernel-row cache bytes (0 = param.cache_size MB) */
	int iters_per_launch;     /* persistent-kernel iterations per launch (0 = all) */
	int verbose;              /* 1 = print setup and cache statistics to stderr */
	int speculative_rows;     /* speculative kernel rows: on a row miss also compute the rows of the
				     strongest other candidates in the same pass over X; results are
				     identical either way.  svmhip_solve (C rule, full cache): on only
				     when 1 (slower on the measured single-solve configurations).
				     Batched and serving solves (single-workgroup C-rule subproblems
				     with a full cache): on unless -1 */
	int mfma_rows;            /* svmhip_solve on dense rows (not PRECOMPUTED) whose whole kernel
				     matrix fits the row cache: 1 = compute all of it at the start of
				     each solve as Gram tiles on the matrix cores (FP16 hi/lo split
				     products, FP32 accumulation; kernel values within ~2e-6 relative of
				     the exact ones, median 5e-8), then run the SMO loop on cache hits;
				     0 = exact rows (FP64 dots) on demand, the oracle's arithmetic.
				     svmhip_predict_batch, models with one coefficient per support
				     vector (two classes, regression, one-class): 1 = dot products on
				     the same FP16 hi/lo Gram tiles, kernel values and sums in FP64
				     (decision values within ~1e-6 relative); other models, values out
				     of FP16's range and svmhip_predict_batch_noalloc run the FP64
				     kernel */
	int exchange_copies;      /* svmhip_solve: copies of the per-iteration exchange records, 1..4
				     (0 = the launch table, else 4 with 64 or more workgroups and 1
				     below); results are identical for every value */
	double hbm_fraction;      /* batched solves: at most this fraction of the free device memory
				     goes to kernel-row cache arenas (0 = 0.7, and at most 32 GiB in
				     all unless cache_bytes or hbm_fraction is given); svm_train and
				     svm_cross_validation take it from the environment variable
				     SVMHIP_HBM_FRACTION */
} svmhip_options;

typedef struct svmhip_solver svmhip_solver;

/* Per-solve statistics (filled by svmhip_solve / svmhip_get_stats). */
typedef struct {
	int iters;                /* SMO iterations (pairs updated) */
	int optimal;              /* 1 = stopping rule met, 0 = iteration cap */
	double obj;               /* 0.5 * sum alpha_i (G_i + p_i), FP64 on device */
	double rho;               /* Solver::calculate_rho / Solver_NU::calculate_rho */
	double r;                 /* Solver_NU only */
	long long row_hits;       /* kernel-row cache hits */
	long long row_misses;     /* kernel rows computed */
	double solve_ms;          /* device time of the iteration loop (HIP events) */
	double kernel_ms;         /* device time of the persistent-kernel launches alone */
	int launches;             /* persistent-kernel launches of the last solve */
	double setup_ms;          /* device time of G initialisation */
	int cache_slots;          /* rows the cache holds */
	int grid_blocks;          /* workgroups used by the persistent kernel */
	double gram_ms;           /* device time of the Gram-tile kernel matrix (mfma_rows = 1; 0 if the
				     solve computed rows on demand); svmhip_batch_solve: device time of
				     the shared dot products on the call that built them, else 0 */
	long long device_allocs;  /* device allocations made by the handle since svmhip_create (a
				     repeated svmhip_setup + svmhip_solve over the same rows adds none
				     once the first solve has sized the buffers) */
	int resident_vectors;     /* 1 = the persistent kernel kept each workgroup's owned G, alpha, QD,
				     labels and statuses in LDS for the solve (they fit); 0 = in
				     device memory.  Results are identical either way */
	int shared_dots;          /* svmhip_batch_solve: 1 = kernel rows were gathers from the handle's
				     exact dot products of all pairs of base rows (dense problems whose
				     l x l FP64 matrix takes at most a quarter of the free device
				     memory; SVMHIP_SHARED_DOTS=0 turns it off).  Results are
				     identical either way */
} svmhip_stats;

/* CudaSolver(prob, param, quiet) / CudaSolverNU(...)
 * (libsvm/cuda_solver.cpp:434-448, load_problem_parameters :298-432):
 * uploads the rows of prob (features truncated to float32, like the
 * reference) in a SELL-64 layout and the kernel parameters. */
svmhip_status svmhip_create(const struct svm_problem *prob, const struct svm_parameter *param,
			    svmhip_kind kind, const svmhip_options *opt, svmhip_solver **out);

/* CudaSolver::setup_rbf_variables (libsvm/cuda_solver.cpp:241-259): squared
 * row norms for RBF.  Also computes them for every other kernel type as a
 * no-op-safe call; l must equal prob->l. */
svmhip_status svmhip_setup_rbf(svmhip_solver *h, int l);

/* CudaSolver::setup_solver (libsvm/cuda_solver.cpp:159-239): copies y, the
 * initial gradient G (= p, as Solver::Solve passes it, libsvm/svm.cpp:546-557),
 * alpha and alpha_status, computes QD, initialises G = p + Q alpha on the
 * device in FP64 and sizes the kernel-row cache.  Calling it again on the
 * same handle discards the previous solve's state and starts a new one over
 * the already uploaded rows (labels, C and alpha may change); the solver
 * vectors, the kernel-row cache and its bookkeeping stay allocated and are
 * reused when the new solve fits them (the cache's rows are invalidated).
 * n_var is l for classification / one-class and 2l for regression (SVR
 * variable i+l shares kernel row i). */
svmhip_status svmhip_setup(svmhip_solver *h, const signed char *y, const double *G, const double *alpha,
			   const char *alpha_status, double Cp, double Cn, int n_var);

/* CudaSolver::select_working_set / CudaSolverNU::select_working_set
 * (libsvm/cuda_solver.cpp:493-518, libsvm/cuda_solverNU.cpp:152-178).
 * *optimal = 1 when the stopping rule holds.  The pair stays on the device,
 * as in the reference; *out_i / *out_j (may be null) receive it for
 * inspection, which costs one extra small copy. */
svmhip_status svmhip_select_working_set(svmhip_solver *h, int n_var, int *optimal, int *out_i, int *out_j);

/* CudaSolver::compute_alpha (libsvm/cuda_solver.cpp:458-463): the two-variable
 * clipped update in FP64. */
svmhip_status svmhip_compute_alpha(svmhip_solver *h);

/* CudaSolver::update_gradient (libsvm/cuda_solver.cpp:520-531). */
svmhip_status svmhip_update_gradient(svmhip_solver *h, int n_var);

/* CudaSolver::update_alpha_status (libsvm/cuda_solver.cpp:465-470). */
svmhip_status svmhip_update_alpha_status(svmhip_solver *h);

/* CudaSolver::fetch_vectors (libsvm/cuda_solver.cpp:533-555): copies G,
 * alpha (FP64) and alpha_status back; any pointer may be null. */
svmhip_status svmhip_fetch(svmhip_solver *h, double *G, double *alpha, char *alpha_status, int n_var);

/* Fused solve: runs the whole SMO loop on the device (one persistent
 * cooperative launch per chunk of iterations) and computes rho / obj / r in
 * FP64 on the device.  No counterpart method in the reference: it replaces
 * the host-driven loop of Solver::Solve (libsvm/svm.cpp:580-776) when the
 * device solver is active.  max_iter = 0 uses the handle's option. */
svmhip_status svmhip_solve(svmhip_solver *h, int max_iter, svmhip_stats *stats);

/* rho / obj / r of the current state (after a step-wise loop). */
svmhip_status svmhip_finalize(svmhip_solver *h, svmhip_stats *stats);

svmhip_status svmhip_get_stats(svmhip_solver *h, svmhip_stats *stats);

/* Device kernel row K(i, 0..l-1) as the solver caches it (float, no label
 * signs), for testing the row evaluator: the exact row, or after a solve in
 * Gram-tile mode (mfma_rows = 1) the row the Gram tiles left in the cache.
 * out has l floats (host memory). */
svmhip_status svmhip_kernel_row(svmhip_solver *h, int i, float *out);

/* Destructor (libsvm/cuda_solver.cpp:450-454) without the device reset. */
void svmhip_destroy(svmhip_solver *h);

const char *svmhip_last_error(const svmhip_solver *h);

/* ---------------------------------------------------------------------------
 * Batched solves.  No counterpart in the reference, which trains the
 * k(k-1)/2 one-vs-one pairs and every cross-validation fold one after the
 * other (libsvm/svm.cpp:2246-2295, :2491-2521).  Here many independent
 * subproblems over row subsets of one problem are in flight together, each
 * solved by one workgroup or by a team of workgroups, each with its own
 * alpha/G/kernel-row cache.  Options: grid_blocks = largest team (0 = 32,
 * 1 = one workgroup per subproblem); iters_per_launch = iterations per
 * subproblem per launch (0 = sized by the scheduler); cache_bytes = arena
 * per workgroup (0 = an equal share of hbm_fraction, default 70%, of free
 * device memory, at most 32 GiB in all when hbm_fraction is not given either:
 * allocating most of the device memory again after freeing it takes
 * seconds); verbose = 1 prints the arena size chosen.
 * ------------------------------------------------------------------------- */

typedef struct svmhip_batch svmhip_batch;

/* One subproblem: a row set of the batch's problem plus one formulation, given
 * exactly as Solver::Solve receives it (libsvm/svm.cpp:515): n_var = n or 2n
 * variables (2n: regression, variable v+n uses row v), y = +1/-1 per variable,
 * p = linear term, alpha = initial point (overwritten with the solution). */
typedef struct {
	int rowset;               /* id from svmhip_batch_add_rowset */
	int kind;                 /* svmhip_kind */
	int n_var;
	const signed char *y;     /* [n_var] */
	const double *p;          /* [n_var] */
	double *alpha;            /* [n_var] in: initial point, out: solution */
	double Cp, Cn;
	double gamma;             /* kernel gamma of this subproblem */
	double eps;
	/* outputs */
	double rho, obj, r;
	int iters;
	int status;               /* svmhip_status of this subproblem */
} svmhip_subproblem;

/* Uploads every row of prob once (SELL-64, float32 features). */
svmhip_status svmhip_batch_create(const struct svm_problem *prob, const struct svm_parameter *param,
				  const svmhip_options *opt, svmhip_batch **out);

/* Registers a row set (indices into prob, any order, repeats not allowed);
 * its rows are packed on the device.  Returns the id (>= 0) or -status. */
int svmhip_batch_add_rowset(svmhip_batch *b, const int *rows, int n);

/* Solves all subproblems.  The launches run in stages: a subproblem pauses
 * at its stage's iteration cap and resumes in the next stage from the same
 * state (the iterates are those of an uninterrupted solve); unfinished
 * subproblems are re-planned into teams as the set shrinks.  stats (may be
 * null) sums iterations and cache counters; solve_ms is the device time of
 * all stages, launches the number of stages.  The device buffers of a solve
 * (the kernel-row cache arenas, by default 70% of the free device memory,
 * and the pooled solver state) stay with the handle for the next solve and
 * are released by svmhip_batch_destroy. */
svmhip_status svmhip_batch_solve(svmhip_batch *b, svmhip_subproblem *subs, int nsub, svmhip_stats *stats);

void svmhip_batch_destroy(svmhip_batch *b);

const char *svmhip_batch_last_error(const svmhip_batch *b);

/* Decision values of n rows against a trained model on the device: the
 * batched form of svm_predict_values (libsvm/svm.cpp:2569-2643), which the
 * reference only has on the CPU.  dec_values receives n x npair values
 * (npair = nr_class*(nr_class-1)/2, or 1 for one-class / regression),
 * labels (may be null) the predicted label per row as svm_predict returns
 * it.  Kernel values are computed in FP64 from the caller's FP64 features as
 * Kernel::k_function computes them (libsvm/svm.cpp:327-384), and each
 * decision value is summed in svm_predict_values' order.  PRECOMPUTED
 * kernels and rows wider than the prediction kernel's LDS allows (about
 * 19,000 features) return SVMHIP_ERR_ARG.  Runs on the device's null stream
 * and returns when the results are on the host. */
svmhip_status svmhip_predict_batch(const struct svm_model *model, const struct svm_node *const *x, int n,
				   const svmhip_options *opt, double *dec_values, double *labels);

/* Bytes of device scratch held by svmhip_predict_batch_noalloc. */
#define SVMHIP_NOALLOC_SCRATCH_BYTES (128u << 20)

/* svmhip_predict_batch without device allocations.  No counterpart in the
 * reference: for callers that cannot call hipMalloc (tight serving loops,
 * harnesses that forbid allocation).  Same arguments, same kernels, same
 * results as svmhip_predict_batch; the support vectors, coefficients, test
 * rows and decision values are staged in a device scratch area of
 * SVMHIP_NOALLOC_SCRATCH_BYTES (128 MiB) that the library holds on each
 * device it runs this on until the process ends, the test rows in as many
 * chunks as needed.  A model whose support vectors (8 bytes per stored
 * feature, plus 4 for its column when sparse), coefficients and one 64-row
 * slice of test rows do not fit in it returns SVMHIP_ERR_NOMEM, with the
 * sizes on stderr.  All calls on one device share its scratch area and keep
 * nothing in it between calls: the library serialises them with a lock, and
 * device work of the same process that overlaps a call and touches the
 * scratch area (there is none in this library) would race with it.  Host
 * memory is used for packing the rows. */
svmhip_status svmhip_predict_batch_noalloc(const struct svm_model *model, const struct svm_node *const *x, int n,
					   const svmhip_options *opt, double *dec_values, double *labels);

/* ---------------------------------------------------------------------------
 * Serving.  No counterpart in the reference (one solve per process, host
 * driven).  A server keeps many requests on the device at once: binary
 * training solves (C_SVC or NU_SVC, as svm_train_one runs them,
 * libsvm/svm.cpp:1707-1769) and predictions against the models of finished
 * solves.  Requests are admitted first come first served while their device
 * footprint (rows, solver state, kernel-row cache) fits the HBM budget;
 * every svmhip_server_step admits what fits, runs the admitted predictions,
 * then advances every admitted solve by one stage of the batched kernel
 * (about stage_ms of device time each), so requests that arrive while others
 * run join at the next stage.  A solve's state and cache stay on the device
 * between stages and its iterates do not depend on what else is in flight.
 * ------------------------------------------------------------------------- */

typedef struct svmhip_server svmhip_server;

typedef struct {
	int id;                   /* request id */
	int type;                 /* 0 = train, 1 = predict */
	int status;               /* svmhip_status */
	int n;                    /* rows of the request */
	int iters;                /* train: SMO iterations */
	int nSV;                  /* train: support vectors (the resident model) */
	double rho, obj, r;       /* train: as Solve reports them (r: nu only) */
	double device_ms;         /* device time of the stages/launches it took part in */
} svmhip_completion;

/* hbm_budget_bytes = 0: 80% of the free device memory.  opt as for the
 * batched solver; cache_bytes caps each solve's kernel-row cache
 * (0 = its whole matrix up to 2 GiB). */
svmhip_status svmhip_server_create(const svmhip_options *opt, double hbm_budget_bytes, svmhip_server **out);

/* Queues a binary training solve of prob with param (C_SVC or NU_SVC,
 * exactly two labels, kernel other than PRECOMPUTED).  The rows are packed
 * on the host now and uploaded at admission.  Returns the id (>= 0) or
 * -status. */
int svmhip_server_submit_train(svmhip_server *s, const struct svm_problem *prob, const struct svm_parameter *param);

/* Queues the decision values of n rows against the resident model of the
 * finished training request model_id.  Returns the id or -status. */
int svmhip_server_submit_predict(svmhip_server *s, int model_id, const struct svm_node *const *x, int n);

/* One round.  Waits for the previous round's device work and collects its
 * results (completions), then admits what fits, launches the admitted
 * predictions and one stage of every admitted solve, and returns without
 * waiting for them, so the caller can queue requests while the device
 * works.  stage_ms <= 0 uses 10 ms.  *active (requests on the device) and
 * *queued (may be null) receive the counts after the launch; both are 0 once
 * everything submitted has completed. */
svmhip_status svmhip_server_step(svmhip_server *s, double stage_ms, int *active, int *queued);

/* Completions since the last call, at most max; returns their number. */
int svmhip_server_poll(svmhip_server *s, svmhip_completion *out, int max);

/* Results of a finished request: a train request's signed coefficients
 * (alpha_i * y_i, scaled by 1/r for NU_SVC, as svm_train_one leaves them;
 * n = its rows) and labels[2] (label of y = +1, label of y = -1); a predict
 * request's decision values (n = its rows) and predicted labels (may be
 * null). */
svmhip_status svmhip_server_train_result(svmhip_server *s, int id, double *coef, int n, double *labels2);
svmhip_status svmhip_server_predict_result(svmhip_server *s, int id, double *dec_values, double *labels, int n);

/* Frees a finished request (for a train request: its resident model). */
svmhip_status svmhip_server_release(svmhip_server *s, int id);

/* Device bytes held by admitted requests and resident models (freed request
   buffers kept for reuse, up to 2 GiB or a tenth of the budget, are not
   counted). */
double svmhip_server_bytes_in_use(const svmhip_server *s);

/* Totals since creation: iters (SMO iterations), row_hits / row_misses,
 * solve_ms (device time of all stages) and launches (stages). */
svmhip_status svmhip_server_stats(const svmhip_server *s, svmhip_stats *stats);

void svmhip_server_destroy(svmhip_server *s);

const char *svmhip_server_last_error(const svmhip_server *s);

/* Number of usable HIP devices (0 if none); never aborts. */
int svmhip_device_count(void);

#ifdef __cplusplus
}
#endif

#endif /* LIBSVM_MI355X_SVM_HIP_H */
