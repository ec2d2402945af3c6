/*
 * libsvm-mi355x public C API.
 *
 * Byte-compatible with the LIBSVM 3.18 API as extended by libsvm-cuda
 * (libsvm/svm.h:25-122): same struct layouts, same enum values, same
 * function names and signatures, so a program compiled against the
 * reference header links against libsvm_hip.so unchanged.
 *
 * The one field the fork adds, svm_parameter.cuda_flag (libsvm/svm.h:49),
 * stays the FIRST member.  In this library cuda_flag = 1 selects the HIP
 * solver on the MI355X; the library carries no CPU solver, so
 * svm_check_parameter() rejects cuda_flag != 1 with an error string.
 */
#ifndef LIBSVM_MI355X_SVM_H
#define LIBSVM_MI355X_SVM_H

#define LIBSVM_VERSION 318

#ifdef __cplusplus
extern "C" {
#endif

extern int libsvm_version;

/* One sparse feature; a row is terminated by index = -1 (libsvm/svm.h:31-35). */
struct svm_node {
	int index;
	double value;
};

/* Training set: l rows, labels y[l], rows x[l] (libsvm/svm.h:37-42). */
struct svm_problem {
	int l;
	double *y;
	struct svm_node **x;
};

enum { C_SVC, NU_SVC, ONE_CLASS, EPSILON_SVR, NU_SVR };      /* svm_type */
enum { LINEAR, POLY, RBF, SIGMOID, PRECOMPUTED };            /* kernel_type */

/* Solver parameters (libsvm/svm.h:47-67).  cuda_flag must stay first. */
struct svm_parameter {
	int cuda_flag;      /* 1 = run the SMO solver on the GPU */
	int svm_type;
	int kernel_type;
	int degree;         /* poly */
	double gamma;       /* poly / rbf / sigmoid */
	double coef0;       /* poly / sigmoid */

	/* training only */
	double cache_size;  /* kernel-row cache, MB */
	double eps;         /* stopping tolerance on the maximal violating pair */
	double C;           /* C_SVC, EPSILON_SVR, NU_SVR */
	int nr_weight;      /* C_SVC per-class weights */
	int *weight_label;
	double *weight;
	double nu;          /* NU_SVC, ONE_CLASS, NU_SVR */
	double p;           /* EPSILON_SVR epsilon-insensitive width */
	int shrinking;      /* shrinking heuristics (ignored by the GPU solver) */
	int probability;    /* probability estimates */
};

/* Trained model (libsvm/svm.h:72-92). */
struct svm_model {
	struct svm_parameter param;
	int nr_class;       /* number of classes; 2 for regression / one-class */
	int l;              /* total number of support vectors */
	struct svm_node **SV;
	double **sv_coef;   /* sv_coef[nr_class-1][l] */
	double *rho;        /* rho[nr_class*(nr_class-1)/2] */
	double *probA;
	double *probB;
	int *sv_indices;    /* 1-based positions of the SVs in the training set */

	int *label;         /* label[nr_class] */
	int *nSV;           /* nSV[nr_class] */
	int free_sv;        /* 1 if created by svm_load_model */
};

struct svm_model *svm_train(const struct svm_problem *prob, const struct svm_parameter *param);
void svm_cross_validation(const struct svm_problem *prob, const struct svm_parameter *param, int nr_fold, double *target);

int svm_save_model(const char *model_file_name, const struct svm_model *model);
struct svm_model *svm_load_model(const char *model_file_name);

int svm_get_svm_type(const struct svm_model *model);
int svm_get_nr_class(const struct svm_model *model);
void svm_get_labels(const struct svm_model *model, int *label);
void svm_get_sv_indices(const struct svm_model *model, int *sv_indices);
int svm_get_nr_sv(const struct svm_model *model);
double svm_get_svr_probability(const struct svm_model *model);

double svm_predict_values(const struct svm_model *model, const struct svm_node *x, double *dec_values);
double svm_predict(const struct svm_model *model, const struct svm_node *x);
double svm_predict_probability(const struct svm_model *model, const struct svm_node *x, double *prob_estimates);

void svm_free_model_content(struct svm_model *model_ptr);
void svm_free_and_destroy_model(struct svm_model **model_ptr_ptr);
void svm_destroy_param(struct svm_parameter *param);

const char *svm_check_parameter(const struct svm_problem *prob, const struct svm_parameter *param);
int svm_check_probability_model(const struct svm_model *model);

void svm_set_print_string_function(void (*print_func)(const char *));

#ifdef __cplusplus
}
#endif

#endif /* LIBSVM_MI355X_SVM_H */
