// svm-train for libsvm_hip.so: the reference's command line
// (svm-train/svm-train.c: options, svmlight input, model file, -v cross
// validation report) over the kept LIBSVM API of include/svm.h.
//
// This library carries only the device solver, so param.cuda_flag is always
// 1: the reference's -C switch is accepted and changes nothing, and -h
// (shrinking) is accepted and ignored, as on the reference's device path.
#include <ctype.h>
#include <errno.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <string>
#include <vector>

#include "../../include/svm.h"

namespace {

void print_null(const char *) {}

[[noreturn]] void usage() {
  printf(
      "Usage: svm-train [options] training_set_file [model_file]\n"
      "options:\n"
      "-C : device solver (always on: the library has no CPU solver; accepted for compatibility)\n"
      "-s svm_type : set type of SVM (default 0)\n"
      "\t0 -- C-SVC\t\t(multi-class classification)\n"
      "\t1 -- nu-SVC\t\t(multi-class classification)\n"
      "\t2 -- one-class SVM\n"
      "\t3 -- epsilon-SVR\t(regression)\n"
      "\t4 -- nu-SVR\t\t(regression)\n"
      "-t kernel_type : set type of kernel function (default 2)\n"
      "\t0 -- linear: u'*v\n"
      "\t1 -- polynomial: (gamma*u'*v + coef0)^degree\n"
      "\t2 -- radial basis function: exp(-gamma*|u-v|^2)\n"
      "\t3 -- sigmoid: tanh(gamma*u'*v + coef0)\n"
      "\t4 -- precomputed kernel (kernel values in training_set_file)\n"
      "-d degree : set degree in kernel function (default 3)\n"
      "-g gamma : set gamma in kernel function (default 1/num_features)\n"
      "-r coef0 : set coef0 in kernel function (default 0)\n"
      "-c cost : set the parameter C of C-SVC, epsilon-SVR, and nu-SVR (default 1)\n"
      "-n nu : set the parameter nu of nu-SVC, one-class SVM, and nu-SVR (default 0.5)\n"
      "-p epsilon : set the epsilon in loss function of epsilon-SVR (default 0.1)\n"
      "-m cachesize : set cache memory size in MB (default 100)\n"
      "-e epsilon : set tolerance of termination criterion (default 0.001)\n"
      "-h shrinking : whether to use the shrinking heuristics, 0 or 1 (default 1; ignored)\n"
      "-b probability_estimates : whether to train a SVC or SVR model for probability estimates, 0 or 1 "
      "(default 0)\n"
      "-wi weight : set the parameter C of class i to weight*C, for C-SVC (default 1)\n"
      "-v n: n-fold cross validation mode\n"
      "-q : quiet mode (no outputs)\n");
  exit(1);
}

[[noreturn]] void input_error(int line) {
  fprintf(stderr, "Wrong input format at line %d\n", line);
  exit(1);
}

struct Options {
  svm_parameter param{};
  int cv_folds = 0;
  std::string input, model;
  std::vector<int> weight_label;
  std::vector<double> weight;
};

Options parse(int argc, char **argv) {
  Options o;
  svm_parameter &p = o.param;
  p.cuda_flag = 1;
  p.svm_type = C_SVC;
  p.kernel_type = RBF;
  p.degree = 3;
  p.gamma = 0;  // 1/num_features
  p.coef0 = 0;
  p.nu = 0.5;
  p.cache_size = 100;
  p.C = 1;
  p.eps = 1e-3;
  p.p = 0.1;
  p.shrinking = 1;
  p.probability = 0;
  void (*print_func)(const char *) = nullptr;
  int i = 1;
  for (; i < argc && argv[i][0] == '-'; ++i) {
    const char opt = argv[i][1];
    if (opt == 'C') continue;
    if (opt == 'q') {
      print_func = &print_null;
      continue;
    }
    if (i + 1 >= argc) usage();
    const char *v = argv[++i];
    switch (opt) {
      case 's': p.svm_type = atoi(v); break;
      case 't': p.kernel_type = atoi(v); break;
      case 'd': p.degree = atoi(v); break;
      case 'g': p.gamma = atof(v); break;
      case 'r': p.coef0 = atof(v); break;
      case 'n': p.nu = atof(v); break;
      case 'm': p.cache_size = atof(v); break;
      case 'c': p.C = atof(v); break;
      case 'e': p.eps = atof(v); break;
      case 'p': p.p = atof(v); break;
      case 'h': p.shrinking = atoi(v); break;
      case 'b': p.probability = atoi(v); break;
      case 'v':
        o.cv_folds = atoi(v);
        if (o.cv_folds < 2) {
          fprintf(stderr, "n-fold cross validation: n must >= 2\n");
          usage();
        }
        break;
      case 'w':
        o.weight_label.push_back(atoi(&argv[i - 1][2]));
        o.weight.push_back(atof(v));
        break;
      default:
        fprintf(stderr, "Unknown option: -%c\n", opt);
        usage();
    }
  }
  svm_set_print_string_function(print_func);
  if (i >= argc) usage();
  o.input = argv[i];
  if (i < argc - 1) {
    o.model = argv[i + 1];
  } else {
    const char *base = strrchr(argv[i], '/');
    o.model = std::string(base ? base + 1 : argv[i]) + ".model";
  }
  p.nr_weight = (int)o.weight.size();
  p.weight_label = o.weight_label.empty() ? nullptr : o.weight_label.data();
  p.weight = o.weight.empty() ? nullptr : o.weight.data();
  return o;
}

// One svmlight line: "label index:value index:value ...".
bool read_line(FILE *fp, std::string &line) {
  line.clear();
  int c;
  while ((c = fgetc(fp)) != EOF) {
    if (c == '\n') return true;
    line.push_back((char)c);
  }
  return !line.empty();
}

struct Data {
  svm_problem prob{};
  std::vector<double> y;
  std::vector<svm_node> nodes;
  std::vector<svm_node *> rows;
  int max_index = 0;
};

void read_problem(const std::string &path, const svm_parameter &param, Data &D) {
  FILE *fp = fopen(path.c_str(), "r");
  if (!fp) {
    fprintf(stderr, "can't open input file %s\n", path.c_str());
    exit(1);
  }
  std::vector<size_t> start;
  std::string line;
  int lineno = 0;
  while (read_line(fp, line)) {
    ++lineno;
    std::vector<char> buf(line.begin(), line.end());
    buf.push_back('\0');
    char *save = nullptr;
    char *label = strtok_r(buf.data(), " \t\n", &save);
    if (!label) input_error(lineno);  // empty line
    char *end = nullptr;
    errno = 0;
    const double yv = strtod(label, &end);
    if (end == label || *end != '\0') input_error(lineno);
    D.y.push_back(yv);
    start.push_back(D.nodes.size());
    int prev = -1;
    for (;;) {
      char *idx = strtok_r(nullptr, ":", &save);
      char *val = strtok_r(nullptr, " \t", &save);
      if (!val) break;
      while (isspace((unsigned char)*idx)) ++idx;
      errno = 0;
      const long ix = strtol(idx, &end, 10);
      if (end == idx || errno != 0 || *end != '\0' || ix <= prev) input_error(lineno);
      errno = 0;
      const double xv = strtod(val, &end);
      if (end == val || errno != 0 || (*end != '\0' && !isspace((unsigned char)*end))) input_error(lineno);
      D.nodes.push_back(svm_node{(int)ix, xv});
      prev = (int)ix;
    }
    if (prev > D.max_index) D.max_index = prev;
    D.nodes.push_back(svm_node{-1, 0.0});
  }
  fclose(fp);
  const int l = (int)D.y.size();
  D.rows.resize(l);
  for (int r = 0; r < l; ++r) D.rows[r] = &D.nodes[start[r]];
  D.prob.l = l;
  D.prob.y = D.y.data();
  D.prob.x = D.rows.data();
  if (param.kernel_type == PRECOMPUTED) {
    for (int r = 0; r < l; ++r) {
      if (D.rows[r][0].index != 0) {
        fprintf(stderr, "Wrong input format: first column must be 0:sample_serial_number\n");
        exit(1);
      }
      if ((int)D.rows[r][0].value <= 0 || (int)D.rows[r][0].value > D.max_index) {
        fprintf(stderr, "Wrong input format: sample_serial_number out of range\n");
        exit(1);
      }
    }
  }
}

void cross_validation(const Data &D, const svm_parameter &param, int folds) {
  const int l = D.prob.l;
  std::vector<double> target(l);
  svm_cross_validation(&D.prob, &param, folds, target.data());
  if (param.svm_type == EPSILON_SVR || param.svm_type == NU_SVR) {
    double err = 0, sv = 0, sy = 0, svv = 0, syy = 0, svy = 0;
    for (int i = 0; i < l; ++i) {
      const double y = D.y[i], v = target[i];
      err += (v - y) * (v - y);
      sv += v;
      sy += y;
      svv += v * v;
      syy += y * y;
      svy += v * y;
    }
    printf("Cross Validation Mean squared error = %g\n", err / l);
    printf("Cross Validation Squared correlation coefficient = %g\n",
           ((l * svy - sv * sy) * (l * svy - sv * sy)) / ((l * svv - sv * sv) * (l * syy - sy * sy)));
  } else {
    int correct = 0;
    for (int i = 0; i < l; ++i) correct += target[i] == D.y[i];
    printf("Cross Validation Accuracy = %g%%\n", 100.0 * correct / l);
  }
}

}  // namespace

int main(int argc, char **argv) {
  Options o = parse(argc, argv);
  Data D;
  read_problem(o.input, o.param, D);
  if (o.param.gamma == 0 && D.max_index > 0) o.param.gamma = 1.0 / D.max_index;
  if (const char *err = svm_check_parameter(&D.prob, &o.param)) {
    fprintf(stderr, "ERROR: %s\n", err);
    return 1;
  }
  if (o.cv_folds) {
    cross_validation(D, o.param, o.cv_folds);
    return 0;
  }
  svm_model *model = svm_train(&D.prob, &o.param);
  if (!model) {
    fprintf(stderr, "training failed on the device\n");
    return 1;
  }
  if (svm_save_model(o.model.c_str(), model)) {
    fprintf(stderr, "can't save model to file %s\n", o.model.c_str());
    return 1;
  }
  svm_free_and_destroy_model(&model);
  return 0;
}
