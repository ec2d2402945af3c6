// gfx950 prediction kernel and its launch wrappers: decision values of many
// test rows against a model (svm_predict_values, libsvm/svm.cpp:2569-2643).
#include "smo_device.hpp"

namespace smo {

// (BLOCK: the solver's workgroup size, 512 threads)

// Decision values of test rows against a model (svm_predict_values,
// libsvm/svm.cpp:2569-2643).  A workgroup stages tb test rows densely in LDS
// as doubles, then walks the support vectors in blocks of 2 * BLOCK, two per
// thread (SELL-64 rows, so a wave's loads are coalesced):
//   1. each thread evaluates K(t, s) for its support vector s and every
//      staged test row t in FP64, as Kernel::k_function does
//      (libsvm/svm.cpp:327-384): the dot product summed in increasing feature
//      order, and for RBF sum_k (x_k - y_k)^2 directly (a dense-stored
//      support vector visits every feature in order, which is k_function's
//      merge sequence term for term; a sparse one adds
//      (x_k - y_k)^2 - y_k^2 over its own features to the test row's squared
//      norm), each product rounded before it is added;
//   2. the values of the block go to LDS, and each (test row, pair) sum is
//      advanced by one thread over the block's support vectors of the pair's
//      two classes in index order.  The same thread owns that sum for every
//      block, so it is summed in svm_predict_values' order: the first
//      class's support vectors, then the second's.
// Up to PJOBS sums per thread live in registers; beyond that (many classes)
// they accumulate in dec itself, still one owner per sum.
//
// TB (test rows per workgroup) is a compile-time power of two so that the
// inner loops carry no per-row predicate: rows past the end of the set are
// zero rows in LDS whose values are never stored.  Staged rows have
// pred_ld_pad(ld) doubles: at least one zero column past the last feature
// (index ld), rounded up to a multiple of 8 so a row is read as 16-byte
// pairs.  Padding terms add exactly +0.0 to a sum: a dense support vector's
// features past ld are (0 - 0)^2 or 0 * 0, and a sparse one's padding slots
// point at the zero column with value 0.
constexpr int PJOBS = 4;
constexpr int PTB = 16;  // test rows per workgroup at most

__host__ __device__ inline int pred_ld_pad(int ld) { return (ld + 1 + 7) / 8 * 8; }

template <class VT>
__device__ __forceinline__ double pval(const PRows &R, long long i) {
  return (double)static_cast<const VT *>(R.vals)[i];
}

// Models with one coefficient per support vector (two classes, regression,
// one-class): the kernel-value buffer holds the products coef_s * K(t, s),
// rounded as svm_predict_values rounds them, and a sum only adds them up.
__device__ __forceinline__ bool pred_products(const PredictArgs &A) { return A.regression_like || A.k == 2; }

__device__ __forceinline__ double pred_pair_sum(const PredictArgs &A, const double *kt, int sb, int se, int pr,
                                                double acc) {
  // kt[s - sb] = K(t, s) (or coef_s * K(t, s), see pred_products) for s in
  // [sb, se); two-class support vectors are stored class by class, so the
  // two classes' runs together are [sb, se) in order
  if (pred_products(A)) {
#pragma unroll 8
    for (int s = sb; s < se; ++s) acc = acc + kt[s - sb];
    return acc;
  }
  int a = 0, rem = pr;  // pair index -> classes (a, b), a < b, in LIBSVM's order
  while (rem >= A.k - 1 - a) { rem -= A.k - 1 - a; ++a; }
  const int b = a + 1 + rem;
  const int nsv = A.sv.l;
  const double *ca = A.coef + (long long)(b - 1) * nsv, *cb = A.coef + (long long)a * nsv;
  int lo = max(sb, A.cls_start[a]), hi = min(se, A.cls_start[a] + A.cls_count[a]);
#pragma unroll 8
  for (int s = lo; s < hi; ++s) acc = acc + ca[s] * kt[s - sb];
  lo = max(sb, A.cls_start[b]);
  hi = min(se, A.cls_start[b] + A.cls_count[b]);
#pragma unroll 8
  for (int s = lo; s < hi; ++s) acc = acc + cb[s] * kt[s - sb];
  return acc;
}

// acc[j][t] += the terms of features k0 .. k0 + 7 of every staged row t,
// for the values v[j][0..7] of those features of the thread's support
// vector j: each test value read from LDS serves both support vectors.
template <int TB, bool RBF>
__device__ __forceinline__ void pred_dense_step(const double *xt, int ldp, int k0, const double (&v)[2][8],
                                                double (&acc)[2][TB]) {
#pragma unroll
  for (int t = 0; t < TB; ++t) {
    const double2 *xp = reinterpret_cast<const double2 *>(xt + t * ldp + k0);
    double x[8];
#pragma unroll
    for (int h = 0; h < 4; ++h) {
      const double2 w = xp[h];
      x[2 * h] = w.x;
      x[2 * h + 1] = w.y;
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      double a = acc[j][t];
      if (RBF) {
#pragma unroll
        for (int u = 0; u < 8; ++u) {
          const double dd = v[j][u] - x[u];
          a = a + dd * dd;
        }
      } else {
#pragma unroll
        for (int u = 0; u < 8; ++u) a = a + v[j][u] * x[u];
      }
      acc[j][t] = a;
    }
  }
}

// RBF: the kernel is exp(-gamma |x - y|^2), else a function of the dot.
template <class VT, int TB, bool RBF>
__global__ void __launch_bounds__(BLOCK) k_predict(PredictArgs A, int first_row) {
  extern __shared__ double plds[];  // [TB][ldp] test rows, then [TB][BLOCK] kernel values
  __shared__ double tsq[TB];
  const PRows &S = A.sv, &T = A.tv;
  const int ld = A.ld, ldp = pred_ld_pad(ld), tid = threadIdx.x;
  const int t0 = first_row + blockIdx.x * TB;
  const int nt = min(TB, T.l - t0);
  if (nt <= 0) return;
  double *xt = plds, *kb = plds + (size_t)TB * ldp;
  for (int e = tid; e < TB * ldp; e += BLOCK) xt[e] = 0.0;
  __syncthreads();
  for (int t = 0; t < nt; ++t) {
    const int q = t0 + t;
    const int len = T.row_len[q];
    const long long base = T.slice_ptr[q / SLICE] + (q % SLICE);
    for (int p = tid; p < len; p += BLOCK) {
      const int c = T.dense ? p : T.cols[base + (long long)p * SLICE];
      xt[t * ldp + c] = pval<VT>(T, base + (long long)p * SLICE);
    }
  }
  __syncthreads();
  if (tid < TB) {  // squared norms of the test rows, summed in feature order
    double s2 = 0.0;
    for (int k = 0; k < ld; ++k) {
      const double x = xt[tid * ldp + k];
      s2 = s2 + x * x;
    }
    tsq[tid] = s2;
  }
  __syncthreads();
  const int njobs = nt * A.npair;
  const bool prod = pred_products(A);
  const bool in_regs = njobs <= PJOBS * BLOCK;
  double jacc[PJOBS];
#pragma unroll
  for (int u = 0; u < PJOBS; ++u) jacc[u] = 0.0;
  // support vectors in blocks of 2 * BLOCK: thread tid takes sb + tid and
  // sb + BLOCK + tid, and the block's two halves are summed one after the
  // other (half 0's support vectors precede half 1's)
  for (int sb = 0; sb < S.l; sb += 2 * BLOCK) {
    double acc[2][TB];
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int t = 0; t < TB; ++t) acc[j][t] = 0.0;
    const int sj[2] = {sb + tid, sb + BLOCK + tid};
    if (S.dense) {
      // every feature in order; a stored row shorter than ld has zeros past
      // its end, and the test rows' zeros add +0 terms.  A thread whose
      // second support vector is past the end repeats its first (the sums
      // are not used).
      if (sj[0] < S.l) {
        int len[2], last[2];
        long long base[2];
#pragma unroll
        for (int j = 0; j < 2; ++j) {
          const int s = sj[j] < S.l ? sj[j] : sj[0];
          len[j] = S.row_len[s];
          last[j] = max(len[j] - 1, 0);
          base[j] = S.slice_ptr[s / SLICE] + (s % SLICE);
        }
        // features k .. k + 7 of both support vectors: while every lane's
        // rows reach k + 8, one address per row and immediate offsets;
        // otherwise indices clamped to the row (the values past its end are
        // replaced by zeros below)
        const VT *pv[2] = {static_cast<const VT *>(S.vals) + base[0], static_cast<const VT *>(S.vals) + base[1]};
        const int minlen = min(len[0], len[1]);
        auto load8 = [&](int k, VT (&w)[2][8]) {
          if (__all(k + 8 <= minlen)) {
#pragma unroll
            for (int j = 0; j < 2; ++j) {
              const VT *q = pv[j] + (long long)k * SLICE;
#pragma unroll
              for (int u = 0; u < 8; ++u) w[j][u] = q[u * SLICE];
            }
          } else {
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
              for (int u = 0; u < 8; ++u) w[j][u] = pv[j][(long long)min(k + u, last[j]) * SLICE];
          }
        };
        // the next 8 features' loads are issued before this step's
        // arithmetic, so their latency hides behind it
        VT w[2][8];
        load8(0, w);
        for (int k0 = 0; k0 < ldp; k0 += 8) {
          double v[2][8];
#pragma unroll
          for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int u = 0; u < 8; ++u) v[j][u] = k0 + u < len[j] ? (double)w[j][u] : 0.0;
          load8(min(k0 + 8, ldp - 8), w);  // the last step reloads its own features
          __builtin_amdgcn_sched_barrier(0);  // keep the loads ahead of the arithmetic
          pred_dense_step<TB, RBF>(xt, ldp, k0, v, acc);
        }
      }
    } else {
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int s = sj[j];
        if (s < S.l) {
          const int len = S.row_len[s];
          const long long base = S.slice_ptr[s / SLICE] + (s % SLICE);
          const int width = (int)((S.slice_ptr[s / SLICE + 1] - S.slice_ptr[s / SLICE]) / SLICE);
          for (int p0 = 0; p0 < width; p0 += 8) {
            double v[8];
            int c[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) {
              const long long q = min(p0 + u, width - 1);
              const bool in = p0 + u < len;
              const double w = pval<VT>(S, base + q * SLICE);
              const int cc = S.cols[base + q * SLICE];
              v[u] = in ? w : 0.0;
              c[u] = in ? cc : ld;  // the zero column
            }
#pragma unroll
            for (int t = 0; t < TB; ++t) {
              double a = acc[j][t];
#pragma unroll
              for (int u = 0; u < 8; ++u) {
                const double x = xt[t * ldp + c[u]];
                if (RBF) {
                  const double dd = v[u] - x;
                  a = a + (dd * dd - x * x);
                } else {
                  a = a + v[u] * x;
                }
              }
              acc[j][t] = a;
            }
          }
          if (RBF) {
#pragma unroll
            for (int t = 0; t < TB; ++t) acc[j][t] = tsq[t] + acc[j][t];
          }
        }
      }
    }
#pragma unroll
    for (int j = 0; j < 2; ++j) {
      const int hb = sb + j * BLOCK;  // first support vector of this half
      if (hb >= S.l) break;           // uniform over the workgroup
      if (sj[j] < S.l) {
#pragma unroll
        for (int t = 0; t < TB; ++t) kb[t * BLOCK + tid] = acc[j][t];
        // the kernel function of each sum, in a loop of its own (one
        // inlined copy of exp / tanh / powi, not one per row)
#pragma nounroll
        for (int t = 0; t < TB; ++t) {
          const double a = kb[t * BLOCK + tid];
          double K;
          if (RBF) {
            K = exp(-A.gamma * a);
          } else {
            switch (A.ktype) {
              case K_LINEAR: K = a; break;
              case K_POLY: K = powi(A.gamma * a + A.coef0, A.degree); break;
              case K_SIGMOID: K = tanh(A.gamma * a + A.coef0); break;
              default: K = 0.0;
            }
          }
          kb[t * BLOCK + tid] = prod ? A.coef[sj[j]] * K : K;
        }
      }
      __syncthreads();
      const int se = min(S.l, hb + BLOCK);
      if (in_regs) {
#pragma unroll
        for (int u = 0; u < PJOBS; ++u) {
          const int q = tid + u * BLOCK;
          if (q < njobs) jacc[u] = pred_pair_sum(A, kb + (q / A.npair) * BLOCK, hb, se, q % A.npair, jacc[u]);
        }
      } else {
        for (int q = tid; q < njobs; q += BLOCK) {
          double *dp = A.dec + (long long)(t0 + q / A.npair) * A.npair + q % A.npair;
          *dp = pred_pair_sum(A, kb + (q / A.npair) * BLOCK, hb, se, q % A.npair, hb == 0 ? 0.0 : *dp);
        }
      }
      __syncthreads();  // kb is rewritten by the next half
    }
  }
  if (in_regs) {
#pragma unroll
    for (int u = 0; u < PJOBS; ++u) {
      const int j = tid + u * BLOCK;
      if (j < njobs) A.dec[(long long)(t0 + j / A.npair) * A.npair + j % A.npair] = jacc[u] - A.rho[j % A.npair];
    }
  } else {
    for (int j = tid; j < njobs; j += BLOCK) {
      double *dp = A.dec + (long long)(t0 + j / A.npair) * A.npair + j % A.npair;
      *dp = (S.l > 0 ? *dp : 0.0) - A.rho[j % A.npair];
    }
  }
}

}  // namespace smo

template <class VT, int TB, bool RBF>
static hipError_t launch_predict_k(const smo::PredictArgs *A, int first_row, int nblocks, size_t shmem,
                                   hipStream_t st) {
  (void)hipFuncSetAttribute((const void *)smo::k_predict<VT, TB, RBF>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)shmem);
  hipLaunchKernelGGL((smo::k_predict<VT, TB, RBF>), dim3(nblocks), dim3(smo::BLOCK), shmem, st, *A, first_row);
  return hipGetLastError();
}

template <class VT, int TB>
static hipError_t launch_predict_tb(const smo::PredictArgs *A, int first_row, int nblocks, size_t shmem,
                                    hipStream_t st) {
  return A->ktype == smo::K_RBF ? launch_predict_k<VT, TB, true>(A, first_row, nblocks, shmem, st)
                                : launch_predict_k<VT, TB, false>(A, first_row, nblocks, shmem, st);
}

template <class VT>
static hipError_t launch_predict_vt(const smo::PredictArgs *A, int first_row, int nblocks, size_t shmem,
                                    hipStream_t st) {
  switch (A->tb) {
    case 16: return launch_predict_tb<VT, 16>(A, first_row, nblocks, shmem, st);
    case 8: return launch_predict_tb<VT, 8>(A, first_row, nblocks, shmem, st);
    case 4: return launch_predict_tb<VT, 4>(A, first_row, nblocks, shmem, st);
    case 2: return launch_predict_tb<VT, 2>(A, first_row, nblocks, shmem, st);
    case 1: return launch_predict_tb<VT, 1>(A, first_row, nblocks, shmem, st);
    default: return hipErrorInvalidValue;
  }
}

extern "C" {

// Decision values of test rows 0 .. nrows - 1 of A->tv.  A->tb is the most
// test rows per workgroup (from smo_predict_tb for A->ld); whole rounds of
// one such workgroup per compute unit run at that size, and the rows left
// over at the smallest power of two that still covers them with one round,
// so the last round spreads over all compute units instead of leaving most
// of them idle.  Every decision value is computed the same way at any size.
hipError_t smo_launch_predict(const smo::PredictArgs *A, int nrows, int cus, hipStream_t st) {
  if (nrows <= 0) return hipSuccess;
  if (A->sv.f64 != A->tv.f64 || A->tb < 1 || A->tb > smo::PTB || (A->tb & (A->tb - 1)) || cus < 1)
    return hipErrorInvalidValue;
  const size_t per = (size_t)(smo::pred_ld_pad(A->ld) + smo::BLOCK) * sizeof(double);
  auto launch = [&](int tb, int first_row, int rows) {
    smo::PredictArgs B = *A;
    B.tb = tb;
    const int nb = (rows + tb - 1) / tb;
    return B.sv.f64 ? launch_predict_vt<double>(&B, first_row, nb, tb * per, st)
                    : launch_predict_vt<float>(&B, first_row, nb, tb * per, st);
  };
  const long long round = (long long)A->tb * cus;
  const int full = (int)(nrows / round * round);
  if (full == nrows || A->tb == 1) return launch(A->tb, 0, nrows);
  if (full > 0) {
    const hipError_t e = launch(A->tb, 0, full);
    if (e != hipSuccess) return e;
  }
  const int rest = nrows - full, need = (rest + cus - 1) / cus;
  int tb = 1;
  while (tb < need) tb *= 2;
  return launch(tb, full, rest);
}

// Test rows per workgroup of the prediction kernel for staged rows of ld
// features (a power of two, at most 16, with the rows and their kernel
// values within lds_bytes of dynamic LDS; 0 if even one row does not fit),
// and the dynamic LDS that takes.
int smo_predict_tb(int ld, size_t lds_bytes, size_t *shmem) {
  const size_t per = (size_t)(smo::pred_ld_pad(ld) + smo::BLOCK) * sizeof(double);
  int tb = smo::PTB;
  while (tb > 0 && (size_t)tb * per > lds_bytes) tb /= 2;
  if (shmem) *shmem = (size_t)tb * per;
  return tb;
}

}  // extern "C"
