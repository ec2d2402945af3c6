// Device-side layout and phase functions of the MI355X SMO solver.
//
// Data layout in HBM (one solver handle):
//   rows of X     SELL-64: rows grouped in slices of 64; slice s stores its
//                 rows column-major, element p of row 64s+lane at
//                 vals[slice_ptr[s] + 64p + lane] (and its feature column at
//                 cols[...] unless the matrix is stored dense).  One lane owns
//                 one row, so a wave walking p reads 64 consecutive floats:
//                 coalesced, and every row's dot product is summed in its own
//                 feature order, which is LIBSVM's merge order
//                 (libsvm/svm.cpp:305-325) -> bit-identical FP64 dots.
//   G, alpha, p, QD   FP64 [n_var];  y, status int8 [n_var]
//   cache rows    float [S][l] kernel rows K(r, .) without label signs;
//                 Q_ij = y_i y_j K(real(i), real(j)) is formed on use, which is
//                 exact (sign flips), so SVC / one-class / SVR share one cache.
//   Ctl           control record: the selected pair, cache decisions, barrier
//                 counters, iteration count, result scalars.
//
// Work split: the persistent grid's workgroup b owns a contiguous range of
// slices and therefore of real rows; with SVR it handles both variables r and
// r + l of each row it owns.  Every per-iteration pass touches only the
// owner's range, so G and the cached rows never cross workgroups inside a
// pass; the only cross-workgroup data are the per-workgroup partial
// reductions and the scalars the leader reads (G_i, G_j, Q_ij, alpha_i/j).
#pragma once

#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdint.h>

namespace smo {

// Threads of every workgroup that runs the solver's device functions (every
// solver kernel is launched with BLOCK threads), and its waves.  Compile-time
// constants: the block size builtin is read from the dispatch packet by a
// memory load, which inside a loop costs a round trip and a wait on every
// load in flight.
constexpr int BLOCK = 512;
constexpr int NWAVES = BLOCK / 64;

constexpr int SLICE = 64;
constexpr double TAU = 1e-12;           // libsvm/svm_defs.h:106, libsvm/svm.cpp:42
constexpr double INF = __builtin_huge_val();

enum : int8_t { AT_LOWER = 0, AT_UPPER = 1, IS_FREE = 2 };   // libsvm/svm_device.cu:13
enum { KIND_C = 0, KIND_NU = 1 };
enum { K_LINEAR = 0, K_POLY = 1, K_RBF = 2, K_SIGMOID = 3, K_PRECOMPUTED = 4 };
enum { STOP_NONE = 0, STOP_OPTIMAL = 1, STOP_MAXITER = 2, STOP_PAUSE = 3 };
enum { ERR_NONE = 0, ERR_TIMEOUT = 5 };

// Control record.  Written by exactly one thread (the leader's thread 0) at
// a time; read by all workgroups after a barrier.
struct Ctl {
  unsigned arrive;       // grid barrier arrivals
  unsigned gen;          // grid barrier generation
  unsigned arrive_lb;    // last-block counter of the step kernels
  int err;
  int i, j, ip, in;      // selected variables
  int slot_i, slot_j, slot_ip, slot_in;
  int miss_i, miss_j, miss_ip, miss_in;
  int stop, iter, max_iter;
  int ginit_next;        // batched solves: first variable of G initialisation not yet done
  double dai, daj;       // alpha changes of the last step
  double Gmax, Gmax2;    // C rule
  double Gp, Gp2, Gn, Gn2;  // nu rule
  int lru_head, lru_tail;
  unsigned long long hits, misses;
  double rho, obj, r;
  alignas(8) unsigned char dec_bytes[384];   // persistent kernel's Dec between launches
};

// One workgroup's partial reduction (48 bytes).
struct Partial {
  double v0, v1, v2, v3;
  int i0, i1, pad0, pad1;
};

// Kernel argument block, passed by value.
struct View {
  // problem
  int l, d, nslices, dense;
  int ktype, degree;
  double gamma, coef0;
  const long long *slice_ptr;
  const float *vals;
  const int *cols;
  const int *row_len;
  const double *xsq;
  const int *serial;      // PRECOMPUTED: (int)x_r[0].value
  // solver
  int nvar, svr, kind, max_iter;
  double Cp, Cn, eps;
  double *G, *alpha, *p, *QD;
  const int8_t *y;
  int8_t *status;
  // kernel-row cache
  float *rows;
  long long row_stride;   // floats per cached row (>= l)
  int S, full;
  int *slot_of_row, *row_of_slot, *lru_prev, *lru_next;
  int8_t *computed;       // full cache: row already materialised
  // control
  Ctl *ctl;
  Partial *part;
  unsigned long long *xw; // persistent kernel / teams: exchange words, two buffers of XW x nblocks
  int8_t *pm_flags;       // persistent kernel, full cache: nblocks x l row flags
  int *pm_lru;            // persistent kernel, LRU cache: nblocks x pm_stride ints
  long long pm_stride;
  float *xs_global;       // per-workgroup query-row scratch when d is too big for LDS
  int use_lds, nblocks;
  int rank;               // batched solves: this workgroup's owner index (-1: blockIdx.x)
  int resume;             // batched solves: G, alpha, status and QD carry over from an earlier stage
  int iter0;              // batched solves: iterations already done
  long long pause_ticks;  // batched solves: pause after this many device clock ticks (0 = never)
  int ginit_from, ginit_to;  // batched solves: variables [from, to) of G = p + Q alpha to add in this
                             // launch (resume = 0); to < nvar pauses before the first iteration
  int own_cache;          // batched solves: rows/S/full/row_stride and the cache metadata
                          // (nblocks copies, pm_stride apart) are the subproblem's own
  int keep_cache;         // batched solves: the own cache is valid from an earlier stage
  int qpre;               // persistent solves, full cache, rows computed on demand, dense rows of at
                          // most BLOCK features: fetch the query row of a likely miss right after
                          // each exchange
  int spec;               // persistent C-rule solves, full cache: speculative rows on an i miss
  int xrep;               // persistent solves: copies of the exchange records (0 = by the grid size)
  int flag_lds;           // persistent solves, full cache: > 0 = the workgroup's row flags live as
                          // a bitset in dynamic LDS from float offset flag_lds (>= 4, past the
                          // query row) during the launch; 0 = in pm_flags only
  unsigned *drain;        // batched solves: workgroups of the launch that found the queue empty
  int drain_at;           // batched solves: pause once drain reaches this (checked every 32 iterations)
  int res_off;            // persistent solves: > 0 = the owned variables' G, alpha, QD, labels and
                          // statuses, and the owned segments of the iteration's kernel rows, live in
                          // dynamic LDS from this float offset during the launch; 0 = device memory
  int res_m;              // rows each resident array holds per variable half (>= the owned rows)
  const double *gdot;     // batched solves: the exact FP64 dot products of every pair of the base
                          // problem's rows (row-major, gdot_ld per row; null = compute from X)
  long long gdot_ld;
  const int *grow;        // batched solves with gdot: the base-problem row of each local row
};

// Work queue and kernel-row cache arenas of the batched solver.  Item q of
// a launch is (subproblem, rank); the items of teams come first.  A team's
// T consecutive items [q0, q0 + T) use arenas [q0, q0 + T) as one row cache
// (each member keeps its own cache metadata), and the workgroup that ran
// member item q keeps arena q for the single-workgroup items it takes
// afterwards; a workgroup whose first item is a single-workgroup one takes
// a fresh arena from team_items + (*arena_next)++.  Every workgroup thus
// holds one arena, no two hold the same, and at most the grid's size are
// used.
struct BatchPool {
  float *rows;             // K arenas of arena_floats floats
  long long arena_floats;
  int8_t *flags;           // K x flag_stride row flags (full-cache arenas)
  int *lru;                // K x lru_stride ints (LRU arenas)
  long long flag_stride, lru_stride;
  int *queue;              // next item (zeroed before each launch)
  int *arena_next;         // fresh arenas handed out (zeroed before each launch)
  int K, team_items;       // arenas; items belonging to teams (the first team_items)
  int lds_floats;          // dynamic LDS of the launch, in floats
  int spec;                // single-workgroup C-rule items with a full cache: on a row miss, also
                           // compute the rows of the pass's per-wave winners in the same pass
  unsigned *idle;          // workgroups that found the queue empty (zeroed before each launch;
                           // null = no drain rule)
  int drain_at;            // running items pause once idle reaches this: the host then re-plans
                           // the rest of the solves (as teams) for the workgroups that went idle
  unsigned long long *t0;  // launch start on the device clock (zeroed before each launch; null = none)
  long long deadline_ticks;  // > 0: items pause at t0 + deadline_ticks, and single-workgroup items
                             // claimed after it are left for the next launch (their Ctl stays clear)
  int flag_words;          // > 0: single-workgroup full-cache items of up to 32 x flag_words rows keep
                           // their row flags as an LDS bitset of that many words past lds_floats
};

// ---------------------------------------------------------------- helpers

__device__ __forceinline__ int real_of(const View &V, int v) { return v < V.l ? v : v - V.l; }

// Owner index of this workgroup: its rank in a batched solve, else its
// block index in the persistent grid.
__device__ __forceinline__ int my_owner(const View &V) { return V.rank >= 0 ? V.rank : (int)blockIdx.x; }

// Index of this workgroup's private cache metadata within V.pm_flags /
// V.pm_lru (batched solves point those at the workgroup's own arena).
__device__ __forceinline__ long long pm_index(const View &V) { return V.rank >= 0 ? 0 : (long long)blockIdx.x; }

__device__ __forceinline__ double C_of(const View &V, int v) { return V.y[v] > 0 ? V.Cp : V.Cn; }

__device__ __forceinline__ int8_t status_for(const View &V, int v, double a) {
  // Solver::update_alpha_status, libsvm/svm.cpp:425-432
  if (a >= C_of(V, v)) return AT_UPPER;
  if (a <= 0) return AT_LOWER;
  return IS_FREE;
}

// argmax with ties to the larger index (the >= scans of libsvm/svm.cpp:851,860)
__device__ __forceinline__ void max_idx(double &v, int &i, double v2, int i2) {
  if (v2 > v || (v2 == v && i2 > i)) { v = v2; i = i2; }
}
// argmin with ties to the larger index (the <= scans of libsvm/svm.cpp:890,914)
__device__ __forceinline__ void min_idx(double &v, int &i, double v2, int i2) {
  if (v2 < v || (v2 == v && i2 > i)) { v = v2; i = i2; }
}

// Wavefront max / min through DPP lane moves (VALU, no LDS round trips):
// quad_perm xor 1 and xor 2, row_ror 4 and 8 (every lane of a row of 16 then
// holds the row's result), row_bcast 15 and 31 (lane 63 then holds the
// wave's), and a read of lane 63.  Rows masked out of a move keep their own
// value, which max / min leave unchanged.
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ int dpp_i(int x) {
  return __builtin_amdgcn_update_dpp(x, x, CTRL, ROWS, 0xf, false);
}
template <int CTRL, int ROWS = 0xf>
__device__ __forceinline__ double dpp_d(double x) {
  const int lo = dpp_i<CTRL, ROWS>(__double2loint(x));
  const int hi = dpp_i<CTRL, ROWS>(__double2hiint(x));
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double lane63_d(double x) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), 63),
                          __builtin_amdgcn_readlane(__double2loint(x), 63));
}
template <bool MAX>
__device__ __forceinline__ double dmm(double a, double b) { return MAX ? fmax(a, b) : fmin(a, b); }
template <bool MAX>
__device__ __forceinline__ double wave_mm(double x) {
  x = dmm<MAX>(x, dpp_d<0xB1>(x));        // quad_perm [1,0,3,2]
  x = dmm<MAX>(x, dpp_d<0x4E>(x));        // quad_perm [2,3,0,1]
  x = dmm<MAX>(x, dpp_d<0x124>(x));       // row_ror:4
  x = dmm<MAX>(x, dpp_d<0x128>(x));       // row_ror:8
  x = dmm<MAX>(x, dpp_d<0x142, 0xa>(x));  // row_bcast:15 into rows 1, 3
  x = dmm<MAX>(x, dpp_d<0x143, 0xc>(x));  // row_bcast:31 into rows 2, 3
  return lane63_d(x);
}
// Max / min over the first row of 16 lanes (quad_perm xor 1 and xor 2,
// row_ror 4 and 8), read from lane 0.
template <bool MAX>
__device__ __forceinline__ double row_mm(double x) {
  x = dmm<MAX>(x, dpp_d<0xB1>(x));
  x = dmm<MAX>(x, dpp_d<0x4E>(x));
  x = dmm<MAX>(x, dpp_d<0x124>(x));
  x = dmm<MAX>(x, dpp_d<0x128>(x));
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(x), 0),
                          __builtin_amdgcn_readlane(__double2loint(x), 0));
}
__device__ __forceinline__ int wave_max_i(int x) {
  x = max(x, dpp_i<0xB1>(x));
  x = max(x, dpp_i<0x4E>(x));
  x = max(x, dpp_i<0x124>(x));
  x = max(x, dpp_i<0x128>(x));
  x = max(x, dpp_i<0x142, 0xa>(x));
  x = max(x, dpp_i<0x143, 0xc>(x));
  return __builtin_amdgcn_readlane(x, 63);
}
__device__ __forceinline__ double wave_max(double v) { return wave_mm<true>(v); }
// Largest i over the lanes of `lanes` (a uniform ballot): one lane read when
// a single lane holds the extreme value (the usual case), the DPP maximum
// otherwise.
__device__ __forceinline__ int lanes_max_i(unsigned long long lanes, bool mine, int i) {
  if (lanes != 0 && (lanes & (lanes - 1)) == 0) return __builtin_amdgcn_readlane(i, __builtin_ctzll(lanes));
  return wave_max_i(mine ? i : INT_MIN);
}
// (value, index) arg-max / arg-min with ties to the larger index: the extreme
// value first, then the largest index among the lanes holding it (the same
// result as any order of pairwise max_idx / min_idx steps).  Every lane
// receives the result.
__device__ __forceinline__ void wave_max_idx(double &v, int &i) {
  const double m = wave_mm<true>(v);
  const bool mine = v == m;
  i = lanes_max_i(__ballot(mine), mine, i);
  v = m;
}
__device__ __forceinline__ void wave_min_idx(double &v, int &i) {
  const double m = wave_mm<false>(v);
  const bool mine = v == m;
  i = lanes_max_i(__ballot(mine), mine, i);
  v = m;
}
__device__ __forceinline__ double wave_sum(double v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ int wave_sum_i(int v) {
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

template <class T>
__device__ __forceinline__ T ld_acq(const T *p) {
  return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// powi by repeated squaring, libsvm/svm.cpp:30-40
__device__ __forceinline__ double powi(double base, int times) {
  double acc = 1.0, sq = base;
  for (int t = times; t > 0; t /= 2) {
    if (t % 2 == 1) acc = acc * sq;
    sq = sq * sq;
  }
  return acc;
}

// Slice range [sb, se) owned by workgroup b of nb.
__device__ __forceinline__ void owned_slices(const View &V, int b, int nb, int &sb, int &se) {
  sb = (int)(((long long)V.nslices * b) / nb);
  se = (int)(((long long)V.nslices * (b + 1)) / nb);
}

// ------------------------------------------------------------ row kernel

// Stage row q of X densely (by feature column, or by position for
// PRECOMPUTED) into xs[0..d).  Whole workgroup; ends with a barrier.
__device__ __forceinline__ void stage_row(const View &V, int q, float *xs) {
  const int tid = threadIdx.x, nt = BLOCK;
  const int len = V.row_len[q];
  const long long base = V.slice_ptr[q / SLICE] + (q % SLICE);
  if (!V.dense) {
    for (int k = tid; k < V.d; k += nt) xs[k] = 0.0f;
    __syncthreads();
    const bool by_pos = V.ktype == K_PRECOMPUTED;
    for (int p = tid; p < len; p += nt) {
      int c = by_pos ? p : V.cols[base + (long long)p * SLICE];
      xs[c] = V.vals[base + (long long)p * SLICE];
    }
  } else {
    for (int p = tid; p < len; p += nt) xs[p] = V.vals[base + (long long)p * SLICE];
  }
  __syncthreads();
}

// dot(x_q, x_r) with x_q staged in xs; row r is this lane's row.  Products of
// two floats are exact in FP64, so fma == multiply-then-add here and the sum
// is LIBSVM's sequential merge-order sum.  Loads are issued in batches of
// ROW_UNROLL per lane (8 KB in flight per wave) so a row pass is bound by
// HBM bandwidth rather than by one memory round trip per few elements; the
// rest of a row goes in batches of ROW_TAIL, their loads clamped to the
// row's stored entries and the entries past the row's length left out of the
// sum (adding +0.0 leaves the sum unchanged: it is never -0.0).
// With PIPE, a sparse row goes in batches of ROW_PIPE with two batches in
// flight: batch k+1's loads are issued before batch k is summed, so the
// memory round trips of a row overlap pairwise.  Same sum order, so the same
// bits; it needs ~50 more registers, which the multi-workgroup passes have
// and the passes of the batched kernel do not.
constexpr int ROW_UNROLL = 32;
constexpr int ROW_TAIL = 8;
constexpr int ROW_PIPE = 16;

// Sparse row dot with two batches of ROW_PIPE entries in flight (the PIPE
// form of row_dot): vp / cp are the row's first value and column in its
// slice, len its length, width the slice's stored width.
template <class XP>
__device__ __forceinline__ double sparse_dot_pipe(const float *vp, const int *cp, XP xs, int len, int width) {
  constexpr int B = ROW_PIPE;
  double acc = 0.0;
  float a0[B], a1[B];
  int c0[B], c1[B];
  auto load = [&](int p0, float(&a)[B], int(&c)[B]) {
    if (p0 + B <= width) {
      const float *v = vp + (long long)p0 * SLICE;
      const int *ci = cp + (long long)p0 * SLICE;
#pragma unroll
      for (int u = 0; u < B; ++u) {
        a[u] = v[u * SLICE];
        c[u] = ci[u * SLICE];
      }
    } else {  // the row's last batch, clamped to its stored entries
#pragma unroll
      for (int u = 0; u < B; ++u) {
        const long long q = min(p0 + u, width - 1);
        a[u] = vp[q * SLICE];
        c[u] = cp[q * SLICE];
      }
    }
  };
  auto sum = [&](int p0, const float(&a)[B], const int(&c)[B]) {
#pragma unroll
    for (int u = 0; u < B; ++u) {
      double t = (p0 + u < len) ? (double)a[u] * (double)xs[c[u]] : 0.0;
      acc = acc + t;
    }
  };
  if (width > 0) load(0, a0, c0);
  for (int p = 0; p < width; p += 2 * B) {
    if (p + B < width) load(p + B, a1, c1);
    sum(p, a0, c0);
    if (p + B >= width) break;
    if (p + 2 * B < width) load(p + 2 * B, a0, c0);
    sum(p + B, a1, c1);
  }
  return acc;
}

// The query row in LDS, addressed as LDS: its reads are then LDS reads
// (counted apart from device-memory loads) rather than flat loads, which
// count as both and make every wait for one of them a wait for all the
// row's loads in flight.
typedef __attribute__((address_space(3))) const float lds_cfloat;

// Dense dot(x_q, x_r), q staged at xs (LDS or device memory): features in
// order, one fma each (row_dot's sum).  A batch: ROW_UNROLL loads of row r
// issued, then the batch's query values read at once (one LDS wait rather
// than one per pair), then the fmas in feature order.  (Measured on the
// n=100k dense solve: 64-entry batches, or two 32-entry batches in flight,
// are slower: 191 and 184 ms against 178.5.)
template <class XP>
__device__ __forceinline__ double dense_dot(const float *vp, XP xs, int len) {
  constexpr int U = ROW_UNROLL;
  double acc = 0.0;
  int p = 0;
  for (; p + U <= len; p += U) {
    float a[U], q[U];
#pragma unroll
    for (int u = 0; u < U; ++u) a[u] = vp[(long long)(p + u) * SLICE];
#pragma unroll
    for (int u = 0; u < U; ++u) q[u] = xs[p + u];
#pragma unroll
    for (int u = 0; u < U; ++u) acc = fma((double)a[u], (double)q[u], acc);
  }
  for (; p < len; p += ROW_TAIL) {  // the rest, in clamped batches
    float a[ROW_TAIL];
#pragma unroll
    for (int u = 0; u < ROW_TAIL; ++u) a[u] = vp[(long long)min(p + u, len - 1) * SLICE];
#pragma unroll
    for (int u = 0; u < ROW_TAIL; ++u)
      if (p + u < len) acc = fma((double)a[u], (double)xs[p + u], acc);
  }
  return acc;
}

// Sparse dot(x_q, x_r): the slice's common width walked (rows are
// zero-padded to it) so the trip count is uniform across the wave.
template <class XP>
__device__ __forceinline__ double sparse_dot(const float *vp, const int *cp, XP xs, int len, int width) {
  constexpr int U = ROW_UNROLL;
  double acc = 0.0;
  int p = 0;
  for (; p + U <= width; p += U) {
    float a[U];
    int c[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      a[u] = vp[(long long)(p + u) * SLICE];
      c[u] = cp[(long long)(p + u) * SLICE];
    }
#pragma unroll
    for (int u = 0; u < U; ++u) {
      double t = (p + u < len) ? (double)a[u] * (double)xs[c[u]] : 0.0;
      acc = acc + t;
    }
  }
  for (; p < width; p += ROW_TAIL) {  // the rest, in clamped batches
    float a[ROW_TAIL];
    int c[ROW_TAIL];
#pragma unroll
    for (int u = 0; u < ROW_TAIL; ++u) {
      const long long q = min(p + u, width - 1);
      a[u] = vp[q * SLICE];
      c[u] = cp[q * SLICE];
    }
#pragma unroll
    for (int u = 0; u < ROW_TAIL; ++u) {
      double t = (p + u < len) ? (double)a[u] * (double)xs[c[u]] : 0.0;
      acc = acc + t;
    }
  }
  return acc;
}

template <bool PIPE = false>
__device__ __forceinline__ double row_dot(const View &V, const float *xs, int r) {
  const int s = r / SLICE;
  // dense rows are stored at width d in every slice: their offset needs no
  // load of the slice pointers (a round trip before the row's loads)
  const long long base = V.dense ? (long long)s * SLICE * V.d + (r % SLICE) : V.slice_ptr[s] + (r % SLICE);
  const float *vp = V.vals + base;
  if (V.dense) {
#if defined(__HIP_DEVICE_COMPILE__)  // (the builtin exists in the device pass only)
    if (__builtin_amdgcn_is_shared((const void *)xs)) return dense_dot(vp, (lds_cfloat *)xs, V.d);
#endif
    return dense_dot(vp, xs, V.d);
  }
  const int len = V.row_len[r];
  const int width = (int)((V.slice_ptr[s + 1] - V.slice_ptr[s]) / SLICE);
  const int *cp = V.cols + base;
#if defined(__HIP_DEVICE_COMPILE__)
  if (__builtin_amdgcn_is_shared((const void *)xs)) {
    lds_cfloat *xl = (lds_cfloat *)xs;
    return PIPE ? sparse_dot_pipe(vp, cp, xl, len, width) : sparse_dot(vp, cp, xl, len, width);
  }
#endif
  return PIPE ? sparse_dot_pipe(vp, cp, xs, len, width) : sparse_dot(vp, cp, xs, len, width);
}

// dot(x_q(m), x_r) for the M rows staged at xs + m * ld, each summed in the
// same order as row_dot (so each equals row_dot bit for bit): one pass over
// row r serves M query rows.  U entries of row r are loaded per trip (one
// memory round trip each), the last trip clamped like row_dot's.
template <int M, int U = 8>
__device__ __forceinline__ void row_dot_multi(const View &V, const float *xs, int ld, int r, double acc[M]) {
  const int s = r / SLICE;
  const long long base = V.slice_ptr[s] + (r % SLICE);
  const float *vp = V.vals + base;
#pragma unroll
  for (int m = 0; m < M; ++m) acc[m] = 0.0;
  if (V.dense) {
    const int len = V.d;
    int p = 0;
    for (; p + U <= len; p += U) {
      float a[U];
#pragma unroll
      for (int u = 0; u < U; ++u) a[u] = vp[(long long)(p + u) * SLICE];
#pragma unroll
      for (int u = 0; u < U; ++u)
#pragma unroll
        for (int m = 0; m < M; ++m) acc[m] = fma((double)a[u], (double)xs[m * ld + p + u], acc[m]);
    }
    if (p < len) {  // the last, partial trip
      float a[U];
#pragma unroll
      for (int u = 0; u < U; ++u) a[u] = vp[(long long)min(p + u, len - 1) * SLICE];
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (p + u < len) {
#pragma unroll
          for (int m = 0; m < M; ++m) acc[m] = fma((double)a[u], (double)xs[m * ld + p + u], acc[m]);
        }
    }
  } else {
    const int len = V.row_len[r];
    const int width = (int)((V.slice_ptr[s + 1] - V.slice_ptr[s]) / SLICE);
    const int *cp = V.cols + base;
    for (int p = 0; p < width; p += U) {
      float a[U];
      int c[U];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const long long q = min(p + u, width - 1);
        a[u] = vp[q * SLICE];
        c[u] = cp[q * SLICE];
      }
#pragma unroll
      for (int u = 0; u < U; ++u)
        if (p + u < len) {
          const double av = a[u];
#pragma unroll
          for (int m = 0; m < M; ++m) acc[m] = acc[m] + av * (double)xs[m * ld + c[u]];
        }
    }
  }
}

// K(q, r) in FP64 given dot(x_q, x_r) (Kernel::kernel_*, libsvm/svm.cpp:243-255).
// The same, the two squared norms given (RBF; read before the dot so their
// loads are not a round trip after it).
__device__ __forceinline__ double kernel_f64_sq(const View &V, double xq, double xr, double dot) {
  switch (V.ktype) {
    case K_LINEAR: return dot;
    case K_POLY: return powi(V.gamma * dot + V.coef0, V.degree);
    case K_RBF: return exp(-V.gamma * ((xq + xr) - 2 * dot));
    case K_SIGMOID: return tanh(V.gamma * dot + V.coef0);
    default: return 0.0;
  }
}

__device__ __forceinline__ double kernel_f64(const View &V, int q, int r, double dot) {
  return kernel_f64_sq(V, V.ktype == K_RBF ? V.xsq[q] : 0.0, V.ktype == K_RBF ? V.xsq[r] : 0.0, dot);
}

// Cached (float-rounded) kernel value K(q, r); q staged in xs.
template <bool PIPE = false>
__device__ __forceinline__ float kernel_value(const View &V, const float *xs, int q, int r) {
  if (V.ktype == K_PRECOMPUTED) {
    int s = V.serial[r];
    return (s >= 0 && s < V.d) ? xs[s] : 0.0f;
  }
  const bool rbf = V.ktype == K_RBF;
  const double xq = rbf ? V.xsq[q] : 0.0, xr = rbf ? V.xsq[r] : 0.0;
  return (float)kernel_f64_sq(V, xq, xr, row_dot<PIPE>(V, xs, r));
}

// dot(x_q, x_r) from the shared dot products (batched solves over one base
// problem): the same FP64 value row_dot sums.
__device__ __forceinline__ double shared_dot(const View &V, int q, int r) {
  return V.gdot[(long long)V.grow[q] * V.gdot_ld + V.grow[r]];
}

// Materialise the owned part of kernel row q into cache slot `slot`.
template <bool PIPE = false>
__device__ __forceinline__ void compute_row_part(const View &V, int q, int slot, float *xs, int sb, int se) {
  if (V.gdot) {  // a gather of one row of the shared dot products
    float *dst = V.rows + (long long)slot * V.row_stride;
    const double *dq = V.gdot + (long long)V.grow[q] * V.gdot_ld;
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64, nw = NWAVES;
    for (int s = sb + wave; s < se; s += nw) {
      const int r = s * SLICE + lane;
      if (r < V.l) dst[r] = (float)kernel_f64(V, q, r, dq[V.grow[r]]);
    }
    __syncthreads();
    return;
  }
  stage_row(V, q, xs);
  float *dst = V.rows + (long long)slot * V.row_stride;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64, nw = NWAVES;
  for (int s = sb + wave; s < se; s += nw) {
    int r = s * SLICE + lane;
    if (r < V.l) dst[r] = kernel_value<PIPE>(V, xs, q, r);
  }
  __syncthreads();
}

// ------------------------------------------------------- block reductions

struct RedSmem {
  double v[2][16][4];
  int i[2][16][2];
  int last;
  unsigned gen;
};

// Block-wide reduction of up to 4 (value, index) lanes; op per slot.
// kinds: 0 = max_idx, 1 = min_idx, 2 = plain max.  The result is left in
// every thread.  One barrier: the waves' results go through sm buffer `buf`,
// which the caller must not have the next reduction on the same buffer write
// before a later barrier (the passes use buffer 0 and the exchanges buffer 1,
// and every pass or exchange has a barrier between two reductions on one
// buffer).  nact > 0: only waves [0, nact) hold values (the others hold the
// identity and skip the wave step).
template <int N>
__device__ __forceinline__ void block_reduce(RedSmem &sm, double (&v)[4], int (&ix)[2], const int (&kinds)[4],
                                             int buf = 0, int nact = 0) {
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64, nw = NWAVES;
  const int na = nact > 0 && nact < nw ? nact : nw;
  if (wave < na) {
    int slot_i = 0;
#pragma unroll
    for (int k = 0; k < N; ++k) {
      if (kinds[k] == 0) { wave_max_idx(v[k], ix[slot_i]); ++slot_i; }
      else if (kinds[k] == 1) { wave_min_idx(v[k], ix[slot_i]); ++slot_i; }
      else v[k] = wave_max(v[k]);
    }
    if (lane == 0) {
#pragma unroll
      for (int k = 0; k < N; ++k) sm.v[buf][wave][k] = v[k];
      sm.i[buf][wave][0] = ix[0];
      sm.i[buf][wave][1] = ix[1];
    }
  }
  __syncthreads();
  if (na == 1) {  // one wave held values: its result is the block's
#pragma unroll
    for (int k = 0; k < N; ++k) v[k] = sm.v[buf][0][k];
    ix[0] = sm.i[buf][0][0];
    ix[1] = sm.i[buf][0][1];
    return;
  }
  // every wave combines the waves' results: lane w < na reads wave w's (one
  // LDS round trip), then DPP moves within the first row of 16 lanes and a
  // lane read; the other lanes hold the identity
  const bool in = lane < na;
  const int src = in ? lane : 0;
  int si = 0;
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const double x = sm.v[buf][src][k];
    if (kinds[k] == 2) {
      v[k] = row_mm<true>(in ? x : -INF);
      continue;
    }
    const int xi = sm.i[buf][src][si];
    const bool mx = kinds[k] == 0;
    const double xv = in ? x : (mx ? -INF : INF);
    const double m = mx ? row_mm<true>(xv) : row_mm<false>(xv);
    const bool mine = in && xv == m;
    ix[si] = lanes_max_i(__ballot(mine), mine, xi);
    v[k] = m;
    ++si;
  }
}

// --------------------------------------------------------------- the cache

// Where one copy of the cache's bookkeeping lives: the shared copy used by
// the step kernels (head/tail/counters in Ctl) or a workgroup's private copy
// used by the persistent kernel.
struct CacheMeta {
  int8_t *computed;                 // full cache: row already materialised
  unsigned *bits;                   // full cache, flags in LDS: bit r of word r / 32 (else null)
  int *slot_of_row, *row_of_slot, *prev, *next;
  int *head, *tail;
  unsigned long long *hits, *misses;
};

// Moves slot s to the front of the LRU list; head, pv and nx are the list's
// head and s's neighbours as they are before the move.
__device__ __forceinline__ void lru_touch(CacheMeta &M, int s, int head, int pv, int nx) {
  if (head == s) return;
  if (pv >= 0) M.next[pv] = nx;
  if (nx >= 0) M.prev[nx] = pv;
  else *M.tail = pv;
  M.prev[s] = -1;
  M.next[s] = head;
  M.prev[head] = s;
  *M.head = s;
}

// Slot holding kernel row `row`; on a miss the least recently used slot other
// than `exclude` is reassigned (the row is then computed by every workgroup
// for its owned range in the next pass).  Single thread.
// Full cache: sets row's flag; true if it was already set.
__device__ __forceinline__ bool flag_test_set(CacheMeta &M, int row) {
  if (M.bits) {
    const unsigned w = M.bits[row >> 5], b = 1u << (row & 31);
    M.bits[row >> 5] = w | b;
    return (w & b) != 0;
  }
  const bool was = M.computed[row] != 0;
  M.computed[row] = 1;
  return was;
}

// FULL: the caller knows the cache holds every row (V.full), so the LRU
// bookkeeping is left out of its code.
template <bool FULL = false>
__device__ __forceinline__ int cache_lookup(const View &V, CacheMeta &M, int row, int exclude, int *miss) {
  if (FULL || V.full) {
    *miss = flag_test_set(M, row) ? 0 : 1;
    if (*miss) ++*M.misses; else ++*M.hits;
    return row;
  }
  // the loads that do not depend on one another are issued together (the
  // lookup runs on the iteration's critical path, one thread, metadata in
  // device memory): head and tail with the row's slot, then the slot's
  // neighbours (and, on a miss, its old row) in one more round trip
  const int s = M.slot_of_row[row];
  const int head = *M.head, tail = *M.tail;
  if (s >= 0) {
    *miss = 0;
    ++*M.hits;
    lru_touch(M, s, head, M.prev[s], M.next[s]);
    return s;
  }
  int v = tail;
  if (v == exclude) v = M.prev[v];
  const int old = M.row_of_slot[v], pv = M.prev[v], nx = M.next[v];
  if (old >= 0) M.slot_of_row[old] = -1;
  M.row_of_slot[v] = row;
  M.slot_of_row[row] = v;
  lru_touch(M, v, head, pv, nx);
  *M.misses += 1;
  *miss = 1;
  return v;
}

__device__ __forceinline__ CacheMeta shared_meta(const View &V) {
  CacheMeta M;
  M.bits = nullptr;
  M.computed = V.computed;
  M.slot_of_row = V.slot_of_row; M.row_of_slot = V.row_of_slot; M.prev = V.lru_prev; M.next = V.lru_next;
  M.head = &V.ctl->lru_head; M.tail = &V.ctl->lru_tail;
  M.hits = &V.ctl->hits; M.misses = &V.ctl->misses;
  return M;
}

__device__ __forceinline__ int cache_lookup(const View &V, int row, int exclude, int *miss) {
  CacheMeta M = shared_meta(V);
  return cache_lookup(V, M, row, exclude, miss);
}

// ------------------------------------------------------------ the step math

// Two-variable analytic update with clipping, FP64 (libsvm/svm.cpp:636-721,
// libsvm/svm_device.cu:828-927), as a pure function of the pair's scalars.
// Qij = Q_i[j] as cached (float).  Returns the new (alpha_i, alpha_j).
__device__ __forceinline__ void alpha_pair(int yi, int yj, double Ci, double Cj, double QDi, double QDj, double Gi,
                                           double Gj, float Qij, double &ai, double &aj) {
  if (yi != yj) {
    double quad = QDi + QDj + (double)(2.0f * Qij);
    if (quad <= 0) quad = TAU;
    double delta = (-Gi - Gj) / quad;
    double diff = ai - aj;
    ai += delta;
    aj += delta;
    if (diff > 0) {
      if (aj < 0) { aj = 0; ai = diff; }
    } else {
      if (ai < 0) { ai = 0; aj = -diff; }
    }
    if (diff > Ci - Cj) {
      if (ai > Ci) { ai = Ci; aj = Ci - diff; }
    } else {
      if (aj > Cj) { aj = Cj; ai = Cj + diff; }
    }
  } else {
    double quad = QDi + QDj - (double)(2.0f * Qij);
    if (quad <= 0) quad = TAU;
    double delta = (Gi - Gj) / quad;
    double sum = ai + aj;
    ai -= delta;
    aj += delta;
    if (sum > Ci) {
      if (ai > Ci) { ai = Ci; aj = sum - Ci; }
    } else {
      if (aj < 0) { aj = 0; ai = sum; }
    }
    if (sum > Cj) {
      if (aj > Cj) { aj = Cj; ai = sum - Cj; }
    } else {
      if (ai < 0) { ai = 0; aj = sum; }
    }
  }
}

// Step-mode wrapper: reads the pair from the arrays, writes alpha and the
// changes into Ctl.  Single thread.
__device__ void alpha_step(const View &V, int i, int j, float Qij) {
  double ai = V.alpha[i], aj = V.alpha[j];
  const double ai0 = ai, aj0 = aj;
  alpha_pair(V.y[i], V.y[j], C_of(V, i), C_of(V, j), V.QD[i], V.QD[j], V.G[i], V.G[j], Qij, ai, aj);
  V.alpha[i] = ai;
  V.alpha[j] = aj;
  V.ctl->dai = ai - ai0;
  V.ctl->daj = aj - aj0;
}

__device__ __forceinline__ float q_elem(const View &V, int slot, int i, int v) {
  // Q_iv = y_i y_v K(real i, real v); sign flips of a float are exact
  float k = V.rows[(long long)slot * V.row_stride + real_of(V, v)];
  return (V.y[i] == V.y[v]) ? k : -k;
}

// --------------------------------------------------------------- the passes

// Snapshot of the control scalars a pass needs.
struct Snap {
  int i, j, ip, in, slot_i, slot_j, slot_ip, slot_in, miss_i, miss_j, miss_ip, miss_in;
  int stop, err, iter;
  double dai, daj, Gmax, Gp, Gn;
};

__device__ __forceinline__ Snap snap(const Ctl *c) {
  Snap s;
  s.i = ld_acq(&c->i); s.j = ld_acq(&c->j); s.ip = ld_acq(&c->ip); s.in = ld_acq(&c->in);
  s.slot_i = ld_acq(&c->slot_i); s.slot_j = ld_acq(&c->slot_j);
  s.slot_ip = ld_acq(&c->slot_ip); s.slot_in = ld_acq(&c->slot_in);
  s.miss_i = ld_acq(&c->miss_i); s.miss_j = ld_acq(&c->miss_j);
  s.miss_ip = ld_acq(&c->miss_ip); s.miss_in = ld_acq(&c->miss_in);
  s.stop = ld_acq(&c->stop); s.err = ld_acq(&c->err); s.iter = ld_acq(&c->iter);
  s.dai = ld_acq(&c->dai); s.daj = ld_acq(&c->daj);
  s.Gmax = ld_acq(&c->Gmax); s.Gp = ld_acq(&c->Gp); s.Gn = ld_acq(&c->Gn);
  return s;
}

// Pass "select i": per-workgroup partial of the WSS i-step over the owned
// variables (libsvm/svm.cpp:840-863 and the Gmax2 terms of :874-915).
// If update is set, first applies G_v += Q_iv dai + Q_jv daj
// (libsvm/svm.cpp:724-736), computing the owned part of row j on a miss.
// Alpha / status of the last pair, written by the workgroup that owns each
// variable (persistent kernel: every workgroup computes the pair update, only
// the owners store it).
struct Pend {
  int i, j;
  double ai, aj;
  int si, sj;  // their alpha_status (decided with the step, so the owner loads no label)
};

// Slices a wave visits per trip of the pass loops, and (for rows that are
// not SVR) in single-workgroup batched solves, whose waves own whole
// subproblems: more slices per trip cost the persistent kernel, whose waves
// own one or two slices, more than they save.
constexpr int PASS_SLICES = 4;
constexpr int PASS_SLICES_SOLO = 8;

__device__ __forceinline__ bool owns(const View &V, int sb, int se, int v) {
  int r = real_of(V, v);
  return r >= sb * SLICE && r < se * SLICE;
}

template <int KIND, int NV, int PS, bool PIPE>
__device__ __forceinline__ Partial pass_gradient_select_nv(const View &V, const Snap &c, float *xs, RedSmem &sm, bool update,
                                        bool select, bool to_part = true, const Pend *pend = nullptr) {
  int sb, se;
  owned_slices(V, my_owner(V), V.nblocks, sb, se);
  if (pend) {
    if (threadIdx.x == 0) {
      if (owns(V, sb, se, pend->i)) { V.alpha[pend->i] = pend->ai; V.status[pend->i] = (int8_t)pend->si; }
      if (owns(V, sb, se, pend->j)) { V.alpha[pend->j] = pend->aj; V.status[pend->j] = (int8_t)pend->sj; }
    }
    // PIPE: this pass takes the two statuses from pend (below) rather than
    // waiting for the stores; later passes read them after other barriers
    if (!PIPE) __syncthreads();
  }
  if (update && c.miss_j) compute_row_part<PIPE>(V, real_of(V, c.j), c.slot_j, xs, sb, se);
  double v[4] = {-INF, -INF, -INF, -INF};
  int ix[2] = {-1, -1};
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64, nw = NWAVES;
  const int8_t yi = update ? V.y[c.i] : (int8_t)0, yj = update ? V.y[c.j] : (int8_t)0;
  // PS slices per trip, all loads issued before any use, so a wave
  // waits for one memory round trip per PS slices rather than per
  // slice (the selection's ties go to the larger index, so the order of
  // visits does not matter)
  for (int s0 = sb + wave; s0 < se; s0 += nw * PS) {
    float Ki[PS], Kj[PS];
    double g[PS][NV];
    int8_t yk[PS][NV], st[PS][NV];
    bool ok[PS];
    // loads are unconditional (lanes past the owned rows read a clamped,
    // valid row and ignore it), so they issue back to back with one wait
#pragma unroll
    for (int u = 0; u < PS; ++u) {
      const int r = (s0 + u * nw) * SLICE + lane;
      ok[u] = s0 + u * nw < se && r < V.l;
      // slots past the wave's owned slices re-read its last owned slice
      // (lines already being fetched) instead of other owners' rows
      const int rc = min(min(s0 + u * nw, se - 1) * SLICE + lane, V.l - 1);
      Ki[u] = Kj[u] = 0.f;
      if (update) {
        Ki[u] = V.rows[(long long)c.slot_i * V.row_stride + rc];
        Kj[u] = V.rows[(long long)c.slot_j * V.row_stride + rc];
      }
#pragma unroll
      for (int h = 0; h < NV; ++h) {
        const int k = rc + h * V.l;
        yk[u][h] = V.y[k];
        g[u][h] = V.G[k];
        st[u][h] = select ? V.status[k] : (int8_t)0;
      }
    }
#pragma unroll
    for (int u = 0; u < PS; ++u) {
      if (!ok[u]) continue;
      const int r = (s0 + u * nw) * SLICE + lane;
#pragma unroll
      for (int h = 0; h < NV; ++h) {
        const int k = r + h * V.l;
        double gk = g[u][h];
        if (update) {
          float Qik = (yi == yk[u][h]) ? Ki[u] : -Ki[u];
          float Qjk = (yj == yk[u][h]) ? Kj[u] : -Kj[u];
          double t1 = (double)Qik * c.dai;
          double t2 = (double)Qjk * c.daj;
          gk = gk + (t1 + t2);
          V.G[k] = gk;
        }
        if (!select) continue;
        int8_t sk = st[u][h];
        if (PIPE && pend) {
          if (k == pend->i) sk = (int8_t)pend->si;
          else if (k == pend->j) sk = (int8_t)pend->sj;
        }
        if (KIND == KIND_C) {
          if (yk[u][h] == +1) {
            if (sk != AT_UPPER) max_idx(v[0], ix[0], -gk, k);
            if (sk != AT_LOWER) {
              if (V.spec) max_idx(v[1], ix[1], gk, k);
              else v[1] = fmax(v[1], gk);
            }
          } else {
            if (sk != AT_LOWER) max_idx(v[0], ix[0], gk, k);
            if (sk != AT_UPPER) {
              if (V.spec) max_idx(v[1], ix[1], -gk, k);
              else v[1] = fmax(v[1], -gk);
            }
          }
        } else {
          if (yk[u][h] == +1) {
            if (sk != AT_UPPER) max_idx(v[0], ix[0], -gk, k);
            if (sk != AT_LOWER) v[2] = fmax(v[2], gk);
          } else {
            if (sk != AT_LOWER) max_idx(v[1], ix[1], gk, k);
            if (sk != AT_UPPER) v[3] = fmax(v[3], -gk);
          }
        }
      }
    }
  }
  Partial pt;
  if (!select) return pt;
  if (KIND == KIND_C) {
    // i, and (for speculative rows) the argmax of Gmax2
    const int kinds_spec[4] = {0, 0, 2, 2}, kinds[4] = {0, 2, 2, 2};
    if (V.spec) block_reduce<2>(sm, v, ix, kinds_spec);
    else block_reduce<2>(sm, v, ix, kinds);
  } else {
    const int kinds[4] = {0, 0, 2, 2};
    block_reduce<4>(sm, v, ix, kinds);
  }
  pt.v0 = v[0]; pt.v1 = v[1]; pt.v2 = v[2]; pt.v3 = v[3];
  pt.i0 = ix[0]; pt.i1 = ix[1]; pt.pad0 = pt.pad1 = 0;
  if (to_part && threadIdx.x == 0) V.part[blockIdx.x] = pt;
  return pt;
}

// Pass "select j": second-order gain over the owned variables
// (libsvm/svm.cpp:866-917; nu rule :1123-1172), row i (rows ip / in) computed
// for the owned range first on a miss.
template <int KIND, int NV, int PS, bool PIPE>
__device__ __forceinline__ Partial pass_select_j_nv(const View &V, const Snap &c, float *xs, RedSmem &sm, bool to_part = true) {
  int sb, se;
  owned_slices(V, my_owner(V), V.nblocks, sb, se);
  if (KIND == KIND_C) {
    if (c.miss_i) compute_row_part<PIPE>(V, real_of(V, c.i), c.slot_i, xs, sb, se);
  } else {
    if (c.ip >= 0 && c.miss_ip) compute_row_part<PIPE>(V, real_of(V, c.ip), c.slot_ip, xs, sb, se);
    if (c.in >= 0 && c.miss_in) compute_row_part<PIPE>(V, real_of(V, c.in), c.slot_in, xs, sb, se);
  }
  double best = INF;
  int bj = -1;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64, nw = NWAVES;
  // rows and scalars of the candidates i (C rule) / i_p, i_n (nu rule)
  const int qi = KIND == KIND_C ? c.i : c.ip;
  const int qn = KIND == KIND_C ? -1 : c.in;
  const long long ri = qi >= 0 ? (long long)(KIND == KIND_C ? c.slot_i : c.slot_ip) * V.row_stride : 0;
  const long long rn = qn >= 0 ? (long long)c.slot_in * V.row_stride : 0;
  const int8_t yi = qi >= 0 ? V.y[qi] : (int8_t)1, yn = qn >= 0 ? V.y[qn] : (int8_t)1;
  const double QDi = qi >= 0 ? V.QD[qi] : 0.0, QDn = qn >= 0 ? V.QD[qn] : 0.0;
  // PS slices per trip, loads first (ties go to the larger index)
  for (int s0 = sb + wave; s0 < se; s0 += nw * PS) {
    float Ki[PS], Kn[PS];
    double g[PS][NV], qd[PS][NV];
    int8_t yk[PS][NV], st[PS][NV];
    bool ok[PS];
    // unconditional loads from a clamped, valid row (see pass_gradient_select)
#pragma unroll
    for (int u = 0; u < PS; ++u) {
      const int r = (s0 + u * nw) * SLICE + lane;
      ok[u] = s0 + u * nw < se && r < V.l;
      // slots past the wave's owned slices re-read its last owned slice
      // (lines already being fetched) instead of other owners' rows
      const int rc = min(min(s0 + u * nw, se - 1) * SLICE + lane, V.l - 1);
      Ki[u] = qi >= 0 ? V.rows[ri + rc] : 0.f;
      Kn[u] = qn >= 0 ? V.rows[rn + rc] : 0.f;
#pragma unroll
      for (int h = 0; h < NV; ++h) {
        const int k = rc + h * V.l;
        yk[u][h] = V.y[k];
        st[u][h] = V.status[k];
        g[u][h] = V.G[k];
        qd[u][h] = V.QD[k];
      }
    }
#pragma unroll
    for (int u = 0; u < PS; ++u) {
      if (!ok[u]) continue;
      const int r = (s0 + u * nw) * SLICE + lane;
#pragma unroll
      for (int h = 0; h < NV; ++h) {
        const int k = r + h * V.l;
        const int8_t y = yk[u][h], sk = st[u][h];
        const double gk = g[u][h];
        if (KIND == KIND_C) {
          double gd;
          if (y == +1) {
            if (sk == AT_LOWER) continue;
            gd = c.Gmax + gk;
          } else {
            if (sk == AT_UPPER) continue;
            gd = c.Gmax - gk;
          }
          if (gd > 0) {
            float Qik = (yi == y) ? Ki[u] : -Ki[u];
            double t = (2.0 * yi) * (double)Qik;
            double quad = (y == +1) ? (QDi + qd[u][h]) - t : (QDi + qd[u][h]) + t;
            double od = quad > 0 ? -(gd * gd) / quad : -(gd * gd) / TAU;
            min_idx(best, bj, od, k);
          }
        } else {
          if (y == +1) {
            if (sk == AT_LOWER) continue;
            double gd = c.Gp + gk;
            if (gd > 0) {
              float Q = (yi == y) ? Ki[u] : -Ki[u];
              double quad = (QDi + qd[u][h]) - (double)(2.0f * Q);
              double od = quad > 0 ? -(gd * gd) / quad : -(gd * gd) / TAU;
              min_idx(best, bj, od, k);
            }
          } else {
            if (sk == AT_UPPER) continue;
            double gd = c.Gn - gk;
            if (gd > 0) {
              float Q = (yn == y) ? Kn[u] : -Kn[u];
              double quad = (QDn + qd[u][h]) - (double)(2.0f * Q);
              double od = quad > 0 ? -(gd * gd) / quad : -(gd * gd) / TAU;
              min_idx(best, bj, od, k);
            }
          }
        }
      }
    }
  }
  double v[4] = {best, 0, 0, 0};
  int ix[2] = {bj, -1};
  const int kinds[4] = {1, 2, 2, 2};
  block_reduce<1>(sm, v, ix, kinds);
  Partial pt;
  pt.v0 = v[0]; pt.v1 = pt.v2 = pt.v3 = 0;
  pt.i0 = ix[0]; pt.i1 = -1; pt.pad0 = pt.pad1 = 0;
  if (to_part && threadIdx.x == 0) V.part[blockIdx.x] = pt;
  return pt;
}

// The passes by row kind: SVR variables come two per row (both halves of
// the row are visited together); every other kind has one, which leaves the
// registers for PS1 slices per trip.  PIPE: rows missed here are computed
// with two batches of loads in flight (row_dot), for kernels with the
// registers to spare.
template <int KIND, int PS1 = PASS_SLICES, bool PIPE = false>
__device__ __forceinline__ Partial pass_gradient_select(const View &V, const Snap &c, float *xs, RedSmem &sm,
                                                        bool update, bool select, bool to_part = true,
                                                        const Pend *pend = nullptr) {
  if (V.svr) return pass_gradient_select_nv<KIND, 2, PASS_SLICES, PIPE>(V, c, xs, sm, update, select, to_part, pend);
  return pass_gradient_select_nv<KIND, 1, PS1, PIPE>(V, c, xs, sm, update, select, to_part, pend);
}

template <int KIND, int PS1 = PASS_SLICES, bool PIPE = false>
__device__ __forceinline__ Partial pass_select_j(const View &V, const Snap &c, float *xs, RedSmem &sm,
                                                 bool to_part = true) {
  if (V.svr) return pass_select_j_nv<KIND, 2, PASS_SLICES, PIPE>(V, c, xs, sm, to_part);
  return pass_select_j_nv<KIND, 1, PS1, PIPE>(V, c, xs, sm, to_part);
}

// Partials are read with agent-scope loads: they were written by other
// workgroups, possibly on other XCDs, earlier in the same launch.
__device__ __forceinline__ Partial load_partial(const Partial *p) {
  Partial q;
  q.v0 = ld_acq(&p->v0); q.v1 = ld_acq(&p->v1); q.v2 = ld_acq(&p->v2); q.v3 = ld_acq(&p->v3);
  q.i0 = ld_acq(&p->i0); q.i1 = ld_acq(&p->i1); q.pad0 = q.pad1 = 0;
  return q;
}

// ----------------------------------------------------------- leader work
// Decisions taken once per pass by a single thread (thread 0 of the leader
// workgroup) from the grid-wide reduced partial.

enum { PH_I = 0, PH_J = 1 };

template <int KIND, int PH>
__device__ __forceinline__ Partial partial_identity() {
  Partial p;
  if (PH == PH_J) { p.v0 = INF; p.v1 = p.v2 = p.v3 = 0; }
  else { p.v0 = p.v1 = p.v2 = p.v3 = -INF; }
  p.i0 = p.i1 = -1;
  p.pad0 = p.pad1 = 0;
  return p;
}

template <int KIND, int PH>
__device__ __forceinline__ void partial_combine(Partial &a, const Partial &b) {
  if (PH == PH_J) {
    min_idx(a.v0, a.i0, b.v0, b.i0);
  } else if (KIND == KIND_C) {
    max_idx(a.v0, a.i0, b.v0, b.i0);
    a.v1 = fmax(a.v1, b.v1);
  } else {
    max_idx(a.v0, a.i0, b.v0, b.i0);
    max_idx(a.v1, a.i1, b.v1, b.i1);
    a.v2 = fmax(a.v2, b.v2);
    a.v3 = fmax(a.v3, b.v3);
  }
}

// Reduce src[first + t*stride], t < count, over the whole workgroup; the
// result is valid in thread 0.
template <int KIND, int PH>
__device__ Partial reduce_partials(const Partial *src, int first, int count, int stride, RedSmem &sm) {
  Partial a = partial_identity<KIND, PH>();
  for (int t = threadIdx.x; t < count; t += BLOCK) partial_combine<KIND, PH>(a, load_partial(&src[first + t * stride]));
  double v[4] = {a.v0, a.v1, a.v2, a.v3};
  int ix[2] = {a.i0, a.i1};
  if (PH == PH_J) {
    const int kinds[4] = {1, 2, 2, 2};
    block_reduce<1>(sm, v, ix, kinds);
  } else if (KIND == KIND_C) {
    const int kinds[4] = {0, 2, 2, 2};
    block_reduce<2>(sm, v, ix, kinds);
  } else {
    const int kinds[4] = {0, 0, 2, 2};
    block_reduce<4>(sm, v, ix, kinds);
  }
  a.v0 = v[0]; a.v1 = v[1]; a.v2 = v[2]; a.v3 = v[3];
  a.i0 = ix[0]; a.i1 = ix[1];
  return a;
}

// After "select i": i (ip / in), Gmax / Gmax2, the stopping test, row slot(s).
template <int KIND>
__device__ void finish_select_i(const View &V, const Partial &t, bool count_maxiter) {
  Ctl *c = V.ctl;
  int stop = STOP_NONE;
  if (count_maxiter && c->iter >= V.max_iter) stop = STOP_MAXITER;
  if (KIND == KIND_C) {
    c->Gmax = t.v0;
    c->Gmax2 = t.v1;
    c->i = t.i0;
    if (!stop && t.v0 + t.v1 < V.eps) stop = STOP_OPTIMAL;   // libsvm/svm.cpp:919
    if (!stop) {
      int m;
      c->slot_i = cache_lookup(V, real_of(V, t.i0), -1, &m);
      c->miss_i = m;
    }
  } else {
    c->Gp = t.v0; c->Gn = t.v1; c->Gp2 = t.v2; c->Gn2 = t.v3;
    c->ip = t.i0; c->in = t.i1;
    if (!stop && fmax(t.v0 + t.v2, t.v1 + t.v3) < V.eps) stop = STOP_OPTIMAL;  // libsvm/svm.cpp:1174
    if (!stop) {
      int m, sp = -1;
      c->miss_ip = 0; c->miss_in = 0;
      if (t.i0 >= 0) { sp = cache_lookup(V, real_of(V, t.i0), -1, &m); c->slot_ip = sp; c->miss_ip = m; }
      if (t.i1 >= 0) { c->slot_in = cache_lookup(V, real_of(V, t.i1), sp, &m); c->miss_in = m; }
    }
  }
  c->stop = stop;
}

// After "select j": j, its row slot, and (persistent mode) the alpha step
// and the status update of the pair.
template <int KIND>
__device__ void finish_select_j(const View &V, const Partial &t, bool do_step) {
  Ctl *c = V.ctl;
  int j = t.i0;
  if (j < 0) {  // no eligible j: only reachable with eps <= 0 or NaN gradients
    c->stop = STOP_OPTIMAL;
    return;
  }
  int i, slot_i;
  if (KIND == KIND_C) {
    i = c->i;
    slot_i = c->slot_i;
  } else {
    // out_i = y[Gmin_idx] == +1 ? Gmaxp_idx : Gmaxn_idx, libsvm/svm.cpp:1177-1180
    if (V.y[j] == +1) { i = c->ip; slot_i = c->slot_ip; }
    else { i = c->in; slot_i = c->slot_in; }
    c->i = i;
    c->slot_i = slot_i;
  }
  c->j = j;
  int m;
  c->slot_j = cache_lookup(V, real_of(V, j), slot_i, &m);
  c->miss_j = m;
  c->iter = c->iter + 1;
  if (do_step) {
    float Qij = q_elem(V, slot_i, i, j);
    alpha_step(V, i, j, Qij);
    V.status[i] = status_for(V, i, V.alpha[i]);
    V.status[j] = status_for(V, j, V.alpha[j]);
  }
}

template <int KIND>
__device__ void leader_after_select_i(const View &V, RedSmem &sm, bool count_maxiter) {
  Partial t = reduce_partials<KIND, PH_I>(V.part, 0, V.nblocks, 1, sm);
  if (threadIdx.x == 0) finish_select_i<KIND>(V, t, count_maxiter);
}

template <int KIND>
__device__ void leader_after_select_j(const View &V, RedSmem &sm, bool do_step) {
  Partial t = reduce_partials<KIND, PH_J>(V.part, 0, V.nblocks, 1, sm);
  if (threadIdx.x == 0) finish_select_j<KIND>(V, t, do_step);
}

// ---------------------------------------------------------- synchronisation

// Last-arriving workgroup detection for the non-persistent step kernels.
__device__ __forceinline__ bool arrive_last(const View &V, RedSmem &sm) {
  __syncthreads();
  if (threadIdx.x == 0) {
    __threadfence();
    unsigned prev = atomicAdd(&V.ctl->arrive_lb, 1u);
    sm.last = (prev == (unsigned)V.nblocks - 1u);
    if (sm.last) atomicExch(&V.ctl->arrive_lb, 0u);
  }
  __syncthreads();
  if (sm.last) __threadfence();
  return sm.last;
}

constexpr long long SPIN_LIMIT = 1LL << 23;

// ------------------------------------------------ persistent-kernel exchange
//
// The persistent kernel keeps every vector (G, alpha, status, cached kernel
// rows) partitioned by owner workgroup, and lets every workgroup take the
// same decisions from the same data.  The only data that crosses workgroups
// are the per-pass partial records below: each carries a workgroup's best
// candidate together with the scalars the decision needs (G, alpha and, for
// the j pass, Q_ij of that candidate).
//
// A record travels as 64-bit words, each holding the exchange's sequence
// number in its high half and 32 bits of the record in its low half.  An
// aligned 64-bit store is single-copy atomic, so a reader that finds the
// expected number in every word of a record has all of that exchange's
// record: the writer issues its stores back to back with no wait and no
// separate flag, and a reader needs one round trip per poll.  The words are
// stored and loaded with agent-scope (cross-XCD coherent) accesses; no
// read-modify-write atomics, no L2 write-back or invalidation.  Word w of
// owner t in buffer p (the sequence number's parity) is at
// xw[(p * XW + w) * nblocks + t], so the readers' loads of one word are
// contiguous across owners.  With 64 or more owners every record is written
// XREP times, to XREP copies of the buffers, and workgroup b reads copy
// b % XREP: every workgroup reads every record at every exchange, and
// spreading those reads over copies (on different memory channels) shortens
// the exchange.

struct XPart {
  unsigned seq;
  int i0, i1, pad;
  double v0, v1, v2, v3;
  double g0, a0, g1, a1;   // G and alpha of the winner(s)
  double qd0, qd1;         // QD of the winner(s)
  int y0, y1;              // label / sign of the winner(s)
  float q0, fpad;          // j record: Q_ij
};

using XWord = unsigned long long;
constexpr int XW = 23;    // words of the largest record format
constexpr int XREP = 4;   // copies of the buffers (n=100k, 256 owners: 1: 242, 2: 236, 4: 235.5, 8: 240, 16: 249 ms)

// Exchange words of one copy of the two buffers of nb owners, and of all copies.
__host__ __device__ constexpr size_t xcopy_words(int nb) { return 2 * (size_t)XW * (size_t)(nb > 0 ? nb : 0); }
__host__ __device__ constexpr size_t xrec_words(int nb) { return XREP * xcopy_words(nb); }

// Record formats: the j record, the i record of the C rule, the i record of
// the nu rule, and a team's final terms.
enum { XF_J = 0, XF_IC = 1, XF_INU = 2, XF_FIN = 3 };

template <int F>
__host__ __device__ constexpr int xf_words() {
  return F == XF_J ? 11 : F == XF_IC ? 14 : F == XF_INU ? 23 : 18;
}

__device__ __forceinline__ void xput_d(unsigned *w, int &k, double x) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(x);
  w[k++] = (unsigned)b;
  w[k++] = (unsigned)(b >> 32);
}

__device__ __forceinline__ double xget_d(const unsigned *w, int &k) {
  const unsigned long long b = (unsigned long long)w[k] | ((unsigned long long)w[k + 1] << 32);
  k += 2;
  return __longlong_as_double((long long)b);
}

template <int F>
__device__ __forceinline__ void xpack(const XPart &p, unsigned (&w)[XW]) {
  int k = 0;
  w[k++] = (unsigned)p.i0;
  if (F != XF_J) w[k++] = (unsigned)p.i1;
  xput_d(w, k, p.v0);
  if (F != XF_J) xput_d(w, k, p.v1);
  if (F == XF_INU || F == XF_FIN) { xput_d(w, k, p.v2); xput_d(w, k, p.v3); }
  xput_d(w, k, p.g0);
  xput_d(w, k, p.a0);
  if (F == XF_INU || F == XF_FIN) { xput_d(w, k, p.g1); xput_d(w, k, p.a1); }
  // the nu rule's i_p has y = +1 and i_n has y = -1 by construction: their
  // labels are not sent
  if (F != XF_FIN) xput_d(w, k, p.qd0);
  if (F == XF_J || F == XF_IC) w[k++] = (unsigned)p.y0;
  if (F == XF_INU) xput_d(w, k, p.qd1);
  if (F == XF_J) w[k++] = __float_as_uint(p.q0);
  if (F == XF_IC || F == XF_INU) w[k++] = (unsigned)p.pad;
}

template <int F>
__device__ __forceinline__ void xunpack(const unsigned (&w)[XW], XPart &p) {
  int k = 0;
  p.i0 = (int)w[k++];
  if (F != XF_J) p.i1 = (int)w[k++];
  p.v0 = xget_d(w, k);
  if (F != XF_J) p.v1 = xget_d(w, k);
  if (F == XF_INU || F == XF_FIN) { p.v2 = xget_d(w, k); p.v3 = xget_d(w, k); }
  p.g0 = xget_d(w, k);
  p.a0 = xget_d(w, k);
  if (F == XF_INU || F == XF_FIN) { p.g1 = xget_d(w, k); p.a1 = xget_d(w, k); }
  if (F != XF_FIN) p.qd0 = xget_d(w, k);
  if (F == XF_J || F == XF_IC) p.y0 = (int)w[k++];
  if (F == XF_INU) { p.qd1 = xget_d(w, k); p.y0 = +1; p.y1 = -1; }
  if (F == XF_J) p.q0 = __uint_as_float(w[k++]);
  if (F == XF_IC || F == XF_INU) p.pad = (int)w[k++];
}

__device__ __forceinline__ XWord *xword(XWord *X, int nb, unsigned parity, int w, int t) {
  return X + ((size_t)(parity * XW + (unsigned)w) * (size_t)nb + (size_t)t);
}

template <class T>
__device__ __forceinline__ void st_agent(T *p, T v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Copies in use for nb owners (xrep > 0: that many, at most XREP): by
// default XREP with 64 or more owners; with few owners the reads of one copy
// are few too, and the extra stores would cost more than they save.
__device__ __forceinline__ int xcopies(int nb, int xrep) {
  if (xrep > 0) return xrep < XREP ? xrep : XREP;
  return nb >= 64 ? XREP : 1;
}

// Owner t's record of exchange `seq`, into every copy in use (one thread).
template <int F>
__device__ __forceinline__ void xpublish(XWord *X, int nb, int t, const XPart &p, unsigned seq, int xrep = 0) {
  unsigned w[XW];
  xpack<F>(p, w);
  const int nc = xcopies(nb, xrep);
#pragma unroll 1
  for (int c = 0; c < nc; ++c)
#pragma unroll
    for (int k = 0; k < xf_words<F>(); ++k)
      st_agent(xword(X + (size_t)c * xcopy_words(nb), nb, seq & 1u, k, t), ((XWord)seq << 32) | (XWord)w[k]);
}

// xpublish by a whole wave (every lane calls it; lane 0 holds the record):
// lane k stores word k % NW of copy k / NW, so all copies of the record go
// out in one or two store instructions.
template <int F>
__device__ __forceinline__ void xpublish_wave(XWord *X, int nb, int t, const XPart &p, unsigned seq, int xrep = 0) {
  constexpr int NW = xf_words<F>();
  unsigned w[XW];
  xpack<F>(p, w);
  const int lane = threadIdx.x & 63, nc = xcopies(nb, xrep);
#pragma unroll
  for (int k0 = 0; k0 < XREP * NW; k0 += 64) {
    const int k = k0 + lane;
    const int q = k % NW;
    unsigned mine = 0;
#pragma unroll
    for (int u = 0; u < NW; ++u) {
      const unsigned s = __builtin_amdgcn_readfirstlane(w[u]);
      mine = q == u ? s : mine;
    }
    if (k < nc * NW) st_agent(xword(X + (size_t)(k / NW) * xcopy_words(nb), nb, seq & 1u, q, t), ((XWord)seq << 32) | mine);
  }
}

// The copy of the buffers this workgroup reads.
__device__ __forceinline__ XWord *xcopy(XWord *X, int nb, int xrep) {
  return X + (size_t)(blockIdx.x % xcopies(nb, xrep)) * xcopy_words(nb);
}

// Waits for owner t's record of exchange `seq` and unpacks it into q (one
// thread).  The wait polls the record's last word alone (every workgroup
// polls every record, so polling whole records multiplies the traffic on a
// few cache lines and slows every reader), then reads the whole record and
// checks every word.  False if the wait exceeded SPIN_LIMIT polls.
template <int F>
__device__ __forceinline__ bool xgather(XWord *X, int nb, int t, unsigned seq, XPart &q, int xrep = 0) {
  constexpr int NW = xf_words<F>();
  unsigned w[XW];
  long long spins = 0;
  X = xcopy(X, nb, xrep);
  // a record takes about a microsecond to travel: a short sleep before the
  // first poll keeps the readers' polls off the memory pipeline meanwhile
  // (n=100k solve, 256 workgroups: 191 -> 188 ms; Gram-tile loop on 128:
  // 89 -> 82 ms).  Polls are what the exchange is most sensitive to: 12
  // sleep units rather than 8 (Gram-tile loop on 192 workgroups 72.0 ->
  // 70.9 ms; 14, 16 and 20: 71.9, 72.6, 74.3 ms), and a second poll kept in
  // flight, which doubles the polls, 72.0 -> 87 ms
  __builtin_amdgcn_s_sleep(12);
  for (;;) {
    while ((unsigned)(ld_acq(xword(X, nb, seq & 1u, NW - 1, t)) >> 32) != seq) {
      if (++spins > SPIN_LIMIT) return false;
      __builtin_amdgcn_s_sleep(1);
    }
    bool ok = true;
#pragma unroll
    for (int k = 0; k < NW; ++k) {
      const XWord x = ld_acq(xword(X, nb, seq & 1u, k, t));
      w[k] = (unsigned)x;
      ok &= (unsigned)(x >> 32) == seq;
    }
    if (ok) break;
    if (++spins > SPIN_LIMIT) return false;
  }
  xunpack<F>(w, q);
  return true;
}

// Payload word w of owner t in buffer `parity` (a completed exchange).
__device__ __forceinline__ unsigned xpeek(XWord *X, int nb, unsigned parity, int w, int t, int xrep = 0) {
  return (unsigned)ld_acq(xword(xcopy(X, nb, xrep), nb, parity, w, t));
}

// Control state every workgroup holds (identical in all of them); also the
// record block 0 leaves in Ctl between launches.
struct Dec {
  int phase, stop, iter, epoch;
  int i, j, ip, in;
  int slot_i, slot_j, slot_ip, slot_in;
  int miss_i, miss_j, miss_ip, miss_in;
  int pend, err, st_i, st_j;   // st_i, st_j: alpha_status of i and j after the pending step
  double ai_new, aj_new, dai, daj;
  double Gmax, Gmax2, Gp, Gn, Gp2, Gn2;
  double Gi, ai, Gip, aip, Gin, ain;
  double QDi, QDip, QDin;
  int yi, yip, yin, yj;
  unsigned long long hits, misses;
  long long t_start;      // device clock at the start of this launch's share of the solve
};

static_assert(sizeof(Dec) <= sizeof(((Ctl *)nullptr)->dec_bytes), "Dec must fit in Ctl");

__host__ __device__ __forceinline__ Dec &ctl_dec(Ctl *c) { return *reinterpret_cast<Dec *>(c->dec_bytes); }

template <int KIND, int PH>
__device__ __forceinline__ XPart xpart_identity() {
  XPart p;
  p.seq = 0; p.i0 = p.i1 = -1; p.pad = 0;
  if (PH == PH_J) { p.v0 = INF; p.v1 = p.v2 = p.v3 = 0; }
  else { p.v0 = p.v1 = p.v2 = p.v3 = -INF; }
  p.g0 = p.a0 = p.g1 = p.a1 = 0;
  p.qd0 = p.qd1 = 0; p.y0 = p.y1 = 0;
  p.q0 = 0; p.fpad = 0;
  return p;
}

struct XSmem {
  XPart win[1];
  int err;
};

// Publish this workgroup's record and gather everyone's: returns, in every
// thread, the combined record (reduced values; winners' payloads in g0/a0/q0
// and, for the nu rule's second argmax, g1/a1).
//
// hold (one winner: the j exchange, or the i exchange of the C rule): the
// result is complete only in the thread that gathered the winning record
// (thread 0 when there is no winner), and *hold is set in that thread alone;
// the caller decides there, which spares the barrier of handing the payload
// over through LDS.
template <int KIND, int PH>
__device__ __forceinline__ XPart exchange(const View &V, RedSmem &sm, XSmem &xs, const XPart &mine, unsigned seq,
                                          bool *hold = nullptr) {
  constexpr int F = PH == PH_J ? XF_J : (KIND == KIND_C ? XF_IC : XF_INU);
  if (threadIdx.x < 64) xpublish_wave<F>(V.xw, V.nblocks, my_owner(V), mine, seq, V.xrep);
  if (threadIdx.x == 0) xs.err = 0;
  // with more than one wave of readers, the others start polling once the
  // record's stores have issued (their polls would otherwise queue ahead of
  // the stores in the workgroup's memory pipeline: measured 198 -> 189 ms
  // per n=100k solve on 256 workgroups)
  if (V.nblocks > 64) __syncthreads();
  XPart q = xpart_identity<KIND, PH>();
  const int t = threadIdx.x;
  if (t < V.nblocks && !xgather<F>(V.xw, V.nblocks, t, seq, q, V.xrep)) xs.err = 1;
  // i records carry each owner's pause vote: every owner sees the same OR
  // (C rule: the max of reduction slot 2, which its i records leave unused,
  // so no barrier of its own; nu rule: one barrier)
  int pause = 0;
  if (PH == PH_I && KIND != KIND_C) pause = __syncthreads_or(t < V.nblocks ? q.pad : 0);
  double v[4] = {q.v0, q.v1, q.v2, q.v3};
  int ix[2] = {q.i0, q.i1};
  const int nact = (V.nblocks + 63) / 64;  // waves holding records
  if (PH == PH_J) {
    const int kinds[4] = {1, 2, 2, 2};
    block_reduce<1>(sm, v, ix, kinds, 1, nact);
  } else if (KIND == KIND_C) {
    v[2] = (t < V.nblocks && q.pad) ? 1.0 : 0.0;
    const int kinds_spec[4] = {0, 0, 2, 2}, kinds[4] = {0, 2, 2, 2};
    if (V.spec) block_reduce<3>(sm, v, ix, kinds_spec, 1, nact);
    else block_reduce<3>(sm, v, ix, kinds, 1, nact);
    pause = v[2] > 0.0 ? 1 : 0;
  } else {
    const int kinds[4] = {0, 0, 2, 2};
    block_reduce<4>(sm, v, ix, kinds, 1, nact);
  }
  if (hold && (PH == PH_J || KIND == KIND_C)) {
    const bool mine_wins = ix[0] >= 0 ? (t < V.nblocks && q.i0 == ix[0]) : (t == 0);
    XPart r = xpart_identity<KIND, PH>();
    r.v0 = v[0]; r.v1 = v[1]; r.v2 = v[2]; r.v3 = v[3];
    r.i0 = ix[0]; r.i1 = ix[1];
    if (mine_wins && ix[0] >= 0) { r.g0 = q.g0; r.a0 = q.a0; r.q0 = q.q0; r.qd0 = q.qd0; r.y0 = q.y0; }
    r.seq = xs.err;  // set before the reduction's barrier
    r.pad = pause;
    *hold = mine_wins;
    return r;
  }
  // every thread holds the reduced values; the record holding each winning
  // index supplies its payload
  if (t < V.nblocks) {
    if (ix[0] >= 0 && q.i0 == ix[0]) {
      xs.win[0].g0 = q.g0; xs.win[0].a0 = q.a0; xs.win[0].q0 = q.q0; xs.win[0].qd0 = q.qd0; xs.win[0].y0 = q.y0;
    }
    if (KIND == KIND_NU && PH == PH_I && ix[1] >= 0 && q.i1 == ix[1]) {
      xs.win[0].g1 = q.g1; xs.win[0].a1 = q.a1; xs.win[0].qd1 = q.qd1; xs.win[0].y1 = q.y1;
    }
  }
  __syncthreads();
  // (xs is written again only in the next exchange, after the next pass's
  // barrier)
  XPart r = xpart_identity<KIND, PH>();
  r.v0 = v[0]; r.v1 = v[1]; r.v2 = v[2]; r.v3 = v[3];
  r.i0 = ix[0]; r.i1 = ix[1];
  if (ix[0] >= 0) {
    r.g0 = xs.win[0].g0; r.a0 = xs.win[0].a0; r.q0 = xs.win[0].q0; r.qd0 = xs.win[0].qd0; r.y0 = xs.win[0].y0;
  }
  if (KIND == KIND_NU && PH == PH_I && ix[1] >= 0) {
    r.g1 = xs.win[0].g1; r.a1 = xs.win[0].a1; r.qd1 = xs.win[0].qd1; r.y1 = xs.win[0].y1;
  }
  r.seq = xs.err;
  r.pad = pause;
  return r;
}

// This owner's vote to pause the solve at the current iteration boundary:
// the stage's time is up, or (batched launches) enough of the launch's
// workgroups have run out of items.
__device__ __forceinline__ int pause_vote(const View &V, const Dec &d) {
  if (V.pause_ticks > 0 && wall_clock64() - d.t_start >= V.pause_ticks) return 1;
  if (V.drain && (d.iter & 31) == 0 && ld_acq(V.drain) >= (unsigned)V.drain_at) return 1;
  return 0;
}

// Dynamic LDS of the launch (every extern __shared__ array starts here).
extern __shared__ float smo_dyn_lds[];

__device__ __forceinline__ CacheMeta private_meta(const View &V, Dec &d) {
  CacheMeta M;
  M.bits = nullptr;
  if (V.full) {
    M.computed = V.pm_flags + pm_index(V) * V.l;
    if (V.flag_lds > 0) M.bits = reinterpret_cast<unsigned *>(smo_dyn_lds + V.flag_lds);
  } else {
    int *base = V.pm_lru + pm_index(V) * V.pm_stride;
    M.slot_of_row = base;
    M.row_of_slot = base + V.l;
    M.prev = base + V.l + V.S;
    M.next = base + V.l + 2 * V.S;
    M.head = base + V.l + 3 * V.S;
    M.tail = M.head + 1;
  }
  M.hits = &d.hits;
  M.misses = &d.misses;
  return M;
}

__device__ __forceinline__ Snap snap_of(const Dec &d) {
  Snap c;
  c.i = d.i; c.j = d.j; c.ip = d.ip; c.in = d.in;
  c.slot_i = d.slot_i; c.slot_j = d.slot_j; c.slot_ip = d.slot_ip; c.slot_in = d.slot_in;
  c.miss_i = d.miss_i; c.miss_j = d.miss_j; c.miss_ip = d.miss_ip; c.miss_in = d.miss_in;
  c.stop = d.stop; c.err = d.err; c.iter = d.iter;
  c.dai = d.dai; c.daj = d.daj; c.Gmax = d.Gmax; c.Gp = d.Gp; c.Gn = d.Gn;
  return c;
}

// Decisions after the i exchange (every workgroup, thread 0).
template <int KIND, bool FULL = false>
__device__ __forceinline__ void decide_i(const View &V, Dec &d, const XPart &r, bool count_maxiter) {
  CacheMeta M = private_meta(V, d);
  int stop = STOP_NONE;
  if (count_maxiter && d.iter >= V.max_iter) stop = STOP_MAXITER;
  if (KIND == KIND_C) {
    d.Gmax = r.v0; d.Gmax2 = r.v1; d.i = r.i0; d.Gi = r.g0; d.ai = r.a0; d.QDi = r.qd0; d.yi = r.y0;
    if (!stop && r.v0 + r.v1 < V.eps) stop = STOP_OPTIMAL;   // libsvm/svm.cpp:919
    if (!stop && r.pad) stop = STOP_PAUSE;
    if (!stop) d.slot_i = cache_lookup<FULL>(V, M, real_of(V, r.i0), -1, &d.miss_i);
  } else {
    d.Gp = r.v0; d.Gn = r.v1; d.Gp2 = r.v2; d.Gn2 = r.v3;
    d.ip = r.i0; d.in = r.i1;
    d.Gip = r.g0; d.aip = r.a0; d.Gin = r.g1; d.ain = r.a1;
    d.QDip = r.qd0; d.yip = r.y0; d.QDin = r.qd1; d.yin = r.y1;
    if (!stop && fmax(r.v0 + r.v2, r.v1 + r.v3) < V.eps) stop = STOP_OPTIMAL;  // libsvm/svm.cpp:1174
    if (!stop && r.pad) stop = STOP_PAUSE;
    if (!stop) {
      int sp = -1;
      d.miss_ip = d.miss_in = 0;
      if (d.ip >= 0) { sp = cache_lookup<FULL>(V, M, real_of(V, d.ip), -1, &d.miss_ip); d.slot_ip = sp; }
      if (d.in >= 0) d.slot_in = cache_lookup<FULL>(V, M, real_of(V, d.in), sp, &d.miss_in);
    }
  }
  d.stop = stop;
}

// Decisions after the j exchange: the pair, the row-j slot and the alpha
// step, computed identically in every workgroup (thread 0).
template <int KIND, bool FULL = false>
__device__ __forceinline__ void decide_j(const View &V, Dec &d, const XPart &r) {
  CacheMeta M = private_meta(V, d);
  const int j = r.i0;
  if (j < 0) {  // only reachable with eps <= 0 or NaN gradients
    d.stop = STOP_OPTIMAL;
    return;
  }
  double Gi, ai, QDi;
  int yi;
  if (KIND == KIND_C) {
    Gi = d.Gi; ai = d.ai; QDi = d.QDi; yi = d.yi;
  } else if (r.y0 == +1) {  // libsvm/svm.cpp:1177-1180
    d.i = d.ip; d.slot_i = d.slot_ip; Gi = d.Gip; ai = d.aip; QDi = d.QDip; yi = d.yip;
  } else {
    d.i = d.in; d.slot_i = d.slot_in; Gi = d.Gin; ai = d.ain; QDi = d.QDin; yi = d.yin;
  }
  const int yj = r.y0;
  d.j = j;
  d.yi = yi;
  d.yj = yj;
  d.slot_j = cache_lookup<FULL>(V, M, real_of(V, j), d.slot_i, &d.miss_j);
  d.iter += 1;
  double ai2 = ai, aj2 = r.a0;
  alpha_pair(yi, yj, yi > 0 ? V.Cp : V.Cn, yj > 0 ? V.Cp : V.Cn, QDi, r.qd0, Gi, r.g0, r.q0, ai2, aj2);
  d.dai = ai2 - ai;
  d.daj = aj2 - r.a0;
  d.ai_new = ai2;
  d.aj_new = aj2;
  // Solver::update_alpha_status (libsvm/svm.cpp:425-432) with the labels at hand
  const double Ci = yi > 0 ? V.Cp : V.Cn, Cj = yj > 0 ? V.Cp : V.Cn;
  d.st_i = ai2 >= Ci ? AT_UPPER : (ai2 <= 0 ? AT_LOWER : IS_FREE);
  d.st_j = aj2 >= Cj ? AT_UPPER : (aj2 <= 0 ? AT_LOWER : IS_FREE);
  d.pend = 1;
}

// One operand of the prediction kernel: rows in SELL-64 (the solver's
// layout) with float (f64 = 0: the serving runtime's resident models, stored
// as training uploads them) or double (f64 = 1: svmhip_predict_batch, the
// caller's FP64 features as svm_predict_values uses them) values.
struct PRows {
  int l, d, nslices, dense, f64;
  const long long *slice_ptr;
  const void *vals;
  const int *cols;         // null when dense
  const int *row_len;
};

// Arguments of the prediction kernel.
struct PredictArgs {
  PRows sv;                // support vectors (l = nSV)
  PRows tv;                // test rows
  int ktype, degree;
  double gamma, coef0;
  const double *coef;      // [(k-1) x nsv]
  const double *rho;       // [npair]
  const int *cls_start;    // [k] first SV of each class (classification)
  const int *cls_count;    // [k]
  int k, npair, tb, regression_like;
  int ld;                  // doubles per test row staged in LDS (>= both operands' d)
  double *dec;             // [ntest x npair]
};

}  // namespace smo
