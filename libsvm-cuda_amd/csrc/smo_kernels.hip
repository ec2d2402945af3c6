// gfx950 kernels of the SMO solver and their launch wrappers.
//
// Kernel set (each replaces one or more reference kernels, see the comments):
//   k_self_dot      x^2 per row                (cuda_setup_x_square, svm_device.cu:771-777)
//   k_qd            QD_v = K(v, v) in FP64     (cuda_setup_QD, svm_device.cu:779-789)
//   k_ginit_row     G += alpha_v Q_v, one v    (cuda_init_gradient_block, svm_device.cu:717-758)
//   k_select_i      WSS i-step + Gmax2          (cuda_prep_gmax + cuda_find_gmax, :760-821)
//   k_select_j      WSS j-step                  (cuda_compute_obj_diff + cuda_find_min_idx, :476-561)
//   k_alpha         pair update, FP64           (cuda_compute_alpha, :828-927)
//   k_gupdate       G update                    (cuda_update_gradient, :591-626)
//   k_status        status of the pair          (cuda_update_alpha_status, :939-948)
//   k_smo           persistent cooperative kernel running select j -> step ->
//                   G update + select i for many iterations (no reference
//                   counterpart: the reference drives ~9 launches and 2 blocking
//                   copies per iteration from the host, cuda_solver.cpp:493-518)
//   k_finalize      rho / obj / r in FP64       (Solver::calculate_rho, svm.cpp:1013-1049,
//                                                Solver_NU::calculate_rho, :1260-1308)
//   k_kernel_row    one kernel row to a buffer  (test hook)
//   k_smo_batch     many subproblems per launch, each by one workgroup or a
//                   team of them (no reference counterpart: the reference
//                   solves one-vs-one pairs and folds one after another)
//   k_pack_rowset   copies a row subset into its own SELL arrays
#include <stdlib.h>
#include <string.h>

#include "smo_device.hpp"

namespace smo {

__global__ void __launch_bounds__(BLOCK) k_self_dot(View V, double *xsq) {
  int r = blockIdx.x * BLOCK + threadIdx.x;
  if (r >= V.l) return;
  const int len = V.row_len[r];
  const long long base = V.slice_ptr[r / SLICE] + (r % SLICE);
  double acc = 0.0;
  // 8 entries per trip (one memory round trip each), the last trip's loads
  // clamped to the row
  for (int p = 0; p < len; p += 8) {
    float a[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) a[u] = V.vals[base + (long long)min(p + u, len - 1) * SLICE];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const double v = (double)a[u];
      if (p + u < len) acc = fma(v, v, acc);  // v*v is exact: same as multiply-then-add
    }
  }
  xsq[r] = acc;
}

__global__ void __launch_bounds__(BLOCK) k_qd(View V) {
  int v = blockIdx.x * BLOCK + threadIdx.x;
  if (v >= V.nvar) return;
  int r = real_of(V, v);
  double q;
  if (V.ktype == K_PRECOMPUTED) {
    // x_r[serial_r].value, by node position (libsvm/svm.cpp:255)
    int s = V.serial[r];
    const long long base = V.slice_ptr[r / SLICE] + (r % SLICE);
    q = (s >= 0 && s < V.row_len[r]) ? (double)V.vals[base + (long long)s * SLICE] : 0.0;
  } else {
    q = kernel_f64(V, r, r, V.xsq[r]);
  }
  V.QD[v] = q;
}

// G_k += alpha_v * Q_vk for all k: one row of the initial gradient sum
// (libsvm/svm.cpp:555-569).  Launched once per v with alpha_v != 0, in
// increasing v, so every G_k accumulates in LIBSVM's order.
__global__ void __launch_bounds__(BLOCK) k_ginit_row(View V, int v) {
  extern __shared__ float xs_lds[];
  float *xs = V.use_lds ? xs_lds : V.xs_global + (long long)blockIdx.x * V.d;
  int sb, se;
  owned_slices(V, blockIdx.x, gridDim.x, sb, se);
  const int q = real_of(V, v);
  stage_row(V, q, xs);
  const double av = V.alpha[v];
  const int8_t yv = V.y[v];
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64, nw = NWAVES;
  const int nv = V.svr ? 2 : 1;
  for (int s = sb + wave; s < se; s += nw) {
    int r = s * SLICE + lane;
    if (r >= V.l) continue;
    float K = kernel_value(V, xs, q, r);
    for (int h = 0; h < nv; ++h) {
      int k = r + h * V.l;
      float Q = (V.y[k] == yv) ? K : -K;
      V.G[k] = V.G[k] + av * (double)Q;
    }
  }
}

template <int KIND>
__global__ void __launch_bounds__(BLOCK) k_select_i(View V) {
  __shared__ RedSmem sm;
  extern __shared__ float xs_lds[];
  float *xs = V.use_lds ? xs_lds : V.xs_global + (long long)blockIdx.x * V.d;
  Snap c = snap(V.ctl);
  pass_gradient_select<KIND>(V, c, xs, sm, false, true);
  if (arrive_last(V, sm)) leader_after_select_i<KIND>(V, sm, false);
}

template <int KIND>
__global__ void __launch_bounds__(BLOCK) k_select_j(View V) {
  __shared__ RedSmem sm;
  extern __shared__ float xs_lds[];
  float *xs = V.use_lds ? xs_lds : V.xs_global + (long long)blockIdx.x * V.d;
  Snap c = snap(V.ctl);
  pass_select_j<KIND>(V, c, xs, sm);
  if (arrive_last(V, sm)) leader_after_select_j<KIND>(V, sm, false);
}

__global__ void k_alpha(View V) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  Ctl *c = V.ctl;
  int i = c->i, j = c->j;
  alpha_step(V, i, j, q_elem(V, c->slot_i, i, j));
}

template <int KIND>
__global__ void __launch_bounds__(BLOCK) k_gupdate(View V) {
  __shared__ RedSmem sm;
  extern __shared__ float xs_lds[];
  float *xs = V.use_lds ? xs_lds : V.xs_global + (long long)blockIdx.x * V.d;
  Snap c = snap(V.ctl);
  pass_gradient_select<KIND>(V, c, xs, sm, true, false);
}

__global__ void k_status(View V) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  Ctl *c = V.ctl;
  V.status[c->i] = status_for(V, c->i, V.alpha[c->i]);
  V.status[c->j] = status_for(V, c->j, V.alpha[c->j]);
}

// Record of this workgroup's i-pass winner(s) with their G and alpha
// (thread 0; the pass's block reduction has synchronised the workgroup).
template <int KIND>
__device__ __forceinline__ XPart make_i_record(const View &V, const Partial &bp) {
  XPart m = xpart_identity<KIND, PH_I>();
  m.v0 = bp.v0; m.v1 = bp.v1; m.v2 = bp.v2; m.v3 = bp.v3;
  m.i0 = bp.i0; m.i1 = bp.i1;
  if (threadIdx.x == 0) {
    if (bp.i0 >= 0) { m.g0 = V.G[bp.i0]; m.a0 = V.alpha[bp.i0]; m.qd0 = V.QD[bp.i0]; m.y0 = V.y[bp.i0]; }
    if (KIND == KIND_NU && bp.i1 >= 0) {
      m.g1 = V.G[bp.i1]; m.a1 = V.alpha[bp.i1]; m.qd1 = V.QD[bp.i1]; m.y1 = V.y[bp.i1];
    }
  }
  return m;
}

// Record of this workgroup's j-pass winner with Q_ij, G_j and alpha_j.
template <int KIND>
__device__ __forceinline__ XPart make_j_record(const View &V, const Snap &c, const Partial &bp) {
  XPart m = xpart_identity<KIND, PH_J>();
  m.v0 = bp.v0;
  m.i0 = bp.i0;
  if (threadIdx.x == 0 && bp.i0 >= 0) {
    const int j = bp.i0;
    int i = c.i, slot = c.slot_i;
    if (KIND == KIND_NU) {
      if (V.y[j] == +1) { i = c.ip; slot = c.slot_ip; }
      else { i = c.in; slot = c.slot_in; }
    }
    m.q0 = q_elem(V, slot, i, j);
    m.g0 = V.G[j];
    m.a0 = V.alpha[j];
    m.qd0 = V.QD[j];
    m.y0 = V.y[j];
  }
  return m;
}

// First selection of a solve (or of a resumed solve): i from the current G,
// exchanged, then the stop test and the row-i slot.
template <int KIND, bool PIPE = false>
__device__ __forceinline__ void smo_select_first(const View &V, Dec &d, RedSmem &sm, XSmem &xsm, float *xs) {
  Snap c = snap_of(d);
  Partial bp = pass_gradient_select<KIND, PASS_SLICES, PIPE>(V, c, xs, sm, false, true, false, nullptr);
  XPart mine = make_i_record<KIND>(V, bp);
  const unsigned seq = (unsigned)d.epoch + 1u;
  XPart r = exchange<KIND, PH_I>(V, sm, xsm, mine, seq);
  if (threadIdx.x == 0) {
    d.epoch = (int)seq;
    d.phase = 1;
    if (r.seq) d.err = 1;
    else decide_i<KIND>(V, d, r, true);
  }
  __syncthreads();
}

// Speculative kernel rows (C rule, full cache).  When row i misses, the
// rows of the strongest other violators are computed in the same pass over
// X: a violator of one iteration is very likely to be selected in a later
// one, and one pass serves several rows (row_dot_multi).  The candidates are
// the owners' best I_up and best I_low variables of the i exchange just
// made, whose records every owner reads alike (they are overwritten only two
// exchanges later), so every owner marks the same rows as cached.  Every
// kernel value is the one a miss would compute.
constexpr int SPEC_ROWS = 8;  // row i plus up to 7 candidates
constexpr int SPEC_UP = 3, SPEC_LOW = 4;

// Owned parts of the rows in list[0..cnt) (full cache: slot = row).
__device__ __forceinline__ void compute_rows_part(const View &V, const int *list, int cnt, float *xs, int sb,
                                                  int se) {
  if (V.gdot) {  // gathers of the shared dot products
    for (int m = 0; m < cnt; ++m) compute_row_part(V, list[m], list[m], xs, sb, se);
    return;
  }
  for (int m = 0; m < cnt; ++m) stage_row(V, list[m], xs + m * V.d);
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64, nw = NWAVES;
  for (int sl = sb + wave; sl < se; sl += nw) {
    const int r = sl * SLICE + lane;
    if (r >= V.l) continue;
    double acc[SPEC_ROWS];
    row_dot_multi<SPEC_ROWS>(V, xs, V.d, r, acc);
#pragma unroll
    for (int m = 0; m < SPEC_ROWS; ++m)
      if (m < cnt) V.rows[(long long)list[m] * V.row_stride + r] = (float)kernel_f64(V, list[m], r, acc[m]);
  }
  __syncthreads();
}

template <int KIND>
__device__ __forceinline__ void spec_rows(const View &V, Dec &d, RedSmem &sm, float *xs) {
  __shared__ int list[SPEC_ROWS];
  __shared__ int s_cnt;
  const unsigned par = (unsigned)d.epoch & 1u;
  const int t = threadIdx.x;
  double cu = -INF, cl = -INF;
  int iu = -1, il = -1;
  if (t < V.nblocks) {
    // the i record (C rule): i0, i1, v0, v1 in words 0..5
    unsigned w[XW];
#pragma unroll
    for (int k = 0; k < 6; ++k) w[k] = xpeek(V.xw, V.nblocks, par, k, t, V.xrep);
    int k = 2;
    iu = (int)w[0];
    il = (int)w[1];
    cu = xget_d(w, k);
    cl = xget_d(w, k);
    if (iu < 0 || iu == d.i) cu = -INF;
    if (il < 0 || il == d.i) cl = -INF;
  }
  // each wave keeps its own top candidates (wave reductions only), then
  // thread 0 merges the waves' lists
  __shared__ double wv[16][SPEC_UP + SPEC_LOW];
  __shared__ int wi[16][SPEC_UP + SPEC_LOW];
  const int wave = t / 64, lane = t % 64, nw = NWAVES;
#pragma unroll
  for (int k = 0; k < SPEC_UP + SPEC_LOW; ++k) {
    const bool up = k < SPEC_UP;
    double v = up ? cu : cl;
    int ix = up ? iu : il;
    wave_max_idx(v, ix);
    if (lane == 0) {
      wv[wave][k] = v;
      wi[wave][k] = v > -INF ? ix : -1;
    }
    if (v > -INF) {
      if (up && iu == ix) cu = -INF;
      if (!up && il == ix) cl = -INF;
    }
  }
  __syncthreads();
  int picks[SPEC_UP + SPEC_LOW];
  if (threadIdx.x == 0) {
    for (int side = 0; side < 2; ++side) {
      const int k0 = side == 0 ? 0 : SPEC_UP, k1 = side == 0 ? SPEC_UP : SPEC_UP + SPEC_LOW;
      for (int k = k0; k < k1; ++k) {  // the best entry of this side not yet picked
        double bv = -INF;
        int bi = -1, bw = -1, bk = -1;
        for (int w = 0; w < nw; ++w)
          for (int kk = k0; kk < k1; ++kk)
            if (wi[w][kk] >= 0 && (wv[w][kk] > bv || (wv[w][kk] == bv && wi[w][kk] > bi))) {
              bv = wv[w][kk];
              bi = wi[w][kk];
              bw = w;
              bk = kk;
            }
        picks[k] = bi;
        if (bw >= 0) wi[bw][bk] = -1;
      }
    }
  }
  if (threadIdx.x == 0) {
    CacheMeta M = private_meta(V, d);
    int cnt = 0;
    list[cnt++] = real_of(V, d.i);  // decide_i has marked it
    for (int k = 0; k < SPEC_UP + SPEC_LOW; ++k) {
      if (picks[k] < 0) continue;
      const int row = real_of(V, picks[k]);
      bool dup = false;
      for (int m = 0; m < cnt; ++m) dup |= list[m] == row;
      if (dup || flag_test_set(M, row)) continue;
      ++d.misses;
      list[cnt++] = row;
    }
    s_cnt = cnt;
  }
  __syncthreads();
  int sb, se;
  owned_slices(V, my_owner(V), V.nblocks, sb, se);
  compute_rows_part(V, list, s_cnt, xs, sb, se);
  if (threadIdx.x == 0) d.miss_i = 0;
  __syncthreads();
}

// ------------------------------------------ resident vectors (persistent kernel)
//
// With V.res_off > 0 each workgroup of k_smo keeps its owned variables' G,
// alpha, QD, labels and statuses in LDS for the whole launch (read at the
// start; G, alpha and statuses written back at the end), together with the
// owned segments of the iteration's kernel rows.  A pass then waits on one
// memory round trip (the cached row segment) instead of two (the vectors,
// then the winner's scalars for its record), and G is not written to device
// memory inside the loop.  Variable v of owned row r = real(v) sits at index
// (v >= l ? m : 0) + r - base.  Within a pass, every entry is read and
// written by the thread that owns its row in the passes' (wave, lane, slice)
// mapping; the records read the winners' entries after the pass's
// reduction barrier.
struct Res {
  double *G, *A, *QD;
  float *seg[3];  // kernel-row segments: C rule [0] row i, [1] row j; nu rule [0] ip, [1] in, [2] j
  int8_t *y, *st;
  int m, base;
};

__device__ __forceinline__ Res res_view(const View &V, int sb) {
  Res R;
  const int nv = V.svr ? 2 : 1, m = V.res_m;
  double *dp = reinterpret_cast<double *>(smo_dyn_lds + V.res_off);
  R.G = dp;
  R.A = dp + nv * m;
  R.QD = dp + 2 * nv * m;
  float *fp = reinterpret_cast<float *>(dp + 3 * nv * m);
  R.seg[0] = fp;
  R.seg[1] = fp + m;
  R.seg[2] = fp + 2 * m;
  int8_t *bp = reinterpret_cast<int8_t *>(fp + 3 * m);
  R.y = bp;
  R.st = bp + nv * m;
  R.m = m;
  R.base = sb * SLICE;
  return R;
}

__device__ __forceinline__ int res_index(const View &V, const Res &R, int v) {
  return v >= V.l ? R.m + (v - V.l) - R.base : v - R.base;
}

// The owned variables' vectors into LDS (start of a launch).  Ends with a barrier.
__device__ __forceinline__ void res_load(const View &V, const Res &R, int sb, int se) {
  const int r0 = sb * SLICE, r1 = min(se * SLICE, V.l), nv = V.svr ? 2 : 1;
  for (int h = 0; h < nv; ++h)
    for (int r = r0 + (int)threadIdx.x; r < r1; r += BLOCK) {
      const int k = r + h * V.l, x = h * R.m + (r - r0);
      R.G[x] = V.G[k];
      R.A[x] = V.alpha[k];
      R.QD[x] = V.QD[k];
      R.y[x] = V.y[k];
      R.st[x] = V.status[k];
    }
  __syncthreads();
}

// G, alpha and statuses back to device memory (end of a launch, after a barrier).
__device__ __forceinline__ void res_store(const View &V, const Res &R, int sb, int se) {
  const int r0 = sb * SLICE, r1 = min(se * SLICE, V.l), nv = V.svr ? 2 : 1;
  for (int h = 0; h < nv; ++h)
    for (int r = r0 + (int)threadIdx.x; r < r1; r += BLOCK) {
      const int k = r + h * V.l, x = h * R.m + (r - r0);
      V.G[k] = R.G[x];
      V.alpha[k] = R.A[x];
      V.status[k] = R.st[x];
    }
}

// Query rows fetched right after an exchange, before the decisions that
// follow it, for rows the next pass will likely compute (their cache flag
// was clear): dense rows of at most BLOCK features, one value per thread.
// The pass stages a row from here when the tag names it, else from memory
// (the flag test is a hint: the values are X's either way).
struct QPre {
  float v[2];
  int row[2];  // -1: nothing fetched into v[m]
};

__device__ __forceinline__ void prefetch_query(const View &V, int row, float &v, int &tag) {
  tag = -1;
  if (row < 0 || !V.dense || V.d > BLOCK || V.flag_lds <= 0) return;
  const unsigned *bits = reinterpret_cast<const unsigned *>(smo_dyn_lds + V.flag_lds);
  // (the decision may be setting the bit meanwhile: a hint, read atomically)
  const unsigned w = __hip_atomic_load(&bits[row >> 5], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
  if ((w >> (row & 31)) & 1u) return;  // cached: no pass computes it
  // dense rows are stored at width d in every slice
  const long long base = (long long)(row / SLICE) * SLICE * V.d + (row % SLICE);
  v = V.vals[base + (long long)min((int)threadIdx.x, V.d - 1) * SLICE];
  tag = row;
}

// compute_row_part that also keeps the owned segment of the row in LDS; the
// query row from qv when qtag names it (dense rows: the values stage_row
// reads).
template <bool PIPE>
__device__ __forceinline__ void compute_row_seg(const View &V, int q, int slot, float *xs, int sb, int se, float *seg,
                                                int base, float qv = 0.f, int qtag = -1) {
  if (qtag == q) {
    if ((int)threadIdx.x < V.d) xs[threadIdx.x] = qv;
    __syncthreads();
  } else {
    stage_row(V, q, xs);
  }
  float *dst = V.rows + (long long)slot * V.row_stride;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64, nw = NWAVES;
  for (int s = sb + wave; s < se; s += nw) {
    const int r = s * SLICE + lane;
    if (r < V.l) {
      const float kv = kernel_value<PIPE>(V, xs, q, r);
      dst[r] = kv;
      seg[r - base] = kv;
    }
  }
  __syncthreads();
}

// Row segments fetched right after an exchange, before the decisions that
// follow it (full cache: a row's slot is the row itself): the first trip's
// PS slices of this wave's part of cached row `row`.  Used by the next pass
// only if the row turns out to be cached; otherwise discarded (the loads
// read valid cache memory either way).
template <int PS>
struct Pre {
  float k[2][PS];
  int row[2];  // -1: nothing fetched into k[m]
};

template <int PS>
__device__ __forceinline__ void prefetch_seg(const View &V, int row, int sb, int se, float (&k)[PS]) {
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64, nw = NWAVES;
  const long long off = (long long)row * V.row_stride;
#pragma unroll
  for (int u = 0; u < PS; ++u) {
    const int rc = min(min(sb + wave + u * nw, se - 1) * SLICE + lane, V.l - 1);
    k[u] = V.rows[off + rc];
  }
}

// The segment of one cached row for this wave's PS slices of a trip (rows
// clamped to the last owned one, as in pass_select_j_nv), from the fetch
// made after the last exchange, from the cache (the loads issued back to
// back), or from the LDS copy of a row computed in this pass.  The branches
// are uniform over the wave, so no load waits on another.
template <int PS>
__device__ __forceinline__ void seg_fetch(const View &V, const Res &R, int s0, int se, int nw, int lane, bool pre,
                                          const float (&pk)[PS], bool cache, long long off, const float *seg,
                                          float (&K)[PS]) {
  if (pre) {
#pragma unroll
    for (int u = 0; u < PS; ++u) K[u] = pk[u];
  } else if (cache) {
#pragma unroll
    for (int u = 0; u < PS; ++u) K[u] = V.rows[off + min(min(s0 + u * nw, se - 1) * SLICE + lane, V.l - 1)];
  } else if (seg) {
#pragma unroll
    for (int u = 0; u < PS; ++u) K[u] = seg[min(min(s0 + u * nw, se - 1) * SLICE + lane, V.l - 1) - R.base];
  } else {
#pragma unroll
    for (int u = 0; u < PS; ++u) K[u] = 0.f;
  }
}

// Pass "select j" over the resident vectors (the arithmetic of
// pass_select_j_nv, libsvm/svm.cpp:866-917 and :1123-1172): the segments of
// row i (rows ip, in) come from the cache, or are computed on a miss, and
// are kept for the update pass.  The selection is written with selects
// rather than branches (the same comparisons, in the same order).
template <int KIND, int NV, int PS>
__device__ __forceinline__ Partial rpass_select_j(const View &V, const Snap &c, const Dec &d, float *xs, RedSmem &sm,
                                                  const Res &R, int sb, int se, const Pre<PS> &pre,
                                                  const QPre *qp = nullptr) {
  const int qi = KIND == KIND_C ? c.i : c.ip;
  const int qn = KIND == KIND_C ? -1 : c.in;
  const int mi = KIND == KIND_C ? c.miss_i : c.miss_ip;
  const int mn = KIND == KIND_C ? 0 : c.miss_in;
  const int slot_i = KIND == KIND_C ? c.slot_i : c.slot_ip;
  if (qi >= 0 && mi)
    compute_row_seg<true>(V, real_of(V, qi), slot_i, xs, sb, se, R.seg[0], R.base, qp ? qp->v[0] : 0.f,
                          qp ? qp->row[0] : -1);
  if (qn >= 0 && mn)
    compute_row_seg<true>(V, real_of(V, qn), c.slot_in, xs, sb, se, R.seg[1], R.base, qp ? qp->v[1] : 0.f,
                          qp ? qp->row[1] : -1);
  const bool li = qi >= 0 && !mi, ln = qn >= 0 && !mn;  // segments read from the cache here
  const long long ri = qi >= 0 ? (long long)slot_i * V.row_stride : 0;
  const long long rn = qn >= 0 ? (long long)c.slot_in * V.row_stride : 0;
  // segments fetched after the last exchange (the same rows, full cache)
  bool pi = li && V.full && pre.row[0] == real_of(V, qi);
  bool pn = ln && V.full && pre.row[1] == real_of(V, qn);
  // the nu rule's ip has y = +1 and in has y = -1 by construction
  const int8_t yi = KIND == KIND_C ? (int8_t)d.yi : (int8_t)1, yn = (int8_t)-1;
  const double QDi = KIND == KIND_C ? d.QDi : d.QDip, QDn = KIND == KIND_C ? 0.0 : d.QDin;
  double best = INF;
  int bj = -1;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64, nw = NWAVES;
  for (int s0 = sb + wave; s0 < se; s0 += nw * PS) {
    float Ki[PS], Kn[PS];
    seg_fetch<PS>(V, R, s0, se, nw, lane, pi, pre.k[0], li, ri, qi >= 0 ? R.seg[0] : nullptr, Ki);
    if (KIND == KIND_NU) seg_fetch<PS>(V, R, s0, se, nw, lane, pn, pre.k[1], ln, rn, qn >= 0 ? R.seg[1] : nullptr, Kn);
    pi = pn = false;
#pragma unroll
    for (int u = 0; u < PS; ++u) {
      const int r = (s0 + u * nw) * SLICE + lane;
      if (s0 + u * nw >= se || r >= V.l) continue;
      const int lr = r - R.base;
      if (li) R.seg[0][lr] = Ki[u];
      if (KIND == KIND_NU && ln) R.seg[1][lr] = Kn[u];
#pragma unroll
      for (int h = 0; h < NV; ++h) {
        const int k = r + h * V.l, x = h * R.m + lr;
        const int8_t y = R.y[x], sk = R.st[x];
        const double gk = R.G[x], qd = R.QD[x];
        const bool pos = y == +1;
        bool cand;
        double gd, quad;
        if (KIND == KIND_C) {
          cand = pos ? sk != AT_LOWER : sk != AT_UPPER;
          gd = pos ? c.Gmax + gk : c.Gmax - gk;
          const float Qik = (yi == y) ? Ki[u] : -Ki[u];
          const double t = (2.0 * yi) * (double)Qik;
          quad = pos ? (QDi + qd) - t : (QDi + qd) + t;
        } else {
          cand = pos ? sk != AT_LOWER : sk != AT_UPPER;
          gd = pos ? c.Gp + gk : c.Gn - gk;
          const float Q = pos ? ((yi == y) ? Ki[u] : -Ki[u]) : ((yn == y) ? Kn[u] : -Kn[u]);
          quad = ((pos ? QDi : QDn) + qd) - (double)(2.0f * Q);
        }
        const double od = -(gd * gd) / (quad > 0 ? quad : TAU);
        const bool take = cand && gd > 0 && (od < best || (od == best && k > bj));
        best = take ? od : best;
        bj = take ? k : bj;
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
  return pt;
}

// Pass "update G, select i" over the resident vectors (the arithmetic of
// pass_gradient_select_nv, libsvm/svm.cpp:724-736 and :840-863): the owners
// store the pending pair's alpha and statuses, row j's segment comes from
// the cache (or is computed on a miss), row i's is the one the j pass kept.
// Selections as selects, as in rpass_select_j.
template <int KIND, int NV, int PS>
__device__ __forceinline__ Partial rpass_update_select_i(const View &V, const Snap &c, const Dec &d, float *xs,
                                                         RedSmem &sm, const Res &R, int sb, int se, const Pend &pd,
                                                         const Pre<PS> &pre, const QPre *qp = nullptr) {
  if (threadIdx.x == 0) {
    if (owns(V, sb, se, pd.i)) {
      const int x = res_index(V, R, pd.i);
      R.A[x] = pd.ai;
      R.st[x] = (int8_t)pd.si;
    }
    if (owns(V, sb, se, pd.j)) {
      const int x = res_index(V, R, pd.j);
      R.A[x] = pd.aj;
      R.st[x] = (int8_t)pd.sj;
    }
  }
  // (the statuses of i and j are taken from pd below, not from those stores)
  const int sj = KIND == KIND_C ? 1 : 2;
  const int si = KIND == KIND_C ? 0 : (c.i == c.ip ? 0 : 1);
  if (c.miss_j)
    compute_row_seg<true>(V, real_of(V, c.j), c.slot_j, xs, sb, se, R.seg[sj], R.base, qp ? qp->v[0] : 0.f,
                          qp ? qp->row[0] : -1);
  const bool lj = !c.miss_j;
  const long long rj = (long long)c.slot_j * V.row_stride;
  bool pj = lj && V.full && pre.row[0] == real_of(V, c.j);  // fetched after the j exchange
  const int8_t yi = (int8_t)d.yi, yj = (int8_t)d.yj;
  double v[4] = {-INF, -INF, -INF, -INF};
  int ix[2] = {-1, -1};
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64, nw = NWAVES;
  for (int s0 = sb + wave; s0 < se; s0 += nw * PS) {
    float Kj[PS];
    seg_fetch<PS>(V, R, s0, se, nw, lane, pj, pre.k[0], lj, rj, R.seg[sj], Kj);
    pj = false;
#pragma unroll
    for (int u = 0; u < PS; ++u) {
      const int r = (s0 + u * nw) * SLICE + lane;
      if (s0 + u * nw >= se || r >= V.l) continue;
      const int lr = r - R.base;
      const float Ki = R.seg[si][lr];
#pragma unroll
      for (int h = 0; h < NV; ++h) {
        const int k = r + h * V.l, x = h * R.m + lr;
        const int8_t yk = R.y[x];
        double gk = R.G[x];
        float Qik = (yi == yk) ? Ki : -Ki;
        float Qjk = (yj == yk) ? Kj[u] : -Kj[u];
        double t1 = (double)Qik * c.dai;
        double t2 = (double)Qjk * c.daj;
        gk = gk + (t1 + t2);
        R.G[x] = gk;
        int8_t sk = R.st[x];
        sk = k == pd.i ? (int8_t)pd.si : (k == pd.j ? (int8_t)pd.sj : sk);
        const bool pos = yk == +1;
        if (KIND == KIND_C) {
          // I_up: -y G (max_idx); I_low: y G (plain max, or max_idx for speculative rows)
          const bool up = pos ? sk != AT_UPPER : sk != AT_LOWER;
          const bool low = pos ? sk != AT_LOWER : sk != AT_UPPER;
          const double a = pos ? -gk : gk, b = pos ? gk : -gk;
          const bool t0 = up && (a > v[0] || (a == v[0] && k > ix[0]));
          v[0] = t0 ? a : v[0];
          ix[0] = t0 ? k : ix[0];
          if (V.spec) {
            const bool t1s = low && (b > v[1] || (b == v[1] && k > ix[1]));
            v[1] = t1s ? b : v[1];
            ix[1] = t1s ? k : ix[1];
          } else {
            v[1] = low ? fmax(v[1], b) : v[1];
          }
        } else {
          const bool c0 = pos && sk != AT_UPPER, c2 = pos && sk != AT_LOWER;
          const bool c1 = !pos && sk != AT_LOWER, c3 = !pos && sk != AT_UPPER;
          const bool t0 = c0 && (-gk > v[0] || (-gk == v[0] && k > ix[0]));
          v[0] = t0 ? -gk : v[0];
          ix[0] = t0 ? k : ix[0];
          const bool t1 = c1 && (gk > v[1] || (gk == v[1] && k > ix[1]));
          v[1] = t1 ? gk : v[1];
          ix[1] = t1 ? k : ix[1];
          v[2] = c2 ? fmax(v[2], gk) : v[2];
          v[3] = c3 ? fmax(v[3], -gk) : v[3];
        }
      }
    }
  }
  Partial pt;
  if (KIND == KIND_C) {
    const int kinds_spec[4] = {0, 0, 2, 2}, kinds[4] = {0, 2, 2, 2};
    if (V.spec) block_reduce<2>(sm, v, ix, kinds_spec);
    else block_reduce<2>(sm, v, ix, kinds);
  } else {
    const int kinds[4] = {0, 0, 2, 2};
    block_reduce<4>(sm, v, ix, kinds);
  }
  pt.v0 = v[0]; pt.v1 = v[1]; pt.v2 = v[2]; pt.v3 = v[3];
  pt.i0 = ix[0]; pt.i1 = ix[1]; pt.pad0 = pt.pad1 = 0;
  return pt;
}

// make_i_record / make_j_record with the winners' scalars from LDS.
template <int KIND>
__device__ __forceinline__ XPart rmake_i_record(const View &V, const Res &R, const Partial &bp) {
  XPart m = xpart_identity<KIND, PH_I>();
  m.v0 = bp.v0; m.v1 = bp.v1; m.v2 = bp.v2; m.v3 = bp.v3;
  m.i0 = bp.i0; m.i1 = bp.i1;
  if (threadIdx.x == 0) {
    if (bp.i0 >= 0) {
      const int x = res_index(V, R, bp.i0);
      m.g0 = R.G[x]; m.a0 = R.A[x]; m.qd0 = R.QD[x]; m.y0 = R.y[x];
    }
    if (KIND == KIND_NU && bp.i1 >= 0) {
      const int x = res_index(V, R, bp.i1);
      m.g1 = R.G[x]; m.a1 = R.A[x]; m.qd1 = R.QD[x]; m.y1 = R.y[x];
    }
  }
  return m;
}

template <int KIND>
__device__ __forceinline__ XPart rmake_j_record(const View &V, const Res &R, const Dec &d, const Partial &bp) {
  XPart m = xpart_identity<KIND, PH_J>();
  m.v0 = bp.v0;
  m.i0 = bp.i0;
  if (threadIdx.x == 0 && bp.i0 >= 0) {
    const int j = bp.i0, x = res_index(V, R, j);
    const int yj = R.y[x];
    // Q_ij from the segment of row i (nu rule: of ip for y_j = +1, of in
    // for y_j = -1, whose labels equal y_j)
    const int yi = KIND == KIND_C ? d.yi : yj;
    const float k = R.seg[KIND == KIND_C ? 0 : (yj == +1 ? 0 : 1)][real_of(V, j) - R.base];
    m.q0 = yi == yj ? k : -k;
    m.g0 = R.G[x]; m.a0 = R.A[x]; m.qd0 = R.QD[x]; m.y0 = yj;
  }
  return m;
}

// smo_iterations over the resident vectors (same decisions and exchanges).
// PS: slices a wave takes per trip of a pass (1 when each wave owns at most
// one slice: the unrolled form's idle copies cost registers and time).
// FULL: the row cache holds every row (V.full).  QF: query rows of likely
// misses fetched after each exchange (V.qpre).
template <int KIND, int NV, int PS, bool FULL, bool QF = false>
__device__ __forceinline__ void smo_iterations_res(const View &V, Dec &d, RedSmem &sm, XSmem &xsm, float *xs,
                                                   const Res &R, int sb, int se, int iter_limit) {
  Pre<PS> pi, pj;  // segments fetched for the next j pass (rows i / ip, in) and the next update pass (row j)
  pi.row[0] = pi.row[1] = pj.row[0] = pj.row[1] = -1;
  // QF: query rows fetched for the same passes (V.qpre; the other forms,
  // and solves that never miss, would pay more for its registers than the
  // fetch saves)
  QPre qi, qj;
  qi.row[0] = qi.row[1] = qj.row[0] = qj.row[1] = -1;
  qi.v[0] = qi.v[1] = qj.v[0] = qj.v[1] = 0.f;
  for (int n = 0; n < iter_limit; ++n) {
    if (d.stop || d.err) break;
    if (KIND == KIND_C && V.spec && d.miss_i) {
      spec_rows<KIND>(V, d, sm, xs);
      pi.row[0] = -1;  // row i was computed after the fetch
    }
    Snap c = snap_of(d);
    Partial bp = rpass_select_j<KIND, NV, PS>(V, c, d, xs, sm, R, sb, se, pi, QF ? &qi : nullptr);
    XPart mine = rmake_j_record<KIND>(V, R, d, bp);
    unsigned seq = (unsigned)d.epoch + 1u;
    bool hold = threadIdx.x == 0;
    XPart r = exchange<KIND, PH_J>(V, sm, xsm, mine, seq, &hold);
    // (every thread holds the winning index)
    pj.row[0] = V.full && r.i0 >= 0 ? real_of(V, r.i0) : -1;
    if (pj.row[0] >= 0) prefetch_seg<PS>(V, pj.row[0], sb, se, pj.k[0]);
    if (QF) prefetch_query(V, pj.row[0], qj.v[0], qj.row[0]);
    if (hold) {
      d.epoch = (int)seq;
      if (r.seq) d.err = 1;
      else decide_j<KIND, FULL>(V, d, r);
    }
    __syncthreads();
    if (d.stop || d.err) break;
    c = snap_of(d);
    Pend pd;
    pd.i = d.i; pd.j = d.j; pd.ai = d.ai_new; pd.aj = d.aj_new; pd.si = d.st_i; pd.sj = d.st_j;
    Partial bp2 = rpass_update_select_i<KIND, NV, PS>(V, c, d, xs, sm, R, sb, se, pd, pj, QF ? &qj : nullptr);
    XPart mine2 = rmake_i_record<KIND>(V, R, bp2);
    mine2.pad = pause_vote(V, d);
    seq = (unsigned)d.epoch + 1u;
    bool hold2 = threadIdx.x == 0;
    XPart r2 = exchange<KIND, PH_I>(V, sm, xsm, mine2, seq, KIND == KIND_C ? &hold2 : nullptr);
    pi.row[0] = V.full && r2.i0 >= 0 ? real_of(V, r2.i0) : -1;
    pi.row[1] = V.full && KIND == KIND_NU && r2.i1 >= 0 ? real_of(V, r2.i1) : -1;
    if (pi.row[0] >= 0) prefetch_seg<PS>(V, pi.row[0], sb, se, pi.k[0]);
    if (pi.row[1] >= 0) prefetch_seg<PS>(V, pi.row[1], sb, se, pi.k[1]);
    if (QF) {
      prefetch_query(V, pi.row[0], qi.v[0], qi.row[0]);
      prefetch_query(V, pi.row[1], qi.v[1], qi.row[1]);
    }
    if (hold2) {
      d.epoch = (int)seq;
      d.pend = 0;
      if (r2.seq) d.err = 1;
      else decide_i<KIND, FULL>(V, d, r2, true);
    }
    __syncthreads();
  }
}

// The SMO loop across co-resident owners, one iteration per trip:
//   [select j over owned variables, computing row i on a miss] exchange
//   [every workgroup: j, row-j slot, alpha step]
//   [owners store the pair's alpha/status; G update, computing row j on a
//    miss; select i] exchange
//   [every workgroup: i, Gmax/Gmax2, stop test, row-i slot]
template <int KIND, bool SPEC = false, bool PIPE = false>
__device__ __forceinline__ void smo_iterations(const View &V, Dec &d, RedSmem &sm, XSmem &xsm, float *xs,
                                               int iter_limit) {
  for (int n = 0; n < iter_limit; ++n) {
    if (d.stop || d.err) break;
    if (SPEC && KIND == KIND_C && V.spec && d.miss_i) spec_rows<KIND>(V, d, sm, xs);
    Snap c = snap_of(d);
    Partial bp = pass_select_j<KIND, PASS_SLICES, PIPE>(V, c, xs, sm, false);
    XPart mine = make_j_record<KIND>(V, c, bp);
    unsigned seq = (unsigned)d.epoch + 1u;
    // PIPE (k_smo): the thread holding the winning record decides
    bool hold = threadIdx.x == 0;
    XPart r = exchange<KIND, PH_J>(V, sm, xsm, mine, seq, PIPE ? &hold : nullptr);
    if (hold) {
      d.epoch = (int)seq;
      if (r.seq) d.err = 1;
      else decide_j<KIND>(V, d, r);
    }
    __syncthreads();
    if (d.stop || d.err) break;
    c = snap_of(d);
    Pend pd;
    pd.i = d.i; pd.j = d.j; pd.ai = d.ai_new; pd.aj = d.aj_new; pd.si = d.st_i; pd.sj = d.st_j;
    Partial bp2 = pass_gradient_select<KIND, PASS_SLICES, PIPE>(V, c, xs, sm, true, true, false, &pd);
    XPart mine2 = make_i_record<KIND>(V, bp2);
    mine2.pad = pause_vote(V, d);
    seq = (unsigned)d.epoch + 1u;
    bool hold2 = threadIdx.x == 0;
    XPart r2 = exchange<KIND, PH_I>(V, sm, xsm, mine2, seq, PIPE && KIND == KIND_C ? &hold2 : nullptr);
    if (hold2) {
      d.epoch = (int)seq;
      d.pend = 0;
      if (r2.seq) d.err = 1;
      else decide_i<KIND>(V, d, r2, true);
    }
    __syncthreads();
  }
}

// The whole SMO loop on the device, the workgroups of the grid being the
// owners.  Requires co-resident workgroups (cooperative launch).  State that
// must survive to the next launch is left in Ctl by workgroup 0.
template <int KIND>
__global__ void __launch_bounds__(BLOCK) k_smo(View V, int iter_limit) {
  __shared__ RedSmem sm;
  __shared__ XSmem xsm;
  __shared__ Dec d;
  extern __shared__ float xs_lds[];
  float *xs = V.use_lds ? xs_lds : V.xs_global + (long long)blockIdx.x * V.d;
  if (threadIdx.x == 0) d = ctl_dec(V.ctl);
  // full cache: this workgroup's row flags as a bitset in LDS for the launch
  // (the lookups after every exchange are then LDS reads, not memory round
  // trips), from and back to its flag bytes
  unsigned *fbits = reinterpret_cast<unsigned *>(smo_dyn_lds + V.flag_lds);  // as private_meta
  int8_t *fbytes = V.pm_flags + pm_index(V) * V.l;
  if (V.flag_lds > 0) {
    for (int w = threadIdx.x; w < (V.l + 31) / 32; w += BLOCK) fbits[w] = 0u;
    __syncthreads();
    for (int r = threadIdx.x; r < V.l; r += BLOCK)
      if (fbytes[r]) atomicOr(&fbits[r >> 5], 1u << (r & 31));
  }
  __syncthreads();
  // resident vectors: this workgroup's owned variables in LDS for the launch
  int sb, se;
  owned_slices(V, my_owner(V), V.nblocks, sb, se);
  const bool res = V.res_off > 0;
  Res R;
  if (res) {
    R = res_view(V, sb);
    res_load(V, R, sb, se);
  }
  if (d.phase == 0 && !d.stop && !d.err) smo_select_first<KIND, true>(V, d, sm, xsm, xs);
  // one instantiation of the resident loop per case, each carrying only the
  // code its case runs (fewer live registers in the loop): variables per
  // row, slices per trip, full cache (no LRU bookkeeping), query rows
  // fetched after the exchanges.  All give the same iterates.
  if (!res) smo_iterations<KIND, true, true>(V, d, sm, xsm, xs, iter_limit);
  else if (se - sb <= NWAVES) {  // every wave of this workgroup owns at most one slice
    if (V.svr) smo_iterations_res<KIND, 2, 1, false>(V, d, sm, xsm, xs, R, sb, se, iter_limit);
    else if (V.full && V.qpre) smo_iterations_res<KIND, 1, 1, true, true>(V, d, sm, xsm, xs, R, sb, se, iter_limit);
    else if (V.full) smo_iterations_res<KIND, 1, 1, true>(V, d, sm, xsm, xs, R, sb, se, iter_limit);
    else smo_iterations_res<KIND, 1, 1, false>(V, d, sm, xsm, xs, R, sb, se, iter_limit);
  } else if (V.svr) {
    smo_iterations_res<KIND, 2, PASS_SLICES, false>(V, d, sm, xsm, xs, R, sb, se, iter_limit);
  } else {
    smo_iterations_res<KIND, 1, PASS_SLICES, false>(V, d, sm, xsm, xs, R, sb, se, iter_limit);
  }
  if (res) {
    __syncthreads();
    res_store(V, R, sb, se);
  }
  if (V.flag_lds > 0) {
    __syncthreads();
    for (int r = threadIdx.x; r < V.l; r += BLOCK) fbytes[r] = (int8_t)((fbits[r >> 5] >> (r & 31)) & 1u);
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    Ctl *c = V.ctl;
    ctl_dec(c) = d;
    c->iter = d.iter;
    c->stop = d.stop;
    c->err = d.err ? (int)ERR_TIMEOUT : 0;
    c->hits = d.hits;
    c->misses = d.misses;
    c->i = d.i; c->j = d.j;
    c->Gmax = d.Gmax; c->Gmax2 = d.Gmax2;
  }
}

// Resets every workgroup's private LRU bookkeeping (persistent kernel).
__global__ void __launch_bounds__(BLOCK) k_lru_private_init(View V) {
  int *base = V.pm_lru + (long long)blockIdx.x * V.pm_stride;
  for (int r = threadIdx.x; r < V.l; r += BLOCK) base[r] = -1;
  for (int s = threadIdx.x; s < V.S; s += BLOCK) {
    base[V.l + s] = -1;
    base[V.l + V.S + s] = s - 1;
    base[V.l + 2 * V.S + s] = s + 1 < V.S ? s + 1 : -1;
  }
  if (threadIdx.x == 0) {
    base[V.l + 3 * V.S] = 0;
    base[V.l + 3 * V.S + 1] = V.S - 1;
  }
}

// rho, obj (and r for the nu rule) over all variables.
template <int KIND>
__global__ void __launch_bounds__(BLOCK) k_finalize(View V) {
  __shared__ RedSmem sm;
  __shared__ double red[BLOCK / 64][8];
  __shared__ int redi[BLOCK / 64][2];
  // accumulators: C: s, ub, lb | NU: s1, ub1, lb1, s2, ub2, lb2 ; obj
  double s1 = 0, ub1 = INF, lb1 = -INF, s2 = 0, ub2 = INF, lb2 = -INF, ob = 0;
  int n1 = 0, n2 = 0;
  for (int v = blockIdx.x * BLOCK + threadIdx.x; v < V.nvar; v += gridDim.x * BLOCK) {
    double g = V.G[v];
    int8_t yv = V.y[v], st = V.status[v];
    ob += V.alpha[v] * (g + V.p[v]);
    if (KIND == KIND_C) {
      double yG = yv * g;
      if (st == AT_UPPER) {
        if (yv == -1) ub1 = fmin(ub1, yG); else lb1 = fmax(lb1, yG);
      } else if (st == AT_LOWER) {
        if (yv == +1) ub1 = fmin(ub1, yG); else lb1 = fmax(lb1, yG);
      } else {
        ++n1;
        s1 += yG;
      }
    } else {
      if (yv == +1) {
        if (st == AT_UPPER) lb1 = fmax(lb1, g);
        else if (st == AT_LOWER) ub1 = fmin(ub1, g);
        else { ++n1; s1 += g; }
      } else {
        if (st == AT_UPPER) lb2 = fmax(lb2, g);
        else if (st == AT_LOWER) ub2 = fmin(ub2, g);
        else { ++n2; s2 += g; }
      }
    }
  }
  // block reduce
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64, nw = NWAVES;
  s1 = wave_sum(s1); s2 = wave_sum(s2); ob = wave_sum(ob);
  lb1 = wave_max(lb1); lb2 = wave_max(lb2);
  ub1 = -wave_max(-ub1); ub2 = -wave_max(-ub2);
  n1 = wave_sum_i(n1); n2 = wave_sum_i(n2);
  if (lane == 0) {
    red[wave][0] = s1; red[wave][1] = ub1; red[wave][2] = lb1; red[wave][3] = s2;
    red[wave][4] = ub2; red[wave][5] = lb2; red[wave][6] = ob;
    redi[wave][0] = n1; redi[wave][1] = n2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < nw; ++w) {
      s1 += red[w][0]; ub1 = fmin(ub1, red[w][1]); lb1 = fmax(lb1, red[w][2]);
      s2 += red[w][3]; ub2 = fmin(ub2, red[w][4]); lb2 = fmax(lb2, red[w][5]);
      ob += red[w][6]; n1 += redi[w][0]; n2 += redi[w][1];
    }
    Partial a, b;
    a.v0 = s1; a.v1 = ub1; a.v2 = lb1; a.v3 = ob; a.i0 = n1; a.i1 = n2; a.pad0 = a.pad1 = 0;
    b.v0 = s2; b.v1 = ub2; b.v2 = lb2; b.v3 = 0; b.i0 = b.i1 = b.pad0 = b.pad1 = 0;
    V.part[2 * blockIdx.x] = a;
    V.part[2 * blockIdx.x + 1] = b;
  }
  if (!arrive_last(V, sm)) return;
  if (threadIdx.x != 0) return;
  s1 = 0; ub1 = INF; lb1 = -INF; s2 = 0; ub2 = INF; lb2 = -INF; ob = 0; n1 = n2 = 0;
  for (int b = 0; b < (int)gridDim.x; ++b) {
    Partial a = load_partial(&V.part[2 * b]), q = load_partial(&V.part[2 * b + 1]);
    s1 += a.v0; ub1 = fmin(ub1, a.v1); lb1 = fmax(lb1, a.v2); ob += a.v3; n1 += a.i0; n2 += a.i1;
    s2 += q.v0; ub2 = fmin(ub2, q.v1); lb2 = fmax(lb2, q.v2);
  }
  Ctl *c = V.ctl;
  c->obj = ob / 2;
  if (KIND == KIND_C) {
    c->rho = n1 > 0 ? s1 / n1 : (ub1 + lb1) / 2;
    c->r = 0;
  } else {
    double r1 = n1 > 0 ? s1 / n1 : (ub1 + lb1) / 2;
    double r2 = n2 > 0 ? s2 / n2 : (ub2 + lb2) / 2;
    c->r = (r1 + r2) / 2;
    c->rho = (r1 - r2) / 2;
  }
}

__global__ void __launch_bounds__(BLOCK) k_kernel_row(View V, int q, float *out) {
  extern __shared__ float xs_lds[];
  float *xs = V.use_lds ? xs_lds : V.xs_global + (long long)blockIdx.x * V.d;
  int sb, se;
  owned_slices(V, blockIdx.x, gridDim.x, sb, se);
  stage_row(V, q, xs);
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64, nw = NWAVES;
  for (int s = sb + wave; s < se; s += nw) {
    int r = s * SLICE + lane;
    if (r < V.l) out[r] = kernel_value(V, xs, q, r);
  }
}

// ------------------------------------------------------------ batched solves
//
// Many independent subproblems (one-vs-one pairs, cross-validation folds,
// gamma grid points) in one cooperative launch.  The workgroups of the grid
// take items from a queue.  An item is a whole subproblem (rank 0 of a team
// of one: every decision local, no exchange) or one member of a team of T
// co-resident workgroups that owns a subproblem the way the persistent
// kernel's grid does.  A team's T items are consecutive and the host queues
// teams only when every item of the launch fits in the grid at once, so
// every member is resident from the start.  Each item runs up to its
// subproblem's iteration cap (V.max_iter); the host resumes unfinished
// subproblems in a later launch from the G, alpha and status left in memory,
// which continues the same sequence of working sets.

__device__ __forceinline__ void solo_init_cache(View &V) {
  if (V.keep_cache) return;
  if (V.full) {
    for (int r = threadIdx.x; r < V.l; r += BLOCK) V.pm_flags[r] = 0;
  } else {
    int *base = V.pm_lru;
    for (int r = threadIdx.x; r < V.l; r += BLOCK) base[r] = -1;
    for (int t = threadIdx.x; t < V.S; t += BLOCK) {
      base[V.l + t] = -1;
      base[V.l + V.S + t] = t - 1;
      base[V.l + 2 * V.S + t] = t + 1 < V.S ? t + 1 : -1;
    }
    if (threadIdx.x == 0) {
      base[V.l + 3 * V.S] = 0;
      base[V.l + 3 * V.S + 1] = V.S - 1;
    }
  }
}

// QD_v = K(v, v) in FP64 for every variable (gamma is per subproblem).
__device__ __forceinline__ void batch_qd(const View &V) {
  for (int v = threadIdx.x; v < V.nvar; v += BLOCK) {
    int r = real_of(V, v);
    V.QD[v] = kernel_f64(V, r, r, V.xsq[r]);
  }
}

// Query rows staged per pass of batch_ginit (LDS permitting).
constexpr int GINIT_M = 16;

// G += sum alpha_v Q_v over the variables v in [from, to) with alpha_v > 0,
// in increasing v (libsvm/svm.cpp:555-569), for rows [r0, r1) (r0 a
// multiple of the slice height).  Up to GINIT_M such rows are staged in LDS
// at a time and each lane's row is read once for all of them; every kernel
// value is the one kernel_value computes.  Ends with a barrier.
template <int M>
__device__ __forceinline__ void ginit_pass(const View &V, float *xs, int ld, const int *list, int cnt, int r0, int r1) {
  const int nv = V.svr ? 2 : 1;
  for (int r = r0 + threadIdx.x; r < r1; r += BLOCK) {
    // row r of a slice is lane r % 64 of wave (r % 512) / 64 in the SELL layout
    double acc[M];
    if (V.gdot) {
#pragma unroll
      for (int m = 0; m < M; ++m) acc[m] = m < cnt ? shared_dot(V, real_of(V, list[m]), r) : 0.0;
    } else {
      row_dot_multi<M>(V, xs, ld, r, acc);
    }
#pragma unroll
    for (int m = 0; m < M; ++m) {  // static indices keep acc in registers
      if (m >= cnt) continue;
      const int v = list[m];
      const int q = real_of(V, v);
      const float K = (float)kernel_f64(V, q, r, acc[m]);
      const double av = V.alpha[v];
      const int8_t yv = V.y[v];
      for (int h = 0; h < nv; ++h) {
        const int k = r + h * V.l;
        const float Qv = (V.y[k] == yv) ? K : -K;
        V.G[k] = V.G[k] + av * (double)Qv;
      }
    }
  }
}

// (8 staged rows in the batched kernel, whose other paths leave fewer
// registers: 16 there spills 296 bytes per lane across the whole kernel)
constexpr int BATCH_GINIT_M = 8;

__device__ __forceinline__ void batch_ginit(const View &V, float *xs, int lds_floats, int r0, int r1, int from,
                                            int to) {
  __shared__ int list[BATCH_GINIT_M];
  __shared__ int s_cnt, s_next;
  const int fit = lds_floats / max(V.d, 1);
  const int M = fit >= BATCH_GINIT_M ? BATCH_GINIT_M : (fit >= 4 ? 4 : 1);
  int v = from;
  for (;;) {
    if (threadIdx.x == 0) {
      int c = 0, u = v;
      for (; u < to && c < M; ++u)
        if (V.status[u] != AT_LOWER) list[c++] = u;
      s_cnt = c;
      s_next = u;
    }
    __syncthreads();
    const int cnt = s_cnt;
    v = s_next;
    if (cnt == 0) break;
    if (!V.gdot)
      for (int m = 0; m < cnt; ++m) stage_row(V, real_of(V, list[m]), xs + m * V.d);
    if (M == BATCH_GINIT_M) ginit_pass<BATCH_GINIT_M>(V, xs, V.d, list, cnt, r0, r1);
    else if (M >= 4) ginit_pass<4>(V, xs, V.d, list, cnt, r0, r1);
    else ginit_pass<1>(V, xs, V.d, list, cnt, r0, r1);
    __syncthreads();
  }
  __syncthreads();
}

// G_k += alpha_v Q_vk for every v in list[0..count) (the variables with
// alpha_v != 0 in increasing v, libsvm/svm.cpp:555-569) over the rows of this
// workgroup's block of BLOCK rows, one launch for the whole list: the
// single-solve form of batch_ginit.  Up to GINIT_M query rows (as the
// dynamic LDS holds, lds_floats) are staged per pass, each lane's row is
// read once per pass for all of them, and every G_k still accumulates in
// increasing v with the kernel values kernel_value computes, so G is
// bit-identical to one k_ginit_row launch per v.
__global__ void __launch_bounds__(BLOCK) k_ginit_multi(View V, const int *list, int count, int lds_floats) {
  extern __shared__ float xs_lds[];
  __shared__ int sl[GINIT_M];
  const int r0 = blockIdx.x * BLOCK, r1 = min(V.l, r0 + BLOCK);
  const int fit = lds_floats / max(V.d, 1);
  const int M = fit >= GINIT_M ? GINIT_M : (fit >= 4 ? 4 : 1);
  for (int c0 = 0; c0 < count; c0 += M) {
    const int cnt = min(M, count - c0);
    if (threadIdx.x < cnt) sl[threadIdx.x] = list[c0 + threadIdx.x];
    __syncthreads();
    for (int m = 0; m < cnt; ++m) stage_row(V, real_of(V, sl[m]), xs_lds + m * V.d);
    if (M == GINIT_M) ginit_pass<GINIT_M>(V, xs_lds, V.d, sl, cnt, r0, r1);
    else if (M == 4) ginit_pass<4>(V, xs_lds, V.d, sl, cnt, r0, r1);
    else ginit_pass<1>(V, xs_lds, V.d, sl, cnt, r0, r1);
    __syncthreads();  // sl and the staged rows are rewritten by the next pass
  }
}

// Terms of rho / obj / r over the variables of rows [r0, r1), reduced over
// the workgroup into thread 0's f[0..6] (s1, ub1, lb1, s2, ub2, lb2, obj) and
// n[0..1] (Solver::calculate_rho, libsvm/svm.cpp:1013-1049; Solver_NU::calculate_rho, :1260-1308;
// the objective as in Solver::Solve, :794-803).
template <int KIND>
__device__ __forceinline__ void final_terms(const View &V, int r0, int r1, double f[7], int n[2]) {
  __shared__ double red[BLOCK / 64][7];
  __shared__ int redi[BLOCK / 64][2];
  double s1 = 0, ub1 = INF, lb1 = -INF, s2 = 0, ub2 = INF, lb2 = -INF, ob = 0;
  int n1 = 0, n2 = 0;
  const int nv = V.svr ? 2 : 1;
  for (int rr = r0 + threadIdx.x; rr < r1; rr += BLOCK) {
    for (int h = 0; h < nv; ++h) {
      const int v = rr + h * V.l;
      double g = V.G[v];
      int8_t yv = V.y[v], st = V.status[v];
      ob += V.alpha[v] * (g + V.p[v]);
      if (KIND == KIND_C) {
        double yG = yv * g;
        if (st == AT_UPPER) {
          if (yv == -1) ub1 = fmin(ub1, yG); else lb1 = fmax(lb1, yG);
        } else if (st == AT_LOWER) {
          if (yv == +1) ub1 = fmin(ub1, yG); else lb1 = fmax(lb1, yG);
        } else {
          ++n1;
          s1 += yG;
        }
      } else {
        if (yv == +1) {
          if (st == AT_UPPER) lb1 = fmax(lb1, g);
          else if (st == AT_LOWER) ub1 = fmin(ub1, g);
          else { ++n1; s1 += g; }
        } else {
          if (st == AT_UPPER) lb2 = fmax(lb2, g);
          else if (st == AT_LOWER) ub2 = fmin(ub2, g);
          else { ++n2; s2 += g; }
        }
      }
    }
  }
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64, nw = NWAVES;
  s1 = wave_sum(s1); s2 = wave_sum(s2); ob = wave_sum(ob);
  lb1 = wave_max(lb1); lb2 = wave_max(lb2);
  ub1 = -wave_max(-ub1); ub2 = -wave_max(-ub2);
  n1 = wave_sum_i(n1); n2 = wave_sum_i(n2);
  if (lane == 0) {
    red[wave][0] = s1; red[wave][1] = ub1; red[wave][2] = lb1; red[wave][3] = s2;
    red[wave][4] = ub2; red[wave][5] = lb2; red[wave][6] = ob;
    redi[wave][0] = n1; redi[wave][1] = n2;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < nw; ++w) {
      s1 += red[w][0]; ub1 = fmin(ub1, red[w][1]); lb1 = fmax(lb1, red[w][2]);
      s2 += red[w][3]; ub2 = fmin(ub2, red[w][4]); lb2 = fmax(lb2, red[w][5]);
      ob += red[w][6]; n1 += redi[w][0]; n2 += redi[w][1];
    }
    f[0] = s1; f[1] = ub1; f[2] = lb1; f[3] = s2; f[4] = ub2; f[5] = lb2; f[6] = ob;
    n[0] = n1; n[1] = n2;
  }
  __syncthreads();
}

// rho, obj and r from the combined terms into Ctl (thread 0).
template <int KIND>
__device__ __forceinline__ void write_final(Ctl *c, const double f[7], const int n[2]) {
  c->obj = f[6] / 2;
  if (KIND == KIND_C) {
    c->rho = n[0] > 0 ? f[0] / n[0] : (f[1] + f[2]) / 2;
    c->r = 0;
  } else {
    double r1 = n[0] > 0 ? f[0] / n[0] : (f[1] + f[2]) / 2;
    double r2 = n[1] > 0 ? f[3] / n[1] : (f[4] + f[5]) / 2;
    c->r = (r1 + r2) / 2;
    c->rho = (r1 - r2) / 2;
  }
}

// rho, obj and r in FP64 of the finished subproblems views[idx[b]], one
// workgroup each, into each view's Ctl.  The terms are reduced by one
// workgroup of fixed shape over all of a subproblem's variables, so the
// result does not depend on how many workgroups solved it (the serving
// runtime's team sizes change from stage to stage).
__global__ void __launch_bounds__(BLOCK) k_finalize_views(const View *views, const int *idx, int n) {
  if ((int)blockIdx.x >= n) return;
  const View V = views[idx[blockIdx.x]];
  double f[7];
  int nn[2];
  if (V.kind == KIND_C) {
    final_terms<KIND_C>(V, 0, V.l, f, nn);
    if (threadIdx.x == 0) write_final<KIND_C>(V.ctl, f, nn);
  } else {
    final_terms<KIND_NU>(V, 0, V.l, f, nn);
    if (threadIdx.x == 0) write_final<KIND_NU>(V.ctl, f, nn);
  }
}

__device__ __forceinline__ void write_ctl_state(Ctl *c, const Dec &d, int err) {
  c->iter = d.iter;
  c->stop = d.stop;
  c->err = err;
  c->hits = d.hits;
  c->misses = d.misses;
}

// One subproblem owned by this workgroup alone.
// Leaves Ctl as a pause before the first iteration (G initialisation not
// finished: it continues from V.ginit_to in the next launch).
__device__ __forceinline__ void write_ginit_pause(const View &V) {
  Ctl *c = V.ctl;
  c->iter = V.iter0;
  c->stop = STOP_PAUSE;
  c->err = 0;
  c->hits = 0;
  c->misses = 0;
  c->ginit_next = V.ginit_to;
}

// Speculative kernel rows of a single-workgroup solve (C rule, full cache).
// A batched subproblem computes most of its kernel rows before it stops,
// each in its own pass over the subproblem's X by one workgroup; so on a row
// miss the rows of the strongest candidates of the pass just made (the
// per-wave winners block_reduce leaves in sm buffer 0: I_up and I_low for
// the i pass, the best gains for the j pass) that are not computed yet go
// through the same pass (row_dot_multi), up to SPEC_ROWS rows and the LDS
// room for their staged query rows.  Every value is the one a miss would
// compute, so the iterates do not change.
__device__ __forceinline__ void solo_spec_rows(const View &V, Dec &d, RedSmem &sm, float *xs, int lds_floats,
                                               bool jside) {
  __shared__ int list[SPEC_ROWS];
  __shared__ int s_cnt;
  if (threadIdx.x == 0) {
    const int cap = min(SPEC_ROWS, lds_floats / max(V.d, 1));
    const int nw = NWAVES;
    CacheMeta M = private_meta(V, d);
    int cnt = 0;
    list[cnt++] = real_of(V, jside ? d.j : d.i);  // the decision has marked it
    while (cnt < cap) {
      double bv = -INF;
      int bw = -1, bk = 0;
      for (int w = 0; w < nw; ++w) {
        if (sm.i[0][w][0] >= 0) {
          const double v = jside ? -sm.v[0][w][0] : sm.v[0][w][0];
          if (v > bv) { bv = v; bw = w; bk = 0; }
        }
        if (!jside && sm.i[0][w][1] >= 0 && sm.v[0][w][1] > bv) { bv = sm.v[0][w][1]; bw = w; bk = 1; }
      }
      if (bw < 0) break;
      const int row = real_of(V, sm.i[0][bw][bk]);
      sm.i[0][bw][bk] = -1;  // taken (the buffer is rewritten by the next pass)
      if (flag_test_set(M, row)) continue;
      ++d.misses;
      list[cnt++] = row;
    }
    s_cnt = cnt;
  }
  __syncthreads();
  if (s_cnt > 1) compute_rows_part(V, list, s_cnt, xs, 0, V.nslices);
  else compute_row_part(V, list[0], list[0], xs, 0, V.nslices);
  if (threadIdx.x == 0) {
    if (jside) d.miss_j = 0;
    else d.miss_i = 0;
  }
  __syncthreads();
}

template <int KIND>
__device__ __forceinline__ void solo_solve(View &V, Dec &d, RedSmem &sm, float *xs, int lds_floats, long long t0) {
  solo_init_cache(V);
  // full cache in LDS: the row flags as a bitset for this item (the lookup
  // after every selection is then no memory round trip), from and back to
  // the arena's flag bytes
  unsigned *fbits = reinterpret_cast<unsigned *>(smo_dyn_lds + V.flag_lds);  // as private_meta
  if (V.flag_lds > 0) {
    for (int w = threadIdx.x; w < (V.l + 31) / 32; w += BLOCK) fbits[w] = 0u;
    __syncthreads();
    if (V.keep_cache)
      for (int r = threadIdx.x; r < V.l; r += BLOCK)
        if (V.pm_flags[r]) atomicOr(&fbits[r >> 5], 1u << (r & 31));
    __syncthreads();
  }
  if (!V.resume) {
    if (V.ginit_from == 0) batch_qd(V);
    batch_ginit(V, xs, lds_floats, 0, V.l, V.ginit_from, V.ginit_to);
    if (V.ginit_to < V.nvar) {
      if (threadIdx.x == 0) write_ginit_pause(V);
      __syncthreads();
      return;
    }
  }
  if (threadIdx.x == 0) {
    memset(&d, 0, sizeof d);
    d.iter = V.iter0;
    d.t_start = t0 > 0 ? t0 : (long long)wall_clock64();
  }
  __syncthreads();
  {
    Snap c = snap_of(d);
    Partial bp = pass_gradient_select<KIND, PASS_SLICES_SOLO>(V, c, xs, sm, false, true, false, nullptr);
    XPart m = make_i_record<KIND>(V, bp);
    if (threadIdx.x == 0) { d.phase = 1; decide_i<KIND>(V, d, m, true); }
    __syncthreads();
  }
  while (!d.stop) {
    if (KIND == KIND_C && V.spec && d.miss_i) solo_spec_rows(V, d, sm, xs, lds_floats, false);
    Snap c = snap_of(d);
    Partial bp = pass_select_j<KIND, PASS_SLICES_SOLO>(V, c, xs, sm, false);
    XPart m = make_j_record<KIND>(V, c, bp);
    if (threadIdx.x == 0) decide_j<KIND>(V, d, m);
    __syncthreads();
    if (d.stop) break;
    if (KIND == KIND_C && V.spec && d.miss_j) solo_spec_rows(V, d, sm, xs, lds_floats, true);
    c = snap_of(d);
    Pend pd;
    pd.i = d.i; pd.j = d.j; pd.ai = d.ai_new; pd.aj = d.aj_new; pd.si = d.st_i; pd.sj = d.st_j;
    Partial bp2 = pass_gradient_select<KIND, PASS_SLICES_SOLO>(V, c, xs, sm, true, true, false, &pd);
    XPart m2 = make_i_record<KIND>(V, bp2);
    if (threadIdx.x == 0) {
      m2.pad = pause_vote(V, d);
      decide_i<KIND>(V, d, m2, true);
    }
    __syncthreads();
  }
  if (V.flag_lds > 0)
    for (int r = threadIdx.x; r < V.l; r += BLOCK) V.pm_flags[r] = (int8_t)((fbits[r >> 5] >> (r & 31)) & 1u);
  double f[7];
  int n[2];
  final_terms<KIND>(V, 0, V.l, f, n);
  if (threadIdx.x == 0) {
    write_final<KIND>(V.ctl, f, n);
    write_ctl_state(V.ctl, d, 0);
  }
}

// One member (rank V.rank of V.nblocks) of a team owning a subproblem.
template <int KIND>
__device__ __forceinline__ void team_solve(View &V, Dec &d, RedSmem &sm, XSmem &xsm, float *xs, int lds_floats,
                                           long long t0) {
  int sb, se;
  owned_slices(V, V.rank, V.nblocks, sb, se);
  const int r0 = sb * SLICE, r1 = min(se * SLICE, V.l);
  solo_init_cache(V);
  if (!V.resume) {
    // every member writes every QD (identical values): row i's QD is read by all
    if (V.ginit_from == 0) batch_qd(V);
    batch_ginit(V, xs, lds_floats, r0, r1, V.ginit_from, V.ginit_to);
    if (V.ginit_to < V.nvar) {  // every member stops at the same variable
      if (V.rank == 0 && threadIdx.x == 0) write_ginit_pause(V);
      __syncthreads();
      return;
    }
  }
  if (threadIdx.x == 0) {
    memset(&d, 0, sizeof d);
    d.iter = V.iter0;
    d.t_start = t0 > 0 ? t0 : (long long)wall_clock64();
  }
  __syncthreads();
  smo_select_first<KIND>(V, d, sm, xsm, xs);
  smo_iterations<KIND>(V, d, sm, xsm, xs, INT_MAX);
  // combine the members' terms: rank 0 gathers the records in rank order
  double f[7];
  int n[2];
  final_terms<KIND>(V, r0, r1, f, n);
  if (threadIdx.x == 0) {
    const unsigned seq = (unsigned)d.epoch + 1u;
    XPart m = xpart_identity<KIND, PH_I>();
    m.v0 = f[0]; m.v1 = f[1]; m.v2 = f[2]; m.v3 = f[3]; m.g0 = f[4]; m.a0 = f[5]; m.g1 = f[6];
    m.i0 = n[0]; m.i1 = n[1]; m.a1 = d.err ? 1.0 : 0.0;
    xpublish<XF_FIN>(V.xw, V.nblocks, V.rank, m, seq);
    if (V.rank == 0) {
      double F[7] = {0, INF, -INF, 0, INF, -INF, 0};
      int N[2] = {0, 0};
      int err = d.err;
      for (int t = 0; t < V.nblocks && !err; ++t) {
        XPart q;
        if (!xgather<XF_FIN>(V.xw, V.nblocks, t, seq, q)) {
          err = 1;
          break;
        }
        F[0] += q.v0; F[1] = fmin(F[1], q.v1); F[2] = fmax(F[2], q.v2);
        F[3] += q.v3; F[4] = fmin(F[4], q.g0); F[5] = fmax(F[5], q.a0);
        F[6] += q.g1;
        N[0] += q.i0; N[1] += q.i1;
        if (q.a1 != 0.0) err = 1;
      }
      write_final<KIND>(V.ctl, F, N);
      write_ctl_state(V.ctl, d, err ? (int)ERR_TIMEOUT : 0);
    }
  }
  __syncthreads();
}

__global__ void __launch_bounds__(BLOCK) k_smo_batch(const View *views, const int *items, int nitems,
                                                    BatchPool pool) {
  __shared__ RedSmem sm;
  __shared__ XSmem xsm;
  __shared__ Dec d;
  __shared__ int s_item;
  __shared__ long long s_t0;
  extern __shared__ float xs_lds[];
  float *xs = xs_lds;
  // the launch's start: the first workgroup to arrive records it
  if (threadIdx.x == 0) {
    long long t = 0;
    if (pool.t0) {
      const unsigned long long now = (unsigned long long)wall_clock64();
      const unsigned long long old = atomicCAS(pool.t0, 0ull, now);
      t = (long long)(old ? old : now);
    }
    s_t0 = t;
  }
  __syncthreads();
  const long long t0 = s_t0;
  __shared__ long long s_now;
  __shared__ int s_arena;
  int my_arena = -1;  // the arena this workgroup holds (uniform)
  for (;;) {
    if (threadIdx.x == 0) {
      s_item = atomicAdd(pool.queue, 1);
      s_now = (long long)wall_clock64();
    }
    __syncthreads();
    // uniform across the workgroup: keeps the View in scalar registers
    const int q = __builtin_amdgcn_readfirstlane(s_item);
    if (q >= nitems) {
      if (threadIdx.x == 0 && pool.idle) atomicAdd(pool.idle, 1u);
      break;
    }
    const int sub = __builtin_amdgcn_readfirstlane(items[2 * q]);
    View V = views[sub];
    V.drain = pool.idle;
    V.drain_at = pool.drain_at;
    // past the launch's deadline, single-workgroup items are left untouched
    // for the next launch (team members are always run: they are queued
    // first and all claimed at the start)
    if (pool.deadline_ticks > 0 && V.nblocks <= 1 && s_now - t0 >= pool.deadline_ticks) {  // uniform: thread 0's clock
      __syncthreads();
      continue;
    }
    V.rank = __builtin_amdgcn_readfirstlane(items[2 * q + 1]);
    int team;
    if (V.own_cache) {
      // the subproblem's own arena; member r keeps metadata copy r
      team = V.nblocks;
      V.pm_flags += (long long)V.rank * V.l;
      V.pm_lru += (long long)V.rank * V.pm_stride;
    } else {
      if (q < pool.team_items) {
        my_arena = q;
      } else if (my_arena < 0) {
        if (threadIdx.x == 0) s_arena = pool.team_items + atomicAdd(pool.arena_next, 1);
        __syncthreads();
        my_arena = __builtin_amdgcn_readfirstlane(s_arena);
      }
      const long long own = my_arena;
      team = q < pool.team_items ? V.nblocks : 1;
      const long long first = own - (team > 1 ? V.rank : 0);
      V.rows = pool.rows + first * pool.arena_floats;
      V.row_stride = ((long long)V.l + 63) / 64 * 64;
      const long long fit = team * pool.arena_floats / (V.row_stride > 0 ? V.row_stride : 1);
      if (fit >= V.l) { V.full = 1; V.S = V.l > 0 ? V.l : 1; }
      else { V.full = 0; V.S = (int)(fit < 2 ? 2 : fit); }
      V.pm_flags = pool.flags + own * pool.flag_stride;
      V.pm_lru = pool.lru + own * pool.lru_stride;
      V.keep_cache = 0;
    }
    V.pm_stride = 0;
    if (team > 1) {
      if (V.kind == KIND_C) team_solve<KIND_C>(V, d, sm, xsm, xs, pool.lds_floats, t0);
      else team_solve<KIND_NU>(V, d, sm, xsm, xs, pool.lds_floats, t0);
    } else {
      V.nblocks = 1;
      V.rank = 0;
      // (no speculative rows over the shared dot products: a row is a gather)
      V.spec = pool.spec && V.kind == KIND_C && V.full && V.ktype != K_PRECOMPUTED && !V.gdot &&
               pool.lds_floats >= 2 * V.d ? 1 : 0;
      V.flag_lds = V.full && pool.flag_words > 0 && (V.l + 31) / 32 <= pool.flag_words ? pool.lds_floats : 0;
      if (V.kind == KIND_C) solo_solve<KIND_C>(V, d, sm, xs, pool.lds_floats, t0);
      else solo_solve<KIND_NU>(V, d, sm, xs, pool.lds_floats, t0);
    }
    __syncthreads();
  }
}

// Copies the rows of one row set out of the full problem's SELL arrays into
// the row set's own SELL arrays (slice pointers precomputed on the host).
__global__ void __launch_bounds__(BLOCK) k_pack_rowset(View src, const int *rows, int n, const long long *dst_slice_ptr,
                                                      float *dst_vals, int *dst_cols, int *dst_len, double *dst_xsq) {
  int r = blockIdx.x * BLOCK + threadIdx.x;
  if (r >= n) return;
  const int g = rows[r];
  const int len = src.row_len[g];
  const long long sbase = src.slice_ptr[g / SLICE] + (g % SLICE);
  const long long dbase = dst_slice_ptr[r / SLICE] + (r % SLICE);
  for (int p = 0; p < len; ++p) {
    dst_vals[dbase + (long long)p * SLICE] = src.vals[sbase + (long long)p * SLICE];
    if (!src.dense) dst_cols[dbase + (long long)p * SLICE] = src.cols[sbase + (long long)p * SLICE];
  }
  dst_len[r] = len;
  dst_xsq[r] = src.xsq[g];
}

// ------------------------------------------------------------ shared dots
//
// D[q][r] = dot(x_q, x_r) for every pair of rows of a dense problem, each
// summed exactly as row_dot sums it (features in order, fused
// multiply-adds of exact float-by-float products), so a kernel value from D
// is bit-identical to one computed from X.  One 64 x 64 tile of the upper
// triangle per workgroup (one SELL slice per side: a tile's rows are one
// slice, whose feature-p values are 64 consecutive floats), written to both
// halves of D.  Each thread sums a 4 x 4 block; the two slices' features
// are staged in LDS as doubles, 32 at a time.
constexpr int DT = 64, DKC = 32;

__device__ __forceinline__ void dots_tile(long long t, int nt, int &tq, int &tr) {
  const double a = 2.0 * nt + 1.0;
  int q = (int)((a - sqrt(a * a - 8.0 * (double)t)) / 2.0);
  q = max(0, min(q, nt - 1));
  while (q > 0 && (long long)q * (2LL * nt - q + 1) / 2 > t) --q;
  while (q + 1 < nt && (long long)(q + 1) * (2LL * nt - q) / 2 <= t) ++q;
  tq = q;
  tr = q + (int)(t - (long long)q * (2LL * nt - q + 1) / 2);
}

__global__ void __launch_bounds__(256) k_dots_f64(View V, double *D, long long ld, long long ntiles, int nt) {
  __shared__ double A[DKC][DT + 1], B[DKC][DT + 1];
  const long long tile = blockIdx.x;
  if (tile >= ntiles) return;
  int tq, tr;
  dots_tile(tile, nt, tq, tr);
  const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
  const float *va = V.vals + V.slice_ptr[tq], *vb = V.vals + V.slice_ptr[tr];
  double acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = 0.0;
  for (int p0 = 0; p0 < V.d; p0 += DKC) {
    const int pe = min(DKC, V.d - p0);
    for (int e = threadIdx.x; e < DT * DKC; e += 256) {
      const int row = e % DT, p = e / DT;
      A[p][row] = p < pe ? (double)va[(long long)(p0 + p) * SLICE + row] : 0.0;
      B[p][row] = p < pe ? (double)vb[(long long)(p0 + p) * SLICE + row] : 0.0;
    }
    __syncthreads();
    for (int p = 0; p < pe; ++p) {  // features in order
      double a[4], b[4];
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        a[i] = A[p][ty * 4 + i];
        b[i] = B[p][tx * 4 + i];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = fma(a[i], b[j], acc[i][j]);
    }
    __syncthreads();
  }
  const int q0 = tq * DT + ty * 4, r0 = tr * DT + tx * 4;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int q = q0 + i, r = r0 + j;
      if (q < V.l && r < V.l) {
        D[(long long)q * ld + r] = acc[i][j];
        if (tq != tr) D[(long long)r * ld + q] = acc[i][j];
      }
    }
}

// ------------------------------------------------------------ prediction
//
// Decision values of many rows against one model (svm_predict_values,
// libsvm/svm.cpp:2569-2643).  Workgroup b takes test rows [b*TB, b*TB+TB):
// it stages them densely in LDS, every lane walks one support vector's row
// (SELL, coalesced) against all TB staged rows with FP64 accumulation, the
// kernel values go to the workgroup's scratch, and each one-vs-one decision
// value is then summed by one wave in a fixed order (deterministic).



}  // namespace smo

// ------------------------------------------------------------ launchers
// Called from the host runtime (smo_runtime.cpp); every launcher returns
// the launch status and never synchronises.
extern "C" {

using smo::View;
static inline unsigned nb_for(long long n, int bs) { return (unsigned)((n + bs - 1) / bs); }

hipError_t smo_launch_self_dot(const View *V, double *xsq, hipStream_t st) {
  hipLaunchKernelGGL(smo::k_self_dot, dim3(nb_for(V->l, smo::BLOCK)), dim3(smo::BLOCK), 0, st, *V, xsq);
  return hipGetLastError();
}

hipError_t smo_launch_qd(const View *V, hipStream_t st) {
  hipLaunchKernelGGL(smo::k_qd, dim3(nb_for(V->nvar, smo::BLOCK)), dim3(smo::BLOCK), 0, st, *V);
  return hipGetLastError();
}

hipError_t smo_launch_ginit_row(const View *V, int v, size_t shmem, hipStream_t st) {
  hipLaunchKernelGGL(smo::k_ginit_row, dim3(V->nblocks), dim3(smo::BLOCK), shmem, st, *V, v);
  return hipGetLastError();
}

// shmem = 0: as many staged rows (16, 4 or 1) as 64 KiB of LDS holds.
hipError_t smo_launch_ginit_multi(const View *V, const int *list, int count, size_t shmem, hipStream_t st) {
  if (count <= 0 || V->l <= 0) return hipSuccess;
  const size_t row = (size_t)std::max(V->d, 1) * sizeof(float);
  if (shmem == 0) {
    const int M = 16 * row <= 65536 ? 16 : (4 * row <= 65536 ? 4 : 1);
    shmem = (size_t)M * row;
  }
  if (shmem < row || shmem > 65536) return hipErrorInvalidValue;
  hipLaunchKernelGGL(smo::k_ginit_multi, dim3(nb_for(V->l, smo::BLOCK)), dim3(smo::BLOCK), shmem, st, *V, list,
                     count, (int)(shmem / sizeof(float)));
  return hipGetLastError();
}

hipError_t smo_launch_select_i(const View *V, size_t shmem, hipStream_t st) {
  if (V->kind == smo::KIND_C)
    hipLaunchKernelGGL(smo::k_select_i<smo::KIND_C>, dim3(V->nblocks), dim3(smo::BLOCK), shmem, st, *V);
  else
    hipLaunchKernelGGL(smo::k_select_i<smo::KIND_NU>, dim3(V->nblocks), dim3(smo::BLOCK), shmem, st, *V);
  return hipGetLastError();
}

hipError_t smo_launch_select_j(const View *V, size_t shmem, hipStream_t st) {
  if (V->kind == smo::KIND_C)
    hipLaunchKernelGGL(smo::k_select_j<smo::KIND_C>, dim3(V->nblocks), dim3(smo::BLOCK), shmem, st, *V);
  else
    hipLaunchKernelGGL(smo::k_select_j<smo::KIND_NU>, dim3(V->nblocks), dim3(smo::BLOCK), shmem, st, *V);
  return hipGetLastError();
}

hipError_t smo_launch_alpha(const View *V, hipStream_t st) {
  hipLaunchKernelGGL(smo::k_alpha, dim3(1), dim3(64), 0, st, *V);
  return hipGetLastError();
}

hipError_t smo_launch_gupdate(const View *V, size_t shmem, hipStream_t st) {
  if (V->kind == smo::KIND_C)
    hipLaunchKernelGGL(smo::k_gupdate<smo::KIND_C>, dim3(V->nblocks), dim3(smo::BLOCK), shmem, st, *V);
  else
    hipLaunchKernelGGL(smo::k_gupdate<smo::KIND_NU>, dim3(V->nblocks), dim3(smo::BLOCK), shmem, st, *V);
  return hipGetLastError();
}

hipError_t smo_launch_status(const View *V, hipStream_t st) {
  hipLaunchKernelGGL(smo::k_status, dim3(1), dim3(64), 0, st, *V);
  return hipGetLastError();
}

hipError_t smo_launch_lru_private_init(const View *V, hipStream_t st) {
  hipLaunchKernelGGL(smo::k_lru_private_init, dim3(V->nblocks), dim3(smo::BLOCK), 0, st, *V);
  return hipGetLastError();
}

// Launches a grid whose workgroups must all be resident together (they
// exchange records with bounded spins; no grid-wide hardware barrier is
// used).  The grid never exceeds the co-resident capacity.  A cooperative
// launch guarantees residency; the runtime refuses it (hipErrorIllegalState)
// while another process holds the device's cooperative-launch resources, and
// then the same grid is launched plainly: on a device not running other
// work every workgroup is resident at once, and if one is not, the spins of
// the others end with an error (ERR_TIMEOUT) rather than a hang.  With the
// environment variable SVMHIP_LAUNCH=plain every such grid is launched
// plainly (same grid, same code): profilers that mishandle cooperative
// launches can then trace the solver.
static bool plain_launches() {
  static const bool plain = [] {
    const char *e = getenv("SVMHIP_LAUNCH");
    return e && strcmp(e, "plain") == 0;
  }();
  return plain;
}

static hipError_t launch_resident(const void *fn, int grid, void **args, size_t shmem, hipStream_t st) {
  if (plain_launches()) return hipLaunchKernel(fn, dim3(grid), dim3(smo::BLOCK), args, shmem, st);
  hipError_t e = hipLaunchCooperativeKernel(fn, dim3(grid), dim3(smo::BLOCK), args, (unsigned)shmem, st);
  if (e == hipErrorIllegalState) {
    (void)hipGetLastError();
    e = hipLaunchKernel(fn, dim3(grid), dim3(smo::BLOCK), args, shmem, st);
  }
  return e;
}

hipError_t smo_launch_persistent(const View *V, int iter_limit, size_t shmem, hipStream_t st) {
  View v = *V;
  void *args[] = {&v, &iter_limit};
  const void *fn = V->kind == smo::KIND_C ? (const void *)smo::k_smo<smo::KIND_C>
                                          : (const void *)smo::k_smo<smo::KIND_NU>;
  return launch_resident(fn, V->nblocks, args, shmem, st);
}

// Largest co-resident grid for the persistent kernel at this LDS size.
hipError_t smo_persistent_capacity(int kind, size_t shmem, int *blocks_per_cu) {
  const void *fn = kind == smo::KIND_C ? (const void *)smo::k_smo<smo::KIND_C>
                                       : (const void *)smo::k_smo<smo::KIND_NU>;
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, fn, smo::BLOCK, shmem);
}

hipError_t smo_launch_finalize_views(const View *views, const int *idx, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(smo::k_finalize_views, dim3(n), dim3(smo::BLOCK), 0, st, views, idx, n);
  return hipGetLastError();
}

hipError_t smo_launch_finalize(const View *V, hipStream_t st) {
  if (V->kind == smo::KIND_C)
    hipLaunchKernelGGL(smo::k_finalize<smo::KIND_C>, dim3(V->nblocks), dim3(smo::BLOCK), 0, st, *V);
  else
    hipLaunchKernelGGL(smo::k_finalize<smo::KIND_NU>, dim3(V->nblocks), dim3(smo::BLOCK), 0, st, *V);
  return hipGetLastError();
}

hipError_t smo_launch_batch(const View *views, const int *items, int nitems, int grid, const smo::BatchPool *pool,
                            size_t shmem, hipStream_t st) {
  if (nitems <= 0 || grid <= 0) return hipSuccess;
  smo::BatchPool P = *pool;
  void *args[] = {(void *)&views, (void *)&items, (void *)&nitems, (void *)&P};
  return launch_resident((const void *)smo::k_smo_batch, grid, args, shmem, st);
}

hipError_t smo_batch_capacity(size_t shmem, int *blocks_per_cu) {
  return hipOccupancyMaxActiveBlocksPerMultiprocessor(blocks_per_cu, (const void *)smo::k_smo_batch,
                                                      smo::BLOCK, shmem);
}

// The exact dot products of every pair of the problem's rows (dense rows
// only) into D (l rows of ld doubles).
hipError_t smo_launch_dots_f64(const smo::View *V, double *D, long long ld, hipStream_t st) {
  if (!V->dense || V->l <= 0) return hipErrorInvalidValue;
  const int nt = (V->l + smo::DT - 1) / smo::DT;
  const long long ntiles = (long long)nt * (nt + 1) / 2;
  if (ntiles > 0x7fffffffLL) return hipErrorInvalidValue;
  hipLaunchKernelGGL(smo::k_dots_f64, dim3((unsigned)ntiles), dim3(256), 0, st, *V, D, ld, ntiles, nt);
  return hipGetLastError();
}

hipError_t smo_launch_pack_rowset(const View *src, const int *rows, int n, const long long *dst_slice_ptr,
                                  float *dst_vals, int *dst_cols, int *dst_len, double *dst_xsq, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(smo::k_pack_rowset, dim3((n + smo::BLOCK - 1) / smo::BLOCK), dim3(smo::BLOCK), 0, st, *src, rows, n,
                     dst_slice_ptr, dst_vals, dst_cols, dst_len, dst_xsq);
  return hipGetLastError();
}

hipError_t smo_launch_kernel_row(const View *V, int q, float *out, size_t shmem, hipStream_t st) {
  hipLaunchKernelGGL(smo::k_kernel_row, dim3(V->nblocks), dim3(smo::BLOCK), shmem, st, *V, q, out);
  return hipGetLastError();
}

}  // extern "C"
