// Dense kernel matrix as blocked Gram tiles on the matrix cores (gfx950
// MFMA), the solver's split-precision mode for dense problems whose whole
// kernel matrix fits in the row cache.
//
// The reference computes every kernel row on a miss, one row per pass over X
// (libsvm/svm_device.cu:295-474, cuda_evalQ).  A single row is a matrix-
// vector product, bound by reading X; here, when the full matrix fits in
// HBM, all of it is computed up front as l x l tiles of X X^T on MFMA, so
// the SMO loop then only hits the cache.
//
//   k_split_f16  X (SELL-64, float) -> row-major FP16 pairs: each float x is
//                held as hi + lo with hi = fp16(x) and lo = fp16(x - hi),
//                22 significant bits of x's 24.  Per row, groups of 8
//                features as [hi x 8][lo x 8] (32 bytes), zero padded to a
//                multiple of 16 features.
//   k_gram_f16   one 128 x 128 tile of K per workgroup (4 waves of 64 x 64,
//                each 2 x 2 MFMA blocks of 32 x 32), tiles of the upper
//                triangle only (K is symmetric): per 16-feature step,
//                hi*hi + hi*lo + lo*hi on v_mfma_f32_32x32x16_f16 with FP32
//                accumulation (the dropped lo*lo term is below 2^-22 of the
//                product).  Then K(q, r) from the dot as the solver defines
//                it: gamma (x_q^2 + x_r^2 - 2 dot) in FP64 with the exact FP64
//                squared norms, exp in FP32 for RBF; rounded to float into
//                cache row q (Qfloat, libsvm/svm.cpp:16).
//
// Measured against exact FP64 dots on the n=100k dense problem's rows, the
// kernel values differ by a median of 5e-8 relative (the float rounding of
// the cache itself is 6e-8) and at most ~2e-6.
#include <hip/hip_runtime.h>

#include "smo_device.hpp"

namespace smo {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef float float16v __attribute__((ext_vector_type(16)));

constexpr int GT = 128;        // tile edge (rows of q and of r)
constexpr int GBLOCK = 256;    // 4 waves

// Halves per row of the split copy: 2 x (d rounded up to 16).
__host__ __device__ inline long long split_row_halves(int d) { return 2LL * ((d + 15) / 16 * 16); }

// One thread per (row, group of 8 features); lanes of a wave take 64
// consecutive rows of one group, so every load is a coalesced 256-byte read
// of the SELL slice.  *bad is set if a value is out of FP16's range.
__global__ void __launch_bounds__(GBLOCK) k_split_f16(View V, _Float16 *xs, int *bad) {
  const long long t = (long long)blockIdx.x * GBLOCK + threadIdx.x;
  const int groups = (V.d + 15) / 16 * 2;  // groups of 8 features per row, padded
  const long long nrows = (long long)V.nslices * SLICE;
  if (t >= nrows * groups) return;
  const int slice = (int)(t / ((long long)groups * SLICE));
  const int rem = (int)(t % ((long long)groups * SLICE));
  const int g = rem / SLICE, lane = rem % SLICE;
  const int r = slice * SLICE + lane;
  if (r >= V.l) return;
  const long long base = V.slice_ptr[slice] + lane;
  half8 hi, lo;
  int b = 0;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    const int p = g * 8 + j;
    const float x = p < V.d ? V.vals[base + (long long)p * SLICE] : 0.0f;
    b |= !(fabsf(x) < 32768.0f);
    const _Float16 h = (_Float16)x;
    hi[j] = h;
    lo[j] = (_Float16)(x - (float)h);
  }
  half8 *dst = reinterpret_cast<half8 *>(xs + (long long)r * split_row_halves(V.d) + (long long)g * 16);
  dst[0] = hi;
  dst[1] = lo;
  if (b) atomicOr(bad, 1);
}

// Tile t of the upper triangle (tq <= tr) of an nt x nt tile grid, rows of
// the triangle in order: tq is the largest with tq (2 nt - tq + 1) / 2 <= t.
__device__ __forceinline__ void tri_tile(long long t, int nt, int &tq, int &tr) {
  const double a = 2.0 * nt + 1.0;
  int q = (int)((a - sqrt(a * a - 8.0 * (double)t)) / 2.0);
  q = max(0, min(q, nt - 1));
  while (q > 0 && (long long)q * (2LL * nt - q + 1) / 2 > t) --q;
  while (q + 1 < nt && (long long)(q + 1) * (2LL * nt - q) / 2 <= t) ++q;
  tq = q;
  tr = q + (int)(t - (long long)q * (2LL * nt - q + 1) / 2);
}

// Kernel value from the dot product (FP64 arithmetic, FP32 exp for RBF),
// for a kernel type fixed at compile time.
template <int KT>
__device__ __forceinline__ float gram_value_t(const View &V, double xq, double xr, float acc) {
  const double dot = (double)acc;
  if (KT == K_LINEAR) return (float)dot;
  if (KT == K_POLY) return (float)powi(V.gamma * dot + V.coef0, V.degree);
  if (KT == K_RBF) return expf((float)(-V.gamma * ((xq + xr) - 2 * dot)));
  return (float)tanh(V.gamma * dot + V.coef0);
}

// Kernel value from the dot product (FP64 arithmetic, FP32 exp for RBF).
__device__ __forceinline__ float gram_value(const View &V, double xq, double xr, float acc) {
  const double dot = (double)acc;
  switch (V.ktype) {
    case K_LINEAR: return (float)dot;
    case K_POLY: return (float)powi(V.gamma * dot + V.coef0, V.degree);
    case K_RBF: return expf((float)(-V.gamma * ((xq + xr) - 2 * dot)));
    case K_SIGMOID: return (float)tanh(V.gamma * dot + V.coef0);
    default: return 0.0f;
  }
}

// K is symmetric: only tiles with tq <= tr are computed, in XCD order
// (consecutive workgroup ids go round-robin over the 8 XCDs, so workgroup b
// takes a tile from XCD (b mod 8)'s contiguous share of the list, and each
// XCD's L2 serves a band of q rows).  A tile off the diagonal is written to
// rows q (K(q, r)) and, transposed through LDS, to rows r (K(r, q)); every
// store is a run of consecutive floats of one cache row.
//
// Operands go through LDS in chunks of GKC = 16 features, double buffered:
// the workgroup's 128 q rows and 128 r rows of a chunk (64 bytes of hi/lo
// halves per row) are loaded by all 256 threads into registers while the
// MFMAs of the previous chunk run from LDS, then stored to the other buffer;
// one barrier per chunk.  The 16-byte pieces of a row are swizzled in LDS
// (gswz) against bank conflicts of the operand reads.  The MFMAs accumulate 16 features
// at a time in feature order whatever the chunk width, so the chunking does
// not change a value.
constexpr int GKC = 16;                   // features per chunk
constexpr int GROW = GKC * 2 * 2;         // LDS bytes per row of a chunk: hi/lo halves
constexpr int GBUF = 2 * GT * GROW;       // one buffer: 128 q rows + 128 r rows
constexpr int GPPR = GKC / 4;             // 16-byte pieces per row of a chunk
constexpr int GNP = GT * GPPR / GBLOCK;   // pieces per thread of the q (and of the r) rows
constexpr int GTP = 4 * 64 * 33 * 4;      // transpose: each wave's 64 x 32 half-block, padded
constexpr int GLDS = 2 * GBUF > GTP ? 2 * GBUF : GTP;
static_assert((GPPR & (GPPR - 1)) == 0, "pieces per row: a power of two");
// 16-byte piece c of LDS row r sits at position c ^ ((r / 4) mod GPPR): the
// 16 rows one ds_read_b128 group of lanes reads (same piece, consecutive
// rows at a 64-byte stride) then cover all 64 banks once
__device__ __forceinline__ int gswz(int r, int c) { return c ^ ((r >> 2) & (GPPR - 1)); }
// 35 KB of LDS (operands 32 KB + norms 2 KB; the transpose reuses the
// operand buffers) and at most 128 VGPRs: four workgroups per CU, so one's
// MFMA phase overlaps the others' loads and epilogues.  n=100k d=256: 15.5
// ms; 16.5 ms with 4-byte stores of each value straight from the MFMA
// registers (4x the store instructions); 17.6 ms with rows padded to 80
// bytes instead of swizzled, three workgroups per CU; 19.6 ms with
// 32-feature chunks, a 64 x 64 transpose and two workgroups per CU.

template <int KT>
__global__ void __launch_bounds__(GBLOCK, 4) k_gram_f16(View V, const _Float16 *xs, long long ntiles, int nt) {
  __shared__ __attribute__((aligned(16))) unsigned char lds[GLDS];
  __shared__ double sq[2 * GT];  // squared norms of the tile's q rows, then r rows
  const long long per = (ntiles + 7) / 8;
  const long long tile = (blockIdx.x % 8) * per + blockIdx.x / 8;
  if (tile >= ntiles) return;  // uniform over the workgroup
  int tq, tr;
  tri_tile(tile, nt, tq, tr);
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  if (KT == K_RBF) {  // read by the epilogue, after the main loop's barriers
    const int t = (int)threadIdx.x, base = t < GT ? tq * GT : tr * GT - GT;
    sq[t] = V.xsq[min(base + t, V.l - 1)];
  }
  const int q0 = tq * GT + (wave / 2) * 64, r0 = tr * GT + (wave % 2) * 64;
  const long long rh = split_row_halves(V.d);
  const int ngroups = (int)(rh / 16);  // groups of 8 features ([hi x 8][lo x 8], 32 bytes) per row
  const int l = V.l;
  // chunk loads: thread t moves pieces p = t + 256 j (j < GNP) of the q rows
  // and of the r rows: row p / GPPR, bytes 16 (p % GPPR) of the chunk's 4 GKC
  const unsigned char *gx = reinterpret_cast<const unsigned char *>(xs);
  long long src_q[GNP], src_r[GNP];
  int dst[GNP], col[GNP];
#pragma unroll
  for (int j = 0; j < GNP; ++j) {
    const int pc = (int)threadIdx.x + GBLOCK * j, row = pc / GPPR;
    col[j] = pc % GPPR;
    src_q[j] = (long long)min(tq * GT + row, l - 1) * rh * 2 + col[j] * 16;
    src_r[j] = (long long)min(tr * GT + row, l - 1) * rh * 2 + col[j] * 16;
    dst[j] = row * GROW + gswz(row, col[j]) * 16;
  }
  typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
  const int nchunks = (V.d + GKC - 1) / GKC;
  // one chunk's pieces of this thread into registers (groups past the
  // row's padded features read as zeros)
#define GRAM_FETCH(c)                                                                          \
  _Pragma("unroll") for (int j = 0; j < GNP; ++j) {                                            \
    const bool in = (c) * (GKC / 8) + col[j] / 2 < ngroups;                                    \
    const u32x4 z = {0u, 0u, 0u, 0u};                                                          \
    pq[j] = in ? *reinterpret_cast<const u32x4 *>(gx + src_q[j] + (long long)(c) * (GKC * 4)) : z; \
    pr[j] = in ? *reinterpret_cast<const u32x4 *>(gx + src_r[j] + (long long)(c) * (GKC * 4)) : z; \
  }
#define GRAM_STASH(b)                                                                          \
  _Pragma("unroll") for (int j = 0; j < GNP; ++j) {                                            \
    *reinterpret_cast<u32x4 *>(lds + (b) * GBUF + dst[j]) = pq[j];                             \
    *reinterpret_cast<u32x4 *>(lds + (b) * GBUF + GT * GROW + dst[j]) = pr[j];                 \
  }
  u32x4 pq[GNP], pr[GNP];
  float16v acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.0f;
  GRAM_FETCH(0)
  GRAM_STASH(0)
  __syncthreads();
  for (int c = 0; c < nchunks; ++c) {
    if (c + 1 < nchunks) {  // in flight during this chunk's MFMAs
      GRAM_FETCH(c + 1)
    }
    const unsigned char *A = lds + (c & 1) * GBUF, *B = A + GT * GROW;
#pragma unroll
    for (int ss = 0; ss < GKC / 16; ++ss) {
      // lane: row lane % 32 of each 32-row block, features 16 ss + 8 (lane / 32) .. + 7
      // of the chunk (group 2 ss + lane / 32: hi at +0, lo at +16 bytes);
      // the four blocks' products of one term go back to back, so no two
      // consecutive MFMAs accumulate into the same registers
      const int pg = 2 * (2 * ss + lane / 32);  // the group's hi piece; lo is pg + 1
      half8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
      for (int b = 0; b < 2; ++b) {
        const int ia = (wave / 2) * 64 + b * 32 + lane % 32, ib = (wave % 2) * 64 + b * 32 + lane % 32;
        ah[b] = *reinterpret_cast<const half8 *>(A + ia * GROW + gswz(ia, pg) * 16);
        al[b] = *reinterpret_cast<const half8 *>(A + ia * GROW + gswz(ia, pg + 1) * 16);
        bh[b] = *reinterpret_cast<const half8 *>(B + ib * GROW + gswz(ib, pg) * 16);
        bl[b] = *reinterpret_cast<const half8 *>(B + ib * GROW + gswz(ib, pg + 1) * 16);
      }
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[i], bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[i], bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[i], bh[j], acc[i][j], 0, 0, 0);
    }
    if (c + 1 < nchunks) {
      GRAM_STASH((c + 1) & 1)
    }
    __syncthreads();
  }
#undef GRAM_FETCH
#undef GRAM_STASH
  // D[m][n] of block (i, j) in lane element e: m = 8 (e / 4) + 4 (lane / 32)
  // + e % 4, n = lane % 32.  Rows q: K(q, r), lanes on consecutive r; the
  // values also go to LDS (over the operand buffers, after the last
  // barrier) for the transposed store.
  float(*tp)[64][33] = reinterpret_cast<float(*)[64][33]>(lds);
  const bool mirror = tq != tr;
  // a wave's block lies inside the matrix unless it touches row or column
  // l (only the last tile row / column): then every store is checked
  const bool inside = q0 + 64 <= l && r0 + 64 <= l;
  const long long stride = V.row_stride;
  const int qa = q0 + 4 * (lane / 32), rb = r0 + lane % 32;  // this lane's first row and column
  float *dst0 = V.rows + (long long)qa * stride + rb;
  float *tp0 = &tp[wave][4 * (lane / 32)][lane % 32];
  const double *sqq = &sq[(wave / 2) * 64 + 4 * (lane / 32)];
  const int q = q0 + lane;
  typedef float f4 __attribute__((ext_vector_type(4)));
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    // columns 32 j .. 32 j + 31 of the wave's block, through LDS (64 x 32
    // per wave): a block inside the matrix goes out in 16-byte stores, 8
    // rows q (K(q, r)) and 4 rows r (K(r, q) = K(q, r)) per instruction;
    // one on the matrix's edge in checked 4-byte stores
    const double xr = KT == K_RBF ? sq[GT + (wave % 2) * 64 + j * 32 + lane % 32] : 0.0;
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int e = 0; e < 16; ++e) {
        // row m = 4 (lane / 32) + mo of the wave's block, column n = 32 j + lane % 32
        const int mo = i * 32 + 8 * (e / 4) + e % 4;
        const float K = gram_value_t<KT>(V, KT == K_RBF ? sqq[mo] : 0.0, xr, acc[i][j][e]);
        // (non-temporal: the 40 GB of a full matrix are not read back soon)
        if (!inside && qa + mo < l && rb + 32 * j < l) __builtin_nontemporal_store(K, &dst0[(long long)mo * stride + 32 * j]);
        tp0[mo * 33] = K;
      }
    __syncthreads();
    if (inside) {
      // rows q: row 8 it + lane / 8 of the block, columns 4 (lane % 8) .. + 3
      const int c4 = 4 * (lane % 8);
      float *qrow = V.rows + (long long)(q0 + lane / 8) * stride + r0 + 32 * j + c4;
#pragma unroll
      for (int it = 0; it < 8; ++it) {
        const int m = 8 * it + lane / 8;
        const f4 v = {tp[wave][m][c4], tp[wave][m][c4 + 1], tp[wave][m][c4 + 2], tp[wave][m][c4 + 3]};
        __builtin_nontemporal_store(v, reinterpret_cast<f4 *>(qrow + (long long)(8 * it) * stride));
      }
      if (mirror) {
        // rows r: row 4 it + lane / 16 of the half, columns q0 + 4 (lane % 16) .. + 3
        const int m4 = 4 * (lane % 16);
        float *rrow = V.rows + (long long)(r0 + 32 * j + lane / 16) * stride + q0 + m4;
#pragma unroll
        for (int it = 0; it < 8; ++it) {
          const int n = 4 * it + lane / 16;
          const f4 v = {tp[wave][m4][n], tp[wave][m4 + 1][n], tp[wave][m4 + 2][n], tp[wave][m4 + 3][n]};
          __builtin_nontemporal_store(v, reinterpret_cast<f4 *>(rrow + (long long)(4 * it) * stride));
        }
      }
    } else if (mirror) {
      // rows r: K(r, q) = K(q, r), lanes on consecutive q
      float *mp = V.rows + (long long)(r0 + 32 * j) * stride + q;
#pragma unroll 8
      for (int n = 0; n < 32; ++n) {
        if (r0 + 32 * j + n < l && q < l) __builtin_nontemporal_store(tp[wave][lane][n], mp);
        mp += stride;
      }
    }
    __syncthreads();  // before the next half overwrites tp
  }
}


// ------------------------------------------------- prediction Gram tiles
//
// Opt-in fast prediction (svmhip_predict_batch with mfma_rows = 1) for
// models with one coefficient per support vector (two classes, regression,
// one-class): decision values within ~1e-6 relative instead of FP64's
// k_function arithmetic.  The test rows and the support vectors are split
// into dense FP16 hi/lo copies (the layout of k_split_f16, from their FP64
// values); k_predict_gram_f16 computes 128 test rows x 128 support vectors
// of dots per workgroup with the three MFMA products of k_gram_f16, the
// kernel values from the dots and the FP64 squared norms, and the tile's
// sum of coef_s * K(t, s) per test row (FP64, in a fixed order);
// k_predict_gram_sum adds the tiles' partial sums in support-vector order.

// Dense split copy of SELL-64 rows with VT values: row r at out + r * rh
// halves; out must be zero (a sparse row writes only its stored features).
// sq[r] = sum of x^2 over the row in feature order (FP64).  *bad is set if
// a value is out of FP16's range.
template <class VT>
__global__ void __launch_bounds__(GBLOCK) k_split_rows_f16(PRows R, long long rh, _Float16 *out, double *sq,
                                                           int *bad) {
  const int r = blockIdx.x * GBLOCK + threadIdx.x;
  if (r >= R.l) return;
  const int len = R.row_len[r];
  const long long base = R.slice_ptr[r / SLICE] + (r % SLICE);
  const VT *vals = static_cast<const VT *>(R.vals);
  _Float16 *row = out + (long long)r * rh;
  double s2 = 0.0;
  int b = 0;
  for (int p = 0; p < len; ++p) {
    const double x = (double)vals[base + (long long)p * SLICE];
    const int c = R.dense ? p : R.cols[base + (long long)p * SLICE];
    s2 = s2 + x * x;
    b |= !(fabs(x) < 32768.0);
    const _Float16 h = (_Float16)x;
    row[(c / 8) * 16 + c % 8] = h;
    row[(c / 8) * 16 + 8 + c % 8] = (_Float16)(x - (double)h);
  }
  sq[r] = s2;
  if (b) atomicOr(bad, 1);
}

// Kernel value from a dot product and the FP64 squared norms.
template <int KT>
__device__ __forceinline__ double pgram_value(const PredictArgs &A, double xt, double xs, double dot) {
  if (KT == K_LINEAR) return dot;
  if (KT == K_POLY) return powi(A.gamma * dot + A.coef0, A.degree);
  if (KT == K_RBF) return exp(-A.gamma * ((xt + xs) - 2 * dot));
  return tanh(A.gamma * dot + A.coef0);
}

// Workgroup b: test-row tile b % ntt, support-vector tile b / ntt.  Wave w
// takes test rows 64 (w / 2) .. + 63 and support vectors 64 (w % 2) .. + 63
// of the tile, as k_gram_f16's waves take q and r.  partial[bs * nt + t] =
// sum over the tile's support vectors s of coef_s * K(t, s).
template <int KT>
__global__ void __launch_bounds__(GBLOCK) k_predict_gram_f16(PredictArgs A, const _Float16 *xt, const _Float16 *xs,
                                                             long long rh, const double *sqt, const double *sqs,
                                                             double *partial) {
  __shared__ double red[4][64];
  const int nt = A.tv.l, ns = A.sv.l;
  const int ntt = (nt + GT - 1) / GT;
  const int bt = blockIdx.x % ntt, bs = blockIdx.x / ntt;
  const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
  const int q0 = bt * GT + (wave / 2) * 64, r0 = bs * GT + (wave % 2) * 64;
  // rows past the ends read the last row (valid memory); their terms are
  // dropped below
  const half8 *aq[2], *br[2];
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    const int q = min(q0 + b * 32 + lane % 32, nt - 1), r = min(r0 + b * 32 + lane % 32, ns - 1);
    aq[b] = reinterpret_cast<const half8 *>(xt + (long long)q * rh) + 2 * (lane / 32);
    br[b] = reinterpret_cast<const half8 *>(xs + (long long)r * rh) + 2 * (lane / 32);
  }
  float16v acc[2][2];
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j)
#pragma unroll
      for (int e = 0; e < 16; ++e) acc[i][j][e] = 0.0f;
  const int steps = (int)(rh / 32);
  for (int s = 0; s < steps; ++s) {
    half8 ah[2], al[2], bh[2], bl[2];
#pragma unroll
    for (int b = 0; b < 2; ++b) {
      ah[b] = aq[b][4 * s];
      al[b] = aq[b][4 * s + 1];
      bh[b] = br[b][4 * s];
      bl[b] = br[b][4 * s + 1];
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[i], bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[i], bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[i], bh[j], acc[i][j], 0, 0, 0);
  }
  // D[m][n] of block (i, j) in lane element e: m = 8 (e / 4) + 4 (lane / 32)
  // + e % 4 (test row), n = lane % 32 (support vector)
  double w[2], ys[2];
#pragma unroll
  for (int j = 0; j < 2; ++j) {
    const int sv = r0 + j * 32 + lane % 32;
    w[j] = sv < ns ? A.coef[sv] : 0.0;
    ys[j] = sqs[min(sv, ns - 1)];
  }
#pragma unroll
  for (int i = 0; i < 2; ++i)
#pragma unroll
    for (int e = 0; e < 16; ++e) {
      const int m = i * 32 + 8 * (e / 4) + 4 * (lane / 32) + e % 4;
      const double xq = sqt[min(q0 + m, nt - 1)];
      double v = w[0] * pgram_value<KT>(A, xq, ys[0], (double)acc[i][0][e]);
      v = v + w[1] * pgram_value<KT>(A, xq, ys[1], (double)acc[i][1][e]);
      // sum over the 32 support vectors of this half-wave
#pragma unroll
      for (int k = 0; k < 5; ++k) v = v + __shfl_xor(v, 16 >> k, 32);
      if (lane % 32 == 0) red[wave][m] = v;
    }
  __syncthreads();
  // waves 2k and 2k + 1 hold the two support-vector halves of test rows
  // 64k .. 64k + 63
  if (threadIdx.x < 128) {
    const int k = threadIdx.x / 64, m = threadIdx.x % 64, t = bt * GT + k * 64 + m;
    if (t < nt) partial[(long long)bs * nt + t] = red[2 * k][m] + red[2 * k + 1][m];
  }
}

// dec[t] = sum over support-vector tiles (in order) of partial - rho.
__global__ void __launch_bounds__(GBLOCK) k_predict_gram_sum(const double *partial, int nt, int nst, const double *rho,
                                                             double *dec) {
  const int t = blockIdx.x * GBLOCK + threadIdx.x;
  if (t >= nt) return;
  double a = 0.0;
  for (int b = 0; b < nst; ++b) a = a + partial[(long long)b * nt + t];
  dec[t] = a - rho[0];
}

}  // namespace smo

extern "C" {

hipError_t smo_launch_split_f16(const smo::View *V, _Float16 *xs, int *bad, hipStream_t st) {
  const long long groups = (V->d + 15) / 16 * 2;
  const long long n = (long long)V->nslices * smo::SLICE * groups;
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(smo::k_split_f16, dim3((unsigned)((n + smo::GBLOCK - 1) / smo::GBLOCK)), dim3(smo::GBLOCK), 0,
                     st, *V, xs, bad);
  return hipGetLastError();
}

hipError_t smo_launch_gram_f16(const smo::View *V, const _Float16 *xs, hipStream_t st) {
  const int nt = (V->l + smo::GT - 1) / smo::GT;
  const long long ntiles = (long long)nt * (nt + 1) / 2;  // the upper triangle of tiles
  if (ntiles <= 0) return hipSuccess;
  const long long grid = (ntiles + 7) / 8 * 8;  // whole rounds of the XCD order
  if (grid > 0x7fffffffLL) return hipErrorInvalidValue;
  switch (V->ktype) {
    case smo::K_LINEAR:
      hipLaunchKernelGGL(smo::k_gram_f16<smo::K_LINEAR>, dim3((unsigned)grid), dim3(smo::GBLOCK), 0, st, *V, xs, ntiles, nt);
      break;
    case smo::K_POLY:
      hipLaunchKernelGGL(smo::k_gram_f16<smo::K_POLY>, dim3((unsigned)grid), dim3(smo::GBLOCK), 0, st, *V, xs, ntiles, nt);
      break;
    case smo::K_RBF:
      hipLaunchKernelGGL(smo::k_gram_f16<smo::K_RBF>, dim3((unsigned)grid), dim3(smo::GBLOCK), 0, st, *V, xs, ntiles, nt);
      break;
    case smo::K_SIGMOID:
      hipLaunchKernelGGL(smo::k_gram_f16<smo::K_SIGMOID>, dim3((unsigned)grid), dim3(smo::GBLOCK), 0, st, *V, xs, ntiles, nt);
      break;
    default:
      return hipErrorInvalidValue;
  }
  return hipGetLastError();
}

// Prediction Gram tiles (see k_predict_gram_f16): the split copies of
// A->tv and A->sv (each zeroed by the caller, smo_predict_split_halves per
// row), their squared norms, partial sums of (nt x ceil(ns / 128)) doubles.
hipError_t smo_launch_predict_split(const smo::PRows *R, _Float16 *out, double *sq, int *bad, int d,
                                    hipStream_t st) {
  if (R->l <= 0) return hipSuccess;
  const long long rh = smo::split_row_halves(d);
  const unsigned nb = (unsigned)((R->l + smo::GBLOCK - 1) / smo::GBLOCK);
  if (R->f64)
    hipLaunchKernelGGL(smo::k_split_rows_f16<double>, dim3(nb), dim3(smo::GBLOCK), 0, st, *R, rh, out, sq, bad);
  else
    hipLaunchKernelGGL(smo::k_split_rows_f16<float>, dim3(nb), dim3(smo::GBLOCK), 0, st, *R, rh, out, sq, bad);
  return hipGetLastError();
}

hipError_t smo_launch_predict_gram(const smo::PredictArgs *A, const _Float16 *xt, const _Float16 *xs,
                                   const double *sqt, const double *sqs, double *partial, hipStream_t st) {
  const int nt = A->tv.l, ns = A->sv.l;
  if (nt <= 0) return hipSuccess;
  if (ns <= 0 || A->npair != 1) return hipErrorInvalidValue;
  const long long rh = smo::split_row_halves(A->ld);
  const long long ntt = (nt + smo::GT - 1) / smo::GT, nst = (ns + smo::GT - 1) / smo::GT;
  if (ntt * nst > 0x7fffffffLL) return hipErrorInvalidValue;
  const dim3 grid((unsigned)(ntt * nst));
  switch (A->ktype) {
    case smo::K_LINEAR:
      hipLaunchKernelGGL(smo::k_predict_gram_f16<smo::K_LINEAR>, grid, dim3(smo::GBLOCK), 0, st, *A, xt, xs, rh, sqt, sqs,
                         partial);
      break;
    case smo::K_POLY:
      hipLaunchKernelGGL(smo::k_predict_gram_f16<smo::K_POLY>, grid, dim3(smo::GBLOCK), 0, st, *A, xt, xs, rh, sqt, sqs,
                         partial);
      break;
    case smo::K_RBF:
      hipLaunchKernelGGL(smo::k_predict_gram_f16<smo::K_RBF>, grid, dim3(smo::GBLOCK), 0, st, *A, xt, xs, rh, sqt, sqs,
                         partial);
      break;
    case smo::K_SIGMOID:
      hipLaunchKernelGGL(smo::k_predict_gram_f16<smo::K_SIGMOID>, grid, dim3(smo::GBLOCK), 0, st, *A, xt, xs, rh, sqt,
                         sqs, partial);
      break;
    default: return hipErrorInvalidValue;
  }
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(smo::k_predict_gram_sum, dim3((unsigned)((nt + smo::GBLOCK - 1) / smo::GBLOCK)), dim3(smo::GBLOCK),
                     0, st, partial, nt, (int)nst, A->rho, A->dec);
  return hipGetLastError();
}

long long smo_predict_split_halves(int d) { return smo::split_row_halves(d); }

int smo_predict_gram_tiles(int ns) { return (ns + smo::GT - 1) / smo::GT; }

size_t smo_split_f16_bytes(int l, int d) { return (size_t)l * (size_t)smo::split_row_halves(d) * sizeof(_Float16); }

}  // extern "C"
