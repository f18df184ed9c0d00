// Grouped (segmented) Jacobi CG: many independent sparse systems -- the
// requests of a serving batch -- iterate together, four launches per
// iteration for the whole group.
//
// Each system ("segment") keeps its matrix in shared pools at its own offsets
// -- CSR, or SELL-64 with int32 or 16-bit slice-relative columns -- and its
// vectors in shared vector pools; column indices are local to the segment.  A
// block table assigns every block of a launch to one segment and a contiguous
// row range (CSR) or slice range (SELL, one slice per wavefront) of it, so each
// block's partial sums belong to exactly one segment; one-block-per-segment finishers
// reduce each segment's partials in a fixed order (deterministic, and
// independent of which other segments share the launch) and update that
// segment's CG scalars.  Converged segments stop at once: their blocks return
// on entry.  Nothing here allocates, synchronises or issues memsets, so a
// chunk of iterations can be captured in a HIP graph and replayed.
//
// Precision per segment (mife_segment.prec).  MIFE_F64 segments iterate in
// FP64 throughout.  MIFE_F32 segments iterate in FP32 (FP32 values, vectors
// and products, FP64 scalars) with reliable updates in FP64, as the
// single-solve mife_refine: the FP64 solution x accumulates the FP32
// correction x32, and whenever a segment's FP32 residual has fallen by
// inner_rtol since its last replacement (tested every check_every iterations),
// x += x32 and the FP32 residual is replaced by the true residual b - A64 x
// computed with the FP64 values; p and q = A p are kept and
// beta = (r.z)_replaced / (r.z)_before-the-step.  Such a segment is done when
// its true FP64 residual meets its tolerance (or at its iteration limit); the
// FP32 recurrence never stops it by itself.  Both kinds share every launch: a
// block takes the branch of its segment's precision.
#include "mife_common.hpp"
#include "mife_profile.hpp"
#include "mife_sparse.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace mife {

namespace grouped {

constexpr int kL = 8;                  // CSR: lanes per row
constexpr int kRowsPerBlock = 256 / kL;
constexpr int kSlicesPerBlock = 4;     // SELL: one slice per wavefront

struct SegState {
  double rz, pq, alpha, beta, rr, bb, tol2, rr_true;
  double rz_prev;  // r.z before the last step
  double rr_rep;   // mixed precision: FP32 r.r right after the last replacement
  int iters, maxit, done;
  int pause;       // mixed precision: waiting for an FP64 replacement
  int mixed;       // the segment iterates in FP32 with FP64 reliable updates
  int pad;
};

// The vector pools: FP64 segments use the first set, FP32 segments the second
// (both at the segment's row_off).
struct Vecs {
  double *dinv, *r, *z, *p, *q;
  float *dinv32, *r32, *z32, *p32, *q32, *x32;
};

constexpr int kNumParts = 4;      // partial values per block slot
constexpr int kReplaceGrid = 2048;  // grid of the (mostly idle) replacement passes

struct Block {
  int32_t seg;
  int32_t row0;  // first row (CSR) or first slice (SELL), local to the segment
  int32_t rows;  // rows (CSR) or slices (SELL)
  int32_t part;  // index of this block's partial slot
};

struct View {
  const mife_segment* segs;
  const Block* blocks;
  const int32_t* rp;
  const int32_t* col;
  const uint16_t* col16;
  const int32_t* cbase;
  const double* val;
  const float* val32;  // FP32 copy of the values (mixed precision), or null
  SegState* st;
  double* parts;  // kNumParts values per partial slot: parts[v * cap + slot]
  int cap;
};

// y_row = (A x)_row for the rows of this block's segment, 8 lanes per row;
// vpool is the value pool of the precision T.
template <typename T>
__device__ __forceinline__ T row_dot(const View& V, const mife_segment& S, const T* __restrict__ vpool,
                                     const T* __restrict__ x, int row, int sub) {
  const int32_t* rp = V.rp + S.rp_off;
  const int32_t* col = V.col + S.nz_off;
  const T* val = vpool + S.nz_off;
  const T* xs = x + S.row_off;
  const int k0 = rp[row], k1 = rp[row + 1];
  T a = 0;
  for (int k = k0 + sub; k < k1; k += kL) a += val[k] * xs[col[k]];
#pragma unroll
  for (int off = kL / 2; off > 0; off >>= 1) a += __shfl_xor(a, off, kL);
  return a;
}

// SELL: (A x) of lane `lane`'s row of slice s of the segment (complete sum).
template <typename T>
__device__ __forceinline__ T slice_dot(const View& V, const mife_segment& S, const T* __restrict__ vpool,
                                       const T* __restrict__ x, int s, int lane) {
  const int32_t* sp = V.rp + S.rp_off;
  const T* val = vpool + S.nz_off;
  const T* xs = x + S.row_off;
  if (S.cb_off >= 0) return sell16_row<T>(sp, V.col16 + S.nz_off, V.cbase + S.cb_off, val, xs, s, lane);
  return sell_row<T>(sp, V.col + S.nz_off, val, xs, s, lane);
}

// Calls f(row, (A x)_row) for every row of block B (by one lane per row).
template <typename T, typename F>
__device__ __forceinline__ void block_rows(const View& V, const Block& B, const mife_segment& S,
                                           const T* __restrict__ vpool, const T* __restrict__ x, F&& f) {
  if (S.format == MIFE_FMT_SELL3) {
    // node rows: lane k of slice s gives rows 3 (64 s + k) + c; block columns in
    // the 16-bit column pool at the segment's value offset
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (wave < B.rows) {
      const int s = B.row0 + wave;
      const int32_t* sp = V.rp + S.rp_off;
      const int base = sp[s], len = (sp[s + 1] - base) / kSlice;
      T ax[3];
      sell3_sum<T, T>(V.col16 + S.nz_off + base + lane, vpool + S.nz_off + (size_t)9 * base + lane,
                      x + S.row_off + 3 * (size_t)V.cbase[S.cb_off + s], len, 0, 1, ax);
      const int row = 3 * (s * kSlice + lane);
#pragma unroll
      for (int c = 0; c < 3; ++c)
        if (row + c < S.n) f(row + c, ax[c]);
    }
    return;
  }
  if (S.format == MIFE_FMT_SELL) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (wave < B.rows) {
      const int s = B.row0 + wave;
      const T ax = slice_dot<T>(V, S, vpool, x, s, lane);
      const int row = s * kSlice + lane;
      if (row < S.n) f(row, ax);
    }
    return;
  }
  const int lr = threadIdx.x / kL, sub = threadIdx.x % kL;
  for (int row = B.row0 + lr; row < B.row0 + B.rows; row += kRowsPerBlock) {
    const T ax = row_dot<T>(V, S, vpool, x, row, sub);
    if (sub == 0) f(row, ax);
  }
}

// First row and row count of block B (vector kernels: one row per thread).
__device__ __forceinline__ void block_row_range(const Block& B, const mife_segment& S, int& r0, int& r1) {
  if (S.format == MIFE_FMT_SELL3) {
    r0 = B.row0 * 3 * kSlice;
    r1 = min(S.n, (B.row0 + B.rows) * 3 * kSlice);
  } else if (S.format == MIFE_FMT_SELL) {
    r0 = B.row0 * kSlice;
    r1 = min(S.n, (B.row0 + B.rows) * kSlice);
  } else {
    r0 = B.row0;
    r1 = B.row0 + B.rows;
  }
}

template <int NV>
__device__ __forceinline__ void store_partials(const View& V, const Block& B, double (&m)[NV]) {
  __shared__ double red[NV][4];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    const double s = wave_sum(m[v]);
    if ((threadIdx.x & 63) == 0) red[v][threadIdx.x >> 6] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) V.parts[(size_t)v * V.cap + B.part] = red[v][0] + red[v][1] + red[v][2] + red[v][3];
  }
}

}  // namespace grouped

using namespace grouped;

template <typename T>
__device__ __forceinline__ void diag_body(const View& V, const Block& B, const mife_segment& S,
                                          const T* __restrict__ vpool, T* __restrict__ dinv) {
  if (S.format == MIFE_FMT_SELL3) {
    // node row k: its block at node column k (before the padding, whose column
    // is also k), values 0, 4 and 8
    const int32_t* sp = V.rp + S.rp_off;
    const T* val = vpool + S.nz_off;
    for (int t = threadIdx.x; t < B.rows * kSlice; t += blockDim.x) {
      const int s = B.row0 + t / kSlice, lane = t % kSlice;
      const int k = s * kSlice + lane;
      if (3 * k >= S.n) continue;
      const int base = sp[s], len = (sp[s + 1] - base) / kSlice, cb = V.cbase[S.cb_off + s];
      int jd = -1;
      for (int j = 0; j < len; ++j)
        if (cb + (int)V.col16[S.nz_off + base + j * kSlice + lane] == k) {
          jd = j;
          break;
        }
#pragma unroll
      for (int c = 0; c < 3; ++c) {
        const T d = jd >= 0 ? val[(size_t)9 * (base + jd * kSlice) + 4 * c * kSlice + lane] : (T)0;
        dinv[S.row_off + 3 * k + c] = d != (T)0 ? (T)1 / d : (T)1;
      }
    }
    return;
  }
  int r0, r1;
  block_row_range(B, S, r0, r1);
  const int row = r0 + (int)threadIdx.x;
  if (row >= r1) return;
  T d = 0;
  const T* val = vpool + S.nz_off;
  if (S.format == MIFE_FMT_SELL) {
    const int32_t* sp = V.rp + S.rp_off;
    const int s = row / kSlice, lane = row - s * kSlice;
    for (int e = sp[s] + lane; e < sp[s + 1]; e += kSlice) {
      const int c = S.cb_off >= 0 ? V.cbase[S.cb_off + s] + (int)V.col16[S.nz_off + e] : V.col[S.nz_off + e];
      if (c == row && val[e] != (T)0) {
        d = val[e];
        break;
      }
    }
  } else {
    const int32_t* rp = V.rp + S.rp_off;
    const int32_t* col = V.col + S.nz_off;
    for (int k = rp[row]; k < rp[row + 1]; ++k)
      if (col[k] == row) d = val[k];
  }
  dinv[S.row_off + row] = d != (T)0 ? (T)1 / d : (T)1;
}

// Jacobi diagonal D^-1 of the segment's precision
__global__ __launch_bounds__(256) void gcg_diag_kernel(View V, Vecs W) {
  const Block B = V.blocks[blockIdx.x];
  const mife_segment S = V.segs[B.seg];
  if (S.prec == MIFE_F32)
    diag_body<float>(V, B, S, V.val32, W.dinv32);
  else
    diag_body<double>(V, B, S, V.val, W.dinv);
}

// FP64 segment: r = b - A x, z = D^-1 r, p = q = 0.
// FP32 segment: r64 = b - A64 x, r32 = r64, z32 = D^-1 r32, p32 = q32 = x32 = 0.
// Partials r.z, r.r, b.b and (FP32 segments) r64.r64.
__global__ __launch_bounds__(256) void gcg_init_kernel(View V, Vecs W, const double* __restrict__ b,
                                                       const double* __restrict__ x) {
  const Block B = V.blocks[blockIdx.x];
  const mife_segment S = V.segs[B.seg];
  double m[kNumParts] = {0.0, 0.0, 0.0, 0.0};
  if (S.prec == MIFE_F32) {
    block_rows<double>(V, B, S, V.val, x, [&](int row, double ax) {
      const long long g = S.row_off + row;
      const double r64 = b[g] - ax;
      const float ri = (float)r64;
      const float zi = W.dinv32[g] * ri;
      W.r32[g] = ri;
      W.z32[g] = zi;
      W.p32[g] = 0.0f;
      W.q32[g] = 0.0f;
      W.x32[g] = 0.0f;
      m[0] += (double)ri * (double)zi;
      m[1] += (double)ri * (double)ri;
      m[2] += b[g] * b[g];
      m[3] += r64 * r64;
    });
  } else {
    block_rows<double>(V, B, S, V.val, x, [&](int row, double ax) {
      const long long g = S.row_off + row;
      const double ri = b[g] - ax;
      const double zi = W.dinv[g] * ri;
      W.r[g] = ri;
      W.z[g] = zi;
      W.p[g] = 0.0;
      W.q[g] = 0.0;
      m[0] += ri * zi;
      m[1] += ri * ri;
      m[2] += b[g] * b[g];
    });
  }
  store_partials<kNumParts>(V, B, m);
}

// one block per segment: fixed-order sums of the segment's partial slots
// [seg_parts[k], seg_parts[k+1]) (k = position of the segment in the table)
template <int NV>
__device__ __forceinline__ void seg_sums(const View& V, const int32_t* __restrict__ seg_parts, int k,
                                         double (&out)[NV]) {
  __shared__ double red[NV][4];
  const int p0 = seg_parts[k], p1 = seg_parts[k + 1];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    double a = 0.0;
    for (int k = p0 + threadIdx.x; k < p1; k += blockDim.x) a += V.parts[(size_t)v * V.cap + k];
    a = wave_sum(a);
    if ((threadIdx.x & 63) == 0) red[v][threadIdx.x >> 6] = a;
  }
  __syncthreads();
#pragma unroll
  for (int v = 0; v < NV; ++v) out[v] = red[v][0] + red[v][1] + red[v][2] + red[v][3];
}

__global__ __launch_bounds__(256) void gfin_init_kernel(View V, const int32_t* __restrict__ seg_parts,
                                                        const int32_t* __restrict__ seg_ids) {
  const int s = seg_ids[blockIdx.x];
  double out[kNumParts];
  seg_sums<kNumParts>(V, seg_parts, blockIdx.x, out);
  if (threadIdx.x == 0) {
    SegState& g = V.st[s];
    const mife_segment S = V.segs[s];
    const bool mixed = S.prec == MIFE_F32;
    g.rz = out[0];
    g.rr = out[1];
    g.bb = out[2];
    g.beta = 0.0;
    g.alpha = 0.0;
    g.pq = 0.0;
    g.rz_prev = 0.0;
    g.rr_rep = out[1];
    g.iters = 0;
    g.maxit = S.maxit;
    g.tol2 = S.rtol * S.rtol;
    g.rr_true = mixed ? out[3] : -1.0;
    g.pause = 0;
    g.mixed = mixed ? 1 : 0;
    g.done = ((mixed ? out[3] : out[1]) <= g.tol2 * out[2] || S.maxit <= 0) ? 1 : 0;
  }
}

__device__ __forceinline__ bool seg_idle(const SegState& g) { return g.done || g.pause; }

template <typename T>
__device__ __forceinline__ void spmv_body(const View& V, const Block& B, const mife_segment& S,
                                          const T* __restrict__ vpool, const T* __restrict__ z,
                                          T* __restrict__ p, T* __restrict__ q) {
  const T beta = (T)V.st[B.seg].beta;
  double m[1] = {0.0};
  block_rows<T>(V, B, S, vpool, z, [&](int row, T az) {
    const long long g = S.row_off + row;
    const T pi = z[g] + beta * p[g];
    const T qi = az + beta * q[g];
    p[g] = pi;
    q[g] = qi;
    m[0] += (double)pi * (double)qi;
  });
  store_partials<1>(V, B, m);
}

// p = z + beta p ; q = A z + beta q ; partial p.q (in the segment's precision)
__global__ __launch_bounds__(256) void gcg_spmv_kernel(View V, Vecs W) {
  const Block B = V.blocks[blockIdx.x];
  if (seg_idle(V.st[B.seg])) return;
  const mife_segment S = V.segs[B.seg];
  if (S.prec == MIFE_F32)
    spmv_body<float>(V, B, S, V.val32, W.z32, W.p32, W.q32);
  else
    spmv_body<double>(V, B, S, V.val, W.z, W.p, W.q);
}

__global__ __launch_bounds__(256) void gfin_alpha_kernel(View V, const int32_t* __restrict__ seg_parts,
                                                         const int32_t* __restrict__ seg_ids) {
  const int s = seg_ids[blockIdx.x];
  if (seg_idle(V.st[s])) return;
  double out[1];
  seg_sums<1>(V, seg_parts, blockIdx.x, out);
  if (threadIdx.x == 0) {
    SegState& g = V.st[s];
    g.pq = out[0];
    g.alpha = out[0] != 0.0 ? g.rz / out[0] : 0.0;
  }
}

template <typename T>
__device__ __forceinline__ void update_body(const View& V, const Block& B, const mife_segment& S, T* __restrict__ x,
                                            T* __restrict__ r, T* __restrict__ z, const T* __restrict__ p,
                                            const T* __restrict__ q, const T* __restrict__ dinv) {
  const T alpha = (T)V.st[B.seg].alpha;
  double m[2] = {0.0, 0.0};
  int r0, r1;
  block_row_range(B, S, r0, r1);
  for (int row = r0 + threadIdx.x; row < r1; row += 256) {
    const long long g = S.row_off + row;
    x[g] += alpha * p[g];
    const T ri = r[g] - alpha * q[g];
    const T zi = dinv[g] * ri;
    r[g] = ri;
    z[g] = zi;
    m[0] += (double)ri * (double)zi;
    m[1] += (double)ri * (double)ri;
  }
  store_partials<2>(V, B, m);
}

// x += alpha p ; r -= alpha q ; z = D^-1 r ; partials r.z, r.r (x: the FP64
// solution of FP64 segments; FP32 segments update their correction x32)
__global__ __launch_bounds__(256) void gcg_update_kernel(View V, Vecs W, double* __restrict__ x) {
  const Block B = V.blocks[blockIdx.x];
  if (seg_idle(V.st[B.seg])) return;
  const mife_segment S = V.segs[B.seg];
  if (S.prec == MIFE_F32)
    update_body<float>(V, B, S, W.x32, W.r32, W.z32, W.p32, W.q32, W.dinv32);
  else
    update_body<double>(V, B, S, x, W.r, W.z, W.p, W.q, W.dinv);
}

// FP64 segments: done when the recurrence residual meets the tolerance or at
// maxit.  FP32 segments pause for a replacement at maxit and, on a CHECK
// iteration, once r.r has fallen by delta2 since the last replacement or
// below a quarter of the target.
template <bool CHECK>
__global__ __launch_bounds__(256) void gfin_beta_kernel(View V, const int32_t* __restrict__ seg_parts,
                                                        const int32_t* __restrict__ seg_ids, double delta2) {
  const int s = seg_ids[blockIdx.x];
  if (seg_idle(V.st[s])) return;
  double out[2];
  seg_sums<2>(V, seg_parts, blockIdx.x, out);
  if (threadIdx.x == 0) {
    SegState& g = V.st[s];
    g.rz_prev = g.rz;
    g.beta = g.rz != 0.0 ? out[0] / g.rz : 0.0;
    g.rz = out[0];
    g.rr = out[1];
    g.iters += 1;
    if (!g.mixed) {
      g.done = (out[1] <= g.tol2 * g.bb || g.iters >= g.maxit) ? 1 : 0;
    } else if (g.iters >= g.maxit ||
               (CHECK && (out[1] <= delta2 * g.rr_rep || out[1] <= 0.25 * g.tol2 * g.bb))) {
      g.pause = 1;
    }
  }
}

// Reliable update, part 1 (paused segments): x += x32, x32 = 0.  Grid-stride
// over the block table.
__global__ __launch_bounds__(256) void gmx_fold_kernel(View V, Vecs W, int n_blocks, double* __restrict__ x) {
  for (int bi = blockIdx.x; bi < n_blocks; bi += gridDim.x) {
    const Block B = V.blocks[bi];
    const SegState& g = V.st[B.seg];
    if (!g.pause || g.done) continue;
    const mife_segment S = V.segs[B.seg];
    int r0, r1;
    block_row_range(B, S, r0, r1);
    for (int row = r0 + threadIdx.x; row < r1; row += 256) {
      const long long i = S.row_off + row;
      x[i] += (double)W.x32[i];
      W.x32[i] = 0.0f;
    }
  }
}

// Reliable update, part 2 (paused segments): r64 = b - A64 x with the FP64
// values, r32 = r64, z32 = D^-1 r32; partials r32.z32, r32.r32, r64.r64.
// Grid-stride over the block table.
__global__ __launch_bounds__(256) void gmx_replace_kernel(View V, Vecs W, int n_blocks, const double* __restrict__ b,
                                                          const double* __restrict__ x) {
  for (int bi = blockIdx.x; bi < n_blocks; bi += gridDim.x) {
    const Block B = V.blocks[bi];
    const SegState& g = V.st[B.seg];
    if (!g.pause || g.done) continue;
    const mife_segment S = V.segs[B.seg];
    double m[3] = {0.0, 0.0, 0.0};
    block_rows<double>(V, B, S, V.val, x, [&](int row, double ax) {
      const long long i = S.row_off + row;
      const double r64 = b[i] - ax;
      const float ri = (float)r64;
      const float zi = W.dinv32[i] * ri;
      W.r32[i] = ri;
      W.z32[i] = zi;
      m[0] += (double)ri * (double)zi;
      m[1] += (double)ri * (double)ri;
      m[2] += r64 * r64;
    });
    store_partials<3>(V, B, m);
    __syncthreads();  // the block's LDS partials are reused by its next table entry
  }
}

// Reliable update, part 3 (one block per paused segment): beta from the
// replaced r.z (0 when the replaced r.r exceeds the recurrence's by
// kRestartGrowth), the true residual and the stopping test; resumes the segment.
__global__ __launch_bounds__(256) void gfin_replace_kernel(View V, const int32_t* __restrict__ seg_parts,
                                                           const int32_t* __restrict__ seg_ids) {
  const int s = seg_ids[blockIdx.x];
  if (!V.st[s].pause || V.st[s].done) return;
  double out[3];
  seg_sums<3>(V, seg_parts, blockIdx.x, out);
  if (threadIdx.x == 0) {
    SegState& g = V.st[s];
    g.beta = (g.rz_prev != 0.0 && out[1] <= kRestartGrowth * g.rr) ? out[0] / g.rz_prev : 0.0;
    g.rz = out[0];
    g.rr = out[1];
    g.rr_rep = out[1];
    g.rr_true = out[2];
    g.pause = 0;
    g.done = (out[2] <= g.tol2 * g.bb || g.iters >= g.maxit) ? 1 : 0;
  }
}

// true residual ||b - A x||^2 of the listed segments
__global__ __launch_bounds__(256) void gcg_resid_kernel(View V, const double* __restrict__ b,
                                                        const double* __restrict__ x) {
  const Block B = V.blocks[blockIdx.x];
  const mife_segment S = V.segs[B.seg];
  double m[1] = {0.0};
  block_rows<double>(V, B, S, V.val, x, [&](int row, double ax) {
    const double ri = b[S.row_off + row] - ax;
    m[0] += ri * ri;
  });
  store_partials<1>(V, B, m);
}

__global__ __launch_bounds__(256) void gfin_resid_kernel(View V, const int32_t* __restrict__ seg_parts,
                                                         const int32_t* __restrict__ seg_ids) {
  const int s = seg_ids[blockIdx.x];
  double out[1];
  seg_sums<1>(V, seg_parts, blockIdx.x, out);
  if (threadIdx.x == 0) V.st[s].rr_true = out[0];
}

}  // namespace mife

using namespace mife;
using namespace mife::grouped;

namespace {

// Device layout of a table, carved from the caller's work buffer:
//   mife_segment[max_seg] (by slot) | Block[max_blocks] | seg_parts[max_seg+1] |
//   seg_ids[max_seg] | parts[kNumParts * max_blocks]
// The per-slot CG scalars live in a separate state buffer, so tables can be
// rebuilt (segments joining or leaving) while the other segments keep iterating.
struct Layout {
  mife_segment* segs;
  Block* blocks;
  int32_t* seg_parts;
  int32_t* seg_ids;
  double* parts;
};

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

Layout carve(void* work, int max_seg, int max_blocks) {
  char* p = (char*)work;
  Layout L;
  size_t off = 0;
  L.segs = (mife_segment*)(p + off);
  off = align256(off + sizeof(mife_segment) * max_seg);
  L.blocks = (Block*)(p + off);
  off = align256(off + sizeof(Block) * max_blocks);
  L.seg_parts = (int32_t*)(p + off);
  off = align256(off + sizeof(int32_t) * (max_seg + 1));
  L.seg_ids = (int32_t*)(p + off);
  off = align256(off + sizeof(int32_t) * max_seg);
  L.parts = (double*)(p + off);
  return L;
}

}  // namespace

extern "C" {

size_t mife_grouped_state_bytes(int max_segments) { return sizeof(SegState) * (size_t)max_segments + 256; }

size_t mife_grouped_work_bytes(int max_segments, int max_blocks) {
  size_t b = 0;
  b = align256(b + sizeof(mife_segment) * max_segments);
  b = align256(b + sizeof(Block) * max_blocks);
  b = align256(b + sizeof(int32_t) * (max_segments + 1));
  b = align256(b + sizeof(int32_t) * max_segments);
  b += sizeof(double) * kNumParts * (size_t)max_blocks + 256;
  return b;
}

int mife_grouped_blocks(int n_rows) { return std::max(1, div_up(n_rows, 4 * kRowsPerBlock)); }

int mife_grouped_setup(const mife_segment* segs_host, int n_seg, int max_segments, int max_blocks,
                       int* n_blocks_out, void* table_host, void* work, void* stream) {
  MIFE_REQUIRE(segs_host && work && n_blocks_out && table_host, MIFE_ERR_POINTER_IS_NULL, "null pointer argument");
  MIFE_REQUIRE(n_seg >= 0 && n_seg <= max_segments, MIFE_ERR_ARG, "too many segments (%d > %d)", n_seg,
               max_segments);
  // host staging buffer (caller-owned pinned or pageable memory of
  // mife_grouped_work_bytes) holding segs | blocks | seg_parts | seg_ids
  Layout D = carve(work, max_segments, max_blocks);
  Layout H = carve(table_host, max_segments, max_blocks);
  int nb = 0;
  for (int k = 0; k < n_seg; ++k) {
    const mife_segment& S = segs_host[k];
    MIFE_REQUIRE(S.n >= 0, MIFE_ERR_ARG, "segment %d has %d rows", k, S.n);
    MIFE_REQUIRE(S.slot >= 0 && S.slot < max_segments, MIFE_ERR_ARG, "segment slot %d out of range", S.slot);
    MIFE_REQUIRE(S.format == MIFE_FMT_CSR || (S.format == MIFE_FMT_SELL && S.n_slices == div_up(S.n, kSlice)) ||
                     (S.format == MIFE_FMT_SELL3 && S.n % 3 == 0 && S.n_slices == div_up(S.n / 3, kSlice) &&
                      S.cb_off >= 0),
                 MIFE_ERR_ARG, "segment %d: format %d / n_slices %d", k, S.format, S.n_slices);
    MIFE_REQUIRE(S.prec == MIFE_F64 || S.prec == MIFE_F32, MIFE_ERR_ARG, "segment %d: precision %d", k, S.prec);
    H.segs[S.slot] = S;
    H.seg_parts[k] = nb;
    H.seg_ids[k] = S.slot;
    const bool sell = S.format != MIFE_FMT_CSR;
    const int per = sell ? kSlicesPerBlock : 4 * kRowsPerBlock;
    const int units = sell ? S.n_slices : S.n;
    int r0 = 0;
    do {
      MIFE_REQUIRE(nb < max_blocks, MIFE_ERR_ARG, "block table full (%d blocks)", max_blocks);
      H.blocks[nb] = Block{S.slot, r0, std::max(0, std::min(per, units - r0)), nb};
      ++nb;
      r0 += per;
    } while (r0 < units);
  }
  H.seg_parts[n_seg] = nb;
  hipStream_t st = (hipStream_t)stream;
  if (n_seg > 0) {
    MIFE_CHECK_HIP(hipMemcpyAsync(D.segs, H.segs, sizeof(mife_segment) * max_segments, hipMemcpyHostToDevice, st));
    MIFE_CHECK_HIP(hipMemcpyAsync(D.blocks, H.blocks, sizeof(Block) * nb, hipMemcpyHostToDevice, st));
    MIFE_CHECK_HIP(
        hipMemcpyAsync(D.seg_parts, H.seg_parts, sizeof(int32_t) * (n_seg + 1), hipMemcpyHostToDevice, st));
    MIFE_CHECK_HIP(hipMemcpyAsync(D.seg_ids, H.seg_ids, sizeof(int32_t) * n_seg, hipMemcpyHostToDevice, st));
    MIFE_CHECK_HIP(hipStreamSynchronize(st));  // the host staging buffer may be reused at once
  }
  *n_blocks_out = nb;
  return MIFE_OK;
}

static View make_view(const Layout& D, void* state, const mife_pools* P, int cap) {
  return View{D.segs, D.blocks, P->rp, P->col, P->col16, P->cbase, P->val, P->val32, (SegState*)state, D.parts, cap};
}

static Vecs make_vecs(double* vec5, float* vec6, int64_t rows) {
  return Vecs{vec5, vec5 + rows, vec5 + 2 * rows, vec5 + 3 * rows, vec5 + 4 * rows,
              vec6, vec6 + rows, vec6 + 2 * rows, vec6 + 3 * rows, vec6 + 4 * rows, vec6 + 5 * rows};
}

int mife_grouped_cg_init(int n_seg, int n_blocks, int max_segments, int max_blocks, const mife_pools* pools,
                         const double* b, const double* x, double* vec5, float* vec6, int64_t pool_rows, void* work,
                         void* state, void* stream) {
  MIFE_REQUIRE(pools && pools->rp && pools->val && pools->val32 && b && x && vec5 && vec6 && work && state,
               MIFE_ERR_POINTER_IS_NULL, "null pointer argument");
  if (n_seg == 0) return MIFE_OK;
  hipStream_t s = (hipStream_t)stream;
  Layout D = carve(work, max_segments, max_blocks);
  View V = make_view(D, state, pools, max_blocks);
  Vecs W = make_vecs(vec5, vec6, pool_rows);
  gcg_diag_kernel<<<n_blocks, 256, 0, s>>>(V, W);
  gcg_init_kernel<<<n_blocks, 256, 0, s>>>(V, W, b, x);
  gfin_init_kernel<<<n_seg, 256, 0, s>>>(V, D.seg_parts, D.seg_ids);
  MIFE_CHECK_HIP(hipGetLastError());
  return MIFE_OK;
}

int mife_grouped_cg_iterate(int n_seg, int n_blocks, int max_segments, int max_blocks, const mife_pools* pools,
                            const double* b, double* x, double* vec5, float* vec6, int64_t pool_rows,
                            int iterations, int check_every, double inner_rtol, void* work, void* state,
                            void* stream) {
  MIFE_REQUIRE(pools && pools->rp && pools->val && pools->val32 && b && x && vec5 && vec6 && work && state,
               MIFE_ERR_POINTER_IS_NULL, "null pointer argument");
  MIFE_REQUIRE(check_every >= 1 && inner_rtol > 0.0 && inner_rtol < 1.0, MIFE_ERR_ARG,
               "invalid check_every %d / inner_rtol %g", check_every, inner_rtol);
  if (n_seg == 0) return MIFE_OK;
  hipStream_t s = (hipStream_t)stream;
  Layout D = carve(work, max_segments, max_blocks);
  View V = make_view(D, state, pools, max_blocks);
  Vecs W = make_vecs(vec5, vec6, pool_rows);
  const double delta2 = inner_rtol * inner_rtol;
  const int rg = std::min(n_blocks, kReplaceGrid);
  for (int it = 0; it < iterations; ++it) {
    const int tk = prof::begin(s);
    gcg_spmv_kernel<<<n_blocks, 256, 0, s>>>(V, W);
    prof::end(tk, s, prof::kGroupedSpmv);
    gfin_alpha_kernel<<<n_seg, 256, 0, s>>>(V, D.seg_parts, D.seg_ids);
    const int tk2 = prof::begin(s);
    gcg_update_kernel<<<n_blocks, 256, 0, s>>>(V, W, x);
    prof::end(tk2, s, prof::kGroupedUpdate);
    if ((it + 1) % check_every == 0 || it + 1 == iterations) {
      gfin_beta_kernel<true><<<n_seg, 256, 0, s>>>(V, D.seg_parts, D.seg_ids, delta2);
      gmx_fold_kernel<<<rg, 256, 0, s>>>(V, W, n_blocks, x);
      const int tk3 = prof::begin(s);
      gmx_replace_kernel<<<rg, 256, 0, s>>>(V, W, n_blocks, b, x);
      prof::end(tk3, s, prof::kGroupedReplace);
      gfin_replace_kernel<<<n_seg, 256, 0, s>>>(V, D.seg_parts, D.seg_ids);
    } else {
      gfin_beta_kernel<false><<<n_seg, 256, 0, s>>>(V, D.seg_parts, D.seg_ids, delta2);
    }
  }
  MIFE_CHECK_HIP(hipGetLastError());
  return MIFE_OK;
}

int mife_grouped_cg_residual(int n_seg, int n_blocks, int max_segments, int max_blocks, const mife_pools* pools,
                             const double* b, const double* x, void* work, void* state, void* stream) {
  MIFE_REQUIRE(pools && pools->rp && pools->val && b && x && work && state, MIFE_ERR_POINTER_IS_NULL,
               "null pointer argument");
  if (n_seg == 0) return MIFE_OK;
  hipStream_t s = (hipStream_t)stream;
  Layout D = carve(work, max_segments, max_blocks);
  View V = make_view(D, state, pools, max_blocks);
  gcg_resid_kernel<<<n_blocks, 256, 0, s>>>(V, b, x);
  gfin_resid_kernel<<<n_seg, 256, 0, s>>>(V, D.seg_parts, D.seg_ids);
  MIFE_CHECK_HIP(hipGetLastError());
  return MIFE_OK;
}

int mife_grouped_cg_status(int max_segments, const void* state, int* iters, int* done, double* relres,
                           void* stream) {
  MIFE_REQUIRE(state && iters && done && relres, MIFE_ERR_POINTER_IS_NULL, "null pointer argument");
  hipStream_t s = (hipStream_t)stream;
  std::vector<SegState> h(max_segments);
  MIFE_CHECK_HIP(hipMemcpyAsync(h.data(), state, sizeof(SegState) * max_segments, hipMemcpyDeviceToHost, s));
  MIFE_CHECK_HIP(hipStreamSynchronize(s));
  for (int i = 0; i < max_segments; ++i) {
    iters[i] = h[i].iters;
    done[i] = h[i].done;
    const double rr = h[i].rr_true >= 0.0 ? h[i].rr_true : h[i].rr;
    relres[i] = h[i].bb > 0.0 ? sqrt(rr / h[i].bb) : 0.0;
  }
  return MIFE_OK;
}

}  // extern "C"
