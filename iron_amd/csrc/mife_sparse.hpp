// Sparse matrix row loops shared by the solver kernels.
//
// Two storage formats of a solver matrix:
//  * CSR: L consecutive lanes share a row (L in {4,8,16,32,64}); each lane
//    walks the row with stride L, four entries per trip.
//  * SELL-64: rows are grouped in slices of 64 (one wavefront, one row per
//    lane, rows sorted by length inside windows so a slice is nearly
//    rectangular); entry j of the 64 rows of a slice is stored contiguously,
//    so every value/column load of the wavefront is one coalesced 256-byte
//    access and no row pointer is read per row.  Column indices are in the
//    permuted (slice) numbering, so the CG vectors live in that numbering too.
// Values and columns are read once per product and streamed past the caches
// with non-temporal loads; the gathered vector stays cache-resident.
#pragma once
#include "mife_common.hpp"
#include <algorithm>

namespace mife {

constexpr int kSlice = 64;
constexpr int kMaxGrid = 2048;  // resident grid: 8 blocks of 256 threads per CU on 256 CUs
// Largest grid of a scheduled SELL product (one slice per wavefront, 4 per block).
constexpr int kMaxSellGrid = 8192;
constexpr int kXcds = 8;

// Internal format tag of SELL with 16-bit slice-relative columns.
constexpr int kFmtSell16 = 2;
// Node-block SELL (3x3 blocks, 16-bit slice-relative node columns): the public tag.
constexpr int kFmtSell3 = MIFE_FMT_SELL3;

// Reliable updates: a replaced residual whose r.r exceeds the recurrence's by
// more than this factor restarts the recurrence (beta = 0).  Once the FP32
// recurrence has run far below the true residual (an FP32 solve that converged
// early, as on an exactly representable solution), keeping the search
// direction with beta = (r.z)_replaced / (r.z)_before-the-step lets the stale
// direction swamp the new residual and the iteration diverges.
constexpr double kRestartGrowth = 16.0;

// Per-XCD range length of a SELL schedule: ceil(n_slices / 8) rounded up to 4.
inline int sell_order_span(int n_slices) { return 4 * div_up(div_up(n_slices, kXcds), 4); }

struct MatArgs {
  int n;
  int fmt;  // MIFE_FMT_CSR / MIFE_FMT_SELL / kFmtSell16 / kFmtSell3 (n: scalar rows)
  const int32_t* ptr;
  const int32_t* col;  // kFmtSell16, kFmtSell3: uint16 offsets
  int n_slices;
  const int32_t* cbase;  // kFmtSell16, kFmtSell3: per-slice base column (kFmtSell3: node)
  const int32_t* order;  // SELL schedule (8 ranges of `span` slices, -1 padded) or null
  int span;
  int split;  // scheduled SELL: wavefronts per slice (1, 2 or 4)
};

// Shape of a SELL3 descriptor: three rows per node row, ceil(node rows / 64)
// slices, 16-bit node columns.
inline bool sell3_shape_ok(const mife_matrix* A) {
  return A->n % 3 == 0 && A->n_slices == div_up(A->n / 3, kSlice) && A->col_base != nullptr;
}

// Wavefronts per slice of a scheduled SELL product: enough wavefronts to fill
// the chip on small matrices (long rows, few slices), one per slice on large.
inline int sell_split(int n_slices) { return n_slices < 4096 ? 4 : (n_slices < 65536 ? 2 : 1); }

inline MatArgs mat_args(const mife_matrix* A) {
  const bool sell3 = A->format == MIFE_FMT_SELL3;
  const bool sell = A->format == MIFE_FMT_SELL || sell3;
  const bool c16 = sell && A->col_base != nullptr;
  const bool sched = sell && A->slice_order != nullptr && A->order_span > 0;
  return MatArgs{A->n, sell3 ? kFmtSell3 : (c16 ? kFmtSell16 : A->format), A->ptr, A->col, A->n_slices,
                 c16 ? A->col_base : nullptr, sched ? A->slice_order : nullptr, sched ? A->order_span : 0,
                 sched ? sell_split(A->n_slices) : 1};
}

template <typename T>
__device__ __forceinline__ T stream_load(const T* p) {
  return __builtin_nontemporal_load(p);
}

template <typename T, int L>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
  for (int off = L / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// CSR partial sum of entries [k0, k1) held by lane `sub` of its L-lane group
// (not reduced), four entries per trip.
template <typename T, int L, typename TV = T>
__device__ __forceinline__ T row_range_partial(int k0, int k1, const int32_t* __restrict__ col,
                                               const TV* __restrict__ vals, const T* __restrict__ x, int sub) {
  T a0 = 0, a1 = 0;
  int k = k0 + sub;
  for (; k + 3 * L < k1; k += 4 * L) {
    const int c0 = stream_load(&col[k]), c1 = stream_load(&col[k + L]);
    const int c2 = stream_load(&col[k + 2 * L]), c3 = stream_load(&col[k + 3 * L]);
    const T v0 = (T)stream_load(&vals[k]), v1 = (T)stream_load(&vals[k + L]);
    const T v2 = (T)stream_load(&vals[k + 2 * L]), v3 = (T)stream_load(&vals[k + 3 * L]);
    a0 += v0 * x[c0];
    a1 += v1 * x[c1];
    a0 += v2 * x[c2];
    a1 += v3 * x[c3];
  }
  for (; k < k1; k += L) a0 += (T)stream_load(&vals[k]) * x[stream_load(&col[k])];
  return a0 + a1;
}

// CSR partial sum of `row` held by lane `sub` of its L-lane group (not reduced).
template <typename T, int L, typename TV = T>
__device__ __forceinline__ T row_partial(const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                         const TV* __restrict__ vals, const T* __restrict__ x, int row, int sub) {
  return row_range_partial<T, L, TV>(rp[row], rp[row + 1], col, vals, x, sub);
}

// SELL-64 row sum of slice `s`, lane `lane` (one row per lane, complete sum),
// eight entries in flight per lane.
template <typename T, typename TV = T>
__device__ __forceinline__ T sell_row(const int32_t* __restrict__ sp, const int32_t* __restrict__ col,
                                      const TV* __restrict__ vals, const T* __restrict__ x, int s, int lane) {
  const int base = sp[s];
  const int len = (sp[s + 1] - base) / kSlice;
  const int32_t* c = col + base + lane;
  const TV* v = vals + base + lane;
  T a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  int j = 0;
  for (; j + 7 < len; j += 8) {
    int ci[8];
    TV vi[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) ci[u] = stream_load(c + u * kSlice);
#pragma unroll
    for (int u = 0; u < 8; ++u) vi[u] = stream_load(v + u * kSlice);
    a0 += vi[0] * x[ci[0]];
    a1 += vi[1] * x[ci[1]];
    a2 += vi[2] * x[ci[2]];
    a3 += vi[3] * x[ci[3]];
    a0 += vi[4] * x[ci[4]];
    a1 += vi[5] * x[ci[5]];
    a2 += vi[6] * x[ci[6]];
    a3 += vi[7] * x[ci[7]];
    c += 8 * kSlice;
    v += 8 * kSlice;
  }
  for (; j < len; ++j) {
    a0 += (T)stream_load(v) * x[stream_load(c)];
    c += kSlice;
    v += kSlice;
  }
  return (a0 + a1) + (a2 + a3);
}

// As sell_row with 16-bit columns relative to the slice's base column; eight
// entries in flight per lane (the 2-byte columns leave registers for it).
template <typename T, typename TV = T>
__device__ __forceinline__ T sell16_row(const int32_t* __restrict__ sp, const uint16_t* __restrict__ col,
                                        const int32_t* __restrict__ cbase, const TV* __restrict__ vals,
                                        const T* __restrict__ x, int s, int lane) {
  const int base = sp[s];
  const int len = (sp[s + 1] - base) / kSlice;
  const T* xs = x + cbase[s];
  const uint16_t* c = col + base + lane;
  const TV* v = vals + base + lane;
  T a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  int j = 0;
  for (; j + 7 < len; j += 8) {
    int ci[8];
    TV vi[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) ci[u] = stream_load(c + u * kSlice);
#pragma unroll
    for (int u = 0; u < 8; ++u) vi[u] = stream_load(v + u * kSlice);
    a0 += vi[0] * xs[ci[0]];
    a1 += vi[1] * xs[ci[1]];
    a2 += vi[2] * xs[ci[2]];
    a3 += vi[3] * xs[ci[3]];
    a0 += vi[4] * xs[ci[4]];
    a1 += vi[5] * xs[ci[5]];
    a2 += vi[6] * xs[ci[6]];
    a3 += vi[7] * xs[ci[7]];
    c += 8 * kSlice;
    v += 8 * kSlice;
  }
  for (; j < len; ++j) {
    a0 += (T)stream_load(v) * xs[(int)stream_load(c)];
    c += kSlice;
    v += kSlice;
  }
  return (a0 + a1) + (a2 + a3);
}

// Part h of W of a SELL row sum (int32 or 16-bit columns) of the slice whose
// entries start at `base`, `len` per row, base column cb (16-bit columns): the
// wavefront takes chunks h, h + W, h + 2W, ... of eight entries; with W = 1 the
// order of the additions is that of sell_row / sell16_row.
template <typename T, bool C16, typename TV = T>
__device__ __forceinline__ T sell_row_part(const MatArgs& A, const TV* __restrict__ vals, const T* __restrict__ x,
                                           int base, int len, int cb, int lane, int h, int W) {
  constexpr int D = 8;
  const T* xs = C16 ? x + cb : x;
  const uint16_t* c16 = reinterpret_cast<const uint16_t*>(A.col) + base + lane;
  const int32_t* c32 = A.col + base + lane;
  const TV* v = vals + base + lane;
  auto col = [&](int j) -> int {
    if constexpr (C16) return (int)stream_load(c16 + j * kSlice);
    else return stream_load(c32 + j * kSlice);
  };
  T a0 = 0, a1 = 0, a2 = 0, a3 = 0;
  int j = h * D;
  for (; j + D <= len; j += W * D) {
    int ci[D];
    TV vi[D];
#pragma unroll
    for (int u = 0; u < D; ++u) ci[u] = col(j + u);
#pragma unroll
    for (int u = 0; u < D; ++u) vi[u] = stream_load(v + (j + u) * kSlice);
    a0 += vi[0] * xs[ci[0]];
    a1 += vi[1] * xs[ci[1]];
    a2 += vi[2] * xs[ci[2]];
    a3 += vi[3] * xs[ci[3]];
    a0 += vi[4] * xs[ci[4]];
    a1 += vi[5] * xs[ci[5]];
    a2 += vi[6] * xs[ci[6]];
    a3 += vi[7] * xs[ci[7]];
  }
  for (int k = j; k < len && k < j + D; ++k) a0 += (T)stream_load(v + k * kSlice) * xs[col(k)];
  return (a0 + a1) + (a2 + a3);
}

// Part h of W of the three row sums of a SELL3 node row: c16 and v point at
// the row's first block column and first value (lane offset included; block j
// value t at v[(9 j + t) * 64]), xs at x of the slice's base node.  The
// wavefront takes chunks h, h + W, ... of four blocks.  A block adds to its
// rows the dot of its row with the three x values of its node column; blocks
// alternate between two partial sums per row, added at the end.
template <typename T, typename TV = T>
__device__ __forceinline__ void sell3_sum(const uint16_t* __restrict__ c16, const TV* __restrict__ v,
                                          const T* __restrict__ xs, int len, int h, int W, T (&acc)[3]) {
  constexpr int D = 4;
  T a[2][3] = {{0, 0, 0}, {0, 0, 0}};
  auto add = [&](const TV (&vb)[9], int c, T (&ar)[3]) {
    const T* xp = xs + 3 * c;
    const T x0 = xp[0], x1 = xp[1], x2 = xp[2];
#pragma unroll
    for (int r = 0; r < 3; ++r) ar[r] += ((T)vb[3 * r] * x0 + (T)vb[3 * r + 1] * x1) + (T)vb[3 * r + 2] * x2;
  };
  int j = h * D;
  for (; j + D <= len; j += W * D) {
    int ci[D];
    TV vi[D][9];
#pragma unroll
    for (int u = 0; u < D; ++u) ci[u] = (int)stream_load(c16 + (j + u) * kSlice);
#pragma unroll
    for (int u = 0; u < D; ++u)
#pragma unroll
      for (int t = 0; t < 9; ++t) vi[u][t] = stream_load(v + (size_t)(9 * (j + u) + t) * kSlice);
#pragma unroll
    for (int u = 0; u < D; ++u) add(vi[u], ci[u], a[u & 1]);
  }
  for (int k = j; k < len && k < j + D; ++k) {
    TV vb[9];
#pragma unroll
    for (int t = 0; t < 9; ++t) vb[t] = stream_load(v + (size_t)(9 * k + t) * kSlice);
    add(vb, (int)stream_load(c16 + k * kSlice), a[(k - j) & 1]);
  }
#pragma unroll
  for (int r = 0; r < 3; ++r) acc[r] = a[0][r] + a[1][r];
}

// sell3_sum for lane `lane` of the slice whose blocks start at `base` (`len`
// per node row, base node cb) of matrix A.
template <typename T, typename TV = T>
__device__ __forceinline__ void sell3_row_part(const MatArgs& A, const TV* __restrict__ vals,
                                               const T* __restrict__ x, int base, int len, int cb, int lane, int h,
                                               int W, T (&acc)[3]) {
  sell3_sum<T, TV>(reinterpret_cast<const uint16_t*>(A.col) + base + lane, vals + (size_t)9 * base + lane,
                   x + 3 * (size_t)cb, len, h, W, acc);
}

// SELL3 form of for_each_row: f(row, (A x)_row) for the three rows 3k + c of
// each node row k.
template <typename T, typename TV, typename F>
__device__ __forceinline__ void for_each_node_row(const MatArgs& A, const TV* __restrict__ vals,
                                                  const T* __restrict__ x, F&& f) {
  const int wpb = blockDim.x >> 6, lane = threadIdx.x & 63;
  auto emit = [&](int s, const T (&ax)[3]) {
    const int row = 3 * (s * kSlice + lane);
#pragma unroll
    for (int c = 0; c < 3; ++c)
      if (row + c < A.n) f(row + c, ax[c]);
  };
  if (A.order) {
    // scheduled, as the scalar SELL path: W wavefronts share a slice and their
    // partial sums meet in LDS, added in a fixed order by the first of them
    __shared__ T part[3][16][kSlice];
    const int W = min(A.split, wpb), spb = wpb / W;
    const int wave = threadIdx.x >> 6, g = wave / W, h = wave - g * W;
    const int xcd = blockIdx.x % kXcds, bpx = gridDim.x / kXcds;
    const int4* ord = reinterpret_cast<const int4*>(A.order) + (size_t)xcd * A.span;
    const int passes = (A.span + bpx * spb - 1) / (bpx * spb);
    for (int ps = 0; ps < passes; ++ps) {
      const int k = (ps * bpx + blockIdx.x / kXcds) * spb + g;
      int4 slot = make_int4(-1, 0, 0, 0);
      if (k < A.span) slot = ord[k];
      const int s = slot.x;
      T ax[3] = {(T)0, (T)0, (T)0};
      if (s >= 0 && x) sell3_row_part<T, TV>(A, vals, x, slot.y, slot.z, slot.w, lane, W == 1 ? 0 : h, W, ax);
      if (W == 1) {
        if (s >= 0) emit(s, ax);
        continue;
      }
#pragma unroll
      for (int c = 0; c < 3; ++c) part[c][wave][lane] = ax[c];
      __syncthreads();
      if (h == 0 && s >= 0) {
        T t[3];
#pragma unroll
        for (int c = 0; c < 3; ++c) {
          t[c] = part[c][wave][lane];
          for (int q = 1; q < W; ++q) t[c] += part[c][wave + q][lane];
        }
        emit(s, t);
      }
      __syncthreads();
    }
  } else {
    for (int s = blockIdx.x * wpb + (threadIdx.x >> 6); s < A.n_slices; s += gridDim.x * wpb) {
      T ax[3] = {(T)0, (T)0, (T)0};
      if (x) {
        const int base = A.ptr[s];
        sell3_row_part<T, TV>(A, vals, x, base, (A.ptr[s + 1] - base) / kSlice, A.cbase[s], lane, 0, 1, ax);
      }
      emit(s, ax);
    }
  }
}

// Calls f(row, (A x)_row) once per row (by one lane), over the whole grid.
template <typename T, int FMT, int L, typename TV = T, typename F>
__device__ __forceinline__ void for_each_row(const MatArgs& A, const TV* __restrict__ vals, const T* __restrict__ x,
                                             F&& f) {
  if constexpr (FMT == kFmtSell3) {
    for_each_node_row<T, TV>(A, vals, x, f);
  } else if constexpr (FMT == MIFE_FMT_SELL || FMT == kFmtSell16) {
    const int wpb = blockDim.x >> 6, lane = threadIdx.x & 63;
    auto one_slice = [&](int s) {
      T ax = (T)0;
      if (x) {
        if constexpr (FMT == kFmtSell16)
          ax = sell16_row<T, TV>(A.ptr, reinterpret_cast<const uint16_t*>(A.col), A.cbase, vals, x, s, lane);
        else
          ax = sell_row<T, TV>(A.ptr, A.col, vals, x, s, lane);
      }
      const int row = s * kSlice + lane;
      if (row < A.n) f(row, ax);
    };
    if (A.order) {
      // Scheduled: the dispatcher places block b on XCD b % 8, which walks its own
      // range of the order (longest slices first); the grid is a multiple of 8.
      // A slot gives the slice, its first entry, its length and its base column in
      // one 16-byte load.  W wavefronts share a slice (W = 1: one slice per
      // wavefront); their partial row sums meet in LDS and the first of them adds
      // them in a fixed order.  Every wavefront of a block makes the same number of
      // passes, so the barriers are uniform.
      __shared__ T part[16][kSlice];
      const int W = min(A.split, wpb), spb = wpb / W;
      const int wave = threadIdx.x >> 6, g = wave / W, h = wave - g * W;
      const int xcd = blockIdx.x % kXcds, bpx = gridDim.x / kXcds;
      const int4* ord = reinterpret_cast<const int4*>(A.order) + (size_t)xcd * A.span;
      const int passes = (A.span + bpx * spb - 1) / (bpx * spb);
      for (int ps = 0; ps < passes; ++ps) {
        const int k = (ps * bpx + blockIdx.x / kXcds) * spb + g;
        int4 slot = make_int4(-1, 0, 0, 0);
        if (k < A.span) slot = ord[k];
        const int s = slot.x;
        if (W == 1) {
          if (s >= 0) {
            T ax = (T)0;
            if (x) ax = sell_row_part<T, FMT == kFmtSell16, TV>(A, vals, x, slot.y, slot.z, slot.w, lane, 0, 1);
            const int row = s * kSlice + lane;
            if (row < A.n) f(row, ax);
          }
          continue;
        }
        T ax = (T)0;
        if (s >= 0 && x)
          ax = sell_row_part<T, FMT == kFmtSell16, TV>(A, vals, x, slot.y, slot.z, slot.w, lane, h, W);
        part[wave][lane] = ax;
        __syncthreads();
        if (h == 0 && s >= 0) {
          T t = part[wave][lane];
          for (int q = 1; q < W; ++q) t += part[wave + q][lane];
          const int row = s * kSlice + lane;
          if (row < A.n) f(row, t);
        }
        __syncthreads();
      }
    } else {
      for (int s = blockIdx.x * wpb + (threadIdx.x >> 6); s < A.n_slices; s += gridDim.x * wpb) one_slice(s);
    }
  } else {
    const int gpb = blockDim.x / L, sub = threadIdx.x % L;
    for (int row = blockIdx.x * gpb + threadIdx.x / L; row < A.n; row += gridDim.x * gpb) {
      T ax = 0;
      if (x) ax = group_sum<T, L>(row_partial<T, L, TV>(A.ptr, A.col, vals, x, row, sub));
      if (sub == 0) f(row, ax);
    }
  }
}

// Lanes per CSR row from the mean row length.
inline int spmv_lanes(long long nnz, int n) {
  const double avg = n > 0 ? (double)nnz / n : 0.0;
  if (avg <= 8) return 4;
  if (avg <= 24) return 8;
  if (avg <= 96) return 16;
  if (avg <= 192) return 32;
  return 64;
}

// Grid of a row loop launched with 256 threads per block.  A scheduled SELL
// matrix gets 4 / split slices per block, a multiple of 8 blocks.
inline int row_grid(const MatArgs& A, int L) {
  if (A.fmt != MIFE_FMT_CSR && A.order)
    return kXcds * std::max(1, std::min(div_up(A.span, 4 / A.split), kMaxSellGrid / kXcds));
  const int units = A.fmt != MIFE_FMT_CSR ? div_up(A.n_slices, 4) : div_up(A.n, 256 / L);
  return std::max(1, std::min(units, kMaxGrid));
}

// Dispatch helper: runs STMT with constexpr FMT_ and L_ bound.
#define MIFE_DISPATCH_FMT(A, L, STMT)                                            \
  do {                                                                          \
    if ((A).fmt == MIFE_FMT_SELL) {                                             \
      constexpr int FMT_ = MIFE_FMT_SELL, L_ = 64;                              \
      STMT;                                                                     \
    } else if ((A).fmt == kFmtSell16) {                                         \
      constexpr int FMT_ = kFmtSell16, L_ = 64;                                 \
      STMT;                                                                     \
    } else if ((A).fmt == kFmtSell3) {                                          \
      constexpr int FMT_ = kFmtSell3, L_ = 64;                                  \
      STMT;                                                                     \
    } else {                                                                    \
      switch (L) {                                                              \
        case 4: { constexpr int FMT_ = MIFE_FMT_CSR, L_ = 4; STMT; } break;     \
        case 8: { constexpr int FMT_ = MIFE_FMT_CSR, L_ = 8; STMT; } break;     \
        case 16: { constexpr int FMT_ = MIFE_FMT_CSR, L_ = 16; STMT; } break;   \
        case 32: { constexpr int FMT_ = MIFE_FMT_CSR, L_ = 32; STMT; } break;   \
        default: { constexpr int FMT_ = MIFE_FMT_CSR, L_ = 64; STMT; } break;   \
      }                                                                         \
    }                                                                           \
  } while (0)

template <typename T, int FMT, int L>
__global__ __launch_bounds__(256) void spmv_kernel(MatArgs A, const T* __restrict__ vals, const T* __restrict__ x,
                                                   T* __restrict__ y) {
  for_each_row<T, FMT, L>(A, vals, x, [&](int row, T ax) { y[row] = ax; });
}

template <typename T>
inline void launch_spmv(const MatArgs& A, const T* vals, const T* x, T* y, int L, hipStream_t s) {
  MIFE_DISPATCH_FMT(A, L, (spmv_kernel<T, FMT_, L_><<<row_grid(A, L_), 256, 0, s>>>(A, vals, x, y)));
}

}  // namespace mife
