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

struct MatArgs {
  int n;
  int fmt;  // MIFE_FMT_CSR / MIFE_FMT_SELL
  const int32_t* ptr;
  const int32_t* col;
  int n_slices;
};

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

// CSR partial sum of `row` held by lane `sub` of its L-lane group (not reduced).
template <typename T, int L>
__device__ __forceinline__ T row_partial(const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                         const T* __restrict__ vals, const T* __restrict__ x, int row, int sub) {
  const int k0 = rp[row], k1 = rp[row + 1];
  T a0 = 0, a1 = 0;
  int k = k0 + sub;
  for (; k + 3 * L < k1; k += 4 * L) {
    const int c0 = stream_load(&col[k]), c1 = stream_load(&col[k + L]);
    const int c2 = stream_load(&col[k + 2 * L]), c3 = stream_load(&col[k + 3 * L]);
    const T v0 = stream_load(&vals[k]), v1 = stream_load(&vals[k + L]);
    const T v2 = stream_load(&vals[k + 2 * L]), v3 = stream_load(&vals[k + 3 * L]);
    a0 += v0 * x[c0];
    a1 += v1 * x[c1];
    a0 += v2 * x[c2];
    a1 += v3 * x[c3];
  }
  for (; k < k1; k += L) a0 += stream_load(&vals[k]) * x[stream_load(&col[k])];
  return a0 + a1;
}

// SELL-64 row sum of slice `s`, lane `lane` (one row per lane, complete sum).
template <typename T>
__device__ __forceinline__ T sell_row(const int32_t* __restrict__ sp, const int32_t* __restrict__ col,
                                      const T* __restrict__ vals, const T* __restrict__ x, int s, int lane) {
  const int base = sp[s];
  const int len = (sp[s + 1] - base) / kSlice;
  const int32_t* c = col + base + lane;
  const T* v = vals + base + lane;
  T a0 = 0, a1 = 0;
  int j = 0;
  for (; j + 3 < len; j += 4) {
    const int c0 = stream_load(c), c1 = stream_load(c + kSlice), c2 = stream_load(c + 2 * kSlice),
              c3 = stream_load(c + 3 * kSlice);
    const T v0 = stream_load(v), v1 = stream_load(v + kSlice), v2 = stream_load(v + 2 * kSlice),
            v3 = stream_load(v + 3 * kSlice);
    a0 += v0 * x[c0];
    a1 += v1 * x[c1];
    a0 += v2 * x[c2];
    a1 += v3 * x[c3];
    c += 4 * kSlice;
    v += 4 * kSlice;
  }
  for (; j < len; ++j) {
    a0 += stream_load(v) * x[stream_load(c)];
    c += kSlice;
    v += kSlice;
  }
  return a0 + a1;
}

// Calls f(row, (A x)_row) once per row (by one lane), over the whole grid.
template <typename T, int FMT, int L, typename F>
__device__ __forceinline__ void for_each_row(const MatArgs& A, const T* __restrict__ vals, const T* __restrict__ x,
                                             F&& f) {
  if constexpr (FMT == MIFE_FMT_SELL) {
    const int wpb = blockDim.x >> 6, lane = threadIdx.x & 63;
    for (int s = blockIdx.x * wpb + (threadIdx.x >> 6); s < A.n_slices; s += gridDim.x * wpb) {
      const T ax = x ? sell_row<T>(A.ptr, A.col, vals, x, s, lane) : (T)0;
      const int row = s * kSlice + lane;
      if (row < A.n) f(row, ax);
    }
  } else {
    const int gpb = blockDim.x / L, sub = threadIdx.x % L;
    for (int row = blockIdx.x * gpb + threadIdx.x / L; row < A.n; row += gridDim.x * gpb) {
      T ax = 0;
      if (x) ax = group_sum<T, L>(row_partial<T, L>(A.ptr, A.col, vals, x, row, sub));
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

inline int row_grid(const MatArgs& A, int L) {
  const int units = A.fmt == MIFE_FMT_SELL ? div_up(A.n_slices, 4) : div_up(A.n, 256 / L);
  return std::max(1, std::min(units, kMaxGrid));
}

// Dispatch helper: runs STMT with constexpr FMT_ and L_ bound.
#define MIFE_DISPATCH_FMT(A, L, STMT)                                            \
  do {                                                                          \
    if ((A).fmt == MIFE_FMT_SELL) {                                             \
      constexpr int FMT_ = MIFE_FMT_SELL, L_ = 64;                              \
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
