// CSR sparse operations on the solve path: SpMV, solver-matrix gather,
// Dirichlet lifting into the solver RHS, field update and back-substitution.
#include "mife_common.hpp"
#include "mife_sparse.hpp"

#include <algorithm>
#include <vector>

namespace mife {

// Each block copies one contiguous chunk of kGatherChunk entries.  In SELL order
// entries k and k + 64 are consecutive entries of the same 64 CSR rows, so the
// CSR lines a chunk reads are reused by the same block (one CU, one L2); a grid
// stride over the whole array spreads those reuses over the XCDs and fetches each
// line several times.
constexpr int kGatherChunk = 4096;

template <typename S, typename D>
__global__ __launch_bounds__(256) void gather_kernel(long long nnz, const int32_t* __restrict__ src_pos,
                                                     const S* __restrict__ src, D* __restrict__ dst) {
  const long long k0 = (long long)blockIdx.x * kGatherChunk;
  const long long k1 = k0 + kGatherChunk < nnz ? k0 + kGatherChunk : nnz;
  for (long long k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
    const int p = src_pos[k];
    dst[k] = p >= 0 ? (D)src[p] : (D)0;
  }
}

// Both precisions in one pass: dst64[k] = src[src_pos[k]], dst32[k] = (float)dst64[k].
__global__ __launch_bounds__(256) void gather2_kernel(long long nnz, const int32_t* __restrict__ src_pos,
                                                      const double* __restrict__ src, double* __restrict__ dst64,
                                                      float* __restrict__ dst32) {
  const long long k0 = (long long)blockIdx.x * kGatherChunk;
  const long long k1 = k0 + kGatherChunk < nnz ? k0 + kGatherChunk : nnz;
  for (long long k = k0 + threadIdx.x; k < k1; k += blockDim.x) {
    const int p = src_pos[k];
    const double v = p >= 0 ? src[p] : 0.0;
    dst64[k] = v;
    dst32[k] = (float)v;
  }
}

// One wavefront per slice, one row per lane: entry j of slice row k.
__global__ void sell_build_kernel(int n, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                  const int32_t* __restrict__ src_pos, const int32_t* __restrict__ perm,
                                  const int32_t* __restrict__ iperm, int n_slices,
                                  const int32_t* __restrict__ slice_ptr, int32_t* __restrict__ sell_col,
                                  int32_t* __restrict__ sell_src) {
  const int s = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (s >= n_slices) return;
  const int k = s * kSlice + lane;
  const int base = slice_ptr[s];
  const int len = (slice_ptr[s + 1] - base) / kSlice;
  int r0 = 0, rl = 0;
  if (k < n) {
    const int row = perm[k];
    r0 = rp[row];
    rl = rp[row + 1] - r0;
  }
  for (int j = 0; j < len; ++j) {
    const int e = base + j * kSlice + lane;
    if (j < rl) {
      const int p = r0 + j;
      sell_col[e] = iperm[col[p]];
      sell_src[e] = src_pos ? src_pos[p] : p;
    } else {
      sell_col[e] = k < n ? k : n - 1;  // a column inside the slice's range
      sell_src[e] = -1;
    }
  }
}

// One wavefront per free row: b[f] = src + rhs - sum over fixed columns of K_ij u_j,
// skipping |u_j| < ZT or |K_ij| < ZT as SOLVER_MATRICES_STATIC_ASSEMBLE does.
// 64 free rows per wavefront, one per lane: each lane forms src + rhs of its row;
// then the wavefront walks the rows that take a lifting term (a ballot of the
// lanes), all 64 lanes over each such row's entries.  The sums are those of one
// wavefront per row, in the same order.
__global__ void solver_rhs_kernel(int n_free, const int32_t* __restrict__ free_idx,
                                  const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
                                  const double* __restrict__ vals, const int32_t* __restrict__ new_idx,
                                  const double* __restrict__ u, const double* __restrict__ src,
                                  const double* __restrict__ rhs, const uint8_t* __restrict__ rhs_fixed,
                                  double* __restrict__ b) {
  const double ZT = 5.0 * 2.220446049250313e-16;
  const int lane = threadIdx.x & 63;
  const int f = (blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 64 + lane;
  int i = -1;
  double v = 0.0, mine = 0.0;
  bool need = false;
  if (f < n_free) {
    i = free_idx[f];
    if (src) v += 1.0 * src[i];
    if (rhs) v += rhs[i];
    need = !(rhs_fixed && rhs_fixed[i]);
  }
  unsigned long long mask = __ballot(need);
  while (mask) {
    const int l = __ffsll((long long)mask) - 1;
    mask &= mask - 1;
    const int il = __shfl(i, l, 64);
    double lift = 0.0;
    for (int k = row_ptr[il] + lane; k < row_ptr[il + 1]; k += 64) {
      const int j = col[k];
      if (new_idx[j] < 0) {
        const double uj = u[j], kv = vals[k];
        if (fabs(uj) >= ZT && fabs(kv) >= ZT) lift += -1.0 * kv * uj;
      }
    }
    lift = wave_sum(lift);
    if (lane == l) mine = lift;
  }
  if (f < n_free) b[f] = v + mine;
}

__global__ void field_update_kernel(int n_free, const int32_t* __restrict__ free_idx, const double* __restrict__ x,
                                    double* __restrict__ u) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f < n_free) u[free_idx[f]] = x[f];
}

// R_i = sum_k K_ik u_j - s_i for the rows whose RHS DOF is free (R_i = -s_i for
// the others).  kBacksubLanes lanes share a row, each summing a strided part of
// it four entries per trip (the SpMV's row_partial); the parts meet by shuffles
// inside the lane group.  Rows are dealt over a resident grid.
constexpr int kBacksubLanes = 16;

__global__ __launch_bounds__(256) void backsub_kernel(int n, const int32_t* __restrict__ row_ptr,
                                                      const int32_t* __restrict__ col,
                                                      const double* __restrict__ vals, const double* __restrict__ u,
                                                      const double* __restrict__ src,
                                                      const uint8_t* __restrict__ rhs_fixed, int iron_compat,
                                                      double* __restrict__ R) {
  constexpr int L = kBacksubLanes;
  const int gpb = blockDim.x / L, sub = threadIdx.x % L;
  for (int i = blockIdx.x * gpb + threadIdx.x / L; i < n; i += gridDim.x * gpb) {
    // the row bounds, the fixed flag and the source are independent loads; a
    // fixed row becomes an empty one
    const int k0 = row_ptr[i];
    int k1 = row_ptr[i + 1];
    if (rhs_fixed && rhs_fixed[i]) k1 = k0;
    double acc = 0.0;
    if (iron_compat) {
      // equations_set_routines.F90:1806 reads the dependent value at the position
      // within the row; the correct product reads it at the column.
      for (int k = k0 + sub; k < k1; k += L) acc += vals[k] * u[k - k0];
    } else {
      acc = row_range_partial<double, L>(k0, k1, col, vals, u, sub);
    }
    acc = group_sum<double, L>(acc);
    if (sub == 0) R[i] = acc - (src ? src[i] : 0.0);
  }
}

// One wavefront per slice: base = min column of the slice, 16-bit offsets, and
// the slice's span (max - min) into span[s].
__global__ void sell_compress_kernel(int n_slices, const int32_t* __restrict__ slice_ptr,
                                     const int32_t* __restrict__ sell_col, int32_t* __restrict__ col_base,
                                     uint16_t* __restrict__ col16, int* __restrict__ span) {
  const int s = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (s >= n_slices) return;
  const int base = slice_ptr[s], end = slice_ptr[s + 1];
  int lo = 0x7fffffff, hi = -1;
  for (int e = base + lane; e < end; e += kSlice) {
    const int c = sell_col[e];
    lo = min(lo, c);
    hi = max(hi, c);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    lo = min(lo, __shfl_xor(lo, off, 64));
    hi = max(hi, __shfl_xor(hi, off, 64));
  }
  if (end == base) lo = hi = 0;
  if (lane == 0) {
    col_base[s] = lo;
    span[s] = hi - lo;
  }
  for (int e = base + lane; e < end; e += kSlice) col16[e] = (uint16_t)min(sell_col[e] - lo, 65535);
}

}  // namespace mife

using namespace mife;

extern "C" {

int mife_spmv(const mife_matrix* A, const void* vals, const void* x, void* y, int prec, void* stream) {
  MIFE_REQUIRE(A && A->ptr && A->col && vals && x && y, MIFE_ERR_POINTER_IS_NULL, "null pointer argument");
  MIFE_REQUIRE(prec == MIFE_F64 || prec == MIFE_F32, MIFE_ERR_ARG, "invalid precision %d", prec);
  MIFE_REQUIRE(A->format == MIFE_FMT_CSR || A->format == MIFE_FMT_SELL || A->format == MIFE_FMT_SELL3, MIFE_ERR_ARG,
               "invalid matrix format");
  if (A->n == 0) return MIFE_OK;
  MIFE_REQUIRE(A->format != MIFE_FMT_SELL3 || sell3_shape_ok(A), MIFE_ERR_ARG,
               "SELL3 matrix needs n = 3 x node rows, n_slices = ceil(n/192) and col_base");
  MIFE_REQUIRE(!A->slice_order || ((uintptr_t)A->slice_order & 15) == 0, MIFE_ERR_ARG,
               "slice_order must be 16-byte aligned");
  hipStream_t s = (hipStream_t)stream;
  const MatArgs M = mat_args(A);
  const int L = spmv_lanes(A->nnz, A->n);
  if (prec == MIFE_F64)
    launch_spmv<double>(M, (const double*)vals, (const double*)x, (double*)y, L, s);
  else
    launch_spmv<float>(M, (const float*)vals, (const float*)x, (float*)y, L, s);
  MIFE_CHECK_HIP(hipGetLastError());
  return MIFE_OK;
}

}  // extern "C"

namespace mife {

// Row order of a SELL-64 matrix, one window of `sigma` positions per block:
// the window's rows by decreasing length, ties by position (the order a stable
// sort of the window would give).  Each thread counts the rows its own row
// follows, reading the window's lengths from LDS (every lane reads the same
// entry at a time).
__global__ __launch_bounds__(1024) void sell_perm_kernel(int n, const int32_t* __restrict__ row_ptr,
                                                         const int32_t* __restrict__ row_at, int sigma,
                                                         int32_t* __restrict__ perm) {
  __shared__ int len_s[1024];
  const long long w0 = (long long)blockIdx.x * sigma;
  const int t = threadIdx.x;
  const int cnt = (int)min((long long)sigma, (long long)n - w0);
  int r = -1, len = -1;
  if (t < cnt) {
    r = row_at ? row_at[w0 + t] : (int)(w0 + t);
    len = row_ptr[r + 1] - row_ptr[r];
  }
  if (t < sigma) len_s[t] = len;
  __syncthreads();
  if (t < cnt) {
    int rank = 0;
    for (int u = 0; u < cnt; ++u) {
      const int lu = len_s[u];
      rank += (lu > len || (lu == len && u < t)) ? 1 : 0;
    }
    perm[w0 + rank] = r;
  }
}

}  // namespace mife

extern "C" {

int mife_sell_permutation(int n, const int32_t* row_ptr, const int32_t* row_at, int sigma, int32_t* perm,
                          void* stream) {
  MIFE_REQUIRE(n >= 0 && sigma >= 1 && sigma <= 1024, MIFE_ERR_ARG, "n %d / window %d (1..1024)", n, sigma);
  if (n == 0) return MIFE_OK;
  MIFE_REQUIRE(row_ptr && perm, MIFE_ERR_POINTER_IS_NULL, "null pointer argument");
  const int threads = 64 * div_up(sigma, 64);
  sell_perm_kernel<<<div_up(n, sigma), threads, 0, (hipStream_t)stream>>>(n, row_ptr, row_at, sigma, perm);
  MIFE_CHECK_HIP(hipGetLastError());
  return MIFE_OK;
}

int mife_sell_build(int n, const int32_t* row_ptr, const int32_t* col, const int32_t* src_pos, const int32_t* perm,
                    const int32_t* iperm, int n_slices, const int32_t* slice_ptr, int32_t* sell_col,
                    int32_t* sell_src, void* stream) {
  MIFE_REQUIRE(row_ptr && col && perm && iperm && slice_ptr && sell_col && sell_src, MIFE_ERR_POINTER_IS_NULL,
               "null pointer argument");
  MIFE_REQUIRE(n_slices == div_up(n, kSlice), MIFE_ERR_ARG, "n_slices must be ceil(n/64)");
  if (n == 0) return MIFE_OK;
  sell_build_kernel<<<div_up(n_slices, 4), 256, 0, (hipStream_t)stream>>>(n, row_ptr, col, src_pos, perm, iperm,
                                                                          n_slices, slice_ptr, sell_col, sell_src);
  MIFE_CHECK_HIP(hipGetLastError());
  return MIFE_OK;
}

int mife_sell_compress(int n_slices, const int32_t* slice_ptr, const int32_t* sell_col, int32_t* col_base,
                       uint16_t* col16, int* ok, void* stream) {
  MIFE_REQUIRE(slice_ptr && sell_col && col_base && col16 && ok, MIFE_ERR_POINTER_IS_NULL, "null pointer argument");
  *ok = 0;
  if (n_slices == 0) {
    *ok = 1;
    return MIFE_OK;
  }
  hipStream_t s = (hipStream_t)stream;
  int* span = nullptr;
  MIFE_CHECK_HIP(hipMallocAsync((void**)&span, sizeof(int) * (size_t)n_slices, s));
  sell_compress_kernel<<<div_up(n_slices, 4), 256, 0, s>>>(n_slices, slice_ptr, sell_col, col_base, col16, span);
  std::vector<int> h(n_slices);
  MIFE_CHECK_HIP(hipMemcpyAsync(h.data(), span, sizeof(int) * (size_t)n_slices, hipMemcpyDeviceToHost, s));
  MIFE_CHECK_HIP(hipFreeAsync(span, s));
  MIFE_CHECK_HIP(hipStreamSynchronize(s));
  *ok = *std::max_element(h.begin(), h.end()) < 65536 ? 1 : 0;
  return MIFE_OK;
}

int mife_sell_order_len(int n_slices) { return n_slices > 0 ? 4 * 8 * sell_order_span(n_slices) : 0; }

int mife_sell_order(int n_slices, const int32_t* slice_ptr, const int32_t* col_base, int32_t* order,
                    int* order_span, void* stream) {
  MIFE_REQUIRE(slice_ptr && order && order_span, MIFE_ERR_POINTER_IS_NULL, "null pointer argument");
  MIFE_REQUIRE(n_slices >= 0, MIFE_ERR_ARG, "negative slice count");
  *order_span = 0;
  if (n_slices == 0) return MIFE_OK;
  hipStream_t s = (hipStream_t)stream;
  std::vector<int32_t> sp(n_slices + 1), cb(n_slices, 0);
  MIFE_CHECK_HIP(hipMemcpyAsync(sp.data(), slice_ptr, sizeof(int32_t) * sp.size(), hipMemcpyDeviceToHost, s));
  if (col_base)
    MIFE_CHECK_HIP(hipMemcpyAsync(cb.data(), col_base, sizeof(int32_t) * cb.size(), hipMemcpyDeviceToHost, s));
  MIFE_CHECK_HIP(hipStreamSynchronize(s));
  const int range = div_up(n_slices, 8), span = sell_order_span(n_slices);
  std::vector<int32_t> ids((size_t)span);
  std::vector<int32_t> h((size_t)4 * 8 * span, 0);
  for (int k = 0; k < 8; ++k) {
    const int s0 = std::min(k * range, n_slices), s1 = std::min(s0 + range, n_slices);
    for (int i = s0; i < s1; ++i) ids[i - s0] = i;
    std::stable_sort(ids.begin(), ids.begin() + (s1 - s0),
                     [&](int32_t a, int32_t b) { return sp[a + 1] - sp[a] > sp[b + 1] - sp[b]; });
    int32_t* out = h.data() + (size_t)4 * k * span;
    for (int j = 0; j < span; ++j) {
      if (j < s1 - s0) {
        const int32_t sl = ids[j];
        out[4 * j + 0] = sl;
        out[4 * j + 1] = sp[sl];
        out[4 * j + 2] = (sp[sl + 1] - sp[sl]) / kSlice;
        out[4 * j + 3] = cb[sl];
      } else {
        out[4 * j + 0] = -1;
      }
    }
  }
  MIFE_CHECK_HIP(hipMemcpyAsync(order, h.data(), sizeof(int32_t) * h.size(), hipMemcpyHostToDevice, s));
  MIFE_CHECK_HIP(hipStreamSynchronize(s));
  *order_span = span;
  return MIFE_OK;
}

int mife_gather_values(int64_t nnz, const int32_t* src_pos, const void* src, int src_prec, void* dst, int dst_prec,
                       void* stream) {
  if (nnz == 0) return MIFE_OK;  // empty arrays may be passed as NULL
  MIFE_REQUIRE(src_pos && src && dst, MIFE_ERR_POINTER_IS_NULL, "null pointer argument");
  MIFE_REQUIRE((src_prec == MIFE_F64 || src_prec == MIFE_F32) && (dst_prec == MIFE_F64 || dst_prec == MIFE_F32),
               MIFE_ERR_ARG, "invalid precision");
  hipStream_t s = (hipStream_t)stream;
  MIFE_REQUIRE(div_up(nnz, (long long)kGatherChunk) <= INT32_MAX, MIFE_ERR_ARG, "too many entries");
  const int grid = (int)div_up(nnz, (long long)kGatherChunk);
  if (src_prec == MIFE_F64 && dst_prec == MIFE_F64)
    gather_kernel<double, double><<<grid, 256, 0, s>>>(nnz, src_pos, (const double*)src, (double*)dst);
  else if (src_prec == MIFE_F64)
    gather_kernel<double, float><<<grid, 256, 0, s>>>(nnz, src_pos, (const double*)src, (float*)dst);
  else if (dst_prec == MIFE_F32)
    gather_kernel<float, float><<<grid, 256, 0, s>>>(nnz, src_pos, (const float*)src, (float*)dst);
  else
    gather_kernel<float, double><<<grid, 256, 0, s>>>(nnz, src_pos, (const float*)src, (double*)dst);
  MIFE_CHECK_HIP(hipGetLastError());
  return MIFE_OK;
}

int mife_gather_values2(int64_t nnz, const int32_t* src_pos, const double* src, double* dst64, float* dst32,
                        void* stream) {
  if (nnz == 0) return MIFE_OK;
  MIFE_REQUIRE(src_pos && src && dst64 && dst32, MIFE_ERR_POINTER_IS_NULL, "null pointer argument");
  MIFE_REQUIRE(nnz > 0, MIFE_ERR_ARG, "negative entry count");
  hipStream_t s = (hipStream_t)stream;
  MIFE_REQUIRE(div_up(nnz, (long long)kGatherChunk) <= INT32_MAX, MIFE_ERR_ARG, "too many entries");
  const int grid = (int)div_up(nnz, (long long)kGatherChunk);
  gather2_kernel<<<grid, 256, 0, s>>>(nnz, src_pos, src, dst64, dst32);
  MIFE_CHECK_HIP(hipGetLastError());
  return MIFE_OK;
}

int mife_solver_rhs(int n_free, const int32_t* free_idx, const int32_t* row_ptr, const int32_t* col,
                    const double* vals, const int32_t* new_idx, const double* u, const double* src,
                    const double* rhs, const uint8_t* rhs_fixed, double* b, void* stream) {
  if (n_free == 0) return MIFE_OK;  // empty arrays may be passed as NULL
  MIFE_REQUIRE(free_idx && row_ptr && col && vals && new_idx && u && b, MIFE_ERR_POINTER_IS_NULL,
               "null pointer argument");
  hipStream_t s = (hipStream_t)stream;
  solver_rhs_kernel<<<div_up(n_free, 256), 256, 0, s>>>(n_free, free_idx, row_ptr, col, vals, new_idx, u, src, rhs,
                                                        rhs_fixed, b);
  MIFE_CHECK_HIP(hipGetLastError());
  return MIFE_OK;
}

int mife_field_update(int n_free, const int32_t* free_idx, const double* x, double* u, void* stream) {
  if (n_free == 0) return MIFE_OK;  // empty arrays may be passed as NULL
  MIFE_REQUIRE(free_idx && x && u, MIFE_ERR_POINTER_IS_NULL, "null pointer argument");
  field_update_kernel<<<div_up(n_free, 256), 256, 0, (hipStream_t)stream>>>(n_free, free_idx, x, u);
  MIFE_CHECK_HIP(hipGetLastError());
  return MIFE_OK;
}

int mife_backsubstitute(int n, const int32_t* row_ptr, const int32_t* col, const double* vals, const double* u,
                        const double* src, const uint8_t* rhs_fixed, int iron_compat, double* R, void* stream) {
  if (n == 0) return MIFE_OK;  // empty arrays may be passed as NULL
  MIFE_REQUIRE(row_ptr && col && vals && u && R, MIFE_ERR_POINTER_IS_NULL, "null pointer argument");
  const int grid = std::max(1, std::min(div_up(n, 256 / kBacksubLanes), kMaxGrid));
  backsub_kernel<<<grid, 256, 0, (hipStream_t)stream>>>(n, row_ptr, col, vals, u, src, rhs_fixed, iron_compat, R);
  MIFE_CHECK_HIP(hipGetLastError());
  return MIFE_OK;
}

}  // extern "C"
