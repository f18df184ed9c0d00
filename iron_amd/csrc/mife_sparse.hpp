// CSR SpMV building blocks shared by the solver kernels.
#pragma once
#include "mife_common.hpp"
#include <algorithm>

namespace mife {

// L consecutive lanes share a row (L in {4,8,16,32,64}); rows are strided over
// the grid so the grid size stays fixed.
template <typename T, int L>
__global__ __launch_bounds__(256) void spmv_kernel(int n, const int32_t* __restrict__ rp,
                                                   const int32_t* __restrict__ col, const T* __restrict__ vals,
                                                   const T* __restrict__ x, T* __restrict__ y) {
  const int groups_per_block = 256 / L;
  const int sub = threadIdx.x % L;
  for (int row = blockIdx.x * groups_per_block + threadIdx.x / L; row < n; row += gridDim.x * groups_per_block) {
    T acc = 0;
    const int k1 = rp[row + 1];
    for (int k = rp[row] + sub; k < k1; k += L) acc += vals[k] * x[col[k]];
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, 64);
    if (sub == 0) y[row] = acc;
  }
}

// Lanes per row from the mean row length.
inline int spmv_lanes(long long nnz, int n) {
  const double avg = n > 0 ? (double)nnz / n : 0.0;
  if (avg <= 6) return 4;
  if (avg <= 12) return 8;
  if (avg <= 40) return 16;
  if (avg <= 96) return 32;
  return 64;
}

inline int spmv_grid(int n, int L) {
  const int gpb = 256 / L;
  return std::max(1, std::min(div_up(n, gpb), 256 * 16));
}

template <typename T>
inline void launch_spmv(int n, const int32_t* rp, const int32_t* col, const T* vals, const T* x, T* y, int L,
                        hipStream_t s) {
  const int grid = spmv_grid(n, L);
  switch (L) {
    case 4: spmv_kernel<T, 4><<<grid, 256, 0, s>>>(n, rp, col, vals, x, y); break;
    case 8: spmv_kernel<T, 8><<<grid, 256, 0, s>>>(n, rp, col, vals, x, y); break;
    case 16: spmv_kernel<T, 16><<<grid, 256, 0, s>>>(n, rp, col, vals, x, y); break;
    case 32: spmv_kernel<T, 32><<<grid, 256, 0, s>>>(n, rp, col, vals, x, y); break;
    default: spmv_kernel<T, 64><<<grid, 256, 0, s>>>(n, rp, col, vals, x, y); break;
  }
}

}  // namespace mife
