// Shared definitions for the mi_fe HIP library (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <cstdarg>

#include "../../include/mife.h"

namespace mife {

// Thread-local last-error text, returned by mife_last_error().
void set_error(const char* fmt, ...);

#define MIFE_CHECK_HIP(expr)                                                  \
  do {                                                                        \
    hipError_t _e = (expr);                                                   \
    if (_e != hipSuccess) {                                                   \
      ::mife::set_error("%s failed: %s (%s:%d)", #expr, hipGetErrorString(_e), \
                        __FILE__, __LINE__);                                  \
      return MIFE_ERR_HIP;                                                    \
    }                                                                         \
  } while (0)

#define MIFE_REQUIRE(cond, code, ...)  \
  do {                                 \
    if (!(cond)) {                     \
      ::mife::set_error(__VA_ARGS__);  \
      return (code);                   \
    }                                  \
  } while (0)

constexpr int kWave = 64;

// Wavefront sum (64 lanes) with DPP-free shuffles.
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
  return v;
}

// Block-wide sum; `scratch` must hold blockDim.x/64 values. All threads get the result.
template <typename T>
__device__ __forceinline__ T block_sum(T v, T* scratch) {
  v = wave_sum(v);
  const int lane = threadIdx.x & (kWave - 1);
  const int wid = threadIdx.x / kWave;
  __syncthreads();
  if (lane == 0) scratch[wid] = v;
  __syncthreads();
  const int nw = (blockDim.x + kWave - 1) / kWave;
  T s = 0;
  for (int i = 0; i < nw; ++i) s += scratch[i];
  return s;
}

inline int div_up(long long a, long long b) { return (int)((a + b - 1) / b); }

}  // namespace mife
