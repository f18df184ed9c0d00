// Runtime services and per-mesh setup: CSR sparsity, scatter map, colouring and
// the reduced (free x free) solver pattern.  All index work runs on the GPU
// except the greedy colouring, which is a host pass over the element list.
#include "mife_common.hpp"

#include <algorithm>
#include <climits>
#include <vector>

namespace mife {

static thread_local char g_err[1024] = {0};

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

// ------------------------------------------------------------------ scan
// Exclusive scan of int32 counts; out[n] receives the total.  Three passes:
// per-tile sums, a single-block scan of tile sums, then per-tile rescans.
constexpr int kScanTile = 1024;  // 256 threads x 4 items

__global__ void scan_tile_sums(const int32_t* __restrict__ in, int n, int64_t* __restrict__ tile_sums) {
  __shared__ int64_t red[4];
  const int base = blockIdx.x * kScanTile;
  int64_t s = 0;
  for (int i = threadIdx.x; i < kScanTile; i += blockDim.x) {
    const int idx = base + i;
    if (idx < n) s += in[idx];
  }
  s = block_sum(s, red);
  if (threadIdx.x == 0) tile_sums[blockIdx.x] = s;
}

__global__ void scan_tile_offsets(int64_t* __restrict__ tile_sums, int n_tiles, int64_t* __restrict__ total) {
  // one block of 1024 threads walks the tile sums in chunks
  __shared__ int64_t buf[1024];
  int64_t carry = 0;
  for (int start = 0; start < n_tiles; start += 1024) {
    const int i = start + threadIdx.x;
    int64_t v = (i < n_tiles) ? tile_sums[i] : 0;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      int64_t t = (threadIdx.x >= off) ? buf[threadIdx.x - off] : 0;
      __syncthreads();
      buf[threadIdx.x] += t;
      __syncthreads();
    }
    if (i < n_tiles) tile_sums[i] = carry + buf[threadIdx.x] - v;  // exclusive
    carry += buf[1023];
    __syncthreads();
  }
  if (threadIdx.x == 0) *total = carry;
}

__global__ void scan_tile_apply(const int32_t* __restrict__ in, int n, const int64_t* __restrict__ tile_off,
                                int32_t* __restrict__ out, const int64_t* __restrict__ total) {
  __shared__ int64_t buf[256];
  const int base = blockIdx.x * kScanTile;
  const int t = threadIdx.x;
  int64_t v[4];
  int64_t loc = 0;
  for (int k = 0; k < 4; ++k) {
    const int idx = base + t * 4 + k;
    v[k] = (idx < n) ? in[idx] : 0;
    loc += v[k];
  }
  buf[t] = loc;
  __syncthreads();
  for (int off = 1; off < 256; off <<= 1) {
    int64_t x = (t >= off) ? buf[t - off] : 0;
    __syncthreads();
    buf[t] += x;
    __syncthreads();
  }
  int64_t run = tile_off[blockIdx.x] + buf[t] - loc;
  for (int k = 0; k < 4; ++k) {
    const int idx = base + t * 4 + k;
    if (idx < n) out[idx] = (int32_t)run;
    run += v[k];
  }
  if (blockIdx.x == 0 && t == 0) out[n] = (int32_t)(*total);
}

static size_t scan_work_bytes(int n) { return sizeof(int64_t) * ((size_t)div_up(n, kScanTile) + 2); }

// Returns the total through *total_host (synchronises the stream).
static int scan_exclusive(const int32_t* in, int n, int32_t* out, void* work, hipStream_t s,
                          int64_t* total_host) {
  const int tiles = std::max(1, div_up(n, kScanTile));
  int64_t* tile = reinterpret_cast<int64_t*>(work);
  int64_t* total = tile + tiles;
  scan_tile_sums<<<tiles, 256, 0, s>>>(in, n, tile);
  scan_tile_offsets<<<1, 1024, 0, s>>>(tile, tiles, total);
  scan_tile_apply<<<tiles, 256, 0, s>>>(in, n, tile, out, total);
  MIFE_CHECK_HIP(hipGetLastError());
  int64_t t = 0;
  MIFE_CHECK_HIP(hipMemcpyAsync(&t, total, sizeof(int64_t), hipMemcpyDeviceToHost, s));
  MIFE_CHECK_HIP(hipStreamSynchronize(s));
  MIFE_REQUIRE(t <= (int64_t)INT32_MAX, MIFE_ERR_ARG, "index total %lld exceeds the int32 range", (long long)t);
  if (total_host) *total_host = t;
  return MIFE_OK;
}

// ----------------------------------------------------------- incidence
__global__ void incidence_count(const int32_t* __restrict__ edofs, long long n_entries, int n_dofs,
                                int32_t* __restrict__ cnt, int* __restrict__ bad) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_entries) return;
  const int d = edofs[t];
  if (d < 0 || d >= n_dofs) {
    atomicOr(bad, 1);
    return;
  }
  atomicAdd(&cnt[d], 1);
}

__global__ void incidence_fill(const int32_t* __restrict__ edofs, long long n_entries, int n_edof,
                               const int32_t* __restrict__ ptr, int32_t* __restrict__ fill,
                               int32_t* __restrict__ idx) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n_entries) return;
  const int d = edofs[t];
  const int slot = atomicAdd(&fill[d], 1);
  idx[ptr[d] + slot] = (int32_t)(t / n_edof);
}

__global__ void max_reduce(const int32_t* __restrict__ v, int n, int* __restrict__ out) {
  int m = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) m = max(m, v[i]);
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) m = max(m, __shfl_xor(m, off, kWave));
  if ((threadIdx.x & 63) == 0) atomicMax(out, m);
}

// ------------------------------------------------------ row sort-unique
// One 64-lane workgroup per row: gather candidate columns from the incident
// elements into LDS, bitonic sort, then count (pass 1) or emit (pass 2) the
// unique values in ascending order.
template <int P>
__global__ __launch_bounds__(64) void row_union(const int32_t* __restrict__ edofs, int n_edof,
                                                const int32_t* __restrict__ inc_ptr,
                                                const int32_t* __restrict__ inc_idx, int n_dofs,
                                                int32_t* __restrict__ row_cnt,
                                                const int32_t* __restrict__ row_ptr,
                                                int32_t* __restrict__ col) {
  __shared__ int32_t s[P];
  const int d = blockIdx.x;
  if (d >= n_dofs) return;
  const int lane = threadIdx.x;
  const int e0 = inc_ptr[d], ne = inc_ptr[d + 1] - e0;
  const int nc = ne * n_edof;  // <= P by construction of the launch
  for (int i = lane; i < P; i += 64) {
    int v = INT_MAX;
    if (i < nc) {
      const int e = inc_idx[e0 + i / n_edof];
      v = edofs[(long long)e * n_edof + (i % n_edof)];
    }
    s[i] = v;
  }
  __syncthreads();
  for (int k = 2; k <= P; k <<= 1) {
    for (int j = k >> 1; j > 0; j >>= 1) {
      for (int i = lane; i < P; i += 64) {
        const int ixj = i ^ j;
        if (ixj > i) {
          const int a = s[i], b = s[ixj];
          const bool up = ((i & k) == 0);
          if ((a > b) == up) {
            s[i] = b;
            s[ixj] = a;
          }
        }
      }
      __syncthreads();
    }
  }
  // unique: element i is a "first" if it is valid and differs from i-1
  int base = 0;
  const int out0 = col ? row_ptr[d] : 0;
  for (int c0 = 0; c0 < P; c0 += 64) {
    const int i = c0 + lane;
    const int v = s[i];
    const bool first = (v != INT_MAX) && (i == 0 || s[i - 1] != v);
    const unsigned long long mask = __ballot(first);
    if (col && first) {
      const int rank = __popcll(mask & ((1ull << lane) - 1ull));
      col[out0 + base + rank] = v;
    }
    base += __popcll(mask);
  }
  if (!col && lane == 0) row_cnt[d] = base;
}

// ---------------------------------------------------------- scatter map
__global__ void scatter_map_kernel(const int32_t* __restrict__ edofs, int n_elem, int n_edof,
                                   const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
                                   int32_t* __restrict__ pos, int* __restrict__ bad) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long per = (long long)n_edof * n_edof;
  if (t >= per * n_elem) return;
  const long long e = t / per;
  const int ij = (int)(t - e * per);
  const int i = ij / n_edof, j = ij - (ij / n_edof) * n_edof;
  const int row = edofs[e * n_edof + i];
  const int target = edofs[e * n_edof + j];
  int lo = row_ptr[row], hi = row_ptr[row + 1];
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (col[mid] < target) lo = mid + 1; else hi = mid;
  }
  if (lo >= row_ptr[row + 1] || col[lo] != target) {
    atomicOr(bad, 1);
    pos[t] = -1;
  } else {
    pos[t] = lo;
  }
}

// ----------------------------------------------------- reduced pattern
__global__ void reduce_count(int n, const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
                             const int32_t* __restrict__ new_idx, int32_t* __restrict__ cnt) {
  // one wave per row
  const int row = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  const int f = new_idx[row];
  if (f < 0) return;
  int c = 0;
  for (int k = row_ptr[row] + lane; k < row_ptr[row + 1]; k += 64) c += (new_idx[col[k]] >= 0);
  c = wave_sum(c);
  if (lane == 0) cnt[f] = c;
}

__global__ void reduce_fill(int n, const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
                            const int32_t* __restrict__ new_idx, const int32_t* __restrict__ s_row_ptr,
                            int32_t* __restrict__ s_col, int32_t* __restrict__ src_pos) {
  const int row = blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (row >= n) return;
  const int f = new_idx[row];
  if (f < 0) return;
  int base = s_row_ptr[f];
  for (int k0 = row_ptr[row]; k0 < row_ptr[row + 1]; k0 += 64) {
    const int k = k0 + lane;
    int nc = -1;
    if (k < row_ptr[row + 1]) nc = new_idx[col[k]];
    const bool keep = nc >= 0;
    const unsigned long long m = __ballot(keep);
    if (keep) {
      const int r = base + __popcll(m & ((1ull << lane) - 1ull));
      s_col[r] = nc;
      src_pos[r] = k;
    }
    base += __popcll(m);
  }
}

}  // namespace mife

using namespace mife;

extern "C" {

const char* mife_version(void) { return "mi_fe 0.1 (gfx950)"; }

int mife_last_error(char* buf, int n) {
  if (!buf) return MIFE_ERR_POINTER_IS_NULL;
  if (n <= 0) return MIFE_ERR_ARG;
  strncpy(buf, g_err, (size_t)n - 1);
  buf[n - 1] = 0;
  return MIFE_OK;
}

int mife_init(int device) {
  int count = 0;
  MIFE_CHECK_HIP(hipGetDeviceCount(&count));
  MIFE_REQUIRE(device >= 0 && device < count, MIFE_ERR_ARG, "device %d out of range (%d devices)", device, count);
  MIFE_CHECK_HIP(hipSetDevice(device));
  MIFE_CHECK_HIP(hipFree(nullptr));
  return MIFE_OK;
}

// Work layout: [cnt n][fill n][inc_ptr n+1][inc_idx E*k][row_cnt n][ints 4][scan work]
size_t mife_csr_pattern_work_bytes(int n_elem, int n_edof, int n_dofs) {
  const size_t n = (size_t)n_dofs;
  const size_t ints = n + n + (n + 1) + (size_t)n_elem * n_edof + n + 4;
  return ((ints * sizeof(int32_t) + 255) / 256) * 256 + scan_work_bytes(n_dofs + 1) + 256;
}

int mife_csr_pattern(const int32_t* elem_dofs, int n_elem, int n_edof, int n_dofs, int32_t* row_ptr,
                     int32_t* col, int64_t* nnz_out, void* work, void* stream) {
  MIFE_REQUIRE(elem_dofs && row_ptr && work, MIFE_ERR_POINTER_IS_NULL, "null pointer argument");
  MIFE_REQUIRE(n_elem > 0 && n_edof > 0 && n_dofs > 0, MIFE_ERR_ARG, "empty mesh");
  hipStream_t s = (hipStream_t)stream;
  const size_t n = (size_t)n_dofs;
  int32_t* cnt = reinterpret_cast<int32_t*>(work);
  int32_t* fill = cnt + n;
  int32_t* inc_ptr = fill + n;
  int32_t* inc_idx = inc_ptr + n + 1;
  int32_t* row_cnt = inc_idx + (size_t)n_elem * n_edof;
  int* flags = row_cnt + n;  // [0]=bad [1]=max incidence
  const size_t ints = n + n + (n + 1) + (size_t)n_elem * n_edof + n + 4;
  void* scan_work = reinterpret_cast<char*>(work) + ((ints * sizeof(int32_t) + 255) / 256) * 256;
  const long long entries = (long long)n_elem * n_edof;
  const int TB = 256;
  int host_flags[2] = {0, 0};
  if (col == nullptr) {
    MIFE_CHECK_HIP(hipMemsetAsync(cnt, 0, sizeof(int32_t) * 2 * n, s));
    MIFE_CHECK_HIP(hipMemsetAsync(flags, 0, sizeof(int) * 4, s));
    incidence_count<<<div_up(entries, TB), TB, 0, s>>>(elem_dofs, entries, n_dofs, cnt, flags);
    max_reduce<<<std::min(1024, div_up(n_dofs, TB)), TB, 0, s>>>(cnt, n_dofs, flags + 1);
    MIFE_CHECK_HIP(hipMemcpyAsync(host_flags, flags, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
    MIFE_CHECK_HIP(hipStreamSynchronize(s));
    MIFE_REQUIRE(host_flags[0] == 0, MIFE_ERR_ARG, "element DOF outside [0, %d)", n_dofs);
    int rc = scan_exclusive(cnt, n_dofs, inc_ptr, scan_work, s, nullptr);
    if (rc) return rc;
    incidence_fill<<<div_up(entries, TB), TB, 0, s>>>(elem_dofs, entries, n_edof, inc_ptr, fill, inc_idx);
  } else {
    MIFE_CHECK_HIP(hipMemcpyAsync(host_flags, flags, sizeof(int) * 2, hipMemcpyDeviceToHost, s));
    MIFE_CHECK_HIP(hipStreamSynchronize(s));
  }
  const int cand = host_flags[1] * n_edof;
  MIFE_REQUIRE(cand <= 4096, MIFE_ERR_ARG, "row candidate count %d exceeds 4096", cand);
  const int32_t* rp = col ? row_ptr : nullptr;
#define LAUNCH_ROW_UNION(P)                                                                    \
  row_union<P><<<n_dofs, 64, 0, s>>>(elem_dofs, n_edof, inc_ptr, inc_idx, n_dofs, row_cnt, rp, col)
  if (cand <= 128) LAUNCH_ROW_UNION(128);
  else if (cand <= 256) LAUNCH_ROW_UNION(256);
  else if (cand <= 512) LAUNCH_ROW_UNION(512);
  else if (cand <= 1024) LAUNCH_ROW_UNION(1024);
  else if (cand <= 2048) LAUNCH_ROW_UNION(2048);
  else LAUNCH_ROW_UNION(4096);
#undef LAUNCH_ROW_UNION
  MIFE_CHECK_HIP(hipGetLastError());
  if (col == nullptr) {
    int64_t nnz = 0;
    int rc = scan_exclusive(row_cnt, n_dofs, row_ptr, scan_work, s, &nnz);
    if (rc) return rc;
    if (nnz_out) *nnz_out = nnz;
  }
  return MIFE_OK;
}

int mife_scatter_map(const int32_t* elem_dofs, int n_elem, int n_edof, const int32_t* row_ptr,
                     const int32_t* col, int32_t* pos, void* stream) {
  MIFE_REQUIRE(elem_dofs && row_ptr && col && pos, MIFE_ERR_POINTER_IS_NULL, "null pointer argument");
  hipStream_t s = (hipStream_t)stream;
  int* bad = nullptr;
  MIFE_CHECK_HIP(hipMallocAsync((void**)&bad, sizeof(int), s));
  MIFE_CHECK_HIP(hipMemsetAsync(bad, 0, sizeof(int), s));
  const long long total = (long long)n_elem * n_edof * n_edof;
  scatter_map_kernel<<<div_up(total, 256), 256, 0, s>>>(elem_dofs, n_elem, n_edof, row_ptr, col, pos, bad);
  MIFE_CHECK_HIP(hipGetLastError());
  int hb = 0;
  MIFE_CHECK_HIP(hipMemcpyAsync(&hb, bad, sizeof(int), hipMemcpyDeviceToHost, s));
  MIFE_CHECK_HIP(hipFreeAsync(bad, s));
  MIFE_CHECK_HIP(hipStreamSynchronize(s));
  MIFE_REQUIRE(hb == 0, MIFE_ERR_PATTERN, "element entry not found in the CSR pattern");
  return MIFE_OK;
}

int mife_colour_elements(const int32_t* edofs, int n_elem, int n_edof, int n_dofs, int32_t* colour,
                         int* n_colours) {
  MIFE_REQUIRE(edofs && colour && n_colours, MIFE_ERR_POINTER_IS_NULL, "null pointer argument");
  // DOF -> element incidence (host)
  std::vector<int32_t> ptr((size_t)n_dofs + 1, 0), idx((size_t)n_elem * n_edof);
  for (long long t = 0; t < (long long)n_elem * n_edof; ++t) {
    const int d = edofs[t];
    MIFE_REQUIRE(d >= 0 && d < n_dofs, MIFE_ERR_ARG, "element DOF outside [0, %d)", n_dofs);
    ptr[(size_t)d + 1]++;
  }
  for (int d = 0; d < n_dofs; ++d) ptr[(size_t)d + 1] += ptr[d];
  std::vector<int32_t> fill(ptr.begin(), ptr.end() - 1);
  for (int e = 0; e < n_elem; ++e)
    for (int a = 0; a < n_edof; ++a) idx[fill[edofs[(size_t)e * n_edof + a]]++] = e;
  int nc = 0;
  for (int e = 0; e < n_elem; ++e) colour[e] = -1;
  for (int e = 0; e < n_elem; ++e) {
    uint64_t used = 0;
    for (int a = 0; a < n_edof; ++a) {
      const int d = edofs[(size_t)e * n_edof + a];
      for (int k = ptr[d]; k < ptr[(size_t)d + 1]; ++k) {
        const int c = colour[idx[k]];
        if (c >= 0) used |= (1ull << c);
      }
    }
    int c = 0;
    while (c < 64 && (used >> c) & 1ull) ++c;
    MIFE_REQUIRE(c < 64, MIFE_ERR, "element colouring needs more than 64 colours");
    colour[e] = c;
    nc = std::max(nc, c + 1);
  }
  *n_colours = nc;
  return MIFE_OK;
}

int mife_reduce_pattern(int n, const int32_t* row_ptr, const int32_t* col, const int32_t* new_idx, int n_free,
                        int32_t* s_row_ptr, int32_t* s_col, int32_t* src_pos, int64_t* nnz_out, void* work,
                        void* stream) {
  MIFE_REQUIRE(row_ptr && col && new_idx && s_row_ptr && work, MIFE_ERR_POINTER_IS_NULL, "null pointer argument");
  hipStream_t s = (hipStream_t)stream;
  const int wpb = 4;
  if (s_col == nullptr) {
    int32_t* cnt = reinterpret_cast<int32_t*>(work);
    void* scan_work = reinterpret_cast<char*>(work) + (((size_t)n_free * 4 + 255) / 256) * 256;
    MIFE_CHECK_HIP(hipMemsetAsync(cnt, 0, sizeof(int32_t) * (size_t)n_free, s));
    reduce_count<<<div_up(n, wpb), 64 * wpb, 0, s>>>(n, row_ptr, col, new_idx, cnt);
    MIFE_CHECK_HIP(hipGetLastError());
    int64_t nnz = 0;
    int rc = scan_exclusive(cnt, n_free, s_row_ptr, scan_work, s, &nnz);
    if (rc) return rc;
    if (nnz_out) *nnz_out = nnz;
    return MIFE_OK;
  }
  MIFE_REQUIRE(src_pos, MIFE_ERR_POINTER_IS_NULL, "null src_pos");
  reduce_fill<<<div_up(n, wpb), 64 * wpb, 0, s>>>(n, row_ptr, col, new_idx, s_row_ptr, s_col, src_pos);
  MIFE_CHECK_HIP(hipGetLastError());
  return MIFE_OK;
}

}  // extern "C"
