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
  int iters, maxit, done, pad;
};

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
  SegState* st;
  double* parts;  // 3 values per partial slot: parts[v * cap + slot]
  int cap;
};

// y_row = (A x)_row for the rows of this block's segment, 8 lanes per row.
__device__ __forceinline__ double row_dot(const View& V, const mife_segment& S, const double* __restrict__ x,
                                          int row, int sub) {
  const int32_t* rp = V.rp + S.rp_off;
  const int32_t* col = V.col + S.nz_off;
  const double* val = V.val + S.nz_off;
  const double* xs = x + S.row_off;
  const int k0 = rp[row], k1 = rp[row + 1];
  double a = 0.0;
  for (int k = k0 + sub; k < k1; k += kL) a += val[k] * xs[col[k]];
#pragma unroll
  for (int off = kL / 2; off > 0; off >>= 1) a += __shfl_xor(a, off, kL);
  return a;
}

// SELL: (A x) of lane `lane`'s row of slice s of the segment (complete sum).
__device__ __forceinline__ double slice_dot(const View& V, const mife_segment& S, const double* __restrict__ x,
                                            int s, int lane) {
  const int32_t* sp = V.rp + S.rp_off;
  const double* val = V.val + S.nz_off;
  const double* xs = x + S.row_off;
  if (S.cb_off >= 0) return sell16_row<double>(sp, V.col16 + S.nz_off, V.cbase + S.cb_off, val, xs, s, lane);
  return sell_row<double>(sp, V.col + S.nz_off, val, xs, s, lane);
}

// Calls f(row, (A x)_row) for every row of block B (by one lane per row).
template <typename F>
__device__ __forceinline__ void block_rows(const View& V, const Block& B, const mife_segment& S,
                                           const double* __restrict__ x, F&& f) {
  if (S.format == MIFE_FMT_SELL) {
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    if (wave < B.rows) {
      const int s = B.row0 + wave;
      const double ax = slice_dot(V, S, x, s, lane);
      const int row = s * kSlice + lane;
      if (row < S.n) f(row, ax);
    }
    return;
  }
  const int lr = threadIdx.x / kL, sub = threadIdx.x % kL;
  for (int row = B.row0 + lr; row < B.row0 + B.rows; row += kRowsPerBlock) {
    const double ax = row_dot(V, S, x, row, sub);
    if (sub == 0) f(row, ax);
  }
}

// First row and row count of block B (vector kernels: one row per thread).
__device__ __forceinline__ void block_row_range(const Block& B, const mife_segment& S, int& r0, int& r1) {
  if (S.format == MIFE_FMT_SELL) {
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

__global__ __launch_bounds__(256) void gcg_diag_kernel(View V, double* __restrict__ dinv) {
  const Block B = V.blocks[blockIdx.x];
  const mife_segment S = V.segs[B.seg];
  int r0, r1;
  block_row_range(B, S, r0, r1);
  const int row = r0 + (int)threadIdx.x;
  if (row >= r1) return;
  double d = 0.0;
  if (S.format == MIFE_FMT_SELL) {
    const int32_t* sp = V.rp + S.rp_off;
    const int s = row / kSlice, lane = row - s * kSlice;
    const double* val = V.val + S.nz_off;
    for (int e = sp[s] + lane; e < sp[s + 1]; e += kSlice) {
      const int c = S.cb_off >= 0 ? V.cbase[S.cb_off + s] + (int)V.col16[S.nz_off + e] : V.col[S.nz_off + e];
      if (c == row && val[e] != 0.0) {
        d = val[e];
        break;
      }
    }
  } else {
    const int32_t* rp = V.rp + S.rp_off;
    const int32_t* col = V.col + S.nz_off;
    const double* val = V.val + S.nz_off;
    for (int k = rp[row]; k < rp[row + 1]; ++k)
      if (col[k] == row) d = val[k];
  }
  dinv[S.row_off + row] = d != 0.0 ? 1.0 / d : 1.0;
}

// r = b - A x, z = D^-1 r, p = q = 0; partials r.z, r.r, b.b
__global__ __launch_bounds__(256) void gcg_init_kernel(View V, const double* __restrict__ b,
                                                       const double* __restrict__ x, const double* __restrict__ dinv,
                                                       double* __restrict__ r, double* __restrict__ z,
                                                       double* __restrict__ p, double* __restrict__ q) {
  const Block B = V.blocks[blockIdx.x];
  const mife_segment S = V.segs[B.seg];
  double m[3] = {0.0, 0.0, 0.0};
  block_rows(V, B, S, x, [&](int row, double ax) {
    const long long g = S.row_off + row;
    const double ri = b[g] - ax;
    const double zi = dinv[g] * ri;
    r[g] = ri;
    z[g] = zi;
    p[g] = 0.0;
    q[g] = 0.0;
    m[0] += ri * zi;
    m[1] += ri * ri;
    m[2] += b[g] * b[g];
  });
  store_partials<3>(V, B, m);
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
  double out[3];
  seg_sums<3>(V, seg_parts, blockIdx.x, out);
  if (threadIdx.x == 0) {
    SegState& g = V.st[s];
    const mife_segment S = V.segs[s];
    g.rz = out[0];
    g.rr = out[1];
    g.bb = out[2];
    g.beta = 0.0;
    g.alpha = 0.0;
    g.pq = 0.0;
    g.iters = 0;
    g.maxit = S.maxit;
    g.tol2 = S.rtol * S.rtol;
    g.rr_true = -1.0;
    g.done = (out[1] <= g.tol2 * out[2] || S.maxit <= 0) ? 1 : 0;
  }
}

// p = z + beta p ; q = A z + beta q ; partial p.q
__global__ __launch_bounds__(256) void gcg_spmv_kernel(View V, const double* __restrict__ z, double* __restrict__ p,
                                                       double* __restrict__ q) {
  const Block B = V.blocks[blockIdx.x];
  if (V.st[B.seg].done) return;
  const mife_segment S = V.segs[B.seg];
  const double beta = V.st[B.seg].beta;
  double m[1] = {0.0};
  block_rows(V, B, S, z, [&](int row, double az) {
    const long long g = S.row_off + row;
    const double pi = z[g] + beta * p[g];
    const double qi = az + beta * q[g];
    p[g] = pi;
    q[g] = qi;
    m[0] += pi * qi;
  });
  store_partials<1>(V, B, m);
}

__global__ __launch_bounds__(256) void gfin_alpha_kernel(View V, const int32_t* __restrict__ seg_parts,
                                                         const int32_t* __restrict__ seg_ids) {
  const int s = seg_ids[blockIdx.x];
  if (V.st[s].done) return;
  double out[1];
  seg_sums<1>(V, seg_parts, blockIdx.x, out);
  if (threadIdx.x == 0) {
    SegState& g = V.st[s];
    g.pq = out[0];
    g.alpha = out[0] != 0.0 ? g.rz / out[0] : 0.0;
  }
}

// x += alpha p ; r -= alpha q ; z = D^-1 r ; partials r.z, r.r
__global__ __launch_bounds__(256) void gcg_update_kernel(View V, double* __restrict__ x, double* __restrict__ r,
                                                         double* __restrict__ z, const double* __restrict__ p,
                                                         const double* __restrict__ q,
                                                         const double* __restrict__ dinv) {
  const Block B = V.blocks[blockIdx.x];
  if (V.st[B.seg].done) return;
  const mife_segment S = V.segs[B.seg];
  const double alpha = V.st[B.seg].alpha;
  double m[2] = {0.0, 0.0};
  int r0, r1;
  block_row_range(B, S, r0, r1);
  for (int row = r0 + threadIdx.x; row < r1; row += 256) {
    const long long g = S.row_off + row;
    x[g] += alpha * p[g];
    const double ri = r[g] - alpha * q[g];
    const double zi = dinv[g] * ri;
    r[g] = ri;
    z[g] = zi;
    m[0] += ri * zi;
    m[1] += ri * ri;
  }
  store_partials<2>(V, B, m);
}

__global__ __launch_bounds__(256) void gfin_beta_kernel(View V, const int32_t* __restrict__ seg_parts,
                                                        const int32_t* __restrict__ seg_ids) {
  const int s = seg_ids[blockIdx.x];
  if (V.st[s].done) return;
  double out[2];
  seg_sums<2>(V, seg_parts, blockIdx.x, out);
  if (threadIdx.x == 0) {
    SegState& g = V.st[s];
    g.beta = g.rz != 0.0 ? out[0] / g.rz : 0.0;
    g.rz = out[0];
    g.rr = out[1];
    g.iters += 1;
    g.done = (out[1] <= g.tol2 * g.bb || g.iters >= g.maxit) ? 1 : 0;
  }
}

// true residual ||b - A x||^2 of the listed segments
__global__ __launch_bounds__(256) void gcg_resid_kernel(View V, const double* __restrict__ b,
                                                        const double* __restrict__ x) {
  const Block B = V.blocks[blockIdx.x];
  const mife_segment S = V.segs[B.seg];
  double m[1] = {0.0};
  block_rows(V, B, S, x, [&](int row, double ax) {
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
//   seg_ids[max_seg] | parts[3 * max_blocks]
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
  b += sizeof(double) * 3 * (size_t)max_blocks + 256;
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
    MIFE_REQUIRE(S.format == MIFE_FMT_CSR || (S.format == MIFE_FMT_SELL && S.n_slices == div_up(S.n, kSlice)),
                 MIFE_ERR_ARG, "segment %d: format %d / n_slices %d", k, S.format, S.n_slices);
    H.segs[S.slot] = S;
    H.seg_parts[k] = nb;
    H.seg_ids[k] = S.slot;
    const bool sell = S.format == MIFE_FMT_SELL;
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
  return View{D.segs, D.blocks, P->rp, P->col, P->col16, P->cbase, P->val, (SegState*)state, D.parts, cap};
}

int mife_grouped_cg_init(int n_seg, int n_blocks, int max_segments, int max_blocks, const mife_pools* pools,
                         const double* b, double* x, double* vec5, int64_t pool_rows, void* work, void* state,
                         void* stream) {
  MIFE_REQUIRE(pools && pools->rp && pools->val && b && x && vec5 && work && state, MIFE_ERR_POINTER_IS_NULL,
               "null pointer argument");
  if (n_seg == 0) return MIFE_OK;
  hipStream_t s = (hipStream_t)stream;
  Layout D = carve(work, max_segments, max_blocks);
  View V = make_view(D, state, pools, max_blocks);
  double* dinv = vec5;
  double* r = vec5 + pool_rows;
  double* z = vec5 + 2 * pool_rows;
  double* p = vec5 + 3 * pool_rows;
  double* q = vec5 + 4 * pool_rows;
  gcg_diag_kernel<<<n_blocks, 256, 0, s>>>(V, dinv);
  gcg_init_kernel<<<n_blocks, 256, 0, s>>>(V, b, x, dinv, r, z, p, q);
  gfin_init_kernel<<<n_seg, 256, 0, s>>>(V, D.seg_parts, D.seg_ids);
  MIFE_CHECK_HIP(hipGetLastError());
  return MIFE_OK;
}

int mife_grouped_cg_iterate(int n_seg, int n_blocks, int max_segments, int max_blocks, const mife_pools* pools,
                            double* x, double* vec5, int64_t pool_rows, int iterations, void* work, void* state,
                            void* stream) {
  MIFE_REQUIRE(pools && pools->rp && pools->val && x && vec5 && work && state, MIFE_ERR_POINTER_IS_NULL,
               "null pointer argument");
  if (n_seg == 0) return MIFE_OK;
  hipStream_t s = (hipStream_t)stream;
  Layout D = carve(work, max_segments, max_blocks);
  View V = make_view(D, state, pools, max_blocks);
  double* dinv = vec5;
  double* r = vec5 + pool_rows;
  double* z = vec5 + 2 * pool_rows;
  double* p = vec5 + 3 * pool_rows;
  double* q = vec5 + 4 * pool_rows;
  for (int it = 0; it < iterations; ++it) {
    const int tk = prof::begin(s);
    gcg_spmv_kernel<<<n_blocks, 256, 0, s>>>(V, z, p, q);
    prof::end(tk, s, prof::kGroupedSpmv);
    gfin_alpha_kernel<<<n_seg, 256, 0, s>>>(V, D.seg_parts, D.seg_ids);
    const int tk2 = prof::begin(s);
    gcg_update_kernel<<<n_blocks, 256, 0, s>>>(V, x, r, z, p, q, dinv);
    prof::end(tk2, s, prof::kGroupedUpdate);
    gfin_beta_kernel<<<n_seg, 256, 0, s>>>(V, D.seg_parts, D.seg_ids);
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
