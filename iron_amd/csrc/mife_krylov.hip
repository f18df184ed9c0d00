// Krylov solvers: Jacobi-preconditioned CG (FP64 or FP32), mixed-precision CG
// with FP64 residual replacement ("reliable updates"), and right-Jacobi
// restarted GMRES (FP64).
//
// CG runs two fused kernels per iteration and keeps every scalar on the
// device.  With q = A p maintained by the recurrence q = A z + beta q, the
// search direction update folds into the SpMV kernel:
//   K_A: p = z + beta p ; q = A z + beta q ; partial (p.q)
//   K_B: alpha = rz / pq ; x += alpha p ; r -= alpha q ; z = D^-1 r ;
//        partials (r.z), (r.r)                                -> beta (finisher)
// Each bandwidth kernel leaves one partial sum per block.  K_B's blocks (few and
// large, one round on the chip) each sum K_A's partials themselves in one fixed
// order, so alpha needs no launch of its own; a one-block finisher reduces K_B's
// partials (deterministic) and updates the device-side state for the next K_A.
// No device-scope fence runs inside the bandwidth kernels, the host only looks
// at the state every few iterations, and kernels return at once once the state
// says stop.
#include "mife_common.hpp"
#include "mife_profile.hpp"
#include "mife_sparse.hpp"

extern "C" int mife_profile_enable(int on);

#include <algorithm>
#include <cmath>
#include <cstddef>
#include <cstring>
#include <mutex>
#include <vector>

namespace mife {

constexpr int kRedBlocks = kMaxSellGrid;  // partial slots per reduced value (largest row grid)
constexpr int kMaxRestart = 256;
constexpr int kFinThreads = 1024;  // one-block finishers: two partials per thread
// CG update blocks: 1024 threads, so that the blocks are few (one round) and each
// sums K_A's partials once
constexpr int kUpdateThreads = 1024;
constexpr int kMaxIterChunk = 256;  // CG iterations launched between two host reads of the state

struct CgState {
  double rz, pq, alpha, beta, rr, bb, tol2, scale;
  double rz_prev, maxrr, rr_seg, rr64;  // rr_seg: recurrence r.r when a segment ended
  // Segmented (mixed-precision) solves: a segment pauses once r.r <= seg_rr, set at
  // each replacement to seg_drop x the new r.r (seg_drop = 0: no pausing).
  double seg_drop, seg_rr;
  int iters, done, maxit, pause;
  unsigned int cnt_a, cnt_b, cnt_c, pad2;
  // per-block partials of up to 4 reduced values each (value v of block b at v*kRedBlocks+b)
  double part_a[4 * kRedBlocks];
  double part_b[4 * kRedBlocks];
};

// Block-level sums of up to 4 values; thread 0 stores them as this block's
// partials (value v of block b at parts[v * kRedBlocks + b]).
template <int NV>
__device__ __forceinline__ void block_partials(double (&mine)[NV], double* parts) {
  static_assert(NV <= 4, "at most 4 reduced values per kernel");
  __shared__ double red[NV][16];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    double s = wave_sum(mine[v]);
    if ((threadIdx.x & 63) == 0) red[v][threadIdx.x >> 6] = s;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
#pragma unroll
    for (int v = 0; v < NV; ++v) {
      double s = 0.0;
      for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[v][w];
      parts[v * kRedBlocks + blockIdx.x] = s;
    }
  }
}

// Fixed-order sums of a producer's nb partials, the result in every thread (a
// one-block finisher, or every block of a consumer kernel).
template <int NV>
__device__ __forceinline__ void finish_sums(const double* __restrict__ parts, int nb, double (&out)[NV]) {
  __shared__ double red[NV][16];
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) s += parts[v * kRedBlocks + b];
    s = wave_sum(s);
    if ((threadIdx.x & 63) == 0) red[v][threadIdx.x >> 6] = s;
  }
  __syncthreads();
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[v][w];
    out[v] = t;
  }
}

// Jacobi diagonal: dinv[i] = 1 / a_ii, or 1 where the diagonal is absent or zero.
// SELL storage (FMT): W wavefronts per slice, max(4 / W, 1) slices per block.
// Wavefront g of a slice scans chunks g, g + W, ... (eight column positions each,
// the eight loads issued together) of the slice's 64 rows and keeps the value at
// column i; a row's columns are distinct and padding entries carry value 0, so
// at most one nonzero candidate exists and the first nonzero one in wavefront
// order is it.  CSR: bisection in the sorted row, one thread per row.
constexpr int kDiagMaxWaves = 16;

template <typename T, typename TV, int FMT, int W>
__global__ __launch_bounds__(64 * (W < 4 ? 4 : W)) void diag_inv_sell_kernel(MatArgs A, const TV* __restrict__ vals,
                                                                            T* __restrict__ dinv) {
  constexpr int WPB = W < 4 ? 4 : W, SPB = WPB / W;
  __shared__ T cand[WPB][kSlice];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, g = wave % W;
  const int s = blockIdx.x * SPB + wave / W;  // slices per block: SPB, the grid covers them
  const int i = s * kSlice + lane;
  T d = 0;
  if (s < A.n_slices && i < A.n) {
    const int base = A.ptr[s], len = (A.ptr[s + 1] - base) / kSlice;
    const uint16_t* c16 = reinterpret_cast<const uint16_t*>(A.col);
    const int cb = FMT == kFmtSell16 ? A.cbase[s] : 0;
    for (int j0 = 8 * g; j0 < len; j0 += 8 * W) {
      int cj[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = base + min(j0 + u, len - 1) * kSlice + lane;  // in the slice: j0 < len
        cj[u] = FMT == kFmtSell16 ? cb + (int)c16[e] : A.col[e];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u)
        if (d == (T)0 && j0 + u < len && cj[u] == i) d = (T)vals[base + (j0 + u) * kSlice + lane];
    }
  }
  cand[wave][lane] = d;
  __syncthreads();
  if (g == 0 && s < A.n_slices && i < A.n) {
    for (int h = 1; h < W && d == (T)0; ++h) d = cand[wave + h][lane];
    dinv[i] = (d != (T)0) ? (T)1 / d : (T)1;
  }
}

// SELL3: one thread per node row k (lane of a slice) finds its block at node
// column k (eight columns in flight; the real block comes before the padding,
// whose column is also k) and takes its values 0, 4 and 8.
template <typename T, typename TV>
__global__ __launch_bounds__(256) void diag_inv_sell3_kernel(MatArgs A, const TV* __restrict__ vals,
                                                             T* __restrict__ dinv) {
  const int lane = threadIdx.x & 63, s = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (s >= A.n_slices) return;
  const int k = s * kSlice + lane;
  if (3 * k >= A.n) return;
  const int base = A.ptr[s], len = (A.ptr[s + 1] - base) / kSlice, cb = A.cbase[s];
  const uint16_t* c16 = reinterpret_cast<const uint16_t*>(A.col);
  int jd = -1;
  for (int j0 = 0; j0 < len && jd < 0; j0 += 8) {
    int cj[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) cj[u] = cb + (int)c16[base + min(j0 + u, len - 1) * kSlice + lane];
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (jd < 0 && j0 + u < len && cj[u] == k) jd = j0 + u;
  }
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    const T d = jd >= 0 ? (T)vals[(size_t)9 * (base + jd * kSlice) + 4 * c * kSlice + lane] : (T)0;
    dinv[3 * k + c] = (d != (T)0) ? (T)1 / d : (T)1;
  }
}

template <typename T, typename TV = T>
__global__ void diag_inv_csr_kernel(MatArgs A, const TV* __restrict__ vals, T* __restrict__ dinv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= A.n) return;
  T d = 0;
  int lo = A.ptr[i], hi = A.ptr[i + 1];
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (A.col[mid] < i) lo = mid + 1; else hi = mid;
  }
  if (lo < A.ptr[i + 1] && A.col[lo] == i) d = (T)vals[lo];
  dinv[i] = (d != (T)0) ? (T)1 / d : (T)1;
}

// Wavefronts per slice from the mean stored row length (`stored` entries over the
// slices): about two chunks of eight per wavefront, at most kDiagMaxWaves.
template <typename T, typename TV, int FMT>
static void launch_diag_inv_sell(const MatArgs& A, long long stored, const TV* vals, T* dinv, hipStream_t s) {
  const long long chunks = stored / (8LL * kSlice * std::max(A.n_slices, 1));
  const int w = chunks >= 24 ? 16 : chunks >= 12 ? 8 : chunks >= 6 ? 4 : chunks >= 3 ? 2 : 1;
  auto go = [&](auto wc) {
    constexpr int W = decltype(wc)::value, SPB = (W < 4 ? 4 : W) / W;
    diag_inv_sell_kernel<T, TV, FMT, W><<<div_up(A.n_slices, SPB), 64 * (W < 4 ? 4 : W), 0, s>>>(A, vals, dinv);
  };
  switch (w) {
    case 16: go(std::integral_constant<int, 16>{}); break;
    case 8: go(std::integral_constant<int, 8>{}); break;
    case 4: go(std::integral_constant<int, 4>{}); break;
    case 2: go(std::integral_constant<int, 2>{}); break;
    default: go(std::integral_constant<int, 1>{}); break;
  }
}

template <typename T, typename TV = T>
static void launch_diag_inv(const MatArgs& A, long long stored, const TV* vals, T* dinv, hipStream_t s) {
  if (A.fmt == kFmtSell3)
    diag_inv_sell3_kernel<T, TV><<<div_up(A.n_slices, 4), 256, 0, s>>>(A, vals, dinv);
  else if (A.fmt == kFmtSell16)
    launch_diag_inv_sell<T, TV, kFmtSell16>(A, stored, vals, dinv, s);
  else if (A.fmt == MIFE_FMT_SELL)
    launch_diag_inv_sell<T, TV, MIFE_FMT_SELL>(A, stored, vals, dinv, s);
  else
    diag_inv_csr_kernel<T, TV><<<div_up(A.n, 256), 256, 0, s>>>(A, vals, dinv);
}

// r = b - A x (x == NULL: x = 0), z = D^-1 r, p = q = 0; reduces rz, rr, bb and
// initialises the state.
template <typename T, int FMT, int L>
__global__ __launch_bounds__(256) void cg_init_kernel(MatArgs A, const T* __restrict__ vals,
                                                      const T* __restrict__ b, const T* __restrict__ x,
                                                      const T* __restrict__ dinv, T* __restrict__ r,
                                                      T* __restrict__ z, T* __restrict__ p, T* __restrict__ q,
                                                      CgState* st) {
  double m[3] = {0.0, 0.0, 0.0};
  for_each_row<T, FMT, L>(A, vals, x, [&](int row, T ax) {
    const T ri = b[row] - ax;
    const T zi = dinv[row] * ri;
    r[row] = ri;
    z[row] = zi;
    p[row] = 0;
    q[row] = 0;
    m[0] += (double)ri * (double)zi;
    m[1] += (double)ri * (double)ri;
    m[2] += (double)b[row] * (double)b[row];
  });
  block_partials<3>(m, st->part_b);
}

__global__ void fin_init_kernel(CgState* st, int nb, double tol2, int maxit) {
  double out[3];
  finish_sums<3>(st->part_b, nb, out);
  if (threadIdx.x == 0) {
    st->rz = out[0];
    st->rr = out[1];
    st->bb = out[2];
    st->maxrr = out[1];
    st->beta = 0.0;
    st->alpha = 0.0;
    st->iters = 0;
    st->maxit = maxit;
    st->tol2 = tol2;
    st->pause = 0;
    st->done = (out[1] <= tol2 * out[2]) ? 1 : 0;
  }
}

template <typename T, int FMT, int L>
__global__ __launch_bounds__(256) void cg_spmv_kernel(MatArgs A, const T* __restrict__ vals,
                                                      const T* __restrict__ z, T* __restrict__ p,
                                                      T* __restrict__ q, CgState* st) {
  if (st->done || st->pause) return;
  const T beta = (T)st->beta;
  double m[1] = {0.0};
  for_each_row<T, FMT, L>(A, vals, z, [&](int row, T az) {
    const T pi = z[row] + beta * p[row];
    const T qi = az + beta * q[row];
    p[row] = pi;
    q[row] = qi;
    m[0] += (double)pi * (double)qi;
  });
  block_partials<1>(m, st->part_a);
}

// alpha = rz / (p.q) from K_A's nb_a partials (summed in every block), then
// x += alpha p, r -= alpha q, z = D^-1 r and the r.z, r.r partials.  Block 0
// records pq and alpha, which no block of this kernel reads.  Each thread takes
// R = 16 / sizeof(T) consecutive rows through 16-byte vector accesses (R = 1
// when a vector is not 16-byte aligned); rows past the last full group are
// taken one per thread.
template <typename T, int R>
__global__ __launch_bounds__(kUpdateThreads) void cg_update_kernel(int n, int nb_a, T* __restrict__ x,
                                                                   T* __restrict__ r, T* __restrict__ z,
                                                                   const T* __restrict__ p,
                                                                   const T* __restrict__ q,
                                                                   const T* __restrict__ dinv, CgState* st) {
  if (st->done || st->pause) return;
  double pq[1];
  finish_sums<1>(st->part_a, nb_a, pq);
  const double alpha64 = st->rz / pq[0];
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->pq = pq[0];
    st->alpha = alpha64;
  }
  const T alpha = (T)alpha64;
  double m[2] = {0.0, 0.0};
  auto row = [&](T& xi, T& ri, T& zi, T pi, T qi, T di) {
    xi += alpha * pi;
    ri = ri - alpha * qi;
    zi = di * ri;
    m[0] += (double)ri * (double)zi;
    m[1] += (double)ri * (double)ri;
  };
  const int nv = n / R;
  const int tid = blockIdx.x * blockDim.x + threadIdx.x, nth = gridDim.x * blockDim.x;
  if constexpr (R > 1) {
    typedef T VT __attribute__((ext_vector_type(R)));
    for (int v = tid; v < nv; v += nth) {
      VT xv = reinterpret_cast<VT*>(x)[v], rv = reinterpret_cast<VT*>(r)[v], zv;
      const VT pv = reinterpret_cast<const VT*>(p)[v], qv = reinterpret_cast<const VT*>(q)[v];
      const VT dv = reinterpret_cast<const VT*>(dinv)[v];
#pragma unroll
      for (int j = 0; j < R; ++j) {
        T xi = xv[j], ri = rv[j], zi;
        row(xi, ri, zi, pv[j], qv[j], dv[j]);
        xv[j] = xi;
        rv[j] = ri;
        zv[j] = zi;
      }
      reinterpret_cast<VT*>(x)[v] = xv;
      reinterpret_cast<VT*>(r)[v] = rv;
      reinterpret_cast<VT*>(z)[v] = zv;
    }
  }
  for (int i = (R > 1 ? nv * R : 0) + tid; i < n; i += nth) {
    T xi = x[i], ri = r[i], zi;
    row(xi, ri, zi, p[i], q[i], dinv[i]);
    x[i] = xi;
    r[i] = ri;
    z[i] = zi;
  }
  block_partials<2>(m, st->part_b);
}

__global__ void fin_beta_kernel(CgState* st, int nb) {
  if (st->done || st->pause) return;
  double out[2];
  finish_sums<2>(st->part_b, nb, out);
  if (threadIdx.x == 0) {
    st->rz_prev = st->rz;
    st->beta = out[0] / st->rz;
    st->rz = out[0];
    st->rr = out[1];
    st->iters += 1;
    st->maxrr = fmax(st->maxrr, out[1]);
    st->done = (out[1] <= st->tol2 * st->bb || st->iters >= st->maxit) ? 1 : 0;
    st->pause = (st->seg_rr > 0.0 && out[1] <= st->seg_rr) ? 1 : 0;
  }
}

// FP64 residual r = b - A x with ||r||^2 and ||b||^2 (values FP64, or FP32 read
// into FP64 arithmetic).
template <int FMT, int L, typename TV = double>
__global__ __launch_bounds__(256) void residual64_kernel(MatArgs A, const TV* __restrict__ vals,
                                                         const double* __restrict__ b, const double* __restrict__ x,
                                                         double* __restrict__ r, CgState* st) {
  double m[2] = {0.0, 0.0};
  for_each_row<double, FMT, L, TV>(A, vals, x, [&](int row, double ax) {
    const double ri = b[row] - ax;
    if (r) r[row] = ri;
    m[0] += ri * ri;
    m[1] += b[row] * b[row];
  });
  block_partials<2>(m, st->part_a);
}

__global__ void fin_residual_kernel(CgState* st, int nb) {
  double out[2];
  finish_sums<2>(st->part_a, nb, out);
  if (threadIdx.x == 0) {
    st->rr64 = out[0];
    st->rr = out[0];
    st->bb = out[1];
    st->scale = sqrt(out[0]);
  }
}

// dst[k] = src[perm[k]] (gather into slice numbering); identity when perm == NULL
template <typename T>
__global__ void permute_in_kernel(int n, const int32_t* __restrict__ perm, const T* __restrict__ src,
                                  T* __restrict__ dst) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
    dst[k] = src[perm ? perm[k] : k];
}

// dst[perm[k]] = src[k]
template <typename T>
__global__ void permute_out_kernel(int n, const int32_t* __restrict__ perm, const T* __restrict__ src,
                                   T* __restrict__ dst) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
    dst[perm ? perm[k] : k] = src[k];
}

// Reliable update, part 1: x64 += x32 ; x32 = 0.
__global__ void fold_correction_kernel(int n, double* __restrict__ x64, float* __restrict__ x32) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    x64[i] += (double)x32[i];
    x32[i] = 0.0f;
  }
}

// Reliable update, part 2 (after residual64 into r64 and `ro`):
// r32 = r64, z = D^-1 r32; the FP32 recurrence restarts from the true residual
// while p and q = A p are kept, beta = (r.z)_new / (r.z)_before-the-step (or 0
// when the replaced r.r exceeds the recurrence's by kRestartGrowth).
__global__ __launch_bounds__(256) void replace_residual_kernel(int n, const double* __restrict__ r64,
                                                               float* __restrict__ r, float* __restrict__ z,
                                                               const float* __restrict__ dinv, CgState* st) {
  double m[2] = {0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const float ri = (float)r64[i];
    const float zi = dinv[i] * ri;
    r[i] = ri;
    z[i] = zi;
    m[0] += (double)ri * (double)zi;
    m[1] += (double)ri * (double)ri;
  }
  block_partials<2>(m, st->part_b);
}

// FP64 restart from the true residual r (already in r): z = D^-1 r; partials r.z, r.r.
__global__ __launch_bounds__(256) void restart_residual64_kernel(int n, const double* __restrict__ r,
                                                                 double* __restrict__ z,
                                                                 const double* __restrict__ dinv, CgState* st) {
  double m[2] = {0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double zi = dinv[i] * r[i];
    z[i] = zi;
    m[0] += r[i] * zi;
    m[1] += r[i] * r[i];
  }
  block_partials<2>(m, st->part_b);
}

// Also keeps the recurrence's r.r from before the replacement (rr_seg) and the
// FP64 residual (rr64), so one read of st gives the host the whole cycle.
__global__ void fin_replace_kernel(CgState* st, const CgState* ro, int nb, int first, double tol2) {
  double out[2];
  finish_sums<2>(st->part_b, nb, out);
  if (threadIdx.x == 0) {
    st->rr_seg = st->rr;
    st->rr64 = ro->rr64;
    st->beta = (first || out[1] > kRestartGrowth * st->rr) ? 0.0 : out[0] / st->rz_prev;
    st->rz = out[0];
    st->rr = out[1];
    st->maxrr = out[1];
    st->bb = ro->bb;
    st->pause = 0;
    // the next segment ends when the recurrence has fallen by seg_drop, or at a
    // quarter of the target r.r when that is nearer
    st->seg_rr = st->seg_drop > 0.0 ? fmax(st->seg_drop * out[1], 0.25 * tol2 * ro->bb) : 0.0;
    st->done = (ro->rr64 <= tol2 * ro->bb) ? 1 : 0;
  }
}

template <typename T>
__global__ void fill_kernel(int n, T* __restrict__ v, T val) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v[i] = val;
}

// ------------------------------------------------------------------ GMRES
// Classical Gram-Schmidt of the new vector w against the basis V_0..V_{m-1}
// (columns of V, stride ldv), two passes per iteration.  On the GMRES systems
// here n is a few 10^4 rows and m up to 257, so each pass is a matrix-vector
// product of a few 10 MB that must spread over the whole chip: the dots take a
// block per (basis vector, eighth of the rows), whose partials every consumer
// sums in chunk order, and the update takes one block per 64 rows with its 8
// wavefronts splitting the basis vectors, their partial sums combined in
// wavefront order.
constexpr int kDotChunks = 8;
constexpr int kAxpyWaves = 8;

// parts[j kDotChunks + c] = V_j . w over row chunk c; block (c, j).
__global__ __launch_bounds__(256) void gs_dots_kernel(int n, const double* __restrict__ V, size_t ldv,
                                                      const double* __restrict__ w, double* __restrict__ parts) {
  __shared__ double red[4];
  const int c = blockIdx.x, j = blockIdx.y;
  const int per = (n + kDotChunks - 1) / kDotChunks;
  const int i1 = min(n, (c + 1) * per);
  const double* Vj = V + (size_t)j * ldv;
  double s0 = 0.0, s1 = 0.0;
  int i = c * per + threadIdx.x;
  for (; i + 256 < i1; i += 512) {
    s0 += Vj[i] * w[i];
    s1 += Vj[i + 256] * w[i + 256];
  }
  if (i < i1) s0 += Vj[i] * w[i];
  const double t = block_sum(s0 + s1, red);
  if (threadIdx.x == 0) parts[j * kDotChunks + c] = t;
}

// Coefficient j from its chunk partials (fixed order).
__device__ __forceinline__ double dot_coef(const double* __restrict__ parts, int j) {
  double s = 0.0;
#pragma unroll
  for (int c = 0; c < kDotChunks; ++c) s += parts[j * kDotChunks + c];
  return s;
}

// w -= sum_{j < m} V_j h_j with h from the dot partials hp; NORM: also the block
// partials of the updated w.w into nparts[blockIdx.x] (blocks loop over 64-row
// tiles).
template <bool NORM>
__global__ __launch_bounds__(64 * kAxpyWaves) void gs_axpy_kernel(int n, int m, const double* __restrict__ V,
                                                                  size_t ldv, const double* __restrict__ hp,
                                                                  double* __restrict__ w,
                                                                  double* __restrict__ nparts) {
  __shared__ double hs[kMaxRestart + 1];
  __shared__ double part[kAxpyWaves][64];
  const int lane = threadIdx.x & 63, q = threadIdx.x >> 6;
  for (int j = threadIdx.x; j < m; j += blockDim.x) hs[j] = dot_coef(hp, j);
  __syncthreads();
  // wavefront q takes basis vectors [j0, j1)
  const int per = (m + kAxpyWaves - 1) / kAxpyWaves;
  const int j0 = min(m, q * per), j1 = min(m, j0 + per);
  double nrm = 0.0;
  for (int r0 = blockIdx.x * 64; r0 < n; r0 += gridDim.x * 64) {
    const int row = r0 + lane;
    double a0 = 0.0, a1 = 0.0;
    if (row < n) {
      int j = j0;
      for (; j + 1 < j1; j += 2) {
        a0 += V[(size_t)j * ldv + row] * hs[j];
        a1 += V[(size_t)(j + 1) * ldv + row] * hs[j + 1];
      }
      if (j < j1) a0 += V[(size_t)j * ldv + row] * hs[j];
    }
    part[q][lane] = a0 + a1;
    __syncthreads();
    if (q == 0 && row < n) {
      double acc = 0.0;
#pragma unroll
      for (int k = 0; k < kAxpyWaves; ++k) acc += part[k][lane];
      const double wi = w[row] - acc;
      w[row] = wi;
      nrm += wi * wi;
    }
    __syncthreads();
  }
  if (NORM && q == 0) {
    const double t = wave_sum(nrm);
    if (lane == 0) nparts[blockIdx.x] = t;
  }
}

// Device state of one GMRES restart cycle.
struct GmresCycle {
  double hk1;  // norm of the newest basis vector before scaling (0: breakdown)
  int done;    // converged, broken down or out of iterations: the cycle stops
  int k_used;  // Hessenberg columns formed
};

// Column k of the Hessenberg matrix from the two Gram-Schmidt passes (their dot
// partials h1, h2) and the new
// vector's norm, the previous Givens rotations applied to it, rotation k formed
// and g rotated; H is (m+1) x m, row-major.  The column and the rotations are
// staged in LDS by the block, thread 0 runs the sequential rotations, and the
// block writes the column back.  Sets done when |g[k+1]| <= tol_abs, on
// breakdown, or when this is the last iteration allowed.
__global__ __launch_bounds__(256) void gmres_givens_kernel(int k, int m, const double* __restrict__ h1,
                                                           const double* __restrict__ h2,
                                                           const double* __restrict__ nparts, int nb,
                                                           double* __restrict__ H, double* __restrict__ cs,
                                                           double* __restrict__ sn, double* __restrict__ g,
                                                           GmresCycle* cy, double tol_abs, int last) {
  __shared__ double col[kMaxRestart + 1];
  __shared__ double c_s[kMaxRestart], s_s[kMaxRestart];
  __shared__ double red[4];
  if (cy->done) return;
  // ||w||^2 from the update's per-block partials (fixed order)
  double q = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) q += nparts[b];
  const double nrm2 = block_sum(q, red);
  for (int j = threadIdx.x; j <= k; j += blockDim.x) col[j] = dot_coef(h1, j) + (h2 ? dot_coef(h2, j) : 0.0);
  for (int j = threadIdx.x; j < k; j += blockDim.x) {
    c_s[j] = cs[j];
    s_s[j] = sn[j];
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const double hk1 = sqrt(fmax(nrm2, 0.0));
    for (int j = 0; j < k; ++j) {
      const double a = col[j], b = col[j + 1];
      col[j] = c_s[j] * a + s_s[j] * b;
      col[j + 1] = -s_s[j] * a + c_s[j] * b;
    }
    const double hkk = col[k];
    const double den = hypot(hkk, hk1);
    const double c = den > 0.0 ? hkk / den : 1.0, sv = den > 0.0 ? hk1 / den : 0.0;
    cs[k] = c;
    sn[k] = sv;
    col[k] = den;
    const double gk = g[k];
    g[k + 1] = -sv * gk;
    g[k] = c * gk;
    cy->hk1 = hk1;
    cy->k_used = k + 1;
    if (fabs(sv * gk) <= tol_abs || hk1 == 0.0 || last) cy->done = 1;
  }
  __syncthreads();
  for (int j = threadIdx.x; j <= k; j += blockDim.x) H[j * m + k] = col[j];
}

// Start of a restart cycle: V_0 = r / beta and t = D^-1 V_0 (the first product's
// input), g = (beta, 0, ...), state cleared.
__global__ void gmres_cycle_init_kernel(int n, int m, const double* __restrict__ r, double* __restrict__ V0,
                                        const double* __restrict__ dinv, double* __restrict__ t, double beta,
                                        double* __restrict__ g, GmresCycle* cy) {
  const int tid = blockIdx.x * blockDim.x + threadIdx.x;
  for (int i = tid; i < n; i += gridDim.x * blockDim.x) {
    const double v = r[i] / beta;
    V0[i] = v;
    t[i] = dinv[i] * v;
  }
  for (int j = tid; j <= m; j += gridDim.x * blockDim.x) g[j] = j == 0 ? beta : 0.0;
  if (tid == 0) {
    cy->hk1 = 0.0;
    cy->done = 0;
    cy->k_used = 0;
  }
}

// V_{k+1} = w / hk1 and t = D^-1 V_{k+1} (the next product's input) while the
// cycle runs (nothing once it has stopped).
__global__ void gmres_next_basis_kernel(int n, const double* __restrict__ w, double* __restrict__ v,
                                        const double* __restrict__ dinv, double* __restrict__ t,
                                        const GmresCycle* cy) {
  if (cy->done) return;
  const double a = cy->hk1 > 0.0 ? 1.0 / cy->hk1 : 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double vi = a * w[i];
    v[i] = vi;
    t[i] = dinv[i] * vi;
  }
}

// x += dinv * (V y)
__global__ void gmres_update_kernel(int n, int m, const double* __restrict__ V, size_t ldv,
                                    const double* __restrict__ y, const double* __restrict__ dinv,
                                    double* __restrict__ x) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int j = 0; j < m; ++j) s += V[j * ldv + i] * y[j];
    x[i] += dinv[i] * s;
  }
}

// ------------------------------------------------------------------ MINRES
// Preconditioned MINRES (Paige & Saunders 1975) for symmetric indefinite
// systems with an SPD diagonal preconditioner: two launches per iteration,
// every scalar on the device, no finisher kernels.  Each block of a launch
// reduces the previous launch's per-block partials itself, in the same fixed
// order (the grids are small), so all blocks hold identical scalars; block 0
// stores them.  The scalars are double-buffered by iteration parity: launch k
// reads s[k&1] and block 0 writes s[(k+1)&1], so no block reads what another
// block of the same launch writes.
//   A(k): reduce r2.z -> beta_k; (k>0) Givens rotation of step k-1 and the
//         w/x update of step k-1; v = z/beta_k; t = A v - (beta_k/beta_{k-1}) r1;
//         partial v.t
//   B(k): reduce v.t -> alfa_k; t -= (alfa/beta) r2; r1 = r2; r2 = t;
//         z = D^-1 r2; partial r2.z
struct MinresScalars {
  double beta, oldb, beta1, alfa, dbar, epsln, phibar, cs, sn, tol, bb;
  int iters, maxit, done, pad;
};

struct MinresState {
  MinresScalars s[2];
  double part_a[kRedBlocks];
  double part_b[2 * kRedBlocks];
};

// Fixed-order sum of nb partials, computed by every block (all threads get it).
__device__ __forceinline__ double block_fixed_sum(const double* __restrict__ parts, int nb) {
  __shared__ double red[4];
  double acc = 0.0;
  for (int i = threadIdx.x; i < nb; i += blockDim.x) acc += parts[i];
  return block_sum(acc, red);
}

template <int FMT, int L, typename TV>
__global__ __launch_bounds__(256) void minres_a_kernel(MatArgs A, const TV* __restrict__ vals,
                                                       const double* __restrict__ z, double* __restrict__ v,
                                                       double* __restrict__ t, const double* __restrict__ r1,
                                                       double* __restrict__ w, double* __restrict__ w1,
                                                       double* __restrict__ w2, double* __restrict__ x,
                                                       MinresState* st, int k, int nb) {
  MinresScalars S = st->s[k & 1];
  if (S.done) {  // carry the final state forward so every later launch sees it
    if (blockIdx.x == 0 && threadIdx.x == 0) st->s[(k + 1) & 1] = S;
    return;
  }
  const double rz = block_fixed_sum(st->part_b, nb);
  const double beta = sqrt(fmax(rz, 0.0));
  bool pend = false;
  double phi = 0.0, denom = 0.0, delta = 0.0, oldeps = 0.0;
  if (k == 0) {
    const double bb = block_fixed_sum(st->part_b + kRedBlocks, nb);
    S.beta1 = beta;
    S.oldb = 0.0;
    S.phibar = beta;
    S.cs = -1.0;
    S.sn = 0.0;
    S.dbar = 0.0;
    S.epsln = 0.0;
    S.bb = bb;
    S.iters = 0;
    S.done = (beta == 0.0 || S.maxit <= 0) ? 1 : 0;
  } else {
    oldeps = S.epsln;
    delta = S.cs * S.dbar + S.sn * S.alfa;
    const double gbar = S.sn * S.dbar - S.cs * S.alfa;
    S.epsln = S.sn * beta;
    S.dbar = -S.cs * beta;
    const double gamma = fmax(sqrt(gbar * gbar + beta * beta), 1e-300);
    S.cs = gbar / gamma;
    S.sn = beta / gamma;
    phi = S.cs * S.phibar;
    S.phibar = S.sn * S.phibar;
    denom = 1.0 / gamma;
    S.oldb = S.beta;
    S.iters += 1;
    S.done = (S.phibar <= S.tol * S.beta1 || S.iters >= S.maxit || beta == 0.0) ? 1 : 0;
    pend = true;
  }
  S.beta = beta;
  const bool done = S.done != 0;
  const double ib = beta > 0.0 ? 1.0 / beta : 0.0, rb = S.oldb != 0.0 ? beta / S.oldb : 0.0;
  double m[1] = {0.0};
  for_each_row<double, FMT, L, TV>(A, vals, done ? nullptr : z, [&](int row, double az) {
    if (pend) {
      const double wo = w[row], w2o = w2[row];
      const double wn = (v[row] - oldeps * w2o - delta * wo) * denom;
      w1[row] = w2o;
      w2[row] = wo;
      w[row] = wn;
      x[row] += phi * wn;
    }
    if (!done) {
      const double vi = z[row] * ib;
      const double ti = az * ib - rb * r1[row];
      v[row] = vi;
      t[row] = ti;
      m[0] += vi * ti;
    }
  });
  if (!done) {
    __shared__ double red[4];
    double sm = block_sum(m[0], red);
    if (threadIdx.x == 0) st->part_a[blockIdx.x] = sm;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) st->s[(k + 1) & 1] = S;
}

__global__ __launch_bounds__(256) void minres_b_kernel(int n, double* __restrict__ t, double* __restrict__ r1,
                                                       double* __restrict__ r2, double* __restrict__ z,
                                                       const double* __restrict__ dinv, MinresState* st, int k,
                                                       int nb) {
  const int sdone = st->s[(k + 1) & 1].done;
  const double sbeta = st->s[(k + 1) & 1].beta;
  if (sdone) return;
  const double alfa = block_fixed_sum(st->part_a, nb);
  const double ab = alfa / sbeta;
  double m = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double r2o = r2[i];
    const double ti = t[i] - ab * r2o;
    const double zi = dinv[i] * ti;
    r1[i] = r2o;
    r2[i] = ti;
    z[i] = zi;
    m += ti * zi;
  }
  __shared__ double red[4];
  m = block_sum(m, red);
  if (threadIdx.x == 0) st->part_b[blockIdx.x] = m;
  if (blockIdx.x == 0 && threadIdx.x == 0) st->s[(k + 1) & 1].alfa = alfa;
}

// r1 = r2 = b - A x, z = D^-1 r1, w = w2 = 0 ; partials r1.z and b.b ; scalars reset
template <int FMT, int L, typename TV>
__global__ __launch_bounds__(256) void minres_init_kernel(MatArgs A, const TV* __restrict__ vals,
                                                          const double* __restrict__ b, const double* __restrict__ x,
                                                          const double* __restrict__ dinv, double* __restrict__ r1,
                                                          double* __restrict__ r2, double* __restrict__ z,
                                                          double* __restrict__ w, double* __restrict__ w2,
                                                          MinresState* st, double tol, int maxit) {
  double m0 = 0.0, m1 = 0.0;
  for_each_row<double, FMT, L, TV>(A, vals, x, [&](int row, double ax) {
    const double ri = b[row] - ax;
    const double zi = dinv[row] * ri;
    r1[row] = ri;
    r2[row] = ri;
    z[row] = zi;
    w[row] = 0.0;
    w2[row] = 0.0;
    m0 += ri * zi;
    m1 += b[row] * b[row];
  });
  __shared__ double red[4];
  m0 = block_sum(m0, red);
  m1 = block_sum(m1, red);
  if (threadIdx.x == 0) {
    st->part_b[blockIdx.x] = m0;
    st->part_b[kRedBlocks + blockIdx.x] = m1;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    MinresScalars S{};
    S.tol = tol;
    S.maxit = maxit;
    st->s[0] = S;
    st->s[1] = S;
  }
}

// Preconditioner diagonal: pc[i] != 0 replaces the Jacobi entry of row i (given in
// the caller's row order; permuted to slice numbering here).  absval: |1/a_ii|.
__global__ void pc_override_kernel(int n, const int32_t* __restrict__ perm, const double* __restrict__ pc,
                                   double* __restrict__ dinv, int absval) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x) {
    double d = dinv[k];
    if (pc) {
      const double o = pc[perm ? perm[k] : k];
      if (o != 0.0) d = o;
    }
    dinv[k] = absval ? fabs(d) : d;
  }
}

// Zeroes a solver state (words of 8 bytes) with a kernel rather than a memset
// node, so that the clear is an ordinary kernel on the caller's stream.
__global__ void zero_words_kernel(uint64_t* __restrict__ p, size_t n_words) {
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n_words; i += (size_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

// ------------------------------------------------------------- host side
// Bump allocator over the caller's workspace.  A layout is taken in full before
// any launch; with base == NULL it only counts, which is how
// mife_solver_work_bytes sizes the workspace from the same layout code.
struct Work {
  char* base;
  size_t off;
  size_t cap;
  template <typename T>
  T* take(size_t count) {
    const uintptr_t b = reinterpret_cast<uintptr_t>(base);
    off = ((b + off + 255) & ~uintptr_t(255)) - b;
    T* p = reinterpret_cast<T*>(b + off);
    off += count * sizeof(T);
    return p;
  }
  bool fits() const { return off <= cap; }
};

static int check_fit(const Work& w, const char* who) {
  MIFE_REQUIRE(w.fits(), MIFE_ERR_ARG, "%s needs a workspace of %zu bytes, %zu given (mife_solver_work_bytes)", who,
               w.off, w.cap);
  return MIFE_OK;
}

template <typename T>
static int zero_state(T* st, hipStream_t s) {
  static_assert(sizeof(T) % 8 == 0, "state size is a multiple of 8 bytes");
  zero_words_kernel<<<64, 256, 0, s>>>(reinterpret_cast<uint64_t*>(st), sizeof(T) / 8);
  MIFE_CHECK_HIP(hipGetLastError());
  return MIFE_OK;
}

static int vec_grid(int n) { return std::max(1, std::min(div_up(n, 256), kMaxGrid)); }

static bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15) == 0; }

struct Mat {
  MatArgs a;
  const int32_t* perm;
  int L;
  long long stored;    // stored entries (SELL: including padding)
  bool graphs = true;  // CG chunks may replay cached HIP graphs (not without allocation)
};

static int make_mat(const mife_matrix* A, Mat* m) {
  MIFE_REQUIRE(A && A->ptr && A->col, MIFE_ERR_POINTER_IS_NULL, "null matrix");
  MIFE_REQUIRE(A->format == MIFE_FMT_CSR || A->format == MIFE_FMT_SELL || A->format == MIFE_FMT_SELL3, MIFE_ERR_ARG,
               "invalid matrix format");
  MIFE_REQUIRE(A->n > 0, MIFE_ERR_ARG, "empty matrix");
  MIFE_REQUIRE(A->format != MIFE_FMT_SELL || (A->perm && A->n_slices == div_up(A->n, kSlice)), MIFE_ERR_ARG,
               "SELL matrix needs perm and n_slices = ceil(n/64)");
  MIFE_REQUIRE(A->format != MIFE_FMT_SELL3 || (A->perm && sell3_shape_ok(A)), MIFE_ERR_ARG,
               "SELL3 matrix needs perm, n = 3 x node rows, n_slices = ceil(n/192) and col_base");
  MIFE_REQUIRE(!A->slice_order || ((uintptr_t)A->slice_order & 15) == 0, MIFE_ERR_ARG,
               "slice_order must be 16-byte aligned");
  m->a = mat_args(A);
  m->perm = A->format != MIFE_FMT_CSR ? A->perm : nullptr;
  m->L = spmv_lanes(A->nnz, A->n);
  m->stored = A->nnz;
  return MIFE_OK;
}

template <typename T>
static void launch_cg_init(const Mat& M, const T* vals, const T* b, const T* x, const T* dinv, T* r, T* z, T* p,
                           T* q, CgState* st, double tol2, int maxit, hipStream_t s) {
  const int tk = prof::begin(s);
  const int grid = row_grid(M.a, M.a.fmt != MIFE_FMT_CSR ? 64 : M.L);
  MIFE_DISPATCH_FMT(M.a, M.L, (cg_init_kernel<T, FMT_, L_><<<grid, 256, 0, s>>>(M.a, vals, b, x, dinv, r, z, p, q,
                                                                                  st)));
  fin_init_kernel<<<1, kFinThreads, 0, s>>>(st, grid, tol2, maxit);
  prof::end(tk, s, prof::kCgInit);
}

// Launches one CG iteration (SpMV kernel, update kernel, finisher); with
// profiling on, the two bandwidth kernels' event tickets are returned through
// tk[0..1] (else -1).
template <typename T>
static void launch_cg_iter(const Mat& M, const T* vals, T* x, T* r, T* z, T* p, T* q, const T* dinv, CgState* st,
                           hipStream_t s, int* tk) {
  const int grid = row_grid(M.a, M.a.fmt != MIFE_FMT_CSR ? 64 : M.L);
  constexpr int R = 16 / sizeof(T);
  const bool vec = aligned16(x) && aligned16(r) && aligned16(z) && aligned16(p) && aligned16(q) && aligned16(dinv);
  const int units = vec ? div_up(M.a.n, R) : M.a.n;
  const int ug = std::max(1, std::min(div_up(units, kUpdateThreads), kMaxGrid));
  tk[0] = prof::begin(s);
  MIFE_DISPATCH_FMT(M.a, M.L, (cg_spmv_kernel<T, FMT_, L_><<<grid, 256, 0, s>>>(M.a, vals, z, p, q, st)));
  prof::end(tk[0], s, sizeof(T) == 8 ? prof::kCgSpmv : prof::kCgSpmv32);
  tk[1] = prof::begin(s);
  if (vec)
    cg_update_kernel<T, R><<<ug, kUpdateThreads, 0, s>>>(M.a.n, grid, x, r, z, p, q, dinv, st);
  else
    cg_update_kernel<T, 1><<<ug, kUpdateThreads, 0, s>>>(M.a.n, grid, x, r, z, p, q, dinv, st);
  prof::end(tk[1], s, sizeof(T) == 8 ? prof::kCgUpdate : prof::kCgUpdate32);
  fin_beta_kernel<<<1, kFinThreads, 0, s>>>(st, ug);
}

template <typename TV = double>
static void launch_residual64(const Mat& M, const TV* vals, const double* b, const double* x, double* r,
                              CgState* st, hipStream_t s) {
  const int grid = row_grid(M.a, M.a.fmt != MIFE_FMT_CSR ? 64 : M.L);
  const int tk = prof::begin(s);
  MIFE_DISPATCH_FMT(M.a, M.L, (residual64_kernel<FMT_, L_, TV><<<grid, 256, 0, s>>>(M.a, vals, b, x, r, st)));
  prof::end(tk, s, prof::kResidual64);
  fin_residual_kernel<<<1, kFinThreads, 0, s>>>(st, grid);
}

struct HostState {
  double rz, pq, alpha, beta, rr, bb, tol2, scale;
  double rz_prev, maxrr, rr_seg, rr64;
  double seg_drop, seg_rr;
  int iters, done, maxit, pause;
};

static int read_state(const CgState* st, HostState* h, hipStream_t s) {
  MIFE_CHECK_HIP(hipMemcpyAsync(h, st, sizeof(HostState), hipMemcpyDeviceToHost, s));
  MIFE_CHECK_HIP(hipStreamSynchronize(s));
  return MIFE_OK;
}

// ------------------------------------------------- CG chunks as HIP graphs
// On small systems an iteration's three kernels take a few microseconds each,
// and launching them from the host costs more than running them.  There a
// chunk of kGraphIters iterations is captured once into a HIP graph (on a
// private stream: the caller's may be the null stream, which cannot be
// captured) and replayed with one launch; the instantiated graphs are cached by
// their kernel arguments.  Only kernels are captured (no memset or copy nodes).
constexpr int kGraphIters = 8;
constexpr int kGraphMaxRows = 131072;  // launch-bound below this many rows
constexpr int kGraphCache = 16;

static bool g_graphs_on = true;
// GMRES classical Gram-Schmidt refinement: 0 never (one pass), 1 always (two)
static int g_gmres_refine = 0;

struct GraphKey {
  int dev, tsize, n, fmt, n_slices, span, split, L;
  const void* ptrs[13];
  bool operator==(const GraphKey& o) const { return std::memcmp(this, &o, sizeof(GraphKey)) == 0; }
};

struct GraphEntry {
  GraphKey key;
  hipGraphExec_t exec;
  unsigned long long used;
};

static std::mutex g_graph_mu;
static std::vector<GraphEntry> g_graphs;
static unsigned long long g_graph_clock = 0;
static hipStream_t g_capture_stream[64];

template <typename T>
static int cg_chunk_graph(const Mat& M, const T* vals, T* x, T* r, T* z, T* p, T* q, const T* dinv, CgState* st,
                          hipGraphExec_t* out) {
  GraphKey key;
  std::memset(&key, 0, sizeof(key));
  MIFE_CHECK_HIP(hipGetDevice(&key.dev));
  MIFE_REQUIRE(key.dev >= 0 && key.dev < 64, MIFE_ERR_ARG, "device %d beyond the 64 supported", key.dev);
  key.tsize = (int)sizeof(T);
  key.n = M.a.n;
  key.fmt = M.a.fmt;
  key.n_slices = M.a.n_slices;
  key.span = M.a.span;
  key.split = M.a.split;
  key.L = M.L;
  const void* ps[13] = {M.a.ptr, M.a.col, M.a.cbase, M.a.order, vals, x, r, z, p, q, dinv, st, nullptr};
  std::memcpy(key.ptrs, ps, sizeof(ps));
  std::lock_guard<std::mutex> lock(g_graph_mu);
  for (auto& e : g_graphs)
    if (e.key == key) {
      e.used = ++g_graph_clock;
      *out = e.exec;
      return MIFE_OK;
    }
  hipStream_t& cs = g_capture_stream[key.dev];
  if (!cs) MIFE_CHECK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
  MIFE_CHECK_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
  int tk[2];
  for (int k = 0; k < kGraphIters; ++k) launch_cg_iter<T>(M, vals, x, r, z, p, q, dinv, st, cs, tk);
  hipGraph_t graph = nullptr;
  const hipError_t ce = hipStreamEndCapture(cs, &graph);
  MIFE_REQUIRE(ce == hipSuccess && graph, MIFE_ERR_HIP, "CG chunk capture failed: %s", hipGetErrorString(ce));
  hipGraphExec_t exec = nullptr;
  const hipError_t ie = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  MIFE_REQUIRE(ie == hipSuccess, MIFE_ERR_HIP, "CG chunk graph instantiation failed: %s", hipGetErrorString(ie));
  if ((int)g_graphs.size() >= kGraphCache) {
    auto old = std::min_element(g_graphs.begin(), g_graphs.end(),
                                [](const GraphEntry& a, const GraphEntry& b) { return a.used < b.used; });
    hipGraphExecDestroy(old->exec);
    g_graphs.erase(old);
  }
  g_graphs.push_back(GraphEntry{key, exec, ++g_graph_clock});
  *out = exec;
  return MIFE_OK;
}

// Launches exactly `count` (<= kMaxIterChunk) CG iterations after iteration it0
// (iterations after a device-side stop return at once), no synchronisation.
// Small systems with profiling off replay the cached graph of kGraphIters
// iterations count / kGraphIters times and launch the rest one by one (the
// same kernels, in the same order).  With profiling on, one iteration in 8
// carries event pairs, which keeps the timing overhead off the solve; their
// tickets go to tickets[0 .. 2 count).
template <typename T>
static int launch_iters(const Mat& M, const T* vals, T* x, T* r, T* z, T* p, T* q, const T* dinv, CgState* st,
                        hipStream_t s, int it0, int count, int* tickets) {
  const bool on = prof::enabled();
  if (!on && g_graphs_on && M.graphs && M.a.n <= kGraphMaxRows) {
    hipGraphExec_t exec = nullptr;
    int rc = cg_chunk_graph<T>(M, vals, x, r, z, p, q, dinv, st, &exec);
    if (rc) return rc;
    for (int c = 0; c < count / kGraphIters; ++c) MIFE_CHECK_HIP(hipGraphLaunch(exec, s));
    for (int k = count / kGraphIters * kGraphIters; k < count; ++k)
      launch_cg_iter<T>(M, vals, x, r, z, p, q, dinv, st, s, &tickets[2 * k]);
    for (int k = 0; k < 2 * count; ++k) tickets[k] = -1;
    MIFE_CHECK_HIP(hipGetLastError());
    return MIFE_OK;
  }
  for (int k = 0; k < count; ++k) {
    const bool sample = on && ((it0 + k) % 8) == 0;
    if (on && !sample) mife_profile_enable(0);
    launch_cg_iter<T>(M, vals, x, r, z, p, q, dinv, st, s, &tickets[2 * k]);
    if (on && !sample) mife_profile_enable(1);
  }
  MIFE_CHECK_HIP(hipGetLastError());
  return MIFE_OK;
}

// Drops the profile tickets of launched iterations the device skipped (at or
// after the iteration count `done_at` it reports).
static void discard_skipped(const int* tickets, int it0, int count, int done_at) {
  for (int k = 0; k < count; ++k)
    if (it0 + k >= done_at) {
      prof::discard(tickets[2 * k]);
      prof::discard(tickets[2 * k + 1]);
    }
}

// Launches `count` CG iterations and reads the state.
template <typename T>
static int run_iters(const Mat& M, const T* vals, T* x, T* r, T* z, T* p, T* q, const T* dinv, CgState* st,
                     hipStream_t s, HostState* hs, int count) {
  count = std::min(count, kMaxIterChunk);
  int tickets[2 * kMaxIterChunk];
  const int it0 = hs->iters;
  int rc = launch_iters<T>(M, vals, x, r, z, p, q, dinv, st, s, it0, count, tickets);
  if (rc) return rc;
  rc = read_state(st, hs, s);
  if (rc) return rc;
  discard_skipped(tickets, it0, count, hs->iters);
  return MIFE_OK;
}

// Iterations the CG residual needs to fall from rr to target at the rate
// (log reduction per iteration, < 0) seen so far; clamped to [lo, hi].
static int predict_iters(double rr, double target, double log_rate, int lo, int hi) {
  if (!(log_rate < 0.0) || !(rr > target)) return lo;
  const double need = std::log(target / rr) / log_rate;
  return (int)std::min((double)hi, std::max((double)lo, std::ceil(need)));
}

// Runs CG until the device state says done, in chunks sized from the observed
// convergence rate so that few launches land after convergence.
template <typename T>
static int run_to_convergence(const Mat& M, const T* vals, T* x, T* r, T* z, T* p, T* q, const T* dinv,
                              CgState* st, hipStream_t s, HostState* hs) {
  int chunk = 16;
  double log_rate = 0.0;
  while (!hs->done) {
    const int it0 = hs->iters;
    const double rr0 = hs->rr;
    int rc = run_iters<T>(M, vals, x, r, z, p, q, dinv, st, s, hs, chunk);
    if (rc) return rc;
    const int ran = hs->iters - it0;
    if (ran > 0 && rr0 > 0.0 && hs->rr > 0.0 && hs->rr < rr0) log_rate = std::log(hs->rr / rr0) / ran;
    chunk = predict_iters(hs->rr, hs->tol2 * hs->bb, log_rate, 2, 64);
  }
  return MIFE_OK;
}

// Workspace layouts.  Every buffer of a solve is taken here, before the first
// launch; mife_solver_work_bytes runs the same functions on a counting Work.
template <typename T>
struct CgBufs {
  CgState *st, *st2;
  T *b, *x, *r, *z, *p, *q, *dinv;
};

template <typename T>
static void cg_layout(Work& w, int n, CgBufs<T>& B) {
  B.st = w.take<CgState>(1);
  B.st2 = w.take<CgState>(1);
  B.b = w.take<T>(n);
  B.x = w.take<T>(n);
  B.r = w.take<T>(n);
  B.z = w.take<T>(n);
  B.p = w.take<T>(n);
  B.q = w.take<T>(n);
  B.dinv = w.take<T>(n);
}

struct RefineBufs {
  CgState *st, *sto;
  double *r64, *b, *x;
  float *x32, *r, *z, *p, *q, *dinv;
};

static void refine_layout(Work& w, int n, RefineBufs& B) {
  B.st = w.take<CgState>(1);
  B.sto = w.take<CgState>(1);
  B.r64 = w.take<double>(n);
  B.b = w.take<double>(n);
  B.x = w.take<double>(n);
  B.x32 = w.take<float>(n);
  B.r = w.take<float>(n);
  B.z = w.take<float>(n);
  B.p = w.take<float>(n);
  B.q = w.take<float>(n);
  B.dinv = w.take<float>(n);
}

struct MinresBufs {
  MinresState* st;
  CgState* rs;
  double *dinv, *b, *x, *z, *v, *t, *r1, *r2, *wv, *w1, *w2;
};

static void minres_layout(Work& w, int n, MinresBufs& B) {
  B.st = w.take<MinresState>(1);
  B.rs = w.take<CgState>(1);
  B.dinv = w.take<double>(n);
  B.b = w.take<double>(n);
  B.x = w.take<double>(n);
  B.z = w.take<double>(n);
  B.v = w.take<double>(n);
  B.t = w.take<double>(n);
  B.r1 = w.take<double>(n);
  B.r2 = w.take<double>(n);
  B.wv = w.take<double>(n);
  B.w1 = w.take<double>(n);
  B.w2 = w.take<double>(n);
}

struct GmresBufs {
  CgState* st;
  double *dinv, *r, *t, *wv, *b, *x, *V, *hdev, *ydev, *ndev, *Hdev, *csdev, *sndev, *gdev;
  GmresCycle* cy;
  size_t ldv;
};

static void gmres_layout(Work& w, int n, int m, GmresBufs& B) {
  B.st = w.take<CgState>(1);
  B.dinv = w.take<double>(n);
  B.r = w.take<double>(n);
  B.t = w.take<double>(n);
  B.wv = w.take<double>(n);
  B.b = w.take<double>(n);
  B.x = w.take<double>(n);
  B.ldv = ((size_t)n + 63) & ~size_t(63);
  B.V = w.take<double>(B.ldv * (m + 1));
  B.hdev = w.take<double>((size_t)(m + 2) * kDotChunks);   // dot partials (first pass)
  B.ydev = w.take<double>((size_t)(m + 2) * kDotChunks);   // dot partials (second pass); y of the update
  B.ndev = w.take<double>(kMaxGrid);
  B.Hdev = w.take<double>((size_t)(m + 1) * m);
  B.csdev = w.take<double>(m + 2);
  B.sndev = w.take<double>(m + 2);
  B.gdev = w.take<double>(m + 2);
  B.cy = w.take<GmresCycle>(1);
}

// Bytes a solve of n rows takes from a workspace at any base address (the
// layout is 256-byte aligned from the base upwards, so 255 bytes of slack).
template <typename F>
static size_t layout_bytes(F&& f) {
  Work w{nullptr, 0, 0};
  f(w);
  return w.off + 255;
}

template <typename T>
static int cg_impl(const Mat& M, const T* vals, const T* b_in, T* x_io, double rtol, int maxit,
                   const CgBufs<T>& B, hipStream_t s, int* iters, double* relres, bool* converged,
                   const double* vals64_check) {
  const int n = M.a.n;
  const int vg = vec_grid(n);
  CgState* st = B.st;
  T *b = B.b, *x = B.x, *r = B.r, *z = B.z, *p = B.p, *q = B.q, *dinv = B.dinv;
  permute_in_kernel<T><<<vg, 256, 0, s>>>(n, M.perm, b_in, b);
  permute_in_kernel<T><<<vg, 256, 0, s>>>(n, M.perm, x_io, x);
  int rc = zero_state(st, s);
  if (rc) return rc;
  launch_diag_inv<T>(M.a, M.stored, vals, dinv, s);
  launch_cg_init<T>(M, vals, b, x, dinv, r, z, p, q, st, rtol * rtol, maxit, s);
  HostState hs{};
  rc = read_state(st, &hs, s);
  if (rc) return rc;
  if (!hs.done) rc = run_to_convergence<T>(M, vals, x, r, z, p, q, dinv, st, s, &hs);
  if (rc) return rc;
  double rel = hs.bb > 0 ? sqrt(hs.rr / hs.bb) : 0.0;
  if constexpr (sizeof(T) == 8) if (vals64_check) {
    // explicit FP64 residual at exit; when rounding has let the recurrence drift
    // below the true residual, restart CG from the true residual (a few times)
    CgState* st2 = B.st2;
    for (int restart = 0;; ++restart) {
      rc = zero_state(st2, s);
      if (rc) return rc;
      launch_residual64(M, vals64_check, (const double*)b, (const double*)x, (double*)r, st2, s);
      HostState h2{};
      rc = read_state(st2, &h2, s);
      if (rc) return rc;
      rel = h2.bb > 0 ? sqrt(h2.rr64 / h2.bb) : 0.0;
      if (rel <= rtol || restart == 4 || hs.iters >= maxit) break;
      restart_residual64_kernel<<<vg, 256, 0, s>>>(n, (const double*)r, (double*)z, (const double*)dinv, st);
      fin_replace_kernel<<<1, kFinThreads, 0, s>>>(st, st2, vg, 1, rtol * rtol);
      rc = read_state(st, &hs, s);
      if (rc) return rc;
      if (!hs.done) rc = run_to_convergence<T>(M, vals, x, r, z, p, q, dinv, st, s, &hs);
      if (rc) return rc;
    }
  }
  permute_out_kernel<T><<<vg, 256, 0, s>>>(n, M.perm, x, x_io);
  MIFE_CHECK_HIP(hipStreamSynchronize(s));
  *iters = hs.iters;
  *relres = rel;
  *converged = rel <= rtol;
  return MIFE_OK;
}

// mife_cg on a workspace of `cap` bytes at `work` (the caller's buffer, or the
// library's own scratch array for the no-allocation entry point).
static int cg_entry(bool graphs, const mife_matrix* A, const void* vals, const void* b, void* x, int prec, double rtol,
                    int maxit, void* work, size_t cap, int* iters, double* relres, void* stream) {
  MIFE_REQUIRE(vals && b && x && work, MIFE_ERR_POINTER_IS_NULL, "null pointer argument");
  MIFE_REQUIRE(prec == MIFE_F64 || prec == MIFE_F32, MIFE_ERR_ARG, "invalid precision %d", prec);
  MIFE_REQUIRE(maxit >= 0 && rtol >= 0.0, MIFE_ERR_ARG, "invalid CG arguments");
  Mat M;
  int rc = make_mat(A, &M);
  M.graphs = graphs;  // graph instantiation may allocate
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  Work w{(char*)work, 0, cap};
  int it = 0;
  double rel = 0.0;
  bool conv = false;
  if (prec == MIFE_F64) {
    CgBufs<double> B;
    cg_layout(w, M.a.n, B);
    if ((rc = check_fit(w, "CG"))) return rc;
    rc = cg_impl<double>(M, (const double*)vals, (const double*)b, (double*)x, rtol, maxit, B, s, &it, &rel, &conv,
                         (const double*)vals);
  } else {
    CgBufs<float> B;
    cg_layout(w, M.a.n, B);
    if ((rc = check_fit(w, "CG"))) return rc;
    rc = cg_impl<float>(M, (const float*)vals, (const float*)b, (float*)x, rtol, maxit, B, s, &it, &rel, &conv,
                        nullptr);
  }
  if (rc) return rc;
  if (iters) *iters = it;
  if (relres) *relres = rel;
  if (!conv) {
    set_error("CG stopped at relative residual %.3e after %d iterations (maximum %d)", rel, it, maxit);
    return MIFE_ERR_NOT_CONVERGED;
  }
  return MIFE_OK;
}

// ------------------------------------------ CG without any allocation
// The workspace of mife_cg_noalloc: a file-scope device array, one copy per GPU
// for the life of the process.  Host code never uses the name `scratch` as an
// address; it takes the per-device address from hipGetSymbolAddress once, with
// that device current, and passes that pointer to the kernels.
constexpr size_t kScratchBytes = size_t(64) << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices];

static int scratch_of_current_device(void** out) {
  int dev = -1;
  MIFE_CHECK_HIP(hipGetDevice(&dev));
  MIFE_REQUIRE(dev >= 0 && dev < kMaxDevices, MIFE_ERR_ARG, "device %d beyond the %d supported", dev, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    MIFE_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    MIFE_REQUIRE(p != nullptr, MIFE_ERR_HIP, "no device address for the scratch array on device %d", dev);
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return MIFE_OK;
}

}  // namespace mife

using namespace mife;

extern "C" {

int mife_graphs_enable(int on) {
  g_graphs_on = on != 0;
  return MIFE_OK;
}

int mife_gmres_cgs_refinement(int mode) {
  MIFE_REQUIRE(mode == 0 || mode == 1, MIFE_ERR_ARG, "CGS refinement mode %d (0 never, 1 always)", mode);
  g_gmres_refine = mode;
  return MIFE_OK;
}

int mife_graphs_cached(void) {
  std::lock_guard<std::mutex> lock(g_graph_mu);
  return (int)g_graphs.size();
}

int mife_cg_noalloc(const mife_matrix* A, const void* vals, const void* b, void* x, int prec, double rtol,
                    int maxit, int* iters, double* relres, void* stream) {
  MIFE_REQUIRE(A && vals && b && x, MIFE_ERR_POINTER_IS_NULL, "null pointer argument");
  MIFE_REQUIRE(prec == MIFE_F64 || prec == MIFE_F32, MIFE_ERR_ARG, "invalid precision %d", prec);
  MIFE_REQUIRE(A->n > 0, MIFE_ERR_ARG, "empty matrix");
  // the layout at any 256-byte aligned base, checked before the device is touched
  const size_t need = prec == MIFE_F64
      ? layout_bytes([&](Work& w) { CgBufs<double> B; cg_layout(w, A->n, B); }) - 255
      : layout_bytes([&](Work& w) { CgBufs<float> B; cg_layout(w, A->n, B); }) - 255;
  MIFE_REQUIRE(need <= kScratchBytes, MIFE_ERR_ARG,
               "mife_cg_noalloc: %d rows need %zu bytes of scratch, more than its %zu bytes", A->n, need,
               kScratchBytes);
  void* base = nullptr;
  int rc = scratch_of_current_device(&base);
  if (rc) return rc;
  return cg_entry(false, A, vals, b, x, prec, rtol, maxit, base, kScratchBytes, iters, relres, stream);
}

size_t mife_solver_work_bytes(int n, int restart) {
  const int nn = std::max(n, 1);
  const int m = std::min(std::max(restart, 1), kMaxRestart);
  size_t bytes = 0;
  bytes = std::max(bytes, layout_bytes([&](Work& w) { CgBufs<double> B; cg_layout(w, nn, B); }));
  bytes = std::max(bytes, layout_bytes([&](Work& w) { CgBufs<float> B; cg_layout(w, nn, B); }));
  bytes = std::max(bytes, layout_bytes([&](Work& w) { RefineBufs B; refine_layout(w, nn, B); }));
  bytes = std::max(bytes, layout_bytes([&](Work& w) { MinresBufs B; minres_layout(w, nn, B); }));
  bytes = std::max(bytes, layout_bytes([&](Work& w) { GmresBufs B; gmres_layout(w, nn, m, B); }));
  return bytes;
}

int mife_cg(const mife_matrix* A, const void* vals, const void* b, void* x, int prec, double rtol, int maxit,
            void* work, size_t work_bytes, int* iters, double* relres, void* stream) {
  return cg_entry(true, A, vals, b, x, prec, rtol, maxit, work, work_bytes, iters, relres, stream);
}

int mife_refine(const mife_matrix* A, const double* vals64, const float* vals32, const double* b_in, double* x_io,
                double rtol, double inner_rtol, int maxit, void* work, size_t work_bytes, int* iters, int* outer,
                double* relres, void* stream) {
  MIFE_REQUIRE(vals64 && vals32 && b_in && x_io && work, MIFE_ERR_POINTER_IS_NULL, "null pointer argument");
  MIFE_REQUIRE(rtol > 0.0 && inner_rtol > 0.0 && inner_rtol < 1.0 && maxit > 0, MIFE_ERR_ARG, "invalid arguments");
  Mat M;
  int rc = make_mat(A, &M);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int n = M.a.n;
  Work w{(char*)work, 0, work_bytes};
  RefineBufs B;
  refine_layout(w, n, B);
  if ((rc = check_fit(w, "mixed-precision CG"))) return rc;
  CgState *st = B.st, *sto = B.sto;
  double *r64 = B.r64, *b = B.b, *x = B.x;
  float *x32 = B.x32, *r = B.r, *z = B.z, *p = B.p, *q = B.q, *dinv = B.dinv;
  const int vg = vec_grid(n);
  permute_in_kernel<double><<<vg, 256, 0, s>>>(n, M.perm, b_in, b);
  permute_in_kernel<double><<<vg, 256, 0, s>>>(n, M.perm, x_io, x);
  if ((rc = zero_state(st, s))) return rc;
  if ((rc = zero_state(sto, s))) return rc;
  launch_diag_inv<float>(M.a, M.stored, vals32, dinv, s);
  fill_kernel<float><<<vg, 256, 0, s>>>(n, x32, 0.0f);
  fill_kernel<float><<<vg, 256, 0, s>>>(n, p, 0.0f);
  fill_kernel<float><<<vg, 256, 0, s>>>(n, q, 0.0f);
  // the inner FP32 recurrence never converges on its own (tol2 = 0 on the
  // device): a segment pauses once its r.r has fallen by delta2 (or reached a
  // quarter of the target), and the FP64 residual decides convergence
  const double delta2 = inner_rtol * inner_rtol;  // FP32 reduction per segment
  HostState init{};
  init.tol2 = 0.0;
  init.maxit = maxit;
  init.seg_drop = delta2;
  const double tol2 = rtol * rtol;
  MIFE_CHECK_HIP(hipMemcpyAsync(st, &init, sizeof(HostState), hipMemcpyHostToDevice, s));
  HostState hs{};
  int nrep = 0;
  bool first = true;
  double log_rate = 0.0;
  int seg = 48, it0 = 0, launched = 0;
  double rr0 = 0.0;
  int tickets[2 * kMaxIterChunk];
  // One host read per cycle: the FP32 segment and the FP64 replacement that
  // follows it are launched back to back.
  while (true) {
    // FP64 replacement: x64 += x32, r64 = b - A64 x64, r32 = r64, z = D^-1 r32
    if (!first) fold_correction_kernel<<<vg, 256, 0, s>>>(n, x, x32);
    launch_residual64(M, vals64, b, x, r64, sto, s);
    replace_residual_kernel<<<vg, 256, 0, s>>>(n, r64, r, z, dinv, st);
    fin_replace_kernel<<<1, kFinThreads, 0, s>>>(st, sto, vg, first ? 1 : 0, tol2);
    MIFE_CHECK_HIP(hipGetLastError());
    rc = read_state(st, &hs, s);
    if (rc) return rc;
    if (!first) {
      discard_skipped(tickets, it0, launched, hs.iters);
      const int ran = hs.iters - it0;
      if (ran > 0 && rr0 > 0.0 && hs.rr_seg > 0.0 && hs.rr_seg < rr0) log_rate = std::log(hs.rr_seg / rr0) / ran;
    }
    first = false;
    ++nrep;
    if (hs.done || hs.iters >= maxit) break;
    // the next segment: the device pauses it at its target (seg_rr); the host
    // launches the iterations the observed rate predicts, plus a margin, and the
    // ones past the pause return at once
    if (nrep > 1) {
      seg = predict_iters(hs.rr, hs.seg_rr, log_rate, 4, kMaxIterChunk);
      seg += seg / 4 + 4;
    }
    seg = std::min(std::min(seg, maxit - hs.iters), kMaxIterChunk);
    it0 = hs.iters;
    rr0 = hs.rr;
    launched = seg;
    rc = launch_iters<float>(M, vals32, x32, r, z, p, q, dinv, st, s, it0, seg, tickets);
    if (rc) return rc;
  }
  permute_out_kernel<double><<<vg, 256, 0, s>>>(n, M.perm, x, x_io);
  MIFE_CHECK_HIP(hipStreamSynchronize(s));
  const double rel = hs.bb > 0 ? sqrt(hs.rr64 / hs.bb) : 0.0;
  if (iters) *iters = hs.iters;
  if (outer) *outer = nrep;
  if (relres) *relres = rel;
  if (rel > rtol) {
    set_error("mixed-precision CG stopped at relative residual %.3e after %d iterations", rel, hs.iters);
    return MIFE_ERR_NOT_CONVERGED;
  }
  return MIFE_OK;
}

}  // extern "C"

namespace {

// MINRES with FP64 vectors and scalars on FP64 values, or on FP32 values read
// into FP64 arithmetic (TV = float: the matrix stream, nearly all the bytes of
// an iteration, shrinks from 8 to 4 bytes per value).
template <typename TV>
int minres_impl(const mife_matrix* A, const TV* vals, const double* b_in, double* x_io, const double* pc_inv,
                double rtol, int maxit, void* work, size_t work_bytes, int* iters, double* relres, void* stream) {
  Mat M;
  int rc = make_mat(A, &M);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int n = M.a.n;
  Work w{(char*)work, 0, work_bytes};
  MinresBufs B;
  minres_layout(w, n, B);
  if ((rc = check_fit(w, "MINRES"))) return rc;
  MinresState* st = B.st;
  CgState* rs = B.rs;
  double *dinv = B.dinv, *b = B.b, *x = B.x, *z = B.z, *v = B.v, *t = B.t, *r1 = B.r1, *r2 = B.r2, *wv = B.wv,
         *w1 = B.w1, *w2 = B.w2;
  const int vg = vec_grid(n);
  const int grid = row_grid(M.a, M.a.fmt != MIFE_FMT_CSR ? 64 : M.L);
  permute_in_kernel<double><<<vg, 256, 0, s>>>(n, M.perm, b_in, b);
  permute_in_kernel<double><<<vg, 256, 0, s>>>(n, M.perm, x_io, x);
  launch_diag_inv<double, TV>(M.a, M.stored, vals, dinv, s);
  pc_override_kernel<<<vg, 256, 0, s>>>(n, M.perm, pc_inv, dinv, 1);
  fill_kernel<double><<<vg, 256, 0, s>>>(n, v, 0.0);
  fill_kernel<double><<<vg, 256, 0, s>>>(n, w1, 0.0);
  MIFE_DISPATCH_FMT(M.a, M.L, (minres_init_kernel<FMT_, L_, TV><<<grid, 256, 0, s>>>(M.a, vals, b, x, dinv, r1, r2,
                                                                                       z, wv, w2, st, rtol, maxit)));
  MIFE_CHECK_HIP(hipGetLastError());
  int done = 0, it = 0, k = 0;
  const int chunk = 32;
  while (!done) {
    for (int c = 0; c < chunk; ++c, ++k) {
      const int nb = (k == 0) ? grid : vg;
      const int tk = prof::begin(s);
      MIFE_DISPATCH_FMT(M.a, M.L, (minres_a_kernel<FMT_, L_, TV><<<grid, 256, 0, s>>>(M.a, vals, z, v, t, r1, wv, w1,
                                                                                        w2, x, st, k, nb)));
      prof::end(tk, s, prof::kMinresSpmv);
      minres_b_kernel<<<vg, 256, 0, s>>>(n, t, r1, r2, z, dinv, st, k, grid);
    }
    MIFE_CHECK_HIP(hipGetLastError());
    MinresScalars hsc[2];
    MIFE_CHECK_HIP(hipMemcpyAsync(hsc, st->s, sizeof(hsc), hipMemcpyDeviceToHost, s));
    MIFE_CHECK_HIP(hipStreamSynchronize(s));
    done = hsc[0].done | hsc[1].done;
    it = std::max(hsc[0].iters, hsc[1].iters);
    if (done) {
      const MinresScalars& d = hsc[0].done ? hsc[0] : hsc[1];
      if (d.iters >= d.maxit && d.maxit > 0 && d.phibar > d.tol * d.beta1) done = 2;
    }
  }
  // true residual at exit
  if ((rc = zero_state(rs, s))) return rc;
  launch_residual64<TV>(M, vals, b, x, nullptr, rs, s);
  HostState hs{};
  rc = read_state(rs, &hs, s);
  if (rc) return rc;
  permute_out_kernel<double><<<vg, 256, 0, s>>>(n, M.perm, x, x_io);
  MIFE_CHECK_HIP(hipStreamSynchronize(s));
  const double rel = hs.bb > 0 ? sqrt(hs.rr64 / hs.bb) : 0.0;
  if (iters) *iters = it;
  if (relres) *relres = rel;
  if (done == 2) {
    set_error("MINRES reached the maximum of %d iterations (relative residual %.3e)", maxit, rel);
    return MIFE_ERR_NOT_CONVERGED;
  }
  return MIFE_OK;
}

}  // namespace

extern "C" {

int mife_minres(const mife_matrix* A, const void* vals, int prec, const double* b_in, double* x_io,
                const double* pc_inv, double rtol, int maxit, void* work, size_t work_bytes, int* iters,
                double* relres, void* stream) {
  MIFE_REQUIRE(vals && b_in && x_io && work, MIFE_ERR_POINTER_IS_NULL, "null pointer argument");
  MIFE_REQUIRE(rtol > 0.0 && maxit >= 0, MIFE_ERR_ARG, "invalid MINRES arguments");
  MIFE_REQUIRE(prec == MIFE_F64 || prec == MIFE_F32, MIFE_ERR_ARG, "invalid precision %d", prec);
  if (prec == MIFE_F32)
    return minres_impl<float>(A, (const float*)vals, b_in, x_io, pc_inv, rtol, maxit, work, work_bytes, iters,
                              relres, stream);
  return minres_impl<double>(A, (const double*)vals, b_in, x_io, pc_inv, rtol, maxit, work, work_bytes, iters,
                             relres, stream);
}

int mife_gmres(const mife_matrix* A, const double* vals, const double* b_in, double* x_io, const double* pc_inv,
               double rtol, int restart, int maxit, void* work, size_t work_bytes, int* iters, double* relres,
               void* stream) {
  MIFE_REQUIRE(vals && b_in && x_io && work, MIFE_ERR_POINTER_IS_NULL, "null pointer argument");
  MIFE_REQUIRE(restart >= 1 && restart <= kMaxRestart && rtol > 0.0, MIFE_ERR_ARG, "invalid GMRES arguments");
  Mat M;
  int rc = make_mat(A, &M);
  if (rc) return rc;
  hipStream_t s = (hipStream_t)stream;
  const int n = M.a.n;
  const int m = restart;
  const bool refine = g_gmres_refine != 0;
  Work w{(char*)work, 0, work_bytes};
  GmresBufs B;
  gmres_layout(w, n, m, B);
  if ((rc = check_fit(w, "GMRES"))) return rc;
  CgState* st = B.st;
  double *dinv = B.dinv, *r = B.r, *t = B.t, *wv = B.wv, *b = B.b, *x = B.x, *V = B.V;
  double *hdev = B.hdev, *ydev = B.ydev, *ndev = B.ndev, *Hdev = B.Hdev, *csdev = B.csdev, *sndev = B.sndev,
         *gdev = B.gdev;
  GmresCycle* cy = B.cy;
  const size_t ldv = B.ldv;
  const int vg = vec_grid(n);
  const int ag = std::min(div_up(n, 64), kMaxGrid);  // Gram-Schmidt update: 64-row tiles
  permute_in_kernel<double><<<vg, 256, 0, s>>>(n, M.perm, b_in, b);
  permute_in_kernel<double><<<vg, 256, 0, s>>>(n, M.perm, x_io, x);
  launch_diag_inv<double>(M.a, M.stored, vals, dinv, s);
  if (pc_inv) pc_override_kernel<<<vg, 256, 0, s>>>(n, M.perm, pc_inv, dinv, 0);
  if ((rc = zero_state(st, s))) return rc;
  launch_residual64(M, vals, b, x, r, st, s);
  HostState hs{};
  rc = read_state(st, &hs, s);
  if (rc) return rc;
  const double bn = sqrt(hs.bb);
  double beta = sqrt(hs.rr64);
  int it = 0;
  // Each restart cycle runs on the device in chunks of kChunk iterations: the
  // Givens rotations and the stopping test are device kernels, and the host
  // reads the cycle state once per chunk.  Iterations of a chunk after the stop
  // compute unused columns (their rotation kernel returns at once).
  constexpr int kChunk = 8;
  std::vector<double> H((size_t)(m + 1) * m), g(m + 2);
  while (bn > 0 && beta / bn > rtol && it < maxit) {
    gmres_cycle_init_kernel<<<vg, 256, 0, s>>>(n, m, r, V, dinv, t, beta, gdev, cy);
    GmresCycle hc{};
    int k = 0;
    while (k < m) {
      const int cnt = std::min(kChunk, std::min(m - k, maxit - it));
      for (int c = 0; c < cnt; ++c) {
        const int kk = k + c;
        launch_spmv<double>(M.a, vals, t, wv, M.L, s);   // t = D^-1 V_kk (written with V_kk)
        // classical Gram-Schmidt: coefficients h, then (refinement "always") a
        // second pass with the correction y; the last update also yields the
        // norm partials, which the rotation kernel reduces
        gs_dots_kernel<<<dim3(kDotChunks, kk + 1), 256, 0, s>>>(n, V, ldv, wv, hdev);
        if (refine) {
          gs_axpy_kernel<false><<<ag, 64 * kAxpyWaves, 0, s>>>(n, kk + 1, V, ldv, hdev, wv, nullptr);
          gs_dots_kernel<<<dim3(kDotChunks, kk + 1), 256, 0, s>>>(n, V, ldv, wv, ydev);
          gs_axpy_kernel<true><<<ag, 64 * kAxpyWaves, 0, s>>>(n, kk + 1, V, ldv, ydev, wv, ndev);
        } else {
          gs_axpy_kernel<true><<<ag, 64 * kAxpyWaves, 0, s>>>(n, kk + 1, V, ldv, hdev, wv, ndev);
        }
        gmres_givens_kernel<<<1, 256, 0, s>>>(kk, m, hdev, refine ? ydev : nullptr, ndev, ag, Hdev, csdev, sndev,
                                             gdev, cy, rtol * bn, (it + c + 1 >= maxit || kk + 1 == m) ? 1 : 0);
        gmres_next_basis_kernel<<<vg, 256, 0, s>>>(n, wv, V + (kk + 1) * ldv, dinv, t, cy);
      }
      MIFE_CHECK_HIP(hipGetLastError());
      MIFE_CHECK_HIP(hipMemcpyAsync(&hc, cy, sizeof(GmresCycle), hipMemcpyDeviceToHost, s));
      MIFE_CHECK_HIP(hipStreamSynchronize(s));
      it += hc.k_used - k;
      k = hc.k_used;
      if (hc.done || cnt == 0) break;
    }
    const int k_used = hc.k_used;
    if (k_used == 0) break;
    MIFE_CHECK_HIP(hipMemcpyAsync(H.data(), Hdev, sizeof(double) * (size_t)k_used * m, hipMemcpyDeviceToHost, s));
    MIFE_CHECK_HIP(hipMemcpyAsync(g.data(), gdev, sizeof(double) * (k_used + 1), hipMemcpyDeviceToHost, s));
    MIFE_CHECK_HIP(hipStreamSynchronize(s));
    std::vector<double> y(k_used);
    for (int i = k_used - 1; i >= 0; --i) {
      double sum = g[i];
      for (int j = i + 1; j < k_used; ++j) sum -= H[(size_t)i * m + j] * y[j];
      y[i] = sum / H[(size_t)i * m + i];
    }
    MIFE_CHECK_HIP(hipMemcpyAsync(ydev, y.data(), sizeof(double) * k_used, hipMemcpyHostToDevice, s));
    gmres_update_kernel<<<vg, 256, 0, s>>>(n, k_used, V, ldv, ydev, dinv, x);
    launch_residual64(M, vals, b, x, r, st, s);
    rc = read_state(st, &hs, s);
    if (rc) return rc;
    beta = sqrt(hs.rr64);
  }
  permute_out_kernel<double><<<vg, 256, 0, s>>>(n, M.perm, x, x_io);
  MIFE_CHECK_HIP(hipStreamSynchronize(s));
  if (iters) *iters = it;
  if (relres) *relres = bn > 0 ? beta / bn : 0.0;
  if (bn > 0 && beta / bn > rtol) {
    set_error("GMRES reached the maximum of %d iterations", maxit);
    return MIFE_ERR_NOT_CONVERGED;
  }
  return MIFE_OK;
}

}  // extern "C"
