// Krylov solvers: Jacobi-preconditioned CG (FP64 or FP32), FP64 iterative
// refinement around an FP32 inner CG, and right-Jacobi restarted GMRES (FP64).
//
// CG runs two fused kernels per iteration and keeps every scalar on the
// device.  With q = A p maintained by the recurrence q = A z + beta q, the
// search direction update folds into the SpMV kernel:
//   K_A: p = z + beta p ; q = A z + beta q ; partial (p.q)   -> alpha = rz / pq
//   K_B: x += alpha p ; r -= alpha q ; z = D^-1 r ; partial (r.z), (r.r) -> beta
// The last block of each kernel reduces the per-block partials in a fixed
// order (deterministic), so the host only looks at the residual every few
// iterations.  Kernels return at once after convergence is flagged.
#include "mife_common.hpp"
#include "mife_sparse.hpp"
#include "mife_profile.hpp"

#include <algorithm>
#include <cmath>
#include <vector>

namespace mife {

constexpr int kRedBlocks = 1024;

struct CgState {
  double rz, pq, alpha, beta, rr, bb, tol2, scale;
  int iters, done, maxit, pad;
  unsigned int cnt_a, cnt_b, cnt_c, pad2;
  // per-block partials of up to 4 reduced values each (value v of block b at v*kRedBlocks+b)
  double part_a[4 * kRedBlocks];
  double part_b[4 * kRedBlocks];
};

// Fixed-order reduction of per-block partials by the last block to finish.
// Returns true in the last block (all threads); sums in out[0..n_vals).
template <int NV>
__device__ __forceinline__ bool last_block_reduce(double (&mine)[NV], double* parts, unsigned int* counter,
                                                  double (&out)[NV]) {
  static_assert(NV <= 4, "at most 4 reduced values per kernel");
  __shared__ double red[NV][8];
  __shared__ bool is_last;
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
      __hip_atomic_store(&parts[v * kRedBlocks + blockIdx.x], s, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    __threadfence();
    const unsigned int ticket = atomicAdd(counter, 1u);
    is_last = (ticket == gridDim.x - 1);
  }
  __syncthreads();
  if (!is_last) return false;
  __threadfence();
#pragma unroll
  for (int v = 0; v < NV; ++v) {
    double s = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += blockDim.x) s += __hip_atomic_load(&parts[v * kRedBlocks + b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s = wave_sum(s);
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[v][threadIdx.x >> 6] = s;
    __syncthreads();
    double t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) t += red[v][w];
    out[v] = t;
  }
  if (threadIdx.x == 0) __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  return true;
}

template <typename T>
__global__ void diag_inv_kernel(int n, const int32_t* __restrict__ rp, const int32_t* __restrict__ col,
                                const T* __restrict__ vals, T* __restrict__ dinv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  int lo = rp[i], hi = rp[i + 1];
  while (lo < hi) {
    const int mid = (lo + hi) >> 1;
    if (col[mid] < i) lo = mid + 1; else hi = mid;
  }
  const T d = (lo < rp[i + 1] && col[lo] == i) ? vals[lo] : (T)0;
  dinv[i] = (d != (T)0) ? (T)1 / d : (T)1;
}

// r = b - A x (x may be skipped when zero), z = D^-1 r, p = 0, q = 0; reduces
// rz, rr, bb and initialises the state.
template <typename T, int L>
__global__ __launch_bounds__(256) void cg_init_kernel(int n, const int32_t* __restrict__ rp,
                                                      const int32_t* __restrict__ col, const T* __restrict__ vals,
                                                      const T* __restrict__ b, const T* __restrict__ x,
                                                      const T* __restrict__ dinv, T* __restrict__ r,
                                                      T* __restrict__ z, T* __restrict__ p, T* __restrict__ q,
                                                      CgState* st, double tol2, int maxit) {
  const int gpb = 256 / L, sub = threadIdx.x % L;
  double m[3] = {0.0, 0.0, 0.0};
  for (int row = blockIdx.x * gpb + threadIdx.x / L; row < n; row += gridDim.x * gpb) {
    T ax = 0;
    if (x) {
      for (int k = rp[row] + sub; k < rp[row + 1]; k += L) ax += vals[k] * x[col[k]];
#pragma unroll
      for (int off = L / 2; off > 0; off >>= 1) ax += __shfl_xor(ax, off, 64);
    }
    if (sub == 0) {
      const T ri = b[row] - ax;
      const T zi = dinv[row] * ri;
      r[row] = ri;
      z[row] = zi;
      p[row] = 0;
      q[row] = 0;
      m[0] += (double)ri * (double)zi;
      m[1] += (double)ri * (double)ri;
      m[2] += (double)b[row] * (double)b[row];
    }
  }
  double out[3];
  if (last_block_reduce<3>(m, st->part_b, &st->cnt_b, out) && threadIdx.x == 0) {
    st->rz = out[0];
    st->rr = out[1];
    st->bb = out[2];
    st->beta = 0.0;
    st->alpha = 0.0;
    st->iters = 0;
    st->maxit = maxit;
    st->tol2 = tol2;
    st->done = (out[1] <= tol2 * out[2]) ? 1 : 0;
  }
}

template <typename T, int L>
__global__ __launch_bounds__(256) void cg_spmv_kernel(int n, const int32_t* __restrict__ rp,
                                                      const int32_t* __restrict__ col, const T* __restrict__ vals,
                                                      const T* __restrict__ z, T* __restrict__ p,
                                                      T* __restrict__ q, CgState* st) {
  if (st->done) return;
  const T beta = (T)st->beta;
  const int gpb = 256 / L, sub = threadIdx.x % L;
  double m[1] = {0.0};
  for (int row = blockIdx.x * gpb + threadIdx.x / L; row < n; row += gridDim.x * gpb) {
    T az = 0;
    const int k1 = rp[row + 1];
    for (int k = rp[row] + sub; k < k1; k += L) az += vals[k] * z[col[k]];
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) az += __shfl_xor(az, off, 64);
    if (sub == 0) {
      const T pi = z[row] + beta * p[row];
      const T qi = az + beta * q[row];
      p[row] = pi;
      q[row] = qi;
      m[0] += (double)pi * (double)qi;
    }
  }
  double out[1];
  if (last_block_reduce<1>(m, st->part_a, &st->cnt_a, out) && threadIdx.x == 0) {
    st->pq = out[0];
    st->alpha = st->rz / out[0];
  }
}

template <typename T>
__global__ __launch_bounds__(256) void cg_update_kernel(int n, T* __restrict__ x, T* __restrict__ r,
                                                        T* __restrict__ z, const T* __restrict__ p,
                                                        const T* __restrict__ q, const T* __restrict__ dinv,
                                                        CgState* st) {
  if (st->done) return;
  const T alpha = (T)st->alpha;
  double m[2] = {0.0, 0.0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    x[i] += alpha * p[i];
    const T ri = r[i] - alpha * q[i];
    const T zi = dinv[i] * ri;
    r[i] = ri;
    z[i] = zi;
    m[0] += (double)ri * (double)zi;
    m[1] += (double)ri * (double)ri;
  }
  double out[2];
  if (last_block_reduce<2>(m, st->part_b, &st->cnt_b, out) && threadIdx.x == 0) {
    st->beta = out[0] / st->rz;
    st->rz = out[0];
    st->rr = out[1];
    st->iters += 1;
    st->done = (out[1] <= st->tol2 * st->bb || st->iters >= st->maxit) ? 1 : 0;
  }
}

// FP64 residual r = b - A x with ||r||^2 and ||b||^2 (refinement / exit checks).
template <int L>
__global__ __launch_bounds__(256) void residual64_kernel(int n, const int32_t* __restrict__ rp,
                                                         const int32_t* __restrict__ col,
                                                         const double* __restrict__ vals,
                                                         const double* __restrict__ b, const double* __restrict__ x,
                                                         double* __restrict__ r, CgState* st) {
  const int gpb = 256 / L, sub = threadIdx.x % L;
  double m[2] = {0.0, 0.0};
  for (int row = blockIdx.x * gpb + threadIdx.x / L; row < n; row += gridDim.x * gpb) {
    double ax = 0;
    for (int k = rp[row] + sub; k < rp[row + 1]; k += L) ax += vals[k] * x[col[k]];
#pragma unroll
    for (int off = L / 2; off > 0; off >>= 1) ax += __shfl_xor(ax, off, 64);
    if (sub == 0) {
      const double ri = b[row] - ax;
      if (r) r[row] = ri;
      m[0] += ri * ri;
      m[1] += b[row] * b[row];
    }
  }
  double out[2];
  if (last_block_reduce<2>(m, st->part_a, &st->cnt_a, out) && threadIdx.x == 0) {
    st->rr = out[0];
    st->bb = out[1];
    st->scale = sqrt(out[0]);
  }
}

// r32 = r64 / ||r64||
__global__ void scale_to_f32_kernel(int n, const double* __restrict__ r, float* __restrict__ r32,
                                    const CgState* st) {
  const double inv = st->scale > 0.0 ? 1.0 / st->scale : 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) r32[i] = (float)(r[i] * inv);
}

// x64 += ||r|| * d32
__global__ void accumulate_kernel(int n, double* __restrict__ x, const float* __restrict__ d, const CgState* st) {
  const double s = st->scale;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) x[i] += s * (double)d[i];
}

template <typename T>
__global__ void fill_kernel(int n, T* __restrict__ v, T val) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v[i] = val;
}

// ------------------------------------------------------------------ GMRES
// h[j] partials of V_j . w for j < m over a fixed grid; reduced by gmres_reduce.
__global__ __launch_bounds__(256) void multidot_kernel(int n, int m, const double* __restrict__ V, size_t ldv,
                                                       const double* __restrict__ w, double* __restrict__ parts) {
  __shared__ double red[32][4];
  double acc[32];
  for (int j = 0; j < 32; ++j) acc[j] = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const double wi = w[i];
    for (int j = 0; j < m; ++j) acc[j] += V[j * ldv + i] * wi;
  }
  for (int j = 0; j < m; ++j) {
    const double s = wave_sum(acc[j]);
    if ((threadIdx.x & 63) == 0) red[j][threadIdx.x >> 6] = s;
  }
  __syncthreads();
  if (threadIdx.x < m) {
    double s = 0.0;
    for (int wv = 0; wv < 4; ++wv) s += red[threadIdx.x][wv];
    parts[threadIdx.x * kRedBlocks + blockIdx.x] = s;
  }
}

__global__ void multidot_reduce(int m, int nblocks, const double* __restrict__ parts, double* __restrict__ h,
                                int accumulate) {
  __shared__ double red[4];
  for (int j = 0; j < m; ++j) {
    double s = 0.0;
    for (int b = threadIdx.x; b < nblocks; b += blockDim.x) s += parts[j * kRedBlocks + b];
    s = block_sum(s, red);
    if (threadIdx.x == 0) h[j] = accumulate ? h[j] + s : s;
    __syncthreads();
  }
}

// w -= sum_j V_j h_j   (h = hsrc, m terms)
__global__ void orth_update_kernel(int n, int m, const double* __restrict__ V, size_t ldv,
                                   const double* __restrict__ h, double* __restrict__ w) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    double s = w[i];
    for (int j = 0; j < m; ++j) s -= V[j * ldv + i] * h[j];
    w[i] = s;
  }
}

__global__ void scale_copy_kernel(int n, const double* __restrict__ src, double* __restrict__ dst, double a) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) dst[i] = a * src[i];
}

__global__ void mul_kernel(int n, const double* __restrict__ a, const double* __restrict__ b, double* __restrict__ c) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) c[i] = a[i] * b[i];
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

// ------------------------------------------------------------- host side
struct Work {
  char* base;
  size_t off;
  template <typename T>
  T* take(size_t count) {
    off = (off + 255) & ~size_t(255);
    T* p = reinterpret_cast<T*>(base + off);
    off += count * sizeof(T);
    return p;
  }
};

static int vec_grid(int n) { return std::max(1, std::min(div_up(n, 256), kRedBlocks)); }

template <typename T>
static void launch_cg_init(int n, int L, const int32_t* rp, const int32_t* col, const T* vals, const T* b,
                           const T* x, const T* dinv, T* r, T* z, T* p, T* q, CgState* st, double tol2, int maxit,
                           hipStream_t s) {
  const int grid = std::min(spmv_grid(n, L), kRedBlocks);
#define CASE_L(LL) \
  case LL: cg_init_kernel<T, LL><<<grid, 256, 0, s>>>(n, rp, col, vals, b, x, dinv, r, z, p, q, st, tol2, maxit); break;
  switch (L) { CASE_L(4) CASE_L(8) CASE_L(16) CASE_L(32) default: CASE_L(64) }
#undef CASE_L
}

// Launches one CG iteration; with profiling on, the two kernels' event tickets
// are returned through tk[0..1] (else -1).
template <typename T>
static void launch_cg_iter(int n, int L, const int32_t* rp, const int32_t* col, const T* vals, T* x, T* r, T* z,
                           T* p, T* q, const T* dinv, CgState* st, hipStream_t s, int* tk) {
  const int grid = std::min(spmv_grid(n, L), kRedBlocks);
  tk[0] = prof::begin(s);
#define CASE_L(LL) \
  case LL: cg_spmv_kernel<T, LL><<<grid, 256, 0, s>>>(n, rp, col, vals, z, p, q, st); break;
  switch (L) { CASE_L(4) CASE_L(8) CASE_L(16) CASE_L(32) default: CASE_L(64) }
#undef CASE_L
  prof::end(tk[0], s, sizeof(T) == 8 ? prof::kCgSpmv : prof::kCgSpmv32);
  tk[1] = prof::begin(s);
  cg_update_kernel<T><<<vec_grid(n), 256, 0, s>>>(n, x, r, z, p, q, dinv, st);
  prof::end(tk[1], s, sizeof(T) == 8 ? prof::kCgUpdate : prof::kCgUpdate32);
}

static void launch_residual64(int n, int L, const int32_t* rp, const int32_t* col, const double* vals,
                              const double* b, const double* x, double* r, CgState* st, hipStream_t s) {
  const int grid = std::min(spmv_grid(n, L), kRedBlocks);
  const int tk = prof::begin(s);
#define CASE_L(LL) \
  case LL: residual64_kernel<LL><<<grid, 256, 0, s>>>(n, rp, col, vals, b, x, r, st); break;
  switch (L) { CASE_L(4) CASE_L(8) CASE_L(16) CASE_L(32) default: CASE_L(64) }
#undef CASE_L
  prof::end(tk, s, prof::kResidual64);
}

struct HostState {
  double rz, pq, alpha, beta, rr, bb, tol2, scale;
  int iters, done, maxit, pad;
};

static int read_state(const CgState* st, HostState* h, hipStream_t s) {
  MIFE_CHECK_HIP(hipMemcpyAsync(h, st, sizeof(HostState), hipMemcpyDeviceToHost, s));
  MIFE_CHECK_HIP(hipStreamSynchronize(s));
  return MIFE_OK;
}

static int nnz_of(const int32_t* rp, int n, hipStream_t s, int32_t* nnz) {
  MIFE_CHECK_HIP(hipMemcpyAsync(nnz, rp + n, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  MIFE_CHECK_HIP(hipStreamSynchronize(s));
  return MIFE_OK;
}

// Runs CG iterations until the device flags convergence; checks every `chunk`.
template <typename T>
static int run_cg(int n, int L, const int32_t* rp, const int32_t* col, const T* vals, const T* b, T* x,
                  bool x_zero, double rtol, int maxit, T* r, T* z, T* p, T* q, T* dinv, CgState* st,
                  hipStream_t s, HostState* hs) {
  MIFE_CHECK_HIP(hipMemsetAsync(st, 0, sizeof(CgState), s));
  diag_inv_kernel<T><<<div_up(n, 256), 256, 0, s>>>(n, rp, col, vals, dinv);
  launch_cg_init<T>(n, L, rp, col, vals, b, x_zero ? nullptr : x, dinv, r, z, p, q, st, rtol * rtol, maxit, s);
  MIFE_CHECK_HIP(hipGetLastError());
  int chunk = 8;
  int launched = 0;
  std::vector<int> tickets;
  while (true) {
    tickets.assign(2 * chunk, -1);
    for (int k = 0; k < chunk; ++k)
      launch_cg_iter<T>(n, L, rp, col, vals, x, r, z, p, q, dinv, st, s, &tickets[2 * k]);
    MIFE_CHECK_HIP(hipGetLastError());
    int rc = read_state(st, hs, s);
    if (rc) return rc;
    // iterations at or past the converged count returned at once: not timed
    for (int k = 0; k < chunk; ++k)
      if (launched + k >= hs->iters) {
        prof::discard(tickets[2 * k]);
        prof::discard(tickets[2 * k + 1]);
      }
    launched += chunk;
    if (hs->done || launched >= maxit + chunk) break;
    chunk = std::min(chunk * 2, 256);
  }
  return MIFE_OK;
}

}  // namespace mife

using namespace mife;

extern "C" {

size_t mife_solver_work_bytes(int n, int restart) {
  const size_t nn = (size_t)std::max(n, 1);
  const size_t m = (size_t)std::max(restart, 1) + 1;
  size_t bytes = 0;
  bytes += 2 * (sizeof(CgState) + 256);
  bytes += (8 * nn + 256) * (8 + m);            // FP64 vectors incl. GMRES basis
  bytes += (4 * nn + 256) * 8;                  // FP32 vectors
  bytes += (8 * (size_t)kRedBlocks * 32 + 256); // GMRES multidot partials
  bytes += 8 * 64 * 4 + 65536;
  return bytes;
}

int mife_cg(int n, const int32_t* row_ptr, const int32_t* col, const void* vals, const void* b, void* x, int prec,
            double rtol, int maxit, void* work, int* iters, double* relres, void* stream) {
  MIFE_REQUIRE(row_ptr && col && vals && b && x && work, MIFE_ERR_POINTER_IS_NULL, "null pointer argument");
  MIFE_REQUIRE(prec == MIFE_F64 || prec == MIFE_F32, MIFE_ERR_ARG, "invalid precision %d", prec);
  MIFE_REQUIRE(n > 0 && maxit >= 0 && rtol >= 0.0, MIFE_ERR_ARG, "invalid CG arguments");
  hipStream_t s = (hipStream_t)stream;
  int32_t nnz = 0;
  int rc = nnz_of(row_ptr, n, s, &nnz);
  if (rc) return rc;
  const int L = spmv_lanes(nnz, n);
  Work w{(char*)work, 0};
  CgState* st = w.take<CgState>(1);
  HostState hs{};
  if (prec == MIFE_F64) {
    double *r = w.take<double>(n), *z = w.take<double>(n), *p = w.take<double>(n), *q = w.take<double>(n),
           *dinv = w.take<double>(n);
    rc = run_cg<double>(n, L, row_ptr, col, (const double*)vals, (const double*)b, (double*)x, false, rtol, maxit,
                        r, z, p, q, dinv, st, s, &hs);
    if (rc) return rc;
    // explicit FP64 residual at exit
    CgState* st2 = w.take<CgState>(1);
    MIFE_CHECK_HIP(hipMemsetAsync(st2, 0, sizeof(CgState), s));
    launch_residual64(n, L, row_ptr, col, (const double*)vals, (const double*)b, (const double*)x, nullptr, st2, s);
    HostState h2{};
    rc = read_state(st2, &h2, s);
    if (rc) return rc;
    if (iters) *iters = hs.iters;
    if (relres) *relres = h2.bb > 0 ? sqrt(h2.rr / h2.bb) : 0.0;
  } else {
    float *r = w.take<float>(n), *z = w.take<float>(n), *p = w.take<float>(n), *q = w.take<float>(n),
          *dinv = w.take<float>(n);
    rc = run_cg<float>(n, L, row_ptr, col, (const float*)vals, (const float*)b, (float*)x, false, rtol, maxit, r, z,
                       p, q, dinv, st, s, &hs);
    if (rc) return rc;
    if (iters) *iters = hs.iters;
    if (relres) *relres = hs.bb > 0 ? sqrt(hs.rr / hs.bb) : 0.0;
  }
  if (hs.iters >= maxit && !(hs.rr <= hs.tol2 * hs.bb)) {
    set_error("CG reached the maximum of %d iterations", maxit);
    return MIFE_ERR_NOT_CONVERGED;
  }
  return MIFE_OK;
}

int mife_refine(int n, const int32_t* row_ptr, const int32_t* col, const double* vals64, const float* vals32,
                const double* b, double* x, double rtol, double inner_rtol, int maxit, void* work, int* iters,
                int* outer, double* relres, void* stream) {
  MIFE_REQUIRE(row_ptr && col && vals64 && vals32 && b && x && work, MIFE_ERR_POINTER_IS_NULL,
               "null pointer argument");
  MIFE_REQUIRE(n > 0 && rtol > 0.0 && inner_rtol > 0.0 && inner_rtol < 1.0, MIFE_ERR_ARG, "invalid arguments");
  hipStream_t s = (hipStream_t)stream;
  int32_t nnz = 0;
  int rc = nnz_of(row_ptr, n, s, &nnz);
  if (rc) return rc;
  const int L = spmv_lanes(nnz, n);
  Work w{(char*)work, 0};
  CgState* st = w.take<CgState>(1);
  CgState* sto = w.take<CgState>(1);
  double* r64 = w.take<double>(n);
  float *b32 = w.take<float>(n), *d32 = w.take<float>(n), *r = w.take<float>(n), *z = w.take<float>(n),
        *p = w.take<float>(n), *q = w.take<float>(n), *dinv = w.take<float>(n);
  HostState hs{}, ho{};
  int total = 0, nout = 0;
  MIFE_CHECK_HIP(hipMemsetAsync(sto, 0, sizeof(CgState), s));
  launch_residual64(n, L, row_ptr, col, vals64, b, x, r64, sto, s);
  rc = read_state(sto, &ho, s);
  if (rc) return rc;
  const double bb = ho.bb;
  double rel = bb > 0 ? sqrt(ho.rr / bb) : 0.0;
  while (rel > rtol && total < maxit) {
    scale_to_f32_kernel<<<vec_grid(n), 256, 0, s>>>(n, r64, b32, sto);
    fill_kernel<float><<<vec_grid(n), 256, 0, s>>>(n, d32, 0.0f);
    // the inner tolerance never asks for more than the outer target needs
    const double itol = std::max(inner_rtol, 0.5 * rtol / rel);
    rc = run_cg<float>(n, L, row_ptr, col, vals32, b32, d32, true, itol, maxit - total, r, z, p, q, dinv, st, s,
                       &hs);
    if (rc) return rc;
    total += hs.iters;
    accumulate_kernel<<<vec_grid(n), 256, 0, s>>>(n, x, d32, sto);
    launch_residual64(n, L, row_ptr, col, vals64, b, x, r64, sto, s);
    rc = read_state(sto, &ho, s);
    if (rc) return rc;
    rel = sqrt(ho.rr / bb);
    ++nout;
    if (hs.iters == 0) break;
  }
  if (iters) *iters = total;
  if (outer) *outer = nout;
  if (relres) *relres = rel;
  if (rel > rtol) {
    set_error("mixed-precision refinement stopped at relative residual %.3e", rel);
    return MIFE_ERR_NOT_CONVERGED;
  }
  return MIFE_OK;
}

int mife_gmres(int n, const int32_t* row_ptr, const int32_t* col, const double* vals, const double* b, double* x,
               double rtol, int restart, int maxit, void* work, int* iters, double* relres, void* stream) {
  MIFE_REQUIRE(row_ptr && col && vals && b && x && work, MIFE_ERR_POINTER_IS_NULL, "null pointer argument");
  MIFE_REQUIRE(n > 0 && restart >= 1 && restart <= 31 && rtol > 0.0, MIFE_ERR_ARG, "invalid GMRES arguments");
  hipStream_t s = (hipStream_t)stream;
  int32_t nnz = 0;
  int rc = nnz_of(row_ptr, n, s, &nnz);
  if (rc) return rc;
  const int L = spmv_lanes(nnz, n);
  const int m = restart;
  Work w{(char*)work, 0};
  CgState* st = w.take<CgState>(1);
  double* dinv = w.take<double>(n);
  double* r = w.take<double>(n);
  double* t = w.take<double>(n);
  double* wv = w.take<double>(n);
  const size_t ldv = ((size_t)n + 63) & ~size_t(63);
  double* V = w.take<double>(ldv * (m + 1));
  double* parts = w.take<double>((size_t)kRedBlocks * 32);
  double* hdev = w.take<double>(64);
  double* ydev = w.take<double>(64);
  diag_inv_kernel<double><<<div_up(n, 256), 256, 0, s>>>(n, row_ptr, col, vals, dinv);
  MIFE_CHECK_HIP(hipMemsetAsync(st, 0, sizeof(CgState), s));
  launch_residual64(n, L, row_ptr, col, vals, b, x, r, st, s);
  HostState hs{};
  rc = read_state(st, &hs, s);
  if (rc) return rc;
  const double bn = sqrt(hs.bb);
  double beta = sqrt(hs.rr);
  int it = 0;
  const int vg = vec_grid(n);
  std::vector<double> H((m + 1) * m), cs(m), sn(m), g(m + 1), hcol(m + 2);
  while (bn > 0 && beta / bn > rtol && it < maxit) {
    scale_copy_kernel<<<vg, 256, 0, s>>>(n, r, V, 1.0 / beta);
    std::fill(H.begin(), H.end(), 0.0);
    std::fill(g.begin(), g.end(), 0.0);
    g[0] = beta;
    int k_used = 0;
    for (int k = 0; k < m; ++k) {
      mul_kernel<<<vg, 256, 0, s>>>(n, dinv, V + k * ldv, t);
      launch_spmv<double>(n, row_ptr, col, vals, t, wv, L, s);
      // classical Gram-Schmidt, two passes
      for (int pass = 0; pass < 2; ++pass) {
        multidot_kernel<<<vg, 256, 0, s>>>(n, k + 1, V, ldv, wv, parts);
        multidot_reduce<<<1, 256, 0, s>>>(k + 1, vg, parts, pass == 0 ? hdev : ydev, 0);
        orth_update_kernel<<<vg, 256, 0, s>>>(n, k + 1, V, ldv, pass == 0 ? hdev : ydev, wv);
      }
      multidot_kernel<<<vg, 256, 0, s>>>(n, 1, wv, ldv, wv, parts);
      multidot_reduce<<<1, 256, 0, s>>>(1, vg, parts, hdev + 32, 0);
      MIFE_CHECK_HIP(hipGetLastError());
      double hh[33], h2[32];
      MIFE_CHECK_HIP(hipMemcpyAsync(hh, hdev, sizeof(double) * 33, hipMemcpyDeviceToHost, s));
      MIFE_CHECK_HIP(hipMemcpyAsync(h2, ydev, sizeof(double) * (k + 1), hipMemcpyDeviceToHost, s));
      MIFE_CHECK_HIP(hipStreamSynchronize(s));
      for (int j = 0; j <= k; ++j) hcol[j] = hh[j] + h2[j];
      const double hk1 = sqrt(std::max(hh[32], 0.0));
      hcol[k + 1] = hk1;
      if (hk1 > 0) scale_copy_kernel<<<vg, 256, 0, s>>>(n, wv, V + (k + 1) * ldv, 1.0 / hk1);
      for (int j = 0; j < k; ++j) {
        const double tmp = cs[j] * hcol[j] + sn[j] * hcol[j + 1];
        hcol[j + 1] = -sn[j] * hcol[j] + cs[j] * hcol[j + 1];
        hcol[j] = tmp;
      }
      const double den = std::hypot(hcol[k], hcol[k + 1]);
      cs[k] = hcol[k] / den;
      sn[k] = hcol[k + 1] / den;
      hcol[k] = den;
      hcol[k + 1] = 0.0;
      g[k + 1] = -sn[k] * g[k];
      g[k] = cs[k] * g[k];
      for (int j = 0; j <= k; ++j) H[j * m + k] = hcol[j];
      ++it;
      k_used = k + 1;
      if (std::fabs(g[k + 1]) / bn <= rtol || it >= maxit || hk1 == 0.0) break;
    }
    std::vector<double> y(k_used);
    for (int i = k_used - 1; i >= 0; --i) {
      double sum = g[i];
      for (int j = i + 1; j < k_used; ++j) sum -= H[i * m + j] * y[j];
      y[i] = sum / H[i * m + i];
    }
    MIFE_CHECK_HIP(hipMemcpyAsync(ydev, y.data(), sizeof(double) * k_used, hipMemcpyHostToDevice, s));
    gmres_update_kernel<<<vg, 256, 0, s>>>(n, k_used, V, ldv, ydev, dinv, x);
    launch_residual64(n, L, row_ptr, col, vals, b, x, r, st, s);
    rc = read_state(st, &hs, s);
    if (rc) return rc;
    beta = sqrt(hs.rr);
  }
  if (iters) *iters = it;
  if (relres) *relres = bn > 0 ? beta / bn : 0.0;
  if (bn > 0 && beta / bn > rtol) {
    set_error("GMRES reached the maximum of %d iterations", maxit);
    return MIFE_ERR_NOT_CONVERGED;
  }
  return MIFE_OK;
}

}  // extern "C"
