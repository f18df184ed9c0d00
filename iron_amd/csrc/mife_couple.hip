// Solver mapping with linear DOF constraints.  iron ties a constrained DOF c to one
// free DOF m, u_c = a_c u_m (BoundaryConditions_DofConstraintSet,
// boundary_condition_routines.F90:3204-3313; the coupling lists of
// BoundaryConditions_DofConstraintsCreateFinish, :3320-3472).  The solver row and
// column of m then stand for every equations row / column coupled to it, each with
// its coefficient (solver_mapping_routines.F90:445-520, 1445-1511), so the solver
// system is P^T K P with one entry a_g per non-fixed row g of P:
//   owner[g] = solver DOF of g (-1 when Dirichlet-fixed), coef[g] = a_g (1 for free DOFs).
//
// The solver pattern comes from one stable radix sort of the equations entries by
// (owner[r], owner[c]); a solver entry is a run of equal keys, and its value is the sum
// of the run's (K_rc a_r) a_c in ascending CSR position, which is the order of
// iron's row-by-row adds (solver_matrices_routines.F90:935-960).
#include "mife_common.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>

namespace mife {

namespace {

constexpr size_t kAlign = 256;
inline size_t up(size_t b) { return (b + kAlign - 1) / kAlign * kAlign; }

struct CoupleWork {
  unsigned long long *key_in, *key_out;  // [nnz]
  int32_t *pos_in, *pos_out;             // [nnz]
  int32_t *prow;                         // [nnz] equations row of CSR position k
  int32_t *head, *head_idx;              // [nnz]
  int32_t *row_cnt, *grp_cnt;            // [n_s + 1]
  int32_t *gkey_in, *gkey_out;           // [n]
  int32_t *grow_in, *grow_out;           // [n]
  int32_t* flags;                        // [0] bad owner [1] n_contrib [2] empty group [3] n_grp
  void* temp;
  size_t temp_bytes;
};

size_t cub_temp_bytes(int n, long long nnz, int n_s) {
  size_t a = 0, b = 0, c = 0, d = 0, e = 0;
  const int inz = (int)std::max<long long>(nnz, 1);
  hipcub::DeviceRadixSort::SortPairs(nullptr, a, (unsigned long long*)nullptr, (unsigned long long*)nullptr,
                                     (int32_t*)nullptr, (int32_t*)nullptr, inz, 0, 64);
  hipcub::DeviceRadixSort::SortPairs(nullptr, b, (int32_t*)nullptr, (int32_t*)nullptr, (int32_t*)nullptr,
                                     (int32_t*)nullptr, std::max(n, 1), 0, 32);
  hipcub::DeviceScan::ExclusiveSum(nullptr, c, (int32_t*)nullptr, (int32_t*)nullptr, inz);
  hipcub::DeviceScan::ExclusiveSum(nullptr, d, (int32_t*)nullptr, (int32_t*)nullptr, n_s + 1);
  e = std::max(std::max(a, b), std::max(c, d));
  return e;
}

size_t layout(int n, long long nnz, int n_s, char* base, CoupleWork* w) {
  const size_t z = (size_t)std::max<long long>(nnz, 1), nn = (size_t)std::max(n, 1), ns = (size_t)n_s + 1;
  size_t off = 0;
  auto take = [&](size_t bytes) {
    char* p = base ? base + off : nullptr;
    off += up(bytes);
    return p;
  };
  CoupleWork t{};
  t.key_in = (unsigned long long*)take(8 * z);
  t.key_out = (unsigned long long*)take(8 * z);
  t.pos_in = (int32_t*)take(4 * z);
  t.pos_out = (int32_t*)take(4 * z);
  t.prow = (int32_t*)take(4 * z);
  t.head = (int32_t*)take(4 * z);
  t.head_idx = (int32_t*)take(4 * z);
  t.row_cnt = (int32_t*)take(4 * ns);
  t.grp_cnt = (int32_t*)take(4 * ns);
  t.gkey_in = (int32_t*)take(4 * nn);
  t.gkey_out = (int32_t*)take(4 * nn);
  t.grow_in = (int32_t*)take(4 * nn);
  t.grow_out = (int32_t*)take(4 * nn);
  t.flags = (int32_t*)take(4 * 16);
  t.temp_bytes = cub_temp_bytes(n, nnz, n_s);
  t.temp = take(t.temp_bytes);
  if (w) *w = t;
  return off;
}

int bits_for(unsigned long long v) {
  int b = 1;
  while (b < 64 && (v >> b) != 0ull) ++b;
  return b;
}

}  // namespace

// One wavefront per equations row: entry keys (owner[r], owner[c]) with the sentinel
// n_s^2 where either DOF is fixed, the identity position, and the row of each entry.
__global__ void couple_keys_kernel(int n, const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
                                   const int32_t* __restrict__ owner, int n_s, unsigned long long* __restrict__ key,
                                   int32_t* __restrict__ pos, int32_t* __restrict__ prow,
                                   int32_t* __restrict__ flags) {
  const int r = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (r >= n) return;
  const unsigned long long ns = (unsigned long long)n_s;
  const unsigned long long sentinel = ns * ns;
  const int o_r = owner[r];
  if (lane == 0 && (o_r < -1 || o_r >= n_s)) atomicOr(&flags[0], 1);
  for (int k = row_ptr[r] + lane; k < row_ptr[r + 1]; k += 64) {
    const int c = col[k];
    const int o_c = (c >= 0 && c < n) ? owner[c] : -2;
    if (o_c < -1 || o_c >= n_s) atomicOr(&flags[0], 1);
    const bool ok = o_r >= 0 && o_r < n_s && o_c >= 0 && o_c < n_s;
    key[k] = ok ? (unsigned long long)o_r * ns + (unsigned long long)o_c : sentinel;
    pos[k] = k;
    prow[k] = r;
  }
}

// Group keys: equations row r goes to solver row owner[r] (fixed rows last).
__global__ void couple_group_keys_kernel(int n, const int32_t* __restrict__ owner, int n_s,
                                         int32_t* __restrict__ gkey, int32_t* __restrict__ grow,
                                         int32_t* __restrict__ grp_cnt) {
  const int r = blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= n) return;
  const int o = owner[r];
  const bool ok = o >= 0 && o < n_s;
  gkey[r] = ok ? o : n_s;
  grow[r] = r;
  if (ok) atomicAdd(&grp_cnt[o], 1);
}

// Run heads of the sorted keys, per-row entry counts, the number of valid
// (non-sentinel) items and solver rows that no equations row maps to.
__global__ void couple_heads_kernel(long long nnz, const unsigned long long* __restrict__ key, int n_s,
                                    int32_t* __restrict__ head, int32_t* __restrict__ row_cnt,
                                    const int32_t* __restrict__ grp_cnt, int32_t* __restrict__ flags) {
  const unsigned long long ns = (unsigned long long)n_s;
  const unsigned long long sentinel = ns * ns;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < nnz; i += stride) {
    const unsigned long long k = key[i];
    const bool valid = k < sentinel;
    const bool h = valid && (i == 0 || key[i - 1] != k);
    head[i] = h ? 1 : 0;
    if (h) atomicAdd(&row_cnt[(int)(k / ns)], 1);
    if (valid && (i + 1 == nnz || key[i + 1] >= sentinel)) flags[1] = (int32_t)(i + 1);
  }
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n_s; i += stride)
    if (grp_cnt[i] == 0) atomicOr(&flags[2], 1);
}

__global__ void couple_fill_kernel(long long n_valid, const unsigned long long* __restrict__ key, int n_s,
                                   const int32_t* __restrict__ head, const int32_t* __restrict__ head_idx,
                                   const int32_t* __restrict__ pos, const int32_t* __restrict__ prow,
                                   int32_t* __restrict__ s_col, int32_t* __restrict__ contrib_ptr,
                                   int32_t* __restrict__ contrib_pos, int32_t* __restrict__ contrib_row) {
  const unsigned long long ns = (unsigned long long)n_s;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n_valid; i += stride) {
    const int p = pos[i];
    contrib_pos[i] = p;
    contrib_row[i] = prow[p];
    if (head[i]) {
      const int e = head_idx[i];
      s_col[e] = (int32_t)(key[i] % ns);
      contrib_ptr[e] = (int32_t)i;
    }
  }
}

__global__ void couple_tail_kernel(const int32_t* __restrict__ flags, const int32_t* __restrict__ s_row_ptr, int n_s,
                                   int32_t* __restrict__ contrib_ptr) {
  if (blockIdx.x == 0 && threadIdx.x == 0) contrib_ptr[s_row_ptr[n_s]] = flags[1];
}

// dst[e] = sum over the run of (src[p] a_r) a_c, in FP64, in run order.
template <typename S, typename D>
__global__ void couple_values_kernel(long long nnz_s, const int32_t* __restrict__ contrib_ptr,
                                     const int32_t* __restrict__ contrib_pos, const int32_t* __restrict__ contrib_row,
                                     const int32_t* __restrict__ col, const double* __restrict__ coef,
                                     const S* __restrict__ src, D* __restrict__ dst) {
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < nnz_s; e += stride) {
    double acc = 0.0;
    for (int i = contrib_ptr[e]; i < contrib_ptr[e + 1]; ++i) {
      const int p = contrib_pos[i];
      acc += ((double)src[p] * coef[contrib_row[i]]) * coef[col[p]];
    }
    dst[e] = (D)acc;
  }
}

// One wavefront per solver row: b[s] = sum over its equations rows r (ascending) of
// a_r (src[r] + rhs[r] - lift_r), lift_r over the Dirichlet-fixed columns with the
// ZERO_TOLERANCE tests of SOLVER_MATRICES_STATIC_ASSEMBLE.
__global__ void coupled_rhs_kernel(int n_s, const int32_t* __restrict__ grp_ptr, const int32_t* __restrict__ grp_rows,
                                   const double* __restrict__ coef, const int32_t* __restrict__ row_ptr,
                                   const int32_t* __restrict__ col, const double* __restrict__ vals,
                                   const int32_t* __restrict__ owner, const double* __restrict__ u,
                                   const double* __restrict__ src, const double* __restrict__ rhs,
                                   const uint8_t* __restrict__ rhs_fixed, double* __restrict__ b) {
  const double ZT = 5.0 * 2.220446049250313e-16;
  const int f = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (f >= n_s) return;
  double acc = 0.0;
  for (int g = grp_ptr[f]; g < grp_ptr[f + 1]; ++g) {
    const int i = grp_rows[g];
    double lift = 0.0;
    const bool lifted = !(rhs_fixed && rhs_fixed[i]);
    for (int k = row_ptr[i] + lane; lifted && k < row_ptr[i + 1]; k += 64) {
      const int j = col[k];
      if (owner[j] < 0) {
        const double uj = u[j], kv = vals[k];
        if (fabs(uj) >= ZT && fabs(kv) >= ZT) lift += -1.0 * kv * uj;
      }
    }
    lift = wave_sum(lift);
    double v = 0.0;
    if (src) v += 1.0 * src[i];
    if (rhs) v += rhs[i];
    acc += (v + lift) * coef[i];
  }
  if (lane == 0) b[f] = acc;
}

// u[g] = a_g x[owner[g]] for every non-fixed DOF (SOLVER_VARIABLES_FIELD_UPDATE with
// the solver-to-variable coupling coefficients).
__global__ void coupled_field_update_kernel(int n, const int32_t* __restrict__ owner, const double* __restrict__ coef,
                                            const double* __restrict__ x, double* __restrict__ u) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= n) return;
  const int o = owner[g];
  if (o >= 0) u[g] = x[o] * coef[g];
}

}  // namespace mife

using namespace mife;

extern "C" {

size_t mife_couple_pattern_work_bytes(int n, int64_t nnz, int n_s) {
  return layout(n, nnz, n_s, nullptr, nullptr);
}

int mife_couple_pattern(int n, int64_t nnz, const int32_t* row_ptr, const int32_t* col, const int32_t* owner,
                        int n_s, int32_t* grp_ptr, int32_t* grp_rows, int32_t* s_row_ptr, int32_t* s_col,
                        int32_t* contrib_ptr, int32_t* contrib_pos, int32_t* contrib_row, int64_t* nnz_s_out,
                        int64_t* n_contrib_out, void* work, void* stream) {
  MIFE_REQUIRE(row_ptr && col && owner && s_row_ptr && grp_ptr && work, MIFE_ERR_POINTER_IS_NULL,
               "null pointer argument");
  MIFE_REQUIRE(n > 0 && n_s > 0 && n_s <= n && nnz > 0 && nnz <= (int64_t)INT32_MAX, MIFE_ERR_ARG,
               "invalid sizes n=%d n_s=%d nnz=%lld", n, n_s, (long long)nnz);
  hipStream_t s = (hipStream_t)stream;
  CoupleWork w;
  layout(n, nnz, n_s, reinterpret_cast<char*>(work), &w);
  const unsigned long long ns = (unsigned long long)n_s;
  const int grid = (int)std::min<long long>(div_up(nnz, 256), 8192);
  if (s_col == nullptr) {
    // pass 1: sort, run heads, row counts, groups
    MIFE_CHECK_HIP(hipMemsetAsync(w.flags, 0, sizeof(int32_t) * 16, s));
    MIFE_CHECK_HIP(hipMemsetAsync(w.row_cnt, 0, sizeof(int32_t) * ((size_t)n_s + 1), s));
    MIFE_CHECK_HIP(hipMemsetAsync(w.grp_cnt, 0, sizeof(int32_t) * ((size_t)n_s + 1), s));
    couple_keys_kernel<<<div_up(n, 4), 256, 0, s>>>(n, row_ptr, col, owner, n_s, w.key_in, w.pos_in, w.prow,
                                                     w.flags);
    couple_group_keys_kernel<<<div_up(n, 256), 256, 0, s>>>(n, owner, n_s, w.gkey_in, w.grow_in, w.grp_cnt);
    MIFE_CHECK_HIP(hipGetLastError());
    int hflag = 0;
    MIFE_CHECK_HIP(hipMemcpyAsync(&hflag, w.flags, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    MIFE_CHECK_HIP(hipStreamSynchronize(s));
    MIFE_REQUIRE(hflag == 0, MIFE_ERR_ARG, "owner[] outside [-1, %d) or column outside [0, %d)", n_s, n);
    size_t tb = w.temp_bytes;
    MIFE_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(w.temp, tb, w.key_in, w.key_out, w.pos_in, w.pos_out,
                                                      (int)nnz, 0, bits_for(ns * ns), s));
    tb = w.temp_bytes;
    MIFE_CHECK_HIP(hipcub::DeviceRadixSort::SortPairs(w.temp, tb, w.gkey_in, w.gkey_out, w.grow_in, w.grow_out, n,
                                                      0, bits_for(ns), s));
    couple_heads_kernel<<<grid, 256, 0, s>>>(nnz, w.key_out, n_s, w.head, w.row_cnt, w.grp_cnt, w.flags);
    MIFE_CHECK_HIP(hipGetLastError());
    tb = w.temp_bytes;
    MIFE_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(w.temp, tb, w.head, w.head_idx, (int)nnz, s));
    tb = w.temp_bytes;
    MIFE_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(w.temp, tb, w.row_cnt, s_row_ptr, n_s + 1, s));
    tb = w.temp_bytes;
    MIFE_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(w.temp, tb, w.grp_cnt, grp_ptr, n_s + 1, s));
    int32_t tail[2] = {0, 0}, hf[3] = {0, 0, 0};
    MIFE_CHECK_HIP(hipMemcpyAsync(&tail[0], s_row_ptr + n_s, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    MIFE_CHECK_HIP(hipMemcpyAsync(&tail[1], grp_ptr + n_s, sizeof(int32_t), hipMemcpyDeviceToHost, s));
    MIFE_CHECK_HIP(hipMemcpyAsync(hf, w.flags, sizeof(int32_t) * 3, hipMemcpyDeviceToHost, s));
    MIFE_CHECK_HIP(hipStreamSynchronize(s));
    MIFE_REQUIRE(hf[2] == 0, MIFE_ERR_ARG, "a solver DOF in [0, %d) has no equations row", n_s);
    if (grp_rows && tail[1] > 0)
      MIFE_CHECK_HIP(hipMemcpyAsync(grp_rows, w.grow_out, sizeof(int32_t) * (size_t)tail[1],
                                    hipMemcpyDeviceToDevice, s));
    if (nnz_s_out) *nnz_s_out = tail[0];
    if (n_contrib_out) *n_contrib_out = hf[1];
    return MIFE_OK;
  }
  // pass 2: solver columns and the contribution runs (work holds pass 1's arrays)
  MIFE_REQUIRE(contrib_ptr && contrib_pos && contrib_row, MIFE_ERR_POINTER_IS_NULL, "null contribution arrays");
  int32_t nv = 0;
  MIFE_CHECK_HIP(hipMemcpyAsync(&nv, w.flags + 1, sizeof(int32_t), hipMemcpyDeviceToHost, s));
  MIFE_CHECK_HIP(hipStreamSynchronize(s));
  if (nv > 0)
    couple_fill_kernel<<<std::min(div_up(nv, 256), 8192), 256, 0, s>>>(nv, w.key_out, n_s, w.head, w.head_idx,
                                                                       w.pos_out, w.prow, s_col, contrib_ptr,
                                                                       contrib_pos, contrib_row);
  couple_tail_kernel<<<1, 64, 0, s>>>(w.flags, s_row_ptr, n_s, contrib_ptr);
  MIFE_CHECK_HIP(hipGetLastError());
  return MIFE_OK;
}

int mife_couple_values(int64_t nnz_s, const int32_t* contrib_ptr, const int32_t* contrib_pos,
                       const int32_t* contrib_row, const int32_t* col, const double* coef, const void* src,
                       int src_prec, void* dst, int dst_prec, void* stream) {
  if (nnz_s == 0) return MIFE_OK;
  MIFE_REQUIRE(contrib_ptr && contrib_pos && contrib_row && col && coef && src && dst, MIFE_ERR_POINTER_IS_NULL,
               "null pointer argument");
  MIFE_REQUIRE((src_prec == MIFE_F64 || src_prec == MIFE_F32) && (dst_prec == MIFE_F64 || dst_prec == MIFE_F32),
               MIFE_ERR_ARG, "invalid precision");
  hipStream_t s = (hipStream_t)stream;
  const int grid = (int)std::min<long long>(div_up(nnz_s, 256), 8192);
#define COUPLE(S, D)                                                                                        \
  couple_values_kernel<S, D><<<grid, 256, 0, s>>>(nnz_s, contrib_ptr, contrib_pos, contrib_row, col, coef, \
                                                  (const S*)src, (D*)dst)
  if (src_prec == MIFE_F64 && dst_prec == MIFE_F64) COUPLE(double, double);
  else if (src_prec == MIFE_F64) COUPLE(double, float);
  else if (dst_prec == MIFE_F32) COUPLE(float, float);
  else COUPLE(float, double);
#undef COUPLE
  MIFE_CHECK_HIP(hipGetLastError());
  return MIFE_OK;
}

int mife_solver_rhs_coupled(int n_s, const int32_t* grp_ptr, const int32_t* grp_rows, const double* coef,
                            const int32_t* row_ptr, const int32_t* col, const double* vals, const int32_t* owner,
                            const double* u, const double* src, const double* rhs, const uint8_t* rhs_fixed,
                            double* b, void* stream) {
  if (n_s == 0) return MIFE_OK;
  MIFE_REQUIRE(grp_ptr && grp_rows && coef && row_ptr && col && vals && owner && u && b, MIFE_ERR_POINTER_IS_NULL,
               "null pointer argument");
  coupled_rhs_kernel<<<div_up(n_s, 4), 256, 0, (hipStream_t)stream>>>(n_s, grp_ptr, grp_rows, coef, row_ptr, col,
                                                                      vals, owner, u, src, rhs, rhs_fixed, b);
  MIFE_CHECK_HIP(hipGetLastError());
  return MIFE_OK;
}

int mife_field_update_coupled(int n, const int32_t* owner, const double* coef, const double* x, double* u,
                              void* stream) {
  if (n == 0) return MIFE_OK;
  MIFE_REQUIRE(owner && coef && x && u, MIFE_ERR_POINTER_IS_NULL, "null pointer argument");
  coupled_field_update_kernel<<<div_up(n, 256), 256, 0, (hipStream_t)stream>>>(n, owner, coef, x, u);
  MIFE_CHECK_HIP(hipGetLastError());
  return MIFE_OK;
}

}  // extern "C"
