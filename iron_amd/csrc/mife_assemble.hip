// Element integration + scatter-add into CSR values.
//
// One 64-lane wavefront integrates one element: lanes gather the element's
// node coordinates into LDS, one lane per Gauss point forms dx/dxi, its
// inverse and |J| (always FP64), then the lanes build the physical gradients
// B[(g,i)][a] and contract K = B^T W B in the value precision T, and finally
// add K into the CSR values through the precomputed position map.  Colours
// are launched one after another, so every CSR entry is updated by at most one
// wavefront per launch and the summation order is fixed (deterministic).
#include "mife_common.hpp"
#include "mife_grid.hpp"
#include "mife_profile.hpp"

#include <algorithm>
#include <vector>

namespace mife {

struct Mat3 {
  double a[3][3];
};

// 2x2 / 3x3 inverse by cofactors (maths.F90 InvertFullDP ordering of terms).
template <int DIM>
__device__ __forceinline__ double invert(const double (&A)[3][3], double (&B)[3][3]) {
  if constexpr (DIM == 2) {
    const double det = A[0][0] * A[1][1] - A[1][0] * A[0][1];
    B[0][0] = A[1][1] / det;
    B[0][1] = -A[0][1] / det;
    B[1][0] = -A[1][0] / det;
    B[1][1] = A[0][0] / det;
    return det;
  } else {
    const double det = A[0][0] * A[1][1] * A[2][2] + A[0][1] * A[1][2] * A[2][0] + A[0][2] * A[1][0] * A[2][1] -
                       A[0][0] * A[2][1] * A[1][2] - A[1][0] * A[0][1] * A[2][2] - A[2][0] * A[1][1] * A[0][2];
    B[0][0] = (A[1][1] * A[2][2] - A[2][1] * A[1][2]) / det;
    B[1][0] = (A[1][2] * A[2][0] - A[2][2] * A[1][0]) / det;
    B[2][0] = (A[1][0] * A[2][1] - A[2][0] * A[1][1]) / det;
    B[0][1] = (A[2][1] * A[0][2] - A[0][1] * A[2][2]) / det;
    B[1][1] = (A[2][2] * A[0][0] - A[0][2] * A[2][0]) / det;
    B[2][1] = (A[2][0] * A[0][1] - A[0][0] * A[2][1]) / det;
    B[0][2] = (A[0][1] * A[1][2] - A[1][1] * A[0][2]) / det;
    B[1][2] = (A[0][2] * A[1][0] - A[1][2] * A[0][0]) / det;
    B[2][2] = (A[0][0] * A[1][1] - A[1][0] * A[0][1]) / det;
    return det;
  }
}

template <typename T>
__device__ __forceinline__ void add_value(T* p, T v, bool atomic) {
  if (atomic) atomicAdd(p, v);
  else *p += v;
}

// Stage layout of a structured (grid) assembly.  With row m = (c, a), column
// n = (c2, b), node a = (ax, ay, az) on the element's NB^3 lattice and element
// e = ex + nx (ey + ny ez), entry (m, n) of element e sits at
//   [ez][ey][c][az][ay][c2][bz][by][ex][ax][bx],
// i.e. stripe U = (ez ny + ey) SU + u(c, az, ay, c2, bz, by), SU = NC^2 NB^4 stripes
// per element row, each stripe nx NB^2 values long.  The entries that one x-line
// of CSR rows needs for one y/z neighbour block are then one contiguous stripe.
// The element kernels write the layout block-wise: each wavefront puts its
// element's entries into LDS as [u][wave][NB^2], and after a block barrier the
// block stores stripe by stripe, so the entries of the block's consecutive
// elements along x go out as one contiguous run per stripe.  Blocks are dealt to
// the eight XCDs in contiguous ranges (block b runs on XCD b % 8), so blocks that
// share a cache line share an L2.
template <int NN, int NC>
struct GridStripe {
  static constexpr int NB = NN == 8 ? 2 : 3, B2 = NB * NB, SU = NC * NC * B2 * B2;
  // LDS slot of entry (m, n) of wavefront w's element
  static __device__ __forceinline__ int slot(int m, int n, int w, int wpb) {
    const int c = m / NN, a = m - c * NN, c2 = n / NN, b = n - c2 * NN;
    const int u = ((((c * NB + a / B2) * NB + (a / NB) % NB) * NC + c2) * NB + b / B2) * NB + (b / NB) % NB;
    return (u * wpb + w) * B2 + (a % NB) * NB + b % NB;
  }
  // stage offset of element e's stripe-0 run
  static __device__ __forceinline__ long long base(long long e, int nx, int ny) {
    const long long ex = e % nx, eyz = e / nx;
    return (eyz * SU * nx + ex) * B2;
  }
  // logical block: contiguous ranges per XCD in a grid stage (gridDim.x a multiple of 8)
  static __device__ __forceinline__ int block(int nx) {
    if (nx <= 0) return blockIdx.x;
    return (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  }
  // block-wide store of the LDS slots K (sbase[w] < 0: wavefront w holds no element)
  template <typename T, int WPB>
  static __device__ __forceinline__ void store(const T* K, const long long* sbase, int nx, T* stage) {
    const long long stride = (long long)nx * B2;
    for (int f = threadIdx.x; f < SU * WPB * B2; f += 64 * WPB) {
      const int u = f / (WPB * B2), r = f - u * (WPB * B2), w = r / B2, j = r - w * B2;
      const long long b = sbase[w];
      if (b >= 0) stage[b + u * stride + j] = K[f];
    }
  }
};

// Per-Gauss-point geometry of one element held in LDS by its wavefront.
template <int DIM, int NN, int NG>
struct GeomLDS {
  double x[NN * DIM];
  double ji[NG][DIM][DIM];  // dxi_j/dx_i stored as ji[g][j][i]
  double rwg[NG];           // |J| * w_g
};

template <int DIM, int NN, int NG>
__device__ __forceinline__ void element_geometry(GeomLDS<DIM, NN, NG>& G, const double* __restrict__ coords,
                                                 const int32_t* __restrict__ enodes, long long e,
                                                 const double* __restrict__ dN, const double* __restrict__ w,
                                                 int lane, bool active) {
  if (active && lane < NN) {
    const long long node = enodes[e * NN + lane];
#pragma unroll
    for (int d = 0; d < DIM; ++d) G.x[lane * DIM + d] = coords[node * DIM + d];
  }
  __syncthreads();
  // dx/dxi at Gauss point g = lane % 32: the node sum split between the two
  // half-waves (nodes below and from NN / 2), the halves added across lanes
  // l and l + 32 (NG <= 32), else one lane per Gauss point
  constexpr bool SPLIT = NG <= 32;
  const int g = SPLIT ? (lane & 31) : lane;
  double J[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  if (active && g < NG) {
    const int a0 = SPLIT && lane >= 32 ? NN / 2 : 0, a1 = SPLIT && lane < 32 ? NN / 2 : NN;
    for (int a = a0; a < a1; ++a) {
#pragma unroll
      for (int j = 0; j < DIM; ++j) {
        const double dn = dN[(g * NN + a) * DIM + j];
#pragma unroll
        for (int d = 0; d < DIM; ++d) J[d][j] += G.x[a * DIM + d] * dn;
      }
    }
  }
  if constexpr (SPLIT) {
#pragma unroll
    for (int d = 0; d < DIM; ++d)
#pragma unroll
      for (int j = 0; j < DIM; ++j) J[d][j] += __shfl_xor(J[d][j], 32);
  }
  if (active && lane < NG) {
    double Ji[3][3];
    const double det = invert<DIM>(J, Ji);
#pragma unroll
    for (int j = 0; j < DIM; ++j)
#pragma unroll
      for (int i = 0; i < DIM; ++i) G.ji[g][j][i] = Ji[j][i];
    G.rwg[g] = fabs(det) * w[g];
  }
  __syncthreads();
}

// ------------------------------------------------------------ scalar field
// Inputs of one scalar assembly (one mesh, one set of CSR values).
template <typename T>
struct ScalarJob {
  const double* coords;
  const int32_t* enodes;
  const int32_t* edofs;
  const double* kappa;
  const double* source;
  const int32_t* pos;   // NULL: stage the element matrix at vals[e][NN][NN]
  T* vals;
  double* svec;
  double* sstage;       // non-NULL: stage the element source vector at sstage[e][NN]
  int gnx, gny;         // GRID: elements in x and y (GridStripe layout)
};

// One element per wavefront (element e when `active`); every wavefront of the
// block calls it the same number of times (it holds block barriers).  GRID: the
// element matrix is stored (not added) in the grid stage (GridStripe) for the
// row-gather pass of a structured assembly.
template <int DIM, int NN, int NG, typename T, int WPB, bool GRID = false>
__device__ __forceinline__ void scalar_element(const ScalarJob<T>& J, long long e, bool active,
                                               const double* __restrict__ dN, const double* __restrict__ N,
                                               const double* __restrict__ w, bool atomic) {
  constexpr int KD = NG * DIM;
  static_assert(!GRID || DIM == 3, "the grid assembly is 3-D");
  using GS = GridStripe<(DIM == 3 ? NN : 8), 1>;
  __shared__ T sK[GRID ? WPB * NN * NN : 1];
  __shared__ long long sbase[WPB];
  __shared__ GeomLDS<DIM, NN, NG> geo[WPB];
  __shared__ T sB[WPB][KD * NN];
  __shared__ T sW[WPB][KD];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  GeomLDS<DIM, NN, NG>& G = geo[wv];
  element_geometry<DIM, NN, NG>(G, J.coords, J.enodes, e, dN, w, lane, active);
  const double ke = (active && J.kappa) ? J.kappa[e] : 1.0;
  if (active) {
    for (int t = lane; t < NG * NN; t += 64) {
      const int g = t / NN, a = t - (t / NN) * NN;
#pragma unroll
      for (int i = 0; i < DIM; ++i) {
        double b = 0.0;
#pragma unroll
        for (int j = 0; j < DIM; ++j) b += dN[(g * NN + a) * DIM + j] * G.ji[g][j][i];
        sB[wv][(g * DIM + i) * NN + a] = (T)b;
      }
    }
    for (int k = lane; k < KD; k += 64) sW[wv][k] = (T)(G.rwg[k / DIM] * ke);
  }
  __syncthreads();
  if (active) {
    const T* B = sB[wv];
    const T* W = sW[wv];
    for (int t = lane; t < NN * NN; t += 64) {
      const int m = t / NN, n = t - (t / NN) * NN;
      T acc = 0;
#pragma unroll 9
      for (int k = 0; k < KD; ++k) acc += W[k] * B[k * NN + m] * B[k * NN + n];
      if constexpr (GRID)
        sK[GS::slot(m, n, wv, WPB)] = acc;
      else
        add_value(&J.vals[J.pos[e * (NN * NN) + t]], acc, atomic);
    }
    if (J.source && (J.svec || J.sstage) && lane < NN) {
      // Rounded product and sum, unfused whatever the caller inlines around them
      // (a kernel whose `atomic` is a constant would otherwise fuse them).
#pragma clang fp contract(off)
      double s = 0.0;
      for (int g = 0; g < NG; ++g) s = fma(N[g * NN + lane], G.rwg[g], s);
      s = s * J.source[e];
      if (J.sstage)
        J.sstage[e * NN + lane] = s;
      else
        add_value(&J.svec[J.edofs[e * NN + lane]], s, atomic);
    }
  }
  if constexpr (GRID) {
    if (lane == 0) sbase[wv] = active ? GS::base(e, J.gnx, J.gny) : -1;
    __syncthreads();
    GS::template store<T, WPB>(sK, sbase, J.gnx, J.vals);
  }
}

template <int DIM, int NN, int NG, typename T, int WPB, bool GRID = false>
__global__ __launch_bounds__(64 * WPB) void scalar_assemble_kernel(
    const double* __restrict__ coords, const int32_t* __restrict__ enodes, const int32_t* __restrict__ edofs,
    const double* __restrict__ kappa, const double* __restrict__ source, const int32_t* __restrict__ pos,
    const int32_t* __restrict__ order, int elem_begin, int elem_count, const double* __restrict__ dN,
    const double* __restrict__ N, const double* __restrict__ w, T* __restrict__ vals, double* __restrict__ svec,
    bool atomic, double* __restrict__ sstage = nullptr, int gnx = 0, int gny = 0) {
  const int blk = GRID ? GridStripe<NN, 1>::block(gnx) : (int)blockIdx.x;
  const int idx = blk * WPB + (threadIdx.x >> 6);
  const bool active = idx < elem_count;
  long long e = 0;
  if (active) e = order ? order[elem_begin + idx] : (long long)elem_begin + idx;
  scalar_element<DIM, NN, NG, T, WPB, GRID>(
      ScalarJob<T>{coords, enodes, edofs, kappa, source, pos, vals, svec, sstage, gnx, gny}, e, active, dN, N, w,
      atomic);
}

// ------------------------------------------------ triquadratic, on MFMA
// K = B^T diag(W) B with B[(g,i)][a] = dphi_a/dx_i at Gauss point g (81 x 27,
// padded to 84 x 32) runs as 16x16x4 MFMA tiles: tile (ti, tj) of K takes 21
// steps over the padded (g,i) axis, A operand = B^T rows, B operand = W B.  K is
// symmetric, so tiles (0,0), (0,1), (1,1) are formed and (1,0) is written as
// the transpose of (0,1).  At step t lane l forms the two gradients
// B[4t + l/16][16c + l%16], c = 0, 1, from the reference gradients (LDS) and
// dxi/dx: each is its A-operand entry and, times W, its B-operand entry.
template <typename T>
struct Mfma16;
template <>
struct Mfma16<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mac(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // row of accumulator register i of lane l (column l % 16)
  static __device__ __forceinline__ int row(int lane, int i) { return 4 * i + (lane >> 4); }
};
template <>
struct Mfma16<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 mac(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int i) { return 4 * (lane >> 4) + i; }
};

template <typename T, int WPB>
__global__ __launch_bounds__(64 * WPB) __attribute__((amdgpu_waves_per_eu(4))) void q2_assemble_mfma_kernel(
    const double* __restrict__ coords, const int32_t* __restrict__ enodes, const int32_t* __restrict__ edofs,
    const double* __restrict__ kappa, const double* __restrict__ source, const int32_t* __restrict__ pos,
    const int32_t* __restrict__ order, int elem_begin, int elem_count, const double* __restrict__ dN,
    const double* __restrict__ N, const double* __restrict__ w, T* __restrict__ vals, double* __restrict__ svec,
    bool atomic, double* __restrict__ sstage = nullptr, int gnx = 0, int gny = 0) {
  constexpr int NN = 27, NG = 27, KD = 81, STEPS = 21;
  using M = Mfma16<T>;
  using GS = GridStripe<NN, 1>;
  // the element matrices of a grid stage reuse the LDS of the reference
  // gradients and geometry once every wavefront is past its products
  struct Tables {
    double dN[NG * NN * 3];
    GeomLDS<3, NN, NG> geo[WPB];
  };
  union Shared {
    Tables t;
    T K[WPB * NN * NN];
  };
  static_assert(sizeof(T) * WPB * NN * NN <= sizeof(Tables), "element matrices exceed the tables' LDS");
  __shared__ Shared sh;
  __shared__ long long sbase[WPB];
  double* sdN = sh.t.dN;
  for (int t = threadIdx.x; t < NG * NN * 3; t += blockDim.x) sdN[t] = dN[t];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int idx = GS::block(gnx) * WPB + wv;
  const bool active = idx < elem_count;
  long long e = 0;
  if (active) e = order ? order[elem_begin + idx] : (long long)elem_begin + idx;
  GeomLDS<3, NN, NG>& G = sh.t.geo[wv];
  element_geometry<3, NN, NG>(G, coords, enodes, e, sdN, w, lane, active);
  if (pos && !active) return;  // the scatter holds no further barrier
  const int kq = lane >> 4, col = lane & 15;
  typename M::v4 k00 = {0, 0, 0, 0}, k01 = {0, 0, 0, 0}, k11 = {0, 0, 0, 0};
  double s = 0.0;
  if (active) {
    const double ke = kappa ? kappa[e] : 1.0;
    // lane's two operand positions: a = col and a = 16 + col (the latter is padding for col >= 11)
    auto grad = [&](int k, int a) -> T {
      if (k >= KD || a >= NN) return (T)0;
      const int g = k / 3, i = k - 3 * (k / 3);
      const double* d = &sdN[(g * NN + a) * 3];
      return (T)(d[0] * G.ji[g][0][i] + d[1] * G.ji[g][1][i] + d[2] * G.ji[g][2][i]);
    };
#pragma unroll 3
    for (int t = 0; t < STEPS; ++t) {
      const int k = 4 * t + kq;
      const T wk = k < KD ? (T)(G.rwg[k / 3] * ke) : (T)0;
      const T a0 = grad(k, col), a1 = grad(k, 16 + col);
      k00 = M::mac(a0, a0 * wk, k00);
      k01 = M::mac(a0, a1 * wk, k01);
      k11 = M::mac(a1, a1 * wk, k11);
    }
    if (source && (svec || sstage) && lane < NN) {
      for (int g = 0; g < NG; ++g) s += N[g * NN + lane] * G.rwg[g];
      s *= source[e];
    }
  }
  if (pos) {
    const int32_t* pe = pos + e * (NN * NN);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = M::row(lane, i);
      // tile (0,0): m = r, n = col
      add_value(&vals[pe[r * NN + col]], k00[i], atomic);
      // tile (0,1) and its transpose: m = r, n = 16 + col
      if (16 + col < NN) {
        add_value(&vals[pe[r * NN + 16 + col]], k01[i], atomic);
        add_value(&vals[pe[(16 + col) * NN + r]], k01[i], atomic);
      }
      // tile (1,1): m = 16 + r, n = 16 + col
      if (16 + r < NN && 16 + col < NN) add_value(&vals[pe[(16 + r) * NN + 16 + col]], k11[i], atomic);
    }
    if (source && svec && lane < NN) add_value(&svec[edofs[e * NN + lane]], s, atomic);
    return;
  }
  // grid stage: the element source vector in element order, the matrix in stripes
  if (active && source && sstage && lane < NN) sstage[e * NN + lane] = s;
  __syncthreads();  // every wavefront is done with the tables
  if (active) {
    T* K = sh.K;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = M::row(lane, i);
      K[GS::slot(r, col, wv, WPB)] = k00[i];
      if (16 + col < NN) {
        K[GS::slot(r, 16 + col, wv, WPB)] = k01[i];
        K[GS::slot(16 + col, r, wv, WPB)] = k01[i];
      }
      if (16 + r < NN && 16 + col < NN) K[GS::slot(16 + r, 16 + col, wv, WPB)] = k11[i];
    }
  }
  if (lane == 0) sbase[wv] = active ? GS::base(e, gnx, gny) : -1;
  __syncthreads();
  GS::template store<T, WPB>(sh.K, sbase, gnx, vals);
}

// ------------------------------------------------------- linear elasticity
// Iron's per-block coefficient use (linear_elasticity_routines.F90:1255-1286):
// diagonal block c uses JRWG*diagC[c][d] on dphi_d*dphi_d; off-diagonal block
// (c1<c2) uses Ca*dphi_n,c1*dphi_m,c2 + Cb*dphi_n,c2*dphi_m,c1.
template <int NN, int NG, typename T, int WPB>
__global__ __launch_bounds__(64 * WPB) void elastic_assemble_kernel(
    const double* __restrict__ coords, const int32_t* __restrict__ enodes, const double* __restrict__ material,
    const int32_t* __restrict__ pos, const int32_t* __restrict__ order, int elem_begin, int elem_count,
    const double* __restrict__ dN, const double* __restrict__ w, T* __restrict__ vals, bool atomic) {
  constexpr int ND = 3 * NN;
  __shared__ GeomLDS<3, NN, NG> geo[WPB];
  __shared__ T sB[WPB][NG][NN][3];
  __shared__ T sC[WPB][NG][12];  // per GP: diag[3][3] (9) + off Ca,Cb handled below
  __shared__ T sOff[WPB][NG][6];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int idx = blockIdx.x * WPB + wv;
  const bool active = idx < elem_count;
  long long e = 0;
  if (active) e = order ? order[elem_begin + idx] : (long long)elem_begin + idx;
  GeomLDS<3, NN, NG>& G = geo[wv];
  element_geometry<3, NN, NG>(G, coords, enodes, e, dN, w, lane, active);
  if (active) {
    for (int t = lane; t < NG * NN; t += 64) {
      const int g = t / NN, a = t - (t / NN) * NN;
      // dphidx(a,mi) = sum_ni DXI_DX(mi,ni) dphi/dxi_ni   (linear_elasticity_routines.F90:1245-1252)
#pragma unroll
      for (int mi = 0; mi < 3; ++mi) {
        double b = 0.0;
#pragma unroll
        for (int ni = 0; ni < 3; ++ni) b += G.ji[g][mi][ni] * dN[(g * NN + a) * 3 + ni];
        sB[wv][g][a][mi] = (T)b;
      }
    }
    if (lane < NG) {
      const int g = lane;
      const double* mat = material + e * 6;
      const double E1 = mat[0], E2 = mat[1], E3 = mat[2], v13 = mat[3], v23 = mat[4], v12 = mat[5];
      const double v31 = v13, v32 = v23, v21 = v12;
      const double gama = 1.0 / (1.0 - v12 * v21 - v23 * v32 - v31 * v13 - 2.0 * v21 * v32 * v13);
      const double C11 = E1 * (1.0 - v23 * v32) * gama, C22 = E2 * (1.0 - v13 * v31) * gama;
      const double C33 = E3 * (1.0 - v12 * v21) * gama;
      const double C12 = E1 * (v21 + v31 * v23) * gama, C13 = E1 * (v31 + v21 * v32) * gama;
      const double C23 = E2 * (v32 + v12 * v31) * gama;
      const double C44 = E2 / (2.0 * (1.0 + v23)), C55 = E1 / (2.0 * (1.0 + v13)), C66 = E3 / (2.0 * (1.0 + v12));
      const double jr = G.rwg[g];
      const double diag[9] = {C11, C66, C44, C66, C22, C55, C44, C55, C33};
      for (int k = 0; k < 9; ++k) sC[wv][g][k] = (T)(jr * diag[k]);
      const double off[6] = {C12, C66, C13, C44, C23, C55};
      for (int k = 0; k < 6; ++k) sOff[wv][g][k] = (T)(jr * off[k]);
    }
  }
  __syncthreads();
  if (active) {
    for (int t = lane; t < ND * ND; t += 64) {
      const int r = t / ND, s = t - (t / ND) * ND;
      int c1 = r / NN, n = r - (r / NN) * NN;
      int c2 = s / NN, m = s - (s / NN) * NN;
      T acc = 0;
      if (c1 == c2) {
        for (int g = 0; g < NG; ++g) {
          const T* D = &sC[wv][g][c1 * 3];
          acc += D[0] * sB[wv][g][n][0] * sB[wv][g][m][0] + D[1] * sB[wv][g][n][1] * sB[wv][g][m][1] +
                 D[2] * sB[wv][g][n][2] * sB[wv][g][m][2];
        }
      } else {
        // order the pair so that (c1 < c2); the lower blocks are the transpose
        if (c1 > c2) {
          int tc = c1; c1 = c2; c2 = tc;
          int tn = n; n = m; m = tn;
        }
        const int blk = (c1 == 0) ? (c2 == 1 ? 0 : 1) : 2;
        for (int g = 0; g < NG; ++g) {
          const T Ca = sOff[wv][g][2 * blk], Cb = sOff[wv][g][2 * blk + 1];
          acc += Ca * sB[wv][g][n][c1] * sB[wv][g][m][c2] + Cb * sB[wv][g][n][c2] * sB[wv][g][m][c1];
        }
      }
      const int p = pos[e * (ND * ND) + t];
      add_value(&vals[p], acc, atomic);
    }
  }
}

// Trilinear elasticity on MFMA: K[(i,a)][(j,b)] = sum_g sum_L U[(i,j,a)][(g,L)]
// B[g][b][L], one (9 NN) x (3 NG) by (3 NG) x NN GEMM per element (one
// wavefront per element, 16x16x4 tiles).  U carries iron's per-block
// coefficients: U[(c,c,a)][(g,L)] = w diagC[c][L] B[g][a][L] on the diagonal
// blocks and, for i != j, U[(i,j,a)][(g,L)] = w (Ca B[g][a][i] [L == j] +
// Cb B[g][a][j] [L == i]) with (Ca, Cb) of the pair {i, j}
// (linear_elasticity_routines.F90:1255-1286; the lower blocks are the transposes).
template <typename T>
struct ElasticJob {
  const double* coords;
  const int32_t* enodes;
  const double* material;
  const int32_t* pos;
  T* vals;
  int gnx, gny;  // GRID: elements in x and y (GridStripe layout)
};

// GRID: every wavefront of the block must call it (it ends with block barriers);
// the element matrix goes to the grid stage (GridStripe) instead of the CSR.
template <typename T, int NN, int NG, int WPB, bool GRID = false>
__device__ __forceinline__ void elastic_element(const ElasticJob<T>& J, long long e, bool active,
                                                const double* __restrict__ dN, const double* __restrict__ w,
                                                bool atomic) {
  constexpr int ND = 3 * NN, MROWS = 9 * NN, KD = 3 * NG;
  constexpr int MT = (MROWS + 15) / 16, STEPS = (KD + 3) / 4;
  static_assert(NN <= 16, "one column tile");
  using M = Mfma16<T>;
  using GS = GridStripe<NN, 3>;
  __shared__ T sK[GRID ? WPB * ND * ND : 1];
  __shared__ long long sbase[WPB];
  __shared__ GeomLDS<3, NN, NG> geo[WPB];
  __shared__ T sB[WPB][NG][NN][3];
  __shared__ T sDiag[WPB][NG][9];  // w * diagC[c][d]
  __shared__ T sOff[WPB][NG][6];   // w * (Ca, Cb) of the pairs (0,1), (0,2), (1,2)
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  GeomLDS<3, NN, NG>& G = geo[wv];
  element_geometry<3, NN, NG>(G, J.coords, J.enodes, e, dN, w, lane, active);
  if (active) {
    for (int t = lane; t < NG * NN; t += 64) {
      const int g = t / NN, a = t - (t / NN) * NN;
      // dphidx(a,mi) = sum_ni DXI_DX(mi,ni) dphi/dxi_ni   (linear_elasticity_routines.F90:1245-1252)
#pragma unroll
      for (int mi = 0; mi < 3; ++mi) {
        double b = 0.0;
#pragma unroll
        for (int ni = 0; ni < 3; ++ni) b += G.ji[g][mi][ni] * dN[(g * NN + a) * 3 + ni];
        sB[wv][g][a][mi] = (T)b;
      }
    }
    if (lane < NG) {
      const int g = lane;
      const double* mat = J.material + e * 6;
      const double E1 = mat[0], E2 = mat[1], E3 = mat[2], v13 = mat[3], v23 = mat[4], v12 = mat[5];
      const double v31 = v13, v32 = v23, v21 = v12;
      const double gama = 1.0 / (1.0 - v12 * v21 - v23 * v32 - v31 * v13 - 2.0 * v21 * v32 * v13);
      const double C11 = E1 * (1.0 - v23 * v32) * gama, C22 = E2 * (1.0 - v13 * v31) * gama;
      const double C33 = E3 * (1.0 - v12 * v21) * gama;
      const double C12 = E1 * (v21 + v31 * v23) * gama, C13 = E1 * (v31 + v21 * v32) * gama;
      const double C23 = E2 * (v32 + v12 * v31) * gama;
      const double C44 = E2 / (2.0 * (1.0 + v23)), C55 = E1 / (2.0 * (1.0 + v13)), C66 = E3 / (2.0 * (1.0 + v12));
      const double jr = G.rwg[g];
      const double diag[9] = {C11, C66, C44, C66, C22, C55, C44, C55, C33};
      for (int k = 0; k < 9; ++k) sDiag[wv][g][k] = (T)(jr * diag[k]);
      const double off[6] = {C12, C66, C13, C44, C23, C55};
      for (int k = 0; k < 6; ++k) sOff[wv][g][k] = (T)(jr * off[k]);
    }
  }
  __syncthreads();
  if (!GRID && !active) return;  // the scatter holds no further barrier
  const int32_t* pe = GRID ? nullptr : J.pos + e * (ND * ND);
  const int col = lane & 15;
  for (int mt = 0; active && mt < MT; ++mt) {
    const int mrow = 16 * mt + col;
    const bool rv = mrow < MROWS;
    const int i = rv ? mrow / (3 * NN) : 0, j = rv ? (mrow / NN) % 3 : 0, a = rv ? mrow % NN : 0;
    const int blk = (i + j == 1) ? 0 : ((i + j == 2) ? 1 : 2);  // pairs (0,1), (0,2), (1,2)
    typename M::v4 acc = {0, 0, 0, 0};
#pragma unroll
    for (int st = 0; st < STEPS; ++st) {
      const int k = 4 * st + (lane >> 4);
      const bool kv = k < KD;
      const int g = kv ? k / 3 : 0, L = k - 3 * (k / 3);
      T u = 0, bop = 0;
      if (kv) {
        if (rv) {
          const T* Bg = &sB[wv][g][a][0];
          if (i == j)
            u = sDiag[wv][g][i * 3 + L] * Bg[L];
          else
            u = (L == j ? sOff[wv][g][2 * blk] * Bg[i] : (T)0) + (L == i ? sOff[wv][g][2 * blk + 1] * Bg[j] : (T)0);
        }
        if (col < NN) bop = sB[wv][g][col][L];
      }
      acc = M::mac(u, bop, acc);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int orow = 16 * mt + M::row(lane, r);
      if (orow < MROWS && col < NN) {
        const int oi = orow / (3 * NN), oj = (orow / NN) % 3, oa = orow % NN;
        if constexpr (GRID)
          sK[GS::slot(oi * NN + oa, oj * NN + col, wv, WPB)] = acc[r];
        else
          add_value(&J.vals[pe[(oi * NN + oa) * ND + oj * NN + col]], acc[r], atomic);
      }
    }
  }
  if constexpr (GRID) {
    if (lane == 0) sbase[wv] = active ? GS::base(e, J.gnx, J.gny) : -1;
    __syncthreads();
    GS::template store<T, WPB>(sK, sbase, J.gnx, J.vals);
  }
}

// GRID: a structured assembly's stage (elements in order, gnx x gny per z-layer).
template <typename T, int NN, int NG, int WPB, bool GRID = false>
__global__ __launch_bounds__(64 * WPB) void elastic_assemble_mfma_kernel(
    const double* __restrict__ coords, const int32_t* __restrict__ enodes, const double* __restrict__ material,
    const int32_t* __restrict__ pos, const int32_t* __restrict__ order, int elem_begin, int elem_count,
    const double* __restrict__ dN, const double* __restrict__ w, T* __restrict__ vals, bool atomic, int gnx = 0,
    int gny = 0) {
  const int blk = GRID ? GridStripe<NN, 3>::block(gnx) : (int)blockIdx.x;
  const int idx = blk * WPB + (threadIdx.x >> 6);
  const bool active = idx < elem_count;
  long long e = 0;
  if (active) e = order ? order[elem_begin + idx] : (long long)elem_begin + idx;
  elastic_element<T, NN, NG, WPB, GRID>(ElasticJob<T>{coords, enodes, material, pos, vals, gnx, gny}, e, active,
                                        dN, w, atomic);
}

// ------------------------------------------- many meshes in one launch
// A batch assembles the CSR values of many independent trilinear meshes (the
// requests a server admits together): one launch per colour and element kind
// covers every mesh of the batch.  A launch walks a list of segments -- the
// elements of one colour of one mesh -- as one flattened index space; each
// wavefront finds its segment by bisection and runs the same element code as
// the one-mesh kernels, so a mesh's values are bit for bit those of its own
// assembly.  Blocks loop over the index space (a resident grid).
struct BatchJob {
  const double* coords;
  const int32_t* enodes;
  const int32_t* edofs;
  const double* coef;
  const double* source;
  const int32_t* pos;
  const int32_t* order;
  void* vals;
  double* svec;
};

struct BatchSeg {
  int32_t job;     // index into the job table
  int32_t begin;   // first position of the colour in the job's element order
  int32_t prefix;  // first flattened index of the segment
  int32_t pad;
};

constexpr int kBatchGrid = 2048;

__device__ __forceinline__ const BatchSeg& batch_seg(const BatchSeg* __restrict__ segs, int nseg, int idx) {
  int lo = 0, hi = nseg - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (segs[mid].prefix <= idx)
      lo = mid;
    else
      hi = mid - 1;
  }
  return segs[lo];
}

template <typename T, int WPB>
__global__ __launch_bounds__(64 * WPB) void scalar_batch_kernel(const BatchJob* __restrict__ jobs,
                                                                const BatchSeg* __restrict__ segs, int nseg,
                                                                int total, const double* __restrict__ dN,
                                                                const double* __restrict__ N,
                                                                const double* __restrict__ w) {
  const int wv = threadIdx.x >> 6;
  for (int base = blockIdx.x * WPB; base < total; base += gridDim.x * WPB) {
    const int idx = base + wv;
    const bool active = idx < total;
    ScalarJob<T> J{};
    long long e = 0;
    if (active) {
      const BatchSeg sg = batch_seg(segs, nseg, idx);
      const BatchJob B = jobs[sg.job];
      J = ScalarJob<T>{B.coords, B.enodes, B.edofs, B.coef, B.source, B.pos, (T*)B.vals, B.svec, nullptr, 0, 0};
      e = B.order[sg.begin + idx - sg.prefix];
    }
    scalar_element<3, 8, 8, T, WPB>(J, e, active, dN, N, w, false);
  }
}

template <typename T, int WPB>
__global__ __launch_bounds__(64 * WPB) void elastic_batch_kernel(const BatchJob* __restrict__ jobs,
                                                                 const BatchSeg* __restrict__ segs, int nseg,
                                                                 int total, const double* __restrict__ dN,
                                                                 const double* __restrict__ w) {
  const int wv = threadIdx.x >> 6;
  for (int base = blockIdx.x * WPB; base < total; base += gridDim.x * WPB) {
    const int idx = base + wv;
    const bool active = idx < total;
    ElasticJob<T> J{};
    long long e = 0;
    if (active) {
      const BatchSeg sg = batch_seg(segs, nseg, idx);
      const BatchJob B = jobs[sg.job];
      J = ElasticJob<T>{B.coords, B.enodes, B.coef, B.pos, (T*)B.vals, 0, 0};
      e = B.order[sg.begin + idx - sg.prefix];
    }
    elastic_element<T, 8, 8, WPB>(J, e, active, dN, w, false);
  }
}

template <int DIM, int NN, int NG, typename T>
static int launch_scalar(const mife_element_desc* d, int n_elem, const double* coords, const int32_t* enodes,
                         const int32_t* edofs, const double* kappa, const double* source, const int32_t* pos,
                         const int32_t* order, const int32_t* cptr, int ncol, T* vals, double* svec,
                         hipStream_t s) {
  constexpr int WPB = (sizeof(T) * NG * DIM * NN > 12000) ? 1 : 4;
  constexpr bool kMfma = DIM == 3 && NN == 27 && NG == 27;
  constexpr int WQ = 4;  // waves (elements) per block of the MFMA kernel
  const int tk = prof::begin(s);
  auto launch = [&](const int32_t* ord, int begin, int count, bool atomic) {
    if constexpr (kMfma)
      q2_assemble_mfma_kernel<T, WQ><<<div_up(count, WQ), 64 * WQ, 0, s>>>(
          coords, enodes, edofs, kappa, source, pos, ord, begin, count, d->dN, d->N, d->w, vals, svec, atomic);
    else
      scalar_assemble_kernel<DIM, NN, NG, T, WPB><<<div_up(count, WPB), 64 * WPB, 0, s>>>(
          coords, enodes, edofs, kappa, source, pos, ord, begin, count, d->dN, d->N, d->w, vals, svec, atomic);
  };
  if (!order) {
    launch(nullptr, 0, n_elem, true);
  } else {
    for (int c = 0; c < ncol; ++c) {
      const int cnt = cptr[c + 1] - cptr[c];
      if (cnt > 0) launch(order, cptr[c], cnt, false);
    }
  }
  prof::end(tk, s, sizeof(T) == 8 ? prof::kAssembleScalar : prof::kAssembleScalar32);
  MIFE_CHECK_HIP(hipGetLastError());
  return MIFE_OK;
}

template <int NN, int NG, typename T>
static int launch_elastic(const mife_element_desc* d, int n_elem, const double* coords, const int32_t* enodes,
                          const double* material, const int32_t* pos, const int32_t* order, const int32_t* cptr,
                          int ncol, T* vals, hipStream_t s) {
  constexpr int WPB = 2;
  constexpr bool kMfma = NN <= 16;  // one MFMA column tile per element
  constexpr int WQ = 4;             // elements (wavefronts) per block of the MFMA kernel
  const int tk = prof::begin(s);
  auto launch = [&](const int32_t* ord, int begin, int count, bool atomic) {
    if constexpr (kMfma)
      elastic_assemble_mfma_kernel<T, NN, NG, WQ><<<div_up(count, WQ), 64 * WQ, 0, s>>>(
          coords, enodes, material, pos, ord, begin, count, d->dN, d->w, vals, atomic);
    else
      elastic_assemble_kernel<NN, NG, T, WPB><<<div_up(count, WPB), 64 * WPB, 0, s>>>(
          coords, enodes, material, pos, ord, begin, count, d->dN, d->w, vals, atomic);
  };
  if (!order) {
    launch(nullptr, 0, n_elem, true);
  } else {
    for (int c = 0; c < ncol; ++c) {
      const int cnt = cptr[c + 1] - cptr[c];
      if (cnt > 0) launch(order, cptr[c], cnt, false);
    }
  }
  prof::end(tk, s, sizeof(T) == 8 ? prof::kAssembleElastic : prof::kAssembleElastic32);
  MIFE_CHECK_HIP(hipGetLastError());
  return MIFE_OK;
}

// --------------------------------------------- structured (grid) assembly
// On a generated regular mesh (elements and nodes numbered x fastest, as iron's
// generated mesh numbers them; one or three components in separated DOF order)
// every CSR row is the tensor product of 1-D neighbour ranges, so a row and the
// elements that touch it follow from the node's grid index.  The assembly then
// runs as two passes: the element kernels store every element matrix in the
// GridStripe layout (one launch, no colouring), and this kernel writes each CSR row whole --
// entry by entry the sum of the element entries that meet there, in the order of
// the element colours (parity of the element index per direction), which is the
// order in which the coloured scatter adds them, so the values are bit for bit
// those of the coloured scatter.  No position map is read and no value is read
// back.  The 1-D neighbour arithmetic is in mife_grid.hpp.
//
// One wavefront (one block) per (component c, x-line (iy, iz), run of R
// consecutive nodes ix, R = P (64 / (P^2 + 2P)) so that the run's pairs fit the
// 64 lanes).  Lane l takes one pair (ix, jx): CSR row (c, ix, iy, iz) and its
// x-neighbour offset qx.  The wavefront walks the y/z neighbour blocks
// (c2, qz, qy) of the x-line, which are the same for every lane, so the y/z
// index arithmetic is wave-uniform.  In the GridStripe layout the entries of
// block (c2, qz, qy) for the element slots (sy, sz) are stripe U, nx NB^2
// values, of which lane l reads X = ex NB^2 + ax NB + bx for its element slots
// sx.  Per (c2, qz) the lanes first issue the loads of all qy, then sum each
// entry's (up to eight) element values in colour order -- parity of the element
// index per direction, z slowest, the order of the coloured scatter -- and put
// the sums in LDS, where each row's (c2, qz) piece, cy cx entries, is
// contiguous as in the CSR row; the pieces are then stored row by row.  Units
// are dealt to the eight XCDs in contiguous ranges (block b runs on XCD b % 8),
// so the x-lines that share stripes share an L2.
// stage: element matrices in the GridStripe layout; sstage (NC == 1, may be
// NULL): element source vectors [E][NN].  Every row must hold the full grid
// stencil (the host checks the element DOFs); entries beyond a row's stored
// length are not written.
template <typename T, typename I, int P, int NC>
__global__ __launch_bounds__(64) void grid_gather_kernel(int nx, int ny, int nz, int runs, int n_units,
                                                         const int32_t* __restrict__ row_ptr,
                                                         const T* __restrict__ stage, T* __restrict__ vals,
                                                         const double* __restrict__ sstage,
                                                         double* __restrict__ svec) {
  constexpr int NB = P + 1, NN = NB * NB * NB, PER = P * P + 2 * P, R = P * (64 / PER), Q = 2 * P + 1;
  constexpr int SU = NC * NC * NB * NB * NB * NB;
  __shared__ T piece[Q * 64];
  const int Nx = P * nx + 1, Ny = P * ny + 1, Nz = P * nz + 1;
  const int unit = (blockIdx.x & 7) * (gridDim.x >> 3) + (blockIdx.x >> 3);
  if (unit >= n_units) return;  // whole block: no barrier is reached
  const int lane = threadIdx.x;
  const int run = unit % runs, line = unit / runs;
  const int iy = line % Ny, iz = (line / Ny) % Nz, c = line / (Ny * Nz);
  const int r0 = run * R, r1 = min(r0 + R, Nx);
  const long long rowline = ((long long)c * Nz + iz) * ((long long)Ny * Nx) + (long long)iy * Nx;
  const int t0 = x_pair_start(r0, P, nx), npairs = x_pair_start(r1, P, nx) - t0;
  // wave-uniform y/z ranges
  int ly, hy, lz, hz;
  elem_range(iy, P, ny, ly, hy);
  elem_range(iz, P, nz, lz, hz);
  const int cy = P * (hy - ly) + P + 1, cz = P * (hz - lz) + P + 1;
  // lane's pair: row ix, neighbour jx, x element slots
  const bool act = lane < npairs;
  int ix = r0, jx = r0;
  if (act) x_pair(t0 + lane, P, ix, jx);
  int lx, hx, ljx, hjx;
  elem_range(ix, P, nx, lx, hx);
  elem_range(jx, P, nx, ljx, hjx);
  const int cx = P * (hx - lx) + P + 1, qx = jx - P * lx;
  const int ax = max(lx, ljx), bx = min(hx, hjx);
  int X[2];
  bool xv[2];
#pragma unroll
  for (int s = 0; s < 2; ++s) {
    const int ex = elem_of_parity(ax, bx, s);
    xv[s] = act && ex >= 0;
    X[s] = xv[s] ? (ex * NB + (ix - P * ex)) * NB + (jx - P * ex) : 0;
  }
  // LDS place of the lane's entry qy of a piece: rows' pieces back to back
  const int wslot = cy * (x_pair_start(ix, P, nx) - t0) + qx;
  // store plan: LDS slot k = lane + 64 i holds entry k of the pieces, i.e. of
  // row fix = row of pair t0 + k / cy, at (c2 cz + qz) cy cx(fix) + local
  int fpos[Q], fcx[Q], flen[Q];
#pragma unroll
  for (int i = 0; i < Q; ++i) {
    const int k = lane + 64 * i;
    fpos[i] = -1;
    fcx[i] = 0;
    flen[i] = 0;
    if (k < cy * npairs) {
      int fix, fjx;
      x_pair(t0 + k / cy, P, fix, fjx);
      int flx, fhx;
      elem_range(fix, P, nx, flx, fhx);
      const int rp = row_ptr[rowline + fix];
      fpos[i] = rp + (k - cy * (x_pair_start(fix, P, nx) - t0));
      fcx[i] = P * (fhx - flx) + P + 1;
      flen[i] = row_ptr[rowline + fix + 1];
    }
  }
  const I stripe = (I)nx * (NB * NB);
  for (int c2 = 0; c2 < NC; ++c2) {
    for (int qz = 0; qz < cz; ++qz) {
      const int jz = P * lz + qz;
      int ljz, hjz;
      elem_range(jz, P, nz, ljz, hjz);
      const int az = max(lz, ljz), bz = min(hz, hjz);
      int ez[2];
#pragma unroll
      for (int s = 0; s < 2; ++s) ez[s] = elem_of_parity(az, bz, s);
      // loads of every qy first (no use before all are issued)
      T v[Q][2][2][2];
      bool yv[Q][2];
#pragma unroll
      for (int qy = 0; qy < Q; ++qy) {
        const int jy = P * ly + qy;
        int ljy, hjy;
        elem_range(jy, P, ny, ljy, hjy);
        const int ay = max(ly, ljy), by = min(hy, hjy);
#pragma unroll
        for (int sy = 0; sy < 2; ++sy) {
          const int ey = qy < cy ? elem_of_parity(ay, by, sy) : -1;
          yv[qy][sy] = ey >= 0;
#pragma unroll
          for (int sz = 0; sz < 2; ++sz) {
            if (ey < 0 || ez[sz] < 0) continue;
            const I U = ((I)ez[sz] * ny + ey) * SU +
                        (((((c * NB + (iz - P * ez[sz])) * NB + (iy - P * ey)) * NC + c2) * NB + (jz - P * ez[sz])) *
                             NB +
                         (jy - P * ey));
            const T* st = stage + U * stripe;
#pragma unroll
            for (int sx = 0; sx < 2; ++sx)
              if (xv[sx]) v[qy][sz][sy][sx] = st[X[sx]];
          }
        }
      }
#pragma unroll
      for (int qy = 0; qy < Q; ++qy) {
        if (qy >= cy) continue;
        T sum = (T)0;
#pragma unroll
        for (int sz = 0; sz < 2; ++sz) {
          if (ez[sz] < 0) continue;
#pragma unroll
          for (int sy = 0; sy < 2; ++sy) {
            if (!yv[qy][sy]) continue;
#pragma unroll
            for (int sx = 0; sx < 2; ++sx)
              if (xv[sx]) sum += v[qy][sz][sy][sx];
          }
        }
        if (act) piece[wslot + qy * cx] = sum;
      }
      __syncthreads();
      const int blk = (c2 * cz + qz) * cy;
#pragma unroll
      for (int i = 0; i < Q; ++i) {
        const int at = fpos[i] + blk * fcx[i];
        if (fpos[i] >= 0 && at < flen[i]) vals[at] = piece[lane + 64 * i];
      }
      __syncthreads();
    }
  }
  if (NC == 1 && sstage && act && jx == ix) {
    double sv = 0.0;
#pragma unroll
    for (int col = 0; col < 8; ++col) {
      const int ex = elem_of_parity(lx, hx, col & 1), ey = elem_of_parity(ly, hy, (col >> 1) & 1),
                ez = elem_of_parity(lz, hz, col >> 2);
      if (ex < 0 || ey < 0 || ez < 0) continue;
      const long long e = ex + (long long)nx * (ey + (long long)ny * ez);
      const int a = (ix - P * ex) + NB * ((iy - P * ey) + NB * (iz - P * ez));
      sv += sstage[e * NN + a];
    }
    svec[rowline + ix] = sv;
  }
}

template <typename T>
static size_t grid_stage_bytes(long long E, int nn, int nc) {
  const size_t nd = (size_t)nc * nn;
  return ((size_t)E * nd * nd * sizeof(T) + 255) / 256 * 256 + (nc == 1 ? (size_t)E * nn * sizeof(double) : 0);
}

template <typename T, int P, int NC>
static int launch_grid(const mife_element_desc* d, const int* ne, const double* coords, const int32_t* enodes,
                       const double* coef, const double* source, const int32_t* row_ptr, long long n_rows, T* vals,
                       double* svec, void* stage, hipStream_t s) {
  constexpr int NB = P + 1, NN = NB * NB * NB, ND = NC * NN;
  const long long E = (long long)ne[0] * ne[1] * ne[2];
  T* kst = (T*)stage;
  double* sst = (NC == 1 && source && svec)
      ? (double*)((char*)stage + ((size_t)E * ND * ND * sizeof(T) + 255) / 256 * 256)
      : nullptr;
  const int tk = prof::begin(s);
  constexpr int WQ = 4;
  const int grid_e = div_up(div_up(E, WQ), 8) * 8;  // a multiple of the XCD count
  if constexpr (NC == 1) {
    if constexpr (NN == 27) {
      q2_assemble_mfma_kernel<T, WQ><<<grid_e, 64 * WQ, 0, s>>>(coords, enodes, nullptr, coef, source, nullptr,
                                                                nullptr, 0, (int)E, d->dN, d->N, d->w, kst,
                                                                nullptr, false, sst, ne[0], ne[1]);
    } else {
      scalar_assemble_kernel<3, NN, NN, T, 4, true><<<grid_e, 64 * 4, 0, s>>>(coords, enodes, nullptr, coef, source,
                                                                        nullptr, nullptr, 0, (int)E, d->dN, d->N,
                                                                        d->w, kst, nullptr, false, sst, ne[0],
                                                                        ne[1]);
    }
  } else {
    elastic_assemble_mfma_kernel<T, NN, NN, WQ, true><<<grid_e, 64 * WQ, 0, s>>>(coords, enodes, coef, nullptr,
                                                                                 nullptr,
                                                                           0, (int)E, d->dN, d->w, kst, false, ne[0],
                                                                           ne[1]);
  }
  constexpr int R = P * (64 / (P * P + 2 * P));
  const int runs = div_up(P * ne[0] + 1, R);
  const long long units = (long long)NC * (P * ne[1] + 1) * (P * ne[2] + 1) * runs;
  const int blocks = div_up(units, 8) * 8;  // a multiple of the XCD count
  if ((size_t)E * ND * ND < (size_t(1) << 31))
    grid_gather_kernel<T, int, P, NC><<<blocks, 64, 0, s>>>(ne[0], ne[1], ne[2], runs, (int)units, row_ptr, kst, vals,
                                                           sst, sst ? svec : nullptr);
  else
    grid_gather_kernel<T, long long, P, NC><<<blocks, 64, 0, s>>>(ne[0], ne[1], ne[2], runs, (int)units, row_ptr, kst,
                                                                 vals, sst, sst ? svec : nullptr);
  const int key = NC == 1 ? (sizeof(T) == 8 ? prof::kAssembleScalar : prof::kAssembleScalar32)
                          : (sizeof(T) == 8 ? prof::kAssembleElastic : prof::kAssembleElastic32);
  prof::end(tk, s, key);
  MIFE_CHECK_HIP(hipGetLastError());
  return MIFE_OK;
}

}  // namespace mife

using namespace mife;

extern "C" size_t mife_assemble_grid_stage_bytes(int n_elem, int n_nodes, int n_comp, int prec) {
  if (n_elem <= 0 || n_nodes <= 0 || (n_comp != 1 && n_comp != 3)) return 0;
  return prec == MIFE_F32 ? grid_stage_bytes<float>(n_elem, n_nodes, n_comp)
                          : grid_stage_bytes<double>(n_elem, n_nodes, n_comp);
}

extern "C" int mife_assemble_grid(const mife_element_desc* d, const int32_t* n_elem, int n_comp,
                                  const double* coords, const int32_t* elem_nodes, const double* coef,
                                  const double* source, const int32_t* row_ptr, int64_t n_rows, int prec, void* vals,
                                  double* src_vec, void* stage, size_t stage_bytes, void* stream) {
  MIFE_REQUIRE(d && n_elem && coords && elem_nodes && row_ptr && vals && stage && d->dN && d->w,
               MIFE_ERR_POINTER_IS_NULL, "null pointer argument");
  MIFE_REQUIRE(d->dim == 3, MIFE_ERR_ARG, "the grid assembly is 3-D");
  MIFE_REQUIRE(prec == MIFE_F64 || prec == MIFE_F32, MIFE_ERR_ARG, "invalid precision %d", prec);
  MIFE_REQUIRE(n_comp == 1 || n_comp == 3, MIFE_ERR_ARG, "%d components (1: scalar field, 3: elasticity)", n_comp);
  MIFE_REQUIRE(n_comp == 1 || coef, MIFE_ERR_POINTER_IS_NULL, "elasticity without material");
  MIFE_REQUIRE(n_comp == 1 || !source, MIFE_ERR_ARG, "a source vector is a scalar-field term");
  MIFE_REQUIRE(!source || !src_vec || d->N, MIFE_ERR_POINTER_IS_NULL, "source given without basis values N");
  MIFE_REQUIRE(n_elem[0] > 0 && n_elem[1] > 0 && n_elem[2] > 0, MIFE_ERR_ARG, "empty grid");
  int P = 0;
  if (d->n_nodes == 8 && d->n_gauss == 8) P = 1;
  if (d->n_nodes == 27 && d->n_gauss == 27 && n_comp == 1) P = 2;
  MIFE_REQUIRE(P > 0, MIFE_ERR_ARG, "grid assembly element: %d nodes, %d Gauss points, %d components", d->n_nodes,
               d->n_gauss, n_comp);
  const long long E = (long long)n_elem[0] * n_elem[1] * n_elem[2];
  const long long Nn = (long long)(P * n_elem[0] + 1) * (P * n_elem[1] + 1) * (P * n_elem[2] + 1);
  MIFE_REQUIRE(E < (1LL << 31) && n_rows < (1LL << 31) && n_rows == n_comp * Nn, MIFE_ERR_ARG,
               "%lld rows given, the grid has %lld nodes x %d components", (long long)n_rows, Nn, n_comp);
  const size_t need = prec == MIFE_F32 ? grid_stage_bytes<float>(E, d->n_nodes, n_comp)
                                       : grid_stage_bytes<double>(E, d->n_nodes, n_comp);
  MIFE_REQUIRE(stage_bytes >= need, MIFE_ERR_ARG, "stage of %zu bytes, %zu needed (mife_assemble_grid_stage_bytes)",
               stage_bytes, need);
  hipStream_t s = (hipStream_t)stream;
  const int ne[3] = {n_elem[0], n_elem[1], n_elem[2]};
  if (n_comp == 3) {
    if (prec == MIFE_F64)
      return launch_grid<double, 1, 3>(d, ne, coords, elem_nodes, coef, nullptr, row_ptr, n_rows, (double*)vals,
                                       nullptr, stage, s);
    return launch_grid<float, 1, 3>(d, ne, coords, elem_nodes, coef, nullptr, row_ptr, n_rows, (float*)vals, nullptr,
                                    stage, s);
  }
  if (P == 2) {
    if (prec == MIFE_F64)
      return launch_grid<double, 2, 1>(d, ne, coords, elem_nodes, coef, source, row_ptr, n_rows, (double*)vals,
                                       src_vec, stage, s);
    return launch_grid<float, 2, 1>(d, ne, coords, elem_nodes, coef, source, row_ptr, n_rows, (float*)vals, src_vec,
                                    stage, s);
  }
  if (prec == MIFE_F64)
    return launch_grid<double, 1, 1>(d, ne, coords, elem_nodes, coef, source, row_ptr, n_rows, (double*)vals,
                                     src_vec, stage, s);
  return launch_grid<float, 1, 1>(d, ne, coords, elem_nodes, coef, source, row_ptr, n_rows, (float*)vals, src_vec,
                                  stage, s);
}

extern "C" int mife_assemble_scalar(const mife_element_desc* d, int n_elem, const double* coords,
                                    const int32_t* elem_nodes, const int32_t* elem_dofs, const double* kappa,
                                    const double* source, const int32_t* pos, const int32_t* order,
                                    const int32_t* colour_ptr_host, int n_colours, int prec, void* vals,
                                    double* src_vec, void* stream) {
  MIFE_REQUIRE(d && coords && elem_nodes && elem_dofs && pos && vals && d->dN && d->w, MIFE_ERR_POINTER_IS_NULL,
               "null pointer argument");
  MIFE_REQUIRE(!source || d->N, MIFE_ERR_POINTER_IS_NULL, "source given without basis values N");
  MIFE_REQUIRE(!order || colour_ptr_host, MIFE_ERR_POINTER_IS_NULL, "colour order without colour_ptr");
  MIFE_REQUIRE(prec == MIFE_F64 || prec == MIFE_F32, MIFE_ERR_ARG, "invalid precision %d", prec);
  if (n_elem == 0) return MIFE_OK;
  hipStream_t s = (hipStream_t)stream;
  const int key = d->dim * 10000 + d->n_nodes * 100 + d->n_gauss;
#define SCALAR_CASE(DIM, NN, NG)                                                                            \
  case DIM * 10000 + NN * 100 + NG:                                                                        \
    if (prec == MIFE_F64)                                                                                   \
      return launch_scalar<DIM, NN, NG, double>(d, n_elem, coords, elem_nodes, elem_dofs, kappa, source,   \
                                                pos, order, colour_ptr_host, n_colours, (double*)vals,     \
                                                src_vec, s);                                               \
    return launch_scalar<DIM, NN, NG, float>(d, n_elem, coords, elem_nodes, elem_dofs, kappa, source, pos, \
                                             order, colour_ptr_host, n_colours, (float*)vals, src_vec, s);
  switch (key) {
    SCALAR_CASE(2, 4, 4)
    SCALAR_CASE(2, 9, 9)
    SCALAR_CASE(3, 8, 8)
    SCALAR_CASE(3, 27, 27)
    SCALAR_CASE(3, 8, 27)
    default:
      set_error("unsupported scalar element (dim=%d, nodes=%d, gauss=%d)", d->dim, d->n_nodes, d->n_gauss);
      return MIFE_ERR_ARG;
  }
#undef SCALAR_CASE
}

extern "C" int mife_assemble_elastic(const mife_element_desc* d, int n_elem, const double* coords,
                                     const int32_t* elem_nodes, const int32_t* elem_dofs, const double* material,
                                     const int32_t* pos, const int32_t* order, const int32_t* colour_ptr_host,
                                     int n_colours, int prec, void* vals, void* stream) {
  (void)elem_dofs;
  MIFE_REQUIRE(d && coords && elem_nodes && material && pos && vals && d->dN && d->w, MIFE_ERR_POINTER_IS_NULL,
               "null pointer argument");
  MIFE_REQUIRE(!order || colour_ptr_host, MIFE_ERR_POINTER_IS_NULL, "colour order without colour_ptr");
  MIFE_REQUIRE(d->dim == 3, MIFE_ERR_ARG, "linear elasticity kernel is 3-D only");
  MIFE_REQUIRE(prec == MIFE_F64 || prec == MIFE_F32, MIFE_ERR_ARG, "invalid precision %d", prec);
  if (n_elem == 0) return MIFE_OK;
  hipStream_t s = (hipStream_t)stream;
  if (d->n_nodes == 8 && d->n_gauss == 8) {
    if (prec == MIFE_F64)
      return launch_elastic<8, 8, double>(d, n_elem, coords, elem_nodes, material, pos, order, colour_ptr_host,
                                          n_colours, (double*)vals, s);
    return launch_elastic<8, 8, float>(d, n_elem, coords, elem_nodes, material, pos, order, colour_ptr_host,
                                       n_colours, (float*)vals, s);
  }
  set_error("unsupported elastic element (nodes=%d, gauss=%d)", d->n_nodes, d->n_gauss);
  return MIFE_ERR_ARG;
}

namespace {

size_t align256(size_t v) { return (v + 255) & ~size_t(255); }

}  // namespace

extern "C" size_t mife_assemble_batch_work_bytes(int n_jobs, int max_colours) {
  const size_t nj = (size_t)std::max(n_jobs, 1);
  return align256(sizeof(BatchJob) * nj) + sizeof(BatchSeg) * nj * (size_t)std::max(max_colours, 1) + 256;
}

extern "C" int mife_assemble_batch(const mife_element_desc* scalar_desc, const mife_element_desc* elastic_desc,
                                   int n_jobs, const mife_assembly_job* jobs, int max_colours, int prec,
                                   void* table_host, void* work, void* stream) {
  MIFE_REQUIRE(n_jobs >= 0 && max_colours >= 0, MIFE_ERR_ARG, "invalid job count %d / colour bound %d", n_jobs,
               max_colours);
  if (n_jobs == 0) return MIFE_OK;
  MIFE_REQUIRE(jobs && table_host && work, MIFE_ERR_POINTER_IS_NULL, "null pointer argument");
  MIFE_REQUIRE(prec == MIFE_F64 || prec == MIFE_F32, MIFE_ERR_ARG, "invalid precision %d", prec);
  bool any_scalar = false, any_elastic = false, any_source = false;
  for (int k = 0; k < n_jobs; ++k) {
    const mife_assembly_job& J = jobs[k];
    MIFE_REQUIRE(J.kind == MIFE_JOB_SCALAR || J.kind == MIFE_JOB_ELASTIC, MIFE_ERR_ARG, "job %d: kind %d", k, J.kind);
    MIFE_REQUIRE(J.n_elem >= 0 && J.n_colours >= 0 && J.n_colours <= max_colours, MIFE_ERR_ARG,
                 "job %d: %d elements in %d colours (bound %d)", k, J.n_elem, J.n_colours, max_colours);
    if (J.n_elem == 0) continue;
    MIFE_REQUIRE(J.coords && J.elem_nodes && J.pos && J.order && J.colour_ptr_host && J.vals,
                 MIFE_ERR_POINTER_IS_NULL, "job %d: null pointer", k);
    MIFE_REQUIRE(J.colour_ptr_host[0] == 0 && J.colour_ptr_host[J.n_colours] == J.n_elem, MIFE_ERR_ARG,
                 "job %d: colour pointers do not cover its %d elements", k, J.n_elem);
    if (J.kind == MIFE_JOB_SCALAR) {
      any_scalar = true;
      if (J.source && J.src_vec) {
        MIFE_REQUIRE(J.elem_dofs, MIFE_ERR_POINTER_IS_NULL, "job %d: source without element DOFs", k);
        any_source = true;
      }
    } else {
      any_elastic = true;
      MIFE_REQUIRE(J.coef, MIFE_ERR_POINTER_IS_NULL, "job %d: elasticity without material", k);
    }
  }
  if (any_scalar) {
    MIFE_REQUIRE(scalar_desc && scalar_desc->dN && scalar_desc->w, MIFE_ERR_POINTER_IS_NULL, "scalar basis missing");
    MIFE_REQUIRE(scalar_desc->dim == 3 && scalar_desc->n_nodes == 8 && scalar_desc->n_gauss == 8, MIFE_ERR_ARG,
                 "batched scalar assembly takes trilinear elements with 2x2x2 Gauss points");
    MIFE_REQUIRE(!any_source || scalar_desc->N, MIFE_ERR_POINTER_IS_NULL, "source given without basis values N");
  }
  if (any_elastic) {
    MIFE_REQUIRE(elastic_desc && elastic_desc->dN && elastic_desc->w, MIFE_ERR_POINTER_IS_NULL,
                 "elastic basis missing");
    MIFE_REQUIRE(elastic_desc->dim == 3 && elastic_desc->n_nodes == 8 && elastic_desc->n_gauss == 8, MIFE_ERR_ARG,
                 "batched elastic assembly takes trilinear elements with 2x2x2 Gauss points");
  }
  // host table: jobs | segments of launch 0 | segments of launch 1 | ...
  char* H = (char*)table_host;
  BatchJob* hj = (BatchJob*)H;
  BatchSeg* hs = (BatchSeg*)(H + align256(sizeof(BatchJob) * n_jobs));
  for (int k = 0; k < n_jobs; ++k) {
    const mife_assembly_job& J = jobs[k];
    hj[k] = BatchJob{J.coords, J.elem_nodes, J.elem_dofs, J.coef, J.source, J.pos, J.order, J.vals, J.src_vec};
  }
  struct Launch {
    int kind, seg0, nseg, total;
  };
  std::vector<Launch> launches;
  int ns = 0;
  for (int c = 0; c < max_colours; ++c) {
    for (int kind : {MIFE_JOB_SCALAR, MIFE_JOB_ELASTIC}) {
      Launch L{kind, ns, 0, 0};
      for (int k = 0; k < n_jobs; ++k) {
        const mife_assembly_job& J = jobs[k];
        if (J.kind != kind || J.n_elem == 0 || c >= J.n_colours) continue;
        const int cnt = J.colour_ptr_host[c + 1] - J.colour_ptr_host[c];
        if (cnt <= 0) continue;
        MIFE_REQUIRE((long long)L.total + cnt < (1LL << 31), MIFE_ERR_ARG, "batch too large");
        hs[ns++] = BatchSeg{k, J.colour_ptr_host[c], L.total, 0};
        ++L.nseg;
        L.total += cnt;
      }
      if (L.nseg > 0) launches.push_back(L);
    }
  }
  hipStream_t s = (hipStream_t)stream;
  const size_t bytes = align256(sizeof(BatchJob) * n_jobs) + sizeof(BatchSeg) * (size_t)std::max(ns, 1);
  MIFE_CHECK_HIP(hipMemcpyAsync(work, table_host, bytes, hipMemcpyHostToDevice, s));
  MIFE_CHECK_HIP(hipStreamSynchronize(s));  // the host table may be reused at once
  const BatchJob* dj = (const BatchJob*)work;
  const BatchSeg* ds = (const BatchSeg*)((char*)work + align256(sizeof(BatchJob) * n_jobs));
  constexpr int WPB = 4;
  const int tk = prof::begin(s);
  for (const Launch& L : launches) {
    const int grid = std::max(1, std::min(div_up(L.total, WPB), kBatchGrid));
    if (L.kind == MIFE_JOB_SCALAR) {
      if (prec == MIFE_F64)
        scalar_batch_kernel<double, WPB><<<grid, 64 * WPB, 0, s>>>(dj, ds + L.seg0, L.nseg, L.total,
                                                                   scalar_desc->dN, scalar_desc->N, scalar_desc->w);
      else
        scalar_batch_kernel<float, WPB><<<grid, 64 * WPB, 0, s>>>(dj, ds + L.seg0, L.nseg, L.total,
                                                                  scalar_desc->dN, scalar_desc->N, scalar_desc->w);
    } else {
      if (prec == MIFE_F64)
        elastic_batch_kernel<double, WPB><<<grid, 64 * WPB, 0, s>>>(dj, ds + L.seg0, L.nseg, L.total,
                                                                    elastic_desc->dN, elastic_desc->w);
      else
        elastic_batch_kernel<float, WPB><<<grid, 64 * WPB, 0, s>>>(dj, ds + L.seg0, L.nseg, L.total,
                                                                   elastic_desc->dN, elastic_desc->w);
    }
  }
  prof::end(tk, s, prof::kAssembleBatch);
  MIFE_CHECK_HIP(hipGetLastError());
  return MIFE_OK;
}
