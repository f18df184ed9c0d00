// Mooney-Rivlin finite elasticity: element residual, exact element Jacobian and
// gravity vector, scattered into a global residual, CSR values and RHS vector.
//
// One workgroup of kMrThreads threads integrates one element in FP64:
//   1. gather the element's reference coordinates and current dependent values
//      (deformed coordinates, pressures) into LDS;
//   2. one thread per Gauss point forms dX/dxi, its inverse, |J| w, F = dz/dX,
//      det F, F^-T and the first Piola stress P;
//   3. all threads build the 81-entry tangent A = dP/dF of every Gauss point;
//   4. per Gauss point the threads form T[a][i][j][L] = sum_K dphi_a/dX_K A_iKjL
//      and accumulate the element Jacobian entries they own in registers;
//   5. residual/RHS entries, then the coloured (or atomic) scatter -- or, staged,
//      plain stores of the element's entries into its slots of a stage buffer,
//      which one launch over all elements fills and a fixed-order reduction
//      (mr_reduce_kernel, by a plan sorted by element) adds into the globals.
//
// The residual is written in the reference configuration,
//   R_(i,a) = sum_g w |J_x| dphi_a/dX_K P_iK,   R_q = sum_g w |J_x| (det F - 1) psi_q,
// which equals the Cauchy-stress form dphi/dz . sigma |J_z| w of
// finite_elasticity_routines.F90:1669-1756 (P = det F sigma F^-T).  The nearly
// incompressible law (:3293-3360, :5308-5320) has no pressure DOFs and no R_q.
#include "mife_common.hpp"
#include "mife_profile.hpp"

#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <climits>

namespace mife {

namespace mr {

__device__ __forceinline__ double det3(const double (&A)[3][3]) {
  return A[0][0] * A[1][1] * A[2][2] + A[0][1] * A[1][2] * A[2][0] + A[0][2] * A[1][0] * A[2][1] -
         A[0][0] * A[2][1] * A[1][2] - A[1][0] * A[0][1] * A[2][2] - A[2][0] * A[1][1] * A[0][2];
}

// B = A^-1 by cofactors; returns det A.
__device__ __forceinline__ double inv3(const double (&A)[3][3], double (&B)[3][3]) {
  const double det = det3(A);
  const double r = 1.0 / det;
  B[0][0] = (A[1][1] * A[2][2] - A[2][1] * A[1][2]) * r;
  B[1][0] = (A[1][2] * A[2][0] - A[2][2] * A[1][0]) * r;
  B[2][0] = (A[1][0] * A[2][1] - A[2][0] * A[1][1]) * r;
  B[0][1] = (A[2][1] * A[0][2] - A[0][1] * A[2][2]) * r;
  B[1][1] = (A[2][2] * A[0][0] - A[0][2] * A[2][0]) * r;
  B[2][1] = (A[2][0] * A[0][1] - A[0][0] * A[2][1]) * r;
  B[0][2] = (A[0][1] * A[1][2] - A[1][1] * A[0][2]) * r;
  B[1][2] = (A[0][2] * A[1][0] - A[1][2] * A[0][0]) * r;
  B[2][2] = (A[0][0] * A[1][1] - A[1][0] * A[0][1]) * r;
  return det;
}

// Per-Gauss-point kinematics and stress shared by the element's threads.
struct GaussState {
  double F[9], G[9], C[9], B[9], X[9];  // X: I2,F (deviatoric) or S (cauchy)
  double Ci[9];                          // C^-1 (cauchy)
  double P[9];                           // first Piola stress
  double dPdp[9];                        // dP/dp
  double JG[9];                          // det F * F^-T = d(det F)/dF
  double s[8];                           // I1, I2, J, a, b, p, c1, c2; nearly incompressible:
                                         // I1, I2, J, a, coef, dV, c1, c2
  double xi[9];                          // dxi_m/dX_K as xi[m*3+K]
  double wj;                             // w |det dX/dxi|
};

typedef double d4 __attribute__((ext_vector_type(4)));
// 16x16x4 FP64 MFMA: lane l gives A[l%16][l/16] and B[l/16][l%16]; accumulator
// register r of lane l holds row 4r + l/16, column l%16.
__device__ __forceinline__ d4 mfma(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

}  // namespace mr

// Threads per element: 16 wavefronts, so that an element's latency-bound stages
// (one colour holds few elements) have waves to switch between.
constexpr int kMrThreads = 1024;

// How an element's entries reach the global arrays.
constexpr int kMrColoured = 0;  // += (the launch's elements share no DOF)
constexpr int kMrAtomic = 1;    // atomicAdd
constexpr int kMrStaged = 2;    // store into the element's stage slot (reduced later)

// Entry v of the element: global position gidx (coloured / atomic) or element
// slot sidx of a stage array (staged); `base` is the global or the stage array.
// The mode is a template parameter: one store path per kernel keeps the MFMA
// tile loop free of the other paths' registers.
template <int MODE>
__device__ __forceinline__ void mr_put(double* base, long long gidx, long long sidx, double v) {
  if constexpr (MODE == kMrStaged) base[sidx] = v;
  else if constexpr (MODE == kMrAtomic) atomicAdd(base + gidx, v);
  else base[gidx] += v;
}

template <int NN, int NP, int NG, int FORM>
struct MrLDS {
  double X[NN][3];
  double Z[NN][3];
  double p[NP > 0 ? NP : 1];
  mr::GaussState gs[NG];
  double A[NG][81];
  double GX[NG][NN][3];
};

// A[i,K,j,L] = dP_iK / dF_jL of one Gauss point; idx = ((i*3+K)*3+j)*3+L.
template <int FORM>
__device__ __forceinline__ double tangent_entry(const mr::GaussState& g, int idx) {
  const int L = idx % 3, j = (idx / 3) % 3, K = (idx / 9) % 3, i = idx / 27;
  const double dij = (i == j) ? 1.0 : 0.0, dKL = (K == L) ? 1.0 : 0.0;
  const double* F = g.F;
  const double* G = g.G;
  const double c2 = g.s[7], p = g.s[5];
  if constexpr (FORM == MIFE_MR_DEVIATORIC) {
    const double I1 = g.s[0], I2 = g.s[1], J = g.s[2], a = g.s[3], b = g.s[4], c1 = g.s[6];
    const double GiK = G[i * 3 + K], GjL = G[j * 3 + L], GiL = G[i * 3 + L], GjK = G[j * 3 + K];
    const double FiK = F[i * 3 + K], FjL = F[j * 3 + L], FiL = F[i * 3 + L], FjK = F[j * 3 + K];
    const double* I2F = g.X;
    const double ib1 = -(2.0 / 3.0) * a * GjL * (2.0 * FiK - (2.0 / 3.0) * I1 * GiK) +
                       a * (2.0 * dij * dKL - (4.0 / 3.0) * GiK * FjL + (2.0 / 3.0) * I1 * GiL * GjK);
    const double i2ff = 2.0 * (2.0 * FiK * FjL + I1 * dij * dKL - dij * g.C[L * 3 + K] - FiL * FjK -
                               g.B[i * 3 + j] * dKL);
    const double ib2 = -(4.0 / 3.0) * b * GjL * (I2F[i * 3 + K] - (4.0 / 3.0) * I2 * GiK) +
                       b * (i2ff - (4.0 / 3.0) * GiK * I2F[j * 3 + L] + (4.0 / 3.0) * I2 * GiL * GjK);
    return c1 * ib1 + c2 * ib2 + p * J * (GjL * GiK - GiL * GjK);
  } else if constexpr (FORM == MIFE_MR_NEARLY_INCOMPRESSIBLE) {
    // P = F S with S = 2a T, T = (c1 + c2 I1) I - c2 C - coef C^-1, a = J^(-2/3),
    // coef = (c1 I1 + 2 c2 I2 - 1.5 kappa (J - 1) J^(5/3)) / 3:
    // dP_iK/dF_jL = dij S_LK - 2/3 G_jL P_iK + 2a [F dT/dF_jL]_iK, using dI1 = 2F,
    // dI2 = 2(I1 F - F C), dJ = J G, dC^-1_MK = -(C^-1_ML G_jK + G_jM C^-1_LK), F C^-1 = G.
    const double I1 = g.s[0], a = g.s[3], coef = g.s[4], dV = g.s[5], c1 = g.s[6];
    const double FjL = F[j * 3 + L];
    const double FC = F[j * 3 + 0] * g.C[0 * 3 + L] + F[j * 3 + 1] * g.C[1 * 3 + L] + F[j * 3 + 2] * g.C[2 * 3 + L];
    const double dcoef = (2.0 * c1 * FjL + 4.0 * c2 * (I1 * FjL - FC) - dV * G[j * 3 + L]) * (1.0 / 3.0);
    const double FdT = 2.0 * c2 * FjL * F[i * 3 + K] - c2 * (F[i * 3 + L] * F[j * 3 + K] + g.B[i * 3 + j] * dKL) -
                       dcoef * G[i * 3 + K] + coef * (G[i * 3 + L] * G[j * 3 + K] + dij * g.Ci[L * 3 + K]);
    return dij * g.X[L * 3 + K] - (2.0 / 3.0) * G[j * 3 + L] * g.P[i * 3 + K] + 2.0 * a * FdT;
  } else {
    const double* S = g.X;
    return dij * S[L * 3 + K] +
           2.0 * c2 * (2.0 * F[i * 3 + K] * F[j * 3 + L] - F[i * 3 + L] * F[j * 3 + K] - g.B[i * 3 + j] * dKL) -
           p * (G[i * 3 + L] * G[j * 3 + K] + dij * g.Ci[L * 3 + K]);
  }
}

// TPB threads per element: kMrThreads, or half of it for two elements per CU
// (each wavefront then takes two MFMA row tiles).
template <int NN, int NP, int NG, int FORM, int MODE, int TPB = kMrThreads>
__global__ __launch_bounds__(TPB, 4) void mr_element_kernel(
    const double* __restrict__ coords, const int32_t* __restrict__ enodes, const int32_t* __restrict__ edofs,
    const double* __restrict__ state, const double* __restrict__ material, const double* __restrict__ dN,
    const double* __restrict__ Nu, const double* __restrict__ Np, const double* __restrict__ w, double bfx,
    double bfy, double bfz, int want_rhs, const int32_t* __restrict__ pos, const int32_t* __restrict__ order,
    int elem_begin, int elem_count, double* __restrict__ residual, double* __restrict__ jac,
    double* __restrict__ rhs) {
  constexpr int NE = 3 * NN + NP;
  constexpr int NENT = NE * NE;
  __shared__ MrLDS<NN, NP, NG, FORM> S;
  const int t = threadIdx.x;
  const int idx = blockIdx.x;
  if (idx >= elem_count) return;  // whole block exits together
  const long long e = order ? order[elem_begin + idx] : (long long)elem_begin + idx;
  const int32_t* ed = edofs + e * NE;

  // 1. gather
  if (t < NN) {
    const long long node = enodes[e * NN + t];
#pragma unroll
    for (int d = 0; d < 3; ++d) {
      S.X[t][d] = coords[node * 3 + d];
      S.Z[t][d] = state[ed[d * NN + t]];
    }
  }
  if (t < NP) S.p[t] = state[ed[3 * NN + t]];
  __syncthreads();

  // 2. per-Gauss-point kinematics and stress
  if (t < NG) {
    const int g = t;
    double Jx[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, Jz[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    for (int a = 0; a < NN; ++a) {
#pragma unroll
      for (int m = 0; m < 3; ++m) {
        const double dn = dN[(g * NN + a) * 3 + m];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
          Jx[d][m] += S.X[a][d] * dn;
          Jz[d][m] += S.Z[a][d] * dn;
        }
      }
    }
    double xi[3][3];
    const double detx = mr::inv3(Jx, xi);  // xi[m][K] = dxi_m/dX_K
    double F[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int K = 0; K < 3; ++K) F[i][K] = Jz[i][0] * xi[0][K] + Jz[i][1] * xi[1][K] + Jz[i][2] * xi[2][K];
    double Fi[3][3];
    const double J = mr::inv3(F, Fi);
    double pg = 0.0;
    if constexpr (NP == 1) {
      pg = S.p[0];
    } else {
#pragma unroll
      for (int q = 0; q < NP; ++q) pg += Np[g * NP + q] * S.p[q];
    }
    constexpr int MS = FORM == MIFE_MR_NEARLY_INCOMPRESSIBLE ? 3 : 2;  // material values per element
    const double c1 = material[e * MS + 0], c2 = material[e * MS + 1];
    mr::GaussState& gs = S.gs[g];
    double C[3][3], B[3][3];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        C[i][j] = F[0][i] * F[0][j] + F[1][i] * F[1][j] + F[2][i] * F[2][j];
        B[i][j] = F[i][0] * F[j][0] + F[i][1] * F[j][1] + F[i][2] * F[j][2];
      }
    const double I1 = C[0][0] + C[1][1] + C[2][2];
    double CC = 0.0;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) CC += C[i][j] * C[i][j];
    const double I2 = 0.5 * (I1 * I1 - CC);
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        gs.F[i * 3 + j] = F[i][j];
        gs.G[i * 3 + j] = Fi[j][i];
        gs.C[i * 3 + j] = C[i][j];
        gs.B[i * 3 + j] = B[i][j];
        gs.JG[i * 3 + j] = J * Fi[j][i];
        gs.xi[i * 3 + j] = xi[i][j];
      }
    if constexpr (FORM == MIFE_MR_DEVIATORIC) {
      const double a = pow(J, -2.0 / 3.0), b = a * a;
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int K = 0; K < 3; ++K) {
          const double FC = F[i][0] * C[0][K] + F[i][1] * C[1][K] + F[i][2] * C[2][K];
          const double i2f = 2.0 * (I1 * F[i][K] - FC);
          const double GiK = Fi[K][i];
          gs.X[i * 3 + K] = i2f;
          gs.P[i * 3 + K] = c1 * a * (2.0 * F[i][K] - (2.0 / 3.0) * I1 * GiK) +
                            c2 * b * (i2f - (4.0 / 3.0) * I2 * GiK) + pg * J * GiK;
          gs.dPdp[i * 3 + K] = J * GiK;
        }
      gs.s[3] = a;
      gs.s[4] = b;
      gs.s[5] = pg;
    } else if constexpr (FORM == MIFE_MR_NEARLY_INCOMPRESSIBLE) {
      // S = 2 J^(-2/3) [(c1 + c2 I1) I - c2 C - (c1 I1 + 2 c2 I2 - 1.5 WV' J^(5/3)) / 3 C^-1],
      // WV' = kappa (J - 1)
      const double kappa = material[e * MS + 2];
      const double a = pow(J, -2.0 / 3.0), J23 = 1.0 / a, J53 = J * J23;
      const double coef = (c1 * I1 + 2.0 * c2 * I2 - 1.5 * kappa * (J - 1.0) * J53) * (1.0 / 3.0);
      double Ci[3][3];
      mr::inv3(C, Ci);
      double Sm[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          Sm[i][j] = 2.0 * a * ((i == j ? c1 + c2 * I1 : 0.0) - c2 * C[i][j] - coef * Ci[i][j]);
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int K = 0; K < 3; ++K) {
          gs.X[i * 3 + K] = Sm[i][K];
          gs.Ci[i * 3 + K] = Ci[i][K];
          gs.P[i * 3 + K] = F[i][0] * Sm[0][K] + F[i][1] * Sm[1][K] + F[i][2] * Sm[2][K];
          gs.dPdp[i * 3 + K] = 0.0;
        }
      gs.s[3] = a;
      gs.s[4] = coef;
      // d(1.5 kappa (J - 1) J^(5/3)) / dF_jL = dV G_jL
      gs.s[5] = 1.5 * kappa * J * (J53 + (5.0 / 3.0) * (J - 1.0) * J23);
    } else {
      double Ci[3][3];
      mr::inv3(C, Ci);
      double Sm[3][3];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          Sm[i][j] = 2.0 * (c1 * (i == j ? 1.0 : 0.0) + c2 * ((i == j ? I1 : 0.0) - C[i][j])) + pg * Ci[i][j];
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int K = 0; K < 3; ++K) {
          gs.X[i * 3 + K] = Sm[i][K];
          gs.Ci[i * 3 + K] = Ci[i][K];
          gs.P[i * 3 + K] = F[i][0] * Sm[0][K] + F[i][1] * Sm[1][K] + F[i][2] * Sm[2][K];
          gs.dPdp[i * 3 + K] = Fi[K][i];
        }
      gs.s[3] = 0.0;
      gs.s[4] = 0.0;
      gs.s[5] = pg;
    }
    gs.s[0] = I1;
    gs.s[1] = I2;
    gs.s[2] = J;
    gs.s[6] = c1;
    gs.s[7] = c2;
    gs.wj = fabs(detx) * w[g];
  }
  __syncthreads();

  // 3. tangents and reference gradients
  const bool want_jac = jac != nullptr;
  if (want_jac) {
    for (int k = t; k < NG * 81; k += TPB) {
      const int g = k / 81;
      S.A[g][k - g * 81] = tangent_entry<FORM>(S.gs[g], k - g * 81);
    }
  }
  for (int k = t; k < NG * NN; k += TPB) {
    const int g = k / NN, a = k - g * NN;
    const double* xi = S.gs[g].xi;
#pragma unroll
    for (int K = 0; K < 3; ++K) {
      double v = 0.0;
#pragma unroll
      for (int m = 0; m < 3; ++m) v += dN[(g * NN + a) * 3 + m] * xi[m * 3 + K];
      S.GX[g][a][K] = v;
    }
  }
  __syncthreads();

  // 4. element Jacobian.
  // K_uu: K[(i,a)][(j,b)] = sum_g sum_L U[(i,j,a)][(g,L)] GX[g][b][L] with
  // U[(i,j,a)][(g,L)] = w_g sum_K GX[g][a][K] A_g[(i,K)][(j,L)]: one GEMM of
  // (9 NN) x (3 NG) by (3 NG) x NN on 16x16x4 FP64 MFMA tiles.  Wavefront w owns
  // row tile w and every column tile; U is formed operand by operand from LDS.
  if (want_jac) {
    constexpr int MROWS = 9 * NN, KD = 3 * NG;
    constexpr int MT = (MROWS + 15) / 16, NT = (NN + 15) / 16, STEPS = (KD + 3) / 4;
    static_assert(NT <= 2, "at most two column tiles");
    const int wv = t >> 6, lane = t & 63;
    const int32_t* pe = MODE == kMrStaged ? nullptr : pos + e * (long long)NENT;
    const long long se = e * (long long)NENT;
    for (int tw = wv; tw < MT; tw += TPB / 64) {
      const int mrow = 16 * tw + (lane & 15);
      const bool rv = mrow < MROWS;
      const int ia = rv ? mrow / (3 * NN) : 0, ja = rv ? (mrow / NN) % 3 : 0, aa = rv ? mrow % NN : 0;
      mr::d4 k0 = {0.0, 0.0, 0.0, 0.0}, k1 = {0.0, 0.0, 0.0, 0.0};
      for (int st = 0; st < STEPS; ++st) {
        const int k = 4 * st + (lane >> 4);
        const bool kv = k < KD;
        const int g = kv ? k / 3 : 0, L = k - 3 * (k / 3);
        double u = 0.0, b0 = 0.0, b1 = 0.0;
        if (kv) {
          const double* A = S.A[g];
          if (rv)
            u = S.gs[g].wj * (S.GX[g][aa][0] * A[(ia * 3 + 0) * 9 + ja * 3 + L] +
                              S.GX[g][aa][1] * A[(ia * 3 + 1) * 9 + ja * 3 + L] +
                              S.GX[g][aa][2] * A[(ia * 3 + 2) * 9 + ja * 3 + L]);
          const int bb = lane & 15;
          if (bb < NN) b0 = S.GX[g][bb][L];
          if (NT > 1 && 16 + bb < NN) b1 = S.GX[g][16 + bb][L];
        }
        k0 = mr::mfma(u, b0, k0);
        if (NT > 1) k1 = mr::mfma(u, b1, k1);
      }
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int orow = 16 * tw + 4 * r + (lane >> 4);  // accumulator row of register r
        if (orow >= MROWS) continue;
        const int i = orow / (3 * NN), j = (orow / NN) % 3, a = orow % NN;
        const int b = lane & 15;
        if (b < NN) {
          const int loc = (i * NN + a) * NE + j * NN + b;
          mr_put<MODE>(jac, MODE == kMrStaged ? 0 : pe[loc], se + loc, k0[r]);
        }
        if (NT > 1 && 16 + b < NN) {
          const int loc = (i * NN + a) * NE + j * NN + 16 + b;
          mr_put<MODE>(jac, MODE == kMrStaged ? 0 : pe[loc], se + loc, k1[r]);
        }
      }
    }
    // K_up = GX . dP/dp psi and K_pu = psi JG . GX (K_pp = 0): sums over the Gauss points
    constexpr int NUP = 3 * NN * NP;
    if constexpr (NP > 0)
    for (int ent = t; ent < 2 * NUP; ent += TPB) {
      const bool up = ent < NUP;
      const int k = up ? ent : ent - NUP;
      const int q = k % NP, r = k / NP, i = r / NN, a = r - (r / NN) * NN;
      double v = 0.0;
      for (int g = 0; g < NG; ++g) {
        const double* m = up ? S.gs[g].dPdp : S.gs[g].JG;
        const double psi = (NP == 1) ? 1.0 : Np[g * NP + q];
        v += S.gs[g].wj * psi *
             (S.GX[g][a][0] * m[i * 3 + 0] + S.GX[g][a][1] * m[i * 3 + 1] + S.GX[g][a][2] * m[i * 3 + 2]);
      }
      const int row = up ? i * NN + a : 3 * NN + q, col = up ? 3 * NN + q : i * NN + a;
      mr_put<MODE>(jac, MODE == kMrStaged ? 0 : pe[row * NE + col], se + row * NE + col, v);
    }
  }

  // 5. residual and gravity vector
  if (residual && t < NE) {
    double r = 0.0;
    if (t < 3 * NN) {
      const int i = t / NN, a = t - i * NN;
      for (int g = 0; g < NG; ++g) {
        const double* P = S.gs[g].P;
        r += S.gs[g].wj *
             (S.GX[g][a][0] * P[i * 3 + 0] + S.GX[g][a][1] * P[i * 3 + 1] + S.GX[g][a][2] * P[i * 3 + 2]);
      }
    } else if constexpr (NP > 0) {
      const int q = t - 3 * NN;
      for (int g = 0; g < NG; ++g) {
        const double psi = (NP == 1) ? 1.0 : Np[g * NP + q];
        r += S.gs[g].wj * (S.gs[g].s[2] - 1.0) * psi;
      }
    }
    mr_put<MODE>(residual, ed[t], e * NE + t, r);
  }
  if (want_rhs && rhs && t < 3 * NN) {
    const int i = t / NN, a = t - i * NN;
    const double bf = i == 0 ? bfx : (i == 1 ? bfy : bfz);
    double r = 0.0;
    for (int g = 0; g < NG; ++g) r += Nu[g * NN + a] * S.gs[g].wj;
    mr_put<MODE>(rhs, ed[t], e * NE + t, r * bf);
  }
}

// Staged mode, second pass: out[k] += sum of stage[list[j]], j in [ptr[k], ptr[k+1])
// (the contributions in element order), one thread per global entry.
__global__ void mr_reduce_kernel(long long n, const int32_t* __restrict__ ptr, const int32_t* __restrict__ list,
                                 const double* __restrict__ stage, double* __restrict__ out) {
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (long long)gridDim.x * blockDim.x) {
    double sum = 0.0;
    for (int j = ptr[k]; j < ptr[k + 1]; ++j) sum += stage[list[j]];
    out[k] += sum;
  }
}

// One element pass: coloured launches (order != NULL), one atomic launch, or,
// with a plan, one staged launch over all elements followed by the fixed-order
// reductions into the global arrays.
template <int NN, int NP, int NG, int FORM>
static int launch_mr(const mife_mr_desc* d, int n_elem, const double* coords, const int32_t* enodes,
                     const int32_t* edofs, const double* state, const double* material, const double* body_force,
                     const int32_t* pos, const int32_t* order, const int32_t* cptr, int ncol,
                     const mife_mr_plan* plan, double* stage, double* residual, double* jac, double* rhs,
                     hipStream_t s) {
  constexpr int NE = 3 * NN + NP;
  const double bx = body_force ? body_force[0] : 0.0, by = body_force ? body_force[1] : 0.0,
               bz = body_force ? body_force[2] : 0.0;
  const int want_rhs = (rhs != nullptr && body_force != nullptr) ? 1 : 0;
  const int tk = prof::begin(s);
  if (plan) {
    double* sj = stage;
    double* sr = sj + (size_t)n_elem * NE * NE;
    double* sf = sr + (size_t)n_elem * NE;
    // with the Jacobian, one element per CU (1024 threads: each wavefront one MFMA
    // row tile, no spills); residual only, two per CU (512 threads)
    if (jac)
      mr_element_kernel<NN, NP, NG, FORM, kMrStaged><<<n_elem, kMrThreads, 0, s>>>(
          coords, enodes, edofs, state, material, d->u->dN, d->u->N, d->Np, d->u->w, bx, by, bz, want_rhs, nullptr,
          nullptr, 0, n_elem, residual ? sr : nullptr, sj, want_rhs ? sf : nullptr);
    else
      mr_element_kernel<NN, NP, NG, FORM, kMrStaged, kMrThreads / 2><<<n_elem, kMrThreads / 2, 0, s>>>(
          coords, enodes, edofs, state, material, d->u->dN, d->u->N, d->Np, d->u->w, bx, by, bz, want_rhs, nullptr,
          nullptr, 0, n_elem, residual ? sr : nullptr, nullptr, want_rhs ? sf : nullptr);
    auto grid_of = [](long long n) { return (int)std::max<long long>(1, std::min<long long>(div_up(n, 256), 8192)); };
    if (jac)
      mr_reduce_kernel<<<grid_of(plan->nnz), 256, 0, s>>>(plan->nnz, plan->jac_ptr, plan->jac_list, sj, jac);
    if (residual)
      mr_reduce_kernel<<<grid_of(plan->n_dofs), 256, 0, s>>>(plan->n_dofs, plan->dof_ptr, plan->dof_list, sr,
                                                            residual);
    if (want_rhs)
      mr_reduce_kernel<<<grid_of(plan->n_dofs), 256, 0, s>>>(plan->n_dofs, plan->dof_ptr, plan->dof_list, sf, rhs);
  } else if (!order) {
    mr_element_kernel<NN, NP, NG, FORM, kMrAtomic><<<n_elem, kMrThreads, 0, s>>>(
        coords, enodes, edofs, state, material, d->u->dN, d->u->N, d->Np, d->u->w, bx, by, bz, want_rhs, pos, nullptr,
        0, n_elem, residual, jac, rhs);
  } else {
    for (int c = 0; c < ncol; ++c) {
      const int cnt = cptr[c + 1] - cptr[c];
      if (cnt <= 0) continue;
      mr_element_kernel<NN, NP, NG, FORM, kMrColoured><<<cnt, kMrThreads, 0, s>>>(
          coords, enodes, edofs, state, material, d->u->dN, d->u->N, d->Np, d->u->w, bx, by, bz, want_rhs, pos, order,
          cptr[c], cnt, residual, jac, rhs);
    }
  }
  // residual + Jacobian launches and residual-only launches (line searches,
  // increment starts) are timed under separate keys
  prof::end(tk, s, jac ? prof::kFiniteElasticity : prof::kFiniteElasticityResidual);
  MIFE_CHECK_HIP(hipGetLastError());
  return MIFE_OK;
}

static int mr_check_desc(const mife_mr_desc* d) {
  MIFE_REQUIRE(d && d->u, MIFE_ERR_POINTER_IS_NULL, "null descriptor");
  MIFE_REQUIRE(d->u->dN && d->u->w, MIFE_ERR_POINTER_IS_NULL, "displacement basis without dN or w");
  MIFE_REQUIRE(d->n_pressure <= 1 || d->Np, MIFE_ERR_POINTER_IS_NULL, "node-based pressure without its basis values");
  MIFE_REQUIRE(d->u->dim == 3, MIFE_ERR_ARG, "finite elasticity kernel is 3-D only");
  MIFE_REQUIRE(d->formulation == MIFE_MR_DEVIATORIC || d->formulation == MIFE_MR_CAUCHY ||
                   d->formulation == MIFE_MR_NEARLY_INCOMPRESSIBLE,
               MIFE_ERR_ARG, "invalid formulation %d", d->formulation);
  MIFE_REQUIRE((d->formulation == MIFE_MR_NEARLY_INCOMPRESSIBLE) == (d->n_pressure == 0), MIFE_ERR_ARG,
               "the nearly incompressible law has no pressure DOFs, the others need them (n_pressure %d)",
               d->n_pressure);
  return MIFE_OK;
}

static int mr_dispatch(const mife_mr_desc* d, int n_elem, const double* coords, const int32_t* elem_nodes,
                       const int32_t* elem_dofs, const double* state, const double* material,
                       const double* body_force, const int32_t* pos, const int32_t* order,
                       const int32_t* colour_ptr_host, int n_colours, const mife_mr_plan* plan, double* stage,
                       double* residual, double* jac_vals, double* rhs, hipStream_t s) {
  const int key = d->u->n_nodes * 10000 + d->n_pressure * 100 + d->u->n_gauss;
#define MR_ARGS(FORM_)                                                                                        \
  (d, n_elem, coords, elem_nodes, elem_dofs, state, material, body_force, jac_vals ? pos : nullptr, order,    \
   colour_ptr_host, n_colours, plan, stage, residual, jac_vals, rhs, s)
#define MR_CASE(NN, NP, NG)                                                   \
  case NN * 10000 + NP * 100 + NG:                                           \
    if (d->formulation == MIFE_MR_DEVIATORIC)                                 \
      return launch_mr<NN, NP, NG, MIFE_MR_DEVIATORIC> MR_ARGS(0);            \
    return launch_mr<NN, NP, NG, MIFE_MR_CAUCHY> MR_ARGS(0);
#define MR_NI_CASE(NN, NG) \
  case NN * 10000 + NG:    \
    return launch_mr<NN, 0, NG, MIFE_MR_NEARLY_INCOMPRESSIBLE> MR_ARGS(0);
  switch (key) {
    MR_NI_CASE(8, 8)
    MR_NI_CASE(27, 27)
    MR_CASE(8, 1, 8)
    MR_CASE(8, 8, 8)
    MR_CASE(27, 1, 27)
    MR_CASE(27, 8, 27)
    default:
      set_error("unsupported finite-elasticity element (nodes=%d, pressure dofs=%d, gauss=%d)", d->u->n_nodes,
                d->n_pressure, d->u->n_gauss);
      return MIFE_ERR_ARG;
  }
#undef MR_CASE
#undef MR_NI_CASE
#undef MR_ARGS
}

// ------------------------------------------------------- staged-scatter plan
// For every global entry (CSR position, or DOF) the list of element slots that
// contribute to it, in increasing slot order (= element order).  Built by a
// counting sort: counts, an exclusive scan, an atomic fill, and an insertion
// sort of each (short) segment, so the result does not depend on the fill order.
__global__ void plan_count_kernel(long long n, const int32_t* __restrict__ keys, long long n_keys,
                                  int32_t* __restrict__ cnt, int* __restrict__ bad) {
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
    const int k = keys[t];
    if (k < 0 || k >= n_keys) atomicOr(bad, 1);
    else atomicAdd(&cnt[k], 1);
  }
}

__global__ void plan_fill_kernel(long long n, const int32_t* __restrict__ keys, long long n_keys,
                                 const int32_t* __restrict__ ptr, int32_t* __restrict__ cursor,
                                 int32_t* __restrict__ list) {
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x) {
    const int k = keys[t];
    if (k < 0 || k >= n_keys) continue;
    const int slot = atomicAdd(&cursor[k], 1);
    list[ptr[k] + slot] = (int32_t)t;
  }
}

__global__ void plan_sort_kernel(long long n_keys, const int32_t* __restrict__ ptr, int32_t* __restrict__ list) {
  for (long long k = (long long)blockIdx.x * blockDim.x + threadIdx.x; k < n_keys;
       k += (long long)gridDim.x * blockDim.x) {
    const int b = ptr[k], e = ptr[k + 1];
    for (int i = b + 1; i < e; ++i) {
      const int v = list[i];
      int j = i - 1;
      while (j >= b && list[j] > v) {
        list[j + 1] = list[j];
        --j;
      }
      list[j + 1] = v;
    }
  }
}

__global__ void plan_zero_kernel(long long n, int32_t* __restrict__ a) {
  for (long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (long long)gridDim.x * blockDim.x)
    a[t] = 0;
}

static size_t plan_scan_bytes(long long n_keys) {
  size_t tb = 0;
  hipcub::DeviceScan::ExclusiveSum(nullptr, tb, (int32_t*)nullptr, (int32_t*)nullptr, (int)(n_keys + 1));
  return tb;
}

static size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// One segment list: keys[0..n) in [0, n_keys) -> ptr[n_keys + 1], list[n].
static int plan_one(long long n, const int32_t* keys, long long n_keys, int32_t* ptr, int32_t* list, char* work,
                    int* bad, hipStream_t s) {
  int32_t* cnt = reinterpret_cast<int32_t*>(work);
  int32_t* cursor = reinterpret_cast<int32_t*>(work + align256(sizeof(int32_t) * (size_t)(n_keys + 1)));
  void* temp = work + 2 * align256(sizeof(int32_t) * (size_t)(n_keys + 1));
  size_t tb = plan_scan_bytes(n_keys);
  auto grid_of = [](long long m) { return (int)std::max<long long>(1, std::min<long long>(div_up(m, 256), 8192)); };
  plan_zero_kernel<<<grid_of(n_keys + 1), 256, 0, s>>>(n_keys + 1, cnt);
  plan_zero_kernel<<<grid_of(n_keys + 1), 256, 0, s>>>(n_keys + 1, cursor);
  plan_count_kernel<<<grid_of(n), 256, 0, s>>>(n, keys, n_keys, cnt, bad);
  MIFE_CHECK_HIP(hipcub::DeviceScan::ExclusiveSum(temp, tb, cnt, ptr, (int)(n_keys + 1), s));
  plan_fill_kernel<<<grid_of(n), 256, 0, s>>>(n, keys, n_keys, ptr, cursor, list);
  plan_sort_kernel<<<grid_of(n_keys), 256, 0, s>>>(n_keys, ptr, list);
  MIFE_CHECK_HIP(hipGetLastError());
  return MIFE_OK;
}

}  // namespace mife

using namespace mife;

extern "C" size_t mife_mr_plan_work_bytes(int n_elem, int ne, int64_t nnz, int n_dofs) {
  const long long k = std::max<long long>(nnz, n_dofs);
  (void)n_elem;
  (void)ne;
  return 2 * align256(sizeof(int32_t) * (size_t)(k + 1)) + align256(std::max(plan_scan_bytes(nnz), plan_scan_bytes(n_dofs))) +
         256;
}

extern "C" int mife_mr_plan_build(int n_elem, int ne, const int32_t* pos, int64_t nnz, const int32_t* elem_dofs,
                                  int n_dofs, int32_t* jac_ptr, int32_t* jac_list, int32_t* dof_ptr,
                                  int32_t* dof_list, void* work, size_t work_bytes, void* stream) {
  MIFE_REQUIRE(pos && elem_dofs && jac_ptr && jac_list && dof_ptr && dof_list && work, MIFE_ERR_POINTER_IS_NULL,
               "null pointer argument");
  MIFE_REQUIRE(n_elem >= 0 && ne > 0 && nnz >= 0 && n_dofs >= 0, MIFE_ERR_ARG, "invalid sizes");
  const long long n_ent = (long long)n_elem * ne * ne;
  MIFE_REQUIRE(n_ent <= INT32_MAX && nnz < INT32_MAX, MIFE_ERR_ARG, "element entries exceed the int32 range");
  MIFE_REQUIRE(work_bytes >= mife_mr_plan_work_bytes(n_elem, ne, nnz, n_dofs), MIFE_ERR_ARG,
               "plan workspace of %zu bytes is smaller than the %zu needed", work_bytes,
               mife_mr_plan_work_bytes(n_elem, ne, nnz, n_dofs));
  hipStream_t s = (hipStream_t)stream;
  char* w = reinterpret_cast<char*>(work);
  int* bad = reinterpret_cast<int*>(w + work_bytes - 256);
  plan_zero_kernel<<<1, 64, 0, s>>>(1, bad);
  int rc = plan_one(n_ent, pos, nnz, jac_ptr, jac_list, w, bad, s);
  if (rc) return rc;
  rc = plan_one((long long)n_elem * ne, elem_dofs, n_dofs, dof_ptr, dof_list, w, bad, s);
  if (rc) return rc;
  int hbad = 0;
  MIFE_CHECK_HIP(hipMemcpyAsync(&hbad, bad, sizeof(int), hipMemcpyDeviceToHost, s));
  MIFE_CHECK_HIP(hipStreamSynchronize(s));
  MIFE_REQUIRE(!hbad, MIFE_ERR_ARG, "a position or element DOF lies outside the CSR entries or DOFs");
  return MIFE_OK;
}

extern "C" size_t mife_mr_stage_bytes(const mife_mr_desc* d, int n_elem) {
  if (!d || !d->u || n_elem <= 0) return 0;
  const size_t ne = 3 * (size_t)d->u->n_nodes + (size_t)d->n_pressure;
  return sizeof(double) * (size_t)n_elem * (ne * ne + 2 * ne);
}

extern "C" int mife_mr_assemble(const mife_mr_desc* d, int n_elem, const double* coords, const int32_t* elem_nodes,
                                const int32_t* elem_dofs, const double* state, const double* material,
                                const double* body_force, const int32_t* pos, const int32_t* order,
                                const int32_t* colour_ptr_host, int n_colours, double* residual, double* jac_vals,
                                double* rhs, void* stream) {
  int rc = mr_check_desc(d);
  if (rc) return rc;
  MIFE_REQUIRE(coords && elem_nodes && elem_dofs && state && material, MIFE_ERR_POINTER_IS_NULL,
               "null pointer argument");
  MIFE_REQUIRE(!jac_vals || pos, MIFE_ERR_POINTER_IS_NULL, "Jacobian requested without a position map");
  MIFE_REQUIRE(!rhs || !body_force || d->u->N, MIFE_ERR_POINTER_IS_NULL, "gravity requested without basis values N");
  MIFE_REQUIRE(!order || colour_ptr_host, MIFE_ERR_POINTER_IS_NULL, "colour order without colour_ptr");
  if (n_elem == 0 || (!residual && !jac_vals && !rhs)) return MIFE_OK;
  return mr_dispatch(d, n_elem, coords, elem_nodes, elem_dofs, state, material, body_force, pos, order,
                     colour_ptr_host, n_colours, nullptr, nullptr, residual, jac_vals, rhs, (hipStream_t)stream);
}

extern "C" int mife_mr_assemble_staged(const mife_mr_desc* d, int n_elem, const double* coords,
                                       const int32_t* elem_nodes, const int32_t* elem_dofs, const double* state,
                                       const double* material, const double* body_force, const mife_mr_plan* plan,
                                       double* stage, double* residual, double* jac_vals, double* rhs,
                                       void* stream) {
  int rc = mr_check_desc(d);
  if (rc) return rc;
  MIFE_REQUIRE(coords && elem_nodes && elem_dofs && state && material && plan && stage, MIFE_ERR_POINTER_IS_NULL,
               "null pointer argument");
  MIFE_REQUIRE(!jac_vals || (plan->jac_ptr && plan->jac_list), MIFE_ERR_POINTER_IS_NULL,
               "Jacobian requested without the plan's entry lists");
  MIFE_REQUIRE(!(residual || (rhs && body_force)) || (plan->dof_ptr && plan->dof_list), MIFE_ERR_POINTER_IS_NULL,
               "residual or gravity requested without the plan's DOF lists");
  MIFE_REQUIRE(!rhs || !body_force || d->u->N, MIFE_ERR_POINTER_IS_NULL, "gravity requested without basis values N");
  if (n_elem == 0 || (!residual && !jac_vals && !rhs)) return MIFE_OK;
  return mr_dispatch(d, n_elem, coords, elem_nodes, elem_dofs, state, material, body_force, nullptr, nullptr, nullptr,
                     0, plan, stage, residual, jac_vals, rhs, (hipStream_t)stream);
}
