/*
 * C restatement of the Mooney-Rivlin finite-elasticity element loop (MOONEY_RIVLIN
 * subtype, deviatoric branch) and of a restarted GMRES with a diagonal
 * preconditioner, FP64, for the CPU baseline of the Newton path.
 *
 * TEST INFRASTRUCTURE ONLY: timed by bench.py's cpu_baseline leg and checked
 * against the NumPy oracle (oracle/mr_oracle.py) in tests/.  Never linked into,
 * or called by, the product library.
 *
 * Element residual, per Gauss point, as the reference writes it
 * (FiniteElasticity_FiniteElementResidualEvaluate, finite_elasticity_routines.F90:
 * 1669-1760): deformation gradient F = dz/dxi . dxi/dX
 * (FiniteElasticity_GaussDeformationGradientTensor :5148-5212), the isochoric
 * second Piola tensor of FINITE_ELASTICITY_GAUSS_STRESS_TENSOR (:7017-7058) pushed
 * forward (:1343-1375), its deviatoric part plus p, contracted with dphi/dz and
 * weighted by w |dz/dxi|; the pressure rows w |dX/dxi| (J - 1) psi_q.
 * Jacobian: the exact derivative of that residual, written in the reference
 * configuration, K_uu = w Jx GX A GX with A = d2W/dF2 + p d(J F^-T)/dF, and the
 * symmetric pressure coupling w Jx GX (J F^-T) psi (the quantity iron's element
 * finite-difference Jacobian, equations_set_routines.F90:2839-2962, approximates).
 * Element entries are added into the CSR values by iron's bisection
 * (matrix_vector.F90:1724-1854).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static double det3(const double A[3][3]) {
  return A[0][0] * (A[1][1] * A[2][2] - A[1][2] * A[2][1]) - A[0][1] * (A[1][0] * A[2][2] - A[1][2] * A[2][0]) +
         A[0][2] * (A[1][0] * A[2][1] - A[1][1] * A[2][0]);
}

/* B = A^-1 (cofactor form, as InvertFullDP, maths.F90:2807-2868) */
static void inv3m(const double A[3][3], double B[3][3]) {
  const double d = det3(A);
  B[0][0] = (A[1][1] * A[2][2] - A[2][1] * A[1][2]) / d;
  B[1][0] = (A[1][2] * A[2][0] - A[2][2] * A[1][0]) / d;
  B[2][0] = (A[1][0] * A[2][1] - A[2][0] * A[1][1]) / d;
  B[0][1] = (A[2][1] * A[0][2] - A[0][1] * A[2][2]) / d;
  B[1][1] = (A[2][2] * A[0][0] - A[0][2] * A[2][0]) / d;
  B[2][1] = (A[2][0] * A[0][1] - A[0][0] * A[2][1]) / d;
  B[0][2] = (A[0][1] * A[1][2] - A[1][1] * A[0][2]) / d;
  B[1][2] = (A[0][2] * A[1][0] - A[1][2] * A[0][0]) / d;
  B[2][2] = (A[0][0] * A[1][1] - A[1][0] * A[0][1]) / d;
}

/* dx/dxi, sqrt|det GL| and dxi/dx of nodal coordinates xe[nn][3] at Gauss point g. */
static void metrics(int nn, const double* dNg, const double* xe, double dx[3][3], double* jac, double inv[3][3]) {
  memset(dx, 0, sizeof(double) * 9);
  for (int a = 0; a < nn; ++a)
    for (int j = 0; j < 3; ++j) {
      const double d = dNg[a * 3 + j];
      for (int i = 0; i < 3; ++i) dx[i][j] += xe[a * 3 + i] * d;
    }
  double GL[3][3];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) GL[i][j] = dx[0][i] * dx[0][j] + dx[1][i] * dx[1][j] + dx[2][i] * dx[2][j];
  *jac = sqrt(fabs(det3(GL)));
  inv3m(dx, inv);
}

static const int V2T[6][2] = {{0, 0}, {1, 1}, {2, 2}, {0, 1}, {0, 2}, {1, 2}};
static const int T2V[3][3] = {{0, 3, 4}, {3, 1, 5}, {4, 5, 2}};

/*
 * Residual re[3nn+np] and Jacobian Ke[(3nn+np)^2] (row-major; NULL to skip) of one
 * element.  xe, ze [nn][3] reference and current nodal coordinates, pe[np]
 * pressures, psi[ng][np] pressure basis values at the Gauss points (np = 1: element
 * pressure, psi = 1), dN[ng][nn][3], w[ng].
 */
void oracle_mr_element(int nn, int np, int ng, const double* dN, const double* psi, const double* w,
                       const double* xe, const double* ze, const double* pe, double c1, double c2, double* re,
                       double* Ke) {
  const int n = 3 * nn + np;
  memset(re, 0, sizeof(double) * n);
  if (Ke) memset(Ke, 0, sizeof(double) * (size_t)n * n);
  double* GX = (double*)malloc(sizeof(double) * nn * 3);
  double* T = (double*)malloc(sizeof(double) * 3 * 3 * 3 * nn); /* [i][K][j][b] */
  for (int g = 0; g < ng; ++g) {
    const double* dNg = dN + (size_t)g * nn * 3;
    double dxX[3][3], dxiX[3][3], dz[3][3], dxiZ[3][3], jx, jz;
    metrics(nn, dNg, xe, dxX, &jx, dxiX);
    metrics(nn, dNg, ze, dz, &jz, dxiZ);
    double F[3][3];
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 3; ++k) F[i][k] = dz[i][0] * dxiX[0][k] + dz[i][1] * dxiX[1][k] + dz[i][2] * dxiX[2][k];
    const double J = jz / jx;
    double p = 0.0;
    for (int q = 0; q < np; ++q) p += psi[g * np + q] * pe[q];
    /* isochoric stress, pushed forward, deviatoric part + p (Voigt) */
    const double s13 = pow(J, -1.0 / 3.0);
    double Fm[3][3], azl[3][3];
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 3; ++k) Fm[i][k] = F[i][k] * s13;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) azl[i][j] = Fm[0][i] * Fm[0][j] + Fm[1][i] * Fm[1][j] + Fm[2][i] * Fm[2][j];
    const double i1 = azl[0][0] + azl[1][1] + azl[2][2];
    const double t1 = -2.0 * c2, t2 = 2.0 * (c1 + i1 * c2);
    const double S[6] = {t1 * azl[0][0] + t2, t1 * azl[1][1] + t2, t1 * azl[2][2] + t2,
                         t1 * azl[1][0], t1 * azl[2][0], t1 * azl[2][1]};
    double sig[6];
    for (int i = 0; i < 6; ++i) {
      const int ai = V2T[i][0], bi = V2T[i][1];
      double s = 0.0;
      for (int j = 0; j < 6; ++j) {
        const int aj = V2T[j][0], bj = V2T[j][1];
        double t = Fm[ai][aj] * Fm[bi][bj];
        if (j >= 3) t += Fm[ai][bj] * Fm[bi][aj];
        s += t * S[j];
      }
      sig[i] = s / J;
    }
    const double third = (sig[0] + sig[1] + sig[2]) / 3.0;
    for (int i = 0; i < 3; ++i) sig[i] += -third + p;
    /* residual rows: dphi/dz . sigma, weight jz w */
    const double wz = jz * w[g];
    for (int a = 0; a < nn; ++a) {
      double dphidz[3];
      for (int m = 0; m < 3; ++m)
        dphidz[m] = dxiZ[0][m] * dNg[a * 3 + 0] + dxiZ[1][m] * dNg[a * 3 + 1] + dxiZ[2][m] * dNg[a * 3 + 2];
      for (int m = 0; m < 3; ++m) {
        double s = 0.0;
        for (int k = 0; k < 3; ++k) s += dphidz[k] * sig[T2V[k][m]];
        re[m * nn + a] += s * wz;
      }
    }
    const double tt = jx * w[g] * (J - 1.0);
    for (int q = 0; q < np; ++q) re[3 * nn + q] += tt * psi[g * np + q];
    if (!Ke) continue;
    /* tangent: A = c1 d2(I1bar)/dF2 + c2 d2(I2bar)/dF2 + p d(J F^-T)/dF */
    double G[3][3], C[3][3], B[3][3], FC[3][3];
    double Finv[3][3];
    inv3m(F, Finv);
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 3; ++k) G[i][k] = Finv[k][i];
    const double Jd = det3(F);
    double I1 = 0.0, CC = 0.0;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        C[i][j] = F[0][i] * F[0][j] + F[1][i] * F[1][j] + F[2][i] * F[2][j];
        B[i][j] = F[i][0] * F[j][0] + F[i][1] * F[j][1] + F[i][2] * F[j][2];
        I1 += F[i][j] * F[i][j];
      }
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) CC += C[i][j] * C[i][j];
    const double I2 = 0.5 * (I1 * I1 - CC);
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) FC[i][j] = F[i][0] * C[0][j] + F[i][1] * C[1][j] + F[i][2] * C[2][j];
    const double a = pow(Jd, -2.0 / 3.0), b = pow(Jd, -4.0 / 3.0);
    double I1F[3][3], I2F[3][3], X1[3][3], X2[3][3];
    for (int i = 0; i < 3; ++i)
      for (int k = 0; k < 3; ++k) {
        I1F[i][k] = 2.0 * F[i][k];
        I2F[i][k] = 2.0 * (I1 * F[i][k] - FC[i][k]);
        X1[i][k] = I1F[i][k] - (2.0 / 3.0) * I1 * G[i][k];
        X2[i][k] = I2F[i][k] - (4.0 / 3.0) * I2 * G[i][k];
      }
    double A[3][3][3][3];
    for (int i = 0; i < 3; ++i)
      for (int K = 0; K < 3; ++K)
        for (int j = 0; j < 3; ++j)
          for (int L = 0; L < 3; ++L) {
            const double dd = (i == j && K == L) ? 1.0 : 0.0;
            const double gtg = G[i][L] * G[j][K];
            const double ib1 = -(2.0 / 3.0) * a * X1[i][K] * G[j][L] +
                               a * (2.0 * dd - (2.0 / 3.0) * G[i][K] * I1F[j][L] + (2.0 / 3.0) * I1 * gtg);
            const double i2ff = 2.0 * (2.0 * F[i][K] * F[j][L] + I1 * dd - (i == j ? C[L][K] : 0.0) -
                                       F[i][L] * F[j][K] - (K == L ? B[i][j] : 0.0));
            const double ib2 = -(4.0 / 3.0) * b * X2[i][K] * G[j][L] +
                               b * (i2ff - (4.0 / 3.0) * G[i][K] * I2F[j][L] + (4.0 / 3.0) * I2 * gtg);
            A[i][K][j][L] = c1 * ib1 + c2 * ib2 + p * (Jd * G[j][L] * G[i][K] - Jd * G[i][L] * G[j][K]);
          }
    for (int a2 = 0; a2 < nn; ++a2)
      for (int K = 0; K < 3; ++K)
        GX[a2 * 3 + K] = dNg[a2 * 3 + 0] * dxiX[0][K] + dNg[a2 * 3 + 1] * dxiX[1][K] + dNg[a2 * 3 + 2] * dxiX[2][K];
    const double wj = jx * w[g];
    /* T[i][K][j][b] = sum_L A[i][K][j][L] GX[b][L] */
    for (int i = 0; i < 3; ++i)
      for (int K = 0; K < 3; ++K)
        for (int j = 0; j < 3; ++j)
          for (int bb = 0; bb < nn; ++bb)
            T[((i * 3 + K) * 3 + j) * nn + bb] =
                A[i][K][j][0] * GX[bb * 3] + A[i][K][j][1] * GX[bb * 3 + 1] + A[i][K][j][2] * GX[bb * 3 + 2];
    for (int i = 0; i < 3; ++i)
      for (int aa = 0; aa < nn; ++aa) {
        double* row = Ke + (size_t)(i * nn + aa) * n;
        for (int j = 0; j < 3; ++j)
          for (int bb = 0; bb < nn; ++bb) {
            double s = 0.0;
            for (int K = 0; K < 3; ++K) s += GX[aa * 3 + K] * T[((i * 3 + K) * 3 + j) * nn + bb];
            row[j * nn + bb] += wj * s;
          }
        /* pressure coupling: d P_iK / dp = J G_iK, symmetric */
        double c = 0.0;
        for (int K = 0; K < 3; ++K) c += GX[aa * 3 + K] * Jd * G[i][K];
        for (int q = 0; q < np; ++q) {
          const double v = wj * c * psi[g * np + q];
          row[3 * nn + q] += v;
          Ke[(size_t)(3 * nn + q) * n + i * nn + aa] += v;
        }
      }
  }
  free(GX);
  free(T);
}

static int64_t find_pos(const int32_t* rp, const int32_t* col, int row, int target) {
  int64_t lo = rp[row], hi = (int64_t)rp[row + 1] - 1;
  while (hi - lo + 1 > 10) {
    const int64_t mid = (lo + hi) / 2;
    if (col[mid] == target) return mid;
    if (col[mid] < target) lo = mid + 1; else hi = mid - 1;
  }
  for (int64_t k = lo; k <= hi; ++k)
    if (col[k] == target) return k;
  return -1;
}

/*
 * Residual R[ndofs] and Jacobian CSR values (rp, col, vals; NULL vals: residual
 * only) of elements [e0, e1) added in element order.  coords[nu][3] reference
 * nodes, state[ndofs] (x, y, z over the nu displacement nodes, then pressures),
 * enodes [E][nn] displacement nodes, edofs [E][3nn+np], mat [E][2] (c1, c2).
 * Returns -1 if an entry is missing from the pattern.
 */
int oracle_mr_assemble(int E, int nn, int np, int ng, int nu, const double* dN, const double* psi, const double* w,
                       const double* coords, const double* state, const int32_t* enodes, const int32_t* edofs,
                       const double* mat, const int32_t* rp, const int32_t* col, double* vals, double* R, int e0,
                       int e1) {
  const int n = 3 * nn + np;
  double* xe = (double*)malloc(sizeof(double) * nn * 3);
  double* ze = (double*)malloc(sizeof(double) * nn * 3);
  double* pe = (double*)malloc(sizeof(double) * np);
  double* re = (double*)malloc(sizeof(double) * n);
  double* Ke = vals ? (double*)malloc(sizeof(double) * (size_t)n * n) : NULL;
  int bad = 0;
  for (int e = e0; e < e1 && e < E; ++e) {
    const int32_t* en = enodes + (size_t)e * nn;
    const int32_t* ed = edofs + (size_t)e * n;
    for (int a = 0; a < nn; ++a)
      for (int d = 0; d < 3; ++d) {
        xe[a * 3 + d] = coords[(size_t)en[a] * 3 + d];
        ze[a * 3 + d] = state[(size_t)d * nu + en[a]];
      }
    for (int q = 0; q < np; ++q) pe[q] = state[ed[3 * nn + q]];
    oracle_mr_element(nn, np, ng, dN, psi, w, xe, ze, pe, mat[2 * e], mat[2 * e + 1], re, Ke);
    for (int i = 0; i < n; ++i) {
      R[ed[i]] += re[i];
      if (!vals) continue;
      for (int j = 0; j < n; ++j) {
        const int64_t p = find_pos(rp, col, ed[i], ed[j]);
        if (p < 0) { bad = 1; continue; }
        vals[p] += Ke[(size_t)i * n + j];
      }
    }
  }
  free(xe); free(ze); free(pe); free(re); free(Ke);
  return bad ? -1 : 0;
}

static void csr_mv(int n, const int32_t* rp, const int32_t* col, const double* v, const double* x, double* y,
                   int threads) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
#endif
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) s += v[k] * x[col[k]];
    y[i] = s;
  }
  (void)threads;
}

static double vdot(int n, const double* a, const double* b, int threads) {
  double s = 0.0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : s) num_threads(threads) if (threads > 1)
#endif
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  (void)threads;
  return s;
}

/*
 * Right-preconditioned restarted GMRES(restart) with a diagonal preconditioner
 * (PCJACOBI 1/a_ii, or pc_inv[i] where nonzero), modified Gram-Schmidt and Givens
 * rotations, as KSPGMRES runs it (solver_routines.F90:11320-11369).  Stops on the
 * rotated residual <= rtol ||b|| or after maxit iterations; returns the iteration
 * count and the true relative residual in *relres.
 */
int oracle_gmres(int n, const int32_t* rp, const int32_t* col, const double* vals, const double* pc_inv,
                 const double* b, double* x, double rtol, int restart, int maxit, int threads, double* relres) {
  const int m = restart;
  double* dinv = (double*)malloc(sizeof(double) * n);
  double* V = (double*)malloc(sizeof(double) * (size_t)n * (m + 1));
  double* Z = (double*)malloc(sizeof(double) * n);
  double* wv = (double*)malloc(sizeof(double) * n);
  double* H = (double*)calloc((size_t)(m + 1) * m, sizeof(double));
  double* cs = (double*)malloc(sizeof(double) * m);
  double* sn = (double*)malloc(sizeof(double) * m);
  double* gg = (double*)malloc(sizeof(double) * (m + 1));
  double* y = (double*)malloc(sizeof(double) * m);
  for (int i = 0; i < n; ++i) {
    double d = 0.0;
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k)
      if (col[k] == i) d = vals[k];
    dinv[i] = (pc_inv && pc_inv[i] != 0.0) ? pc_inv[i] : (d != 0.0 ? 1.0 / d : 1.0);
  }
  const double bn = sqrt(vdot(n, b, b, threads));
  int it = 0;
  double beta;
  for (;;) {
    csr_mv(n, rp, col, vals, x, wv, threads);
    for (int i = 0; i < n; ++i) V[i] = b[i] - wv[i];
    beta = sqrt(vdot(n, V, V, threads));
    if (bn == 0.0 || beta <= rtol * bn || it >= maxit) break;
    for (int i = 0; i < n; ++i) V[i] /= beta;
    memset(gg, 0, sizeof(double) * (m + 1));
    gg[0] = beta;
    int k = 0;
    for (; k < m && it < maxit; ++k, ++it) {
      double* vk = V + (size_t)k * n;
      for (int i = 0; i < n; ++i) Z[i] = dinv[i] * vk[i];
      double* vn = V + (size_t)(k + 1) * n;
      csr_mv(n, rp, col, vals, Z, vn, threads);
      for (int j = 0; j <= k; ++j) {
        const double* vj = V + (size_t)j * n;
        const double h = vdot(n, vn, vj, threads);
        H[j * m + k] = h;
        for (int i = 0; i < n; ++i) vn[i] -= h * vj[i];
      }
      const double hn = sqrt(vdot(n, vn, vn, threads));
      H[(k + 1) * m + k] = hn;
      if (hn > 0.0)
        for (int i = 0; i < n; ++i) vn[i] /= hn;
      for (int j = 0; j < k; ++j) {
        const double a1 = H[j * m + k], a2 = H[(j + 1) * m + k];
        H[j * m + k] = cs[j] * a1 + sn[j] * a2;
        H[(j + 1) * m + k] = -sn[j] * a1 + cs[j] * a2;
      }
      const double a1 = H[k * m + k], a2 = H[(k + 1) * m + k];
      const double r = sqrt(a1 * a1 + a2 * a2);
      cs[k] = r > 0 ? a1 / r : 1.0;
      sn[k] = r > 0 ? a2 / r : 0.0;
      H[k * m + k] = r;
      H[(k + 1) * m + k] = 0.0;
      gg[k + 1] = -sn[k] * gg[k];
      gg[k] = cs[k] * gg[k];
      if (fabs(gg[k + 1]) <= rtol * bn || hn == 0.0) { ++k; ++it; break; }
    }
    for (int i = k - 1; i >= 0; --i) {
      double s = gg[i];
      for (int j = i + 1; j < k; ++j) s -= H[i * m + j] * y[j];
      y[i] = s / H[i * m + i];
    }
    memset(wv, 0, sizeof(double) * n);
    for (int j = 0; j < k; ++j) {
      const double* vj = V + (size_t)j * n;
      for (int i = 0; i < n; ++i) wv[i] += y[j] * vj[i];
    }
    for (int i = 0; i < n; ++i) x[i] += dinv[i] * wv[i];
  }
  if (relres) *relres = bn > 0 ? beta / bn : 0.0;
  free(dinv); free(V); free(Z); free(wv); free(H); free(cs); free(sn); free(gg); free(y);
  return it;
}
