/*
 * C restatement of OpenCMISS-Iron's CPU path for the scalar (Laplace /
 * generalised Laplace / constant-source Poisson) and linear-elasticity
 * assembly-and-solve stage, FP64, one core per call unless OpenMP threads are
 * requested.
 *
 * TEST INFRASTRUCTURE ONLY: the timed CPU baseline in bench.py and a checker in
 * tests/.  It is never linked into, or called by, the product library.
 *
 * Follows, per element and per Gauss point, the loops of
 *   LaplaceEquation_FiniteElementCalculate   Laplace_equations_routines.F90:482-579
 *   (materials scaling of the generalised subtype :581-747, isotropic kappa*I)
 *   POISSON constant source vector            Poisson_equations_routines.F90:3407-3415
 *   COORDINATE_METRICS_CALCULATE (GL, GU, |J|) coordinate_routines.F90:639-826
 *   InvertFullDP                               maths.F90:2807-2868
 * and adds every element entry into the CSR values the way
 *   MATRIX_VALUES_ADD_DP2 -> MATRIX_STORAGE_LOCATION_FIND does
 *   (bisection down to 10 entries, then a linear scan; matrix_vector.F90:184,1724-1854,3119).
 * The solve is PETSc-style CG with Jacobi preconditioning on the reduced
 * (free x free) CSR, stopped on the true relative residual.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static double inv3(const double A[3][3], double B[3][3]) {
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

/* MATRIX_STORAGE_LOCATION_FIND, CSR: bisection to <= 10 entries then linear. */
static int64_t location_find(const int32_t* rp, const int32_t* col, int row, int target) {
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

/* Element matrix and source vector of element e (iron's metric path: GL = J^T J,
 * GU = GL^-1, sqrt|det GL|, and the pgmsi . GU . pgnsi contraction of
 * Laplace_equations_routines.F90:482-579).  Ke[nn][nn], se[nn], xe[nn][3] scratch. */
static void element_scalar3(int e, int nn, int ng, const double* dN, const double* N, const double* w,
                            const double* coords, const int32_t* enodes, const double* source, double* Ke,
                            double* se, double* xe) {
  memset(Ke, 0, sizeof(double) * nn * nn);
  memset(se, 0, sizeof(double) * nn);
  for (int a = 0; a < nn; ++a)
    for (int d = 0; d < 3; ++d) xe[a * 3 + d] = coords[(int64_t)enodes[(int64_t)e * nn + a] * 3 + d];
  for (int g = 0; g < ng; ++g) {
    double J[3][3] = {{0}};
    for (int a = 0; a < nn; ++a)
      for (int j = 0; j < 3; ++j) {
        const double dn = dN[((int64_t)g * nn + a) * 3 + j];
        for (int d = 0; d < 3; ++d) J[d][j] += xe[a * 3 + d] * dn;
      }
    double GL[3][3], GU[3][3];
    for (int mi = 0; mi < 3; ++mi)
      for (int ni = 0; ni < 3; ++ni) GL[mi][ni] = J[0][mi] * J[0][ni] + J[1][mi] * J[1][ni] + J[2][mi] * J[2][ni];
    const double detgl = inv3(GL, GU);
    const double rwg = sqrt(fabs(detgl)) * w[g];
    for (int ms = 0; ms < nn; ++ms) {
      const double* pm = &dN[((int64_t)g * nn + ms) * 3];
      for (int ns = 0; ns < nn; ++ns) {
        const double* pn = &dN[((int64_t)g * nn + ns) * 3];
        double sum = 0.0;
        for (int mi = 0; mi < 3; ++mi)
          for (int ni = 0; ni < 3; ++ni) sum += pm[mi] * pn[ni] * GU[mi][ni];
        Ke[ms * nn + ns] += sum * rwg;
      }
      if (source) se[ms] += source[e] * N[(int64_t)g * nn + ms] * rwg;
    }
  }
}

/* Scalar 3-D assembly.  dN[ng][nn][3], N[ng][nn], w[ng]; coords[node][3];
 * enodes/edofs [E][nn]; kappa/source per element (NULL = 1 / none).
 * vals (CSR values) and src (vector) are accumulated into, each entry found by
 * iron's bisection (matrix_vector.F90:1765-1786).  Returns -1 if an entry is
 * missing from the pattern. */
int oracle_assemble_scalar3(int E, int nn, int ng, const double* dN, const double* N, const double* w,
                            const double* coords, const int32_t* enodes, const int32_t* edofs,
                            const double* kappa, const double* source, const int32_t* rp, const int32_t* col,
                            double* vals, double* src, int e_begin, int e_end) {
  double* Ke = (double*)malloc(sizeof(double) * nn * nn);
  double* se = (double*)malloc(sizeof(double) * nn);
  double* xe = (double*)malloc(sizeof(double) * nn * 3);
  int bad = 0;
  for (int e = e_begin; e < e_end && e < E; ++e) {
    element_scalar3(e, nn, ng, dN, N, w, coords, enodes, source, Ke, se, xe);
    const double ke = kappa ? kappa[e] : 1.0;
    for (int ms = 0; ms < nn; ++ms) {
      const int row = edofs[(int64_t)e * nn + ms];
      for (int ns = 0; ns < nn; ++ns) {
        const int64_t p = location_find(rp, col, row, edofs[(int64_t)e * nn + ns]);
        if (p < 0) { bad = 1; continue; }
        vals[p] += ke * Ke[ms * nn + ns];
      }
      if (source && src) src[row] += se[ms];
    }
  }
  free(Ke);
  free(se);
  free(xe);
  return bad ? -1 : 0;
}

/* As oracle_assemble_scalar3 on `threads` OpenMP threads (elements shared out
 * dynamically, CSR and vector adds atomic): the all-cores CPU proxy for iron on
 * several ranks.  The sums equal the serial ones up to the order of additions. */
int oracle_assemble_scalar3_par(int E, int nn, int ng, const double* dN, const double* N, const double* w,
                                const double* coords, const int32_t* enodes, const int32_t* edofs,
                                const double* kappa, const double* source, const int32_t* rp, const int32_t* col,
                                double* vals, double* src, int e_begin, int e_end, int threads) {
  const int e_stop = e_end < E ? e_end : E;
  int bad = 0;
#ifdef _OPENMP
#pragma omp parallel num_threads(threads) reduction(| : bad)
#endif
  {
    double* Ke = (double*)malloc(sizeof(double) * nn * nn);
    double* se = (double*)malloc(sizeof(double) * nn);
    double* xe = (double*)malloc(sizeof(double) * nn * 3);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
    for (int e = e_begin; e < e_stop; ++e) {
      element_scalar3(e, nn, ng, dN, N, w, coords, enodes, source, Ke, se, xe);
      const double ke = kappa ? kappa[e] : 1.0;
      for (int ms = 0; ms < nn; ++ms) {
        const int row = edofs[(int64_t)e * nn + ms];
        for (int ns = 0; ns < nn; ++ns) {
          const int64_t p = location_find(rp, col, row, edofs[(int64_t)e * nn + ns]);
          if (p < 0) { bad = 1; continue; }
          const double v = ke * Ke[ms * nn + ns];
#ifdef _OPENMP
#pragma omp atomic
#endif
          vals[p] += v;
        }
        if (source && src) {
#ifdef _OPENMP
#pragma omp atomic
#endif
          src[row] += se[ms];
        }
      }
    }
    free(Ke);
    free(se);
    free(xe);
  }
  (void)threads;
  return bad ? -1 : 0;
}

/* K_ff extraction + RHS lifting over free rows (SOLVER_MATRICES_STATIC_ASSEMBLE). */
int64_t oracle_reduce(int n, const int32_t* rp, const int32_t* col, const double* vals, const int32_t* new_idx,
                      const double* u, const double* src, int32_t* srp, int32_t* scol, double* svals, double* b) {
  const double ZT = 5.0 * 2.220446049250313e-16;
  int64_t k2 = 0;
  int f = 0;
  srp[0] = 0;
  for (int i = 0; i < n; ++i) {
    if (new_idx[i] < 0) continue;
    double bi = src ? src[i] : 0.0;
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) {
      const int j = col[k];
      if (new_idx[j] >= 0) {
        scol[k2] = new_idx[j];
        svals[k2] = vals[k];
        ++k2;
      } else if (fabs(u[j]) >= ZT && fabs(vals[k]) >= ZT) {
        bi += -1.0 * vals[k] * u[j];
      }
    }
    b[f] = bi;
    srp[++f] = (int32_t)k2;
  }
  return k2;
}

static double dot(int n, const double* a, const double* b, int threads) {
  double s = 0.0;
#ifdef _OPENMP
#pragma omp parallel for reduction(+ : s) num_threads(threads) if (threads > 1)
#endif
  for (int i = 0; i < n; ++i) s += a[i] * b[i];
  (void)threads;
  return s;
}

static void matvec(int n, const int32_t* rp, const int32_t* col, const double* v, const double* x, double* y,
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

/* Jacobi-preconditioned CG, FP64.  Runs at most maxit iterations; returns the
 * iteration count, and the true relative residual in *relres. */
int oracle_cg(int n, const int32_t* rp, const int32_t* col, const double* vals, const double* b, double* x,
              double rtol, int maxit, int threads, double* relres) {
  double* r = (double*)malloc(sizeof(double) * n);
  double* z = (double*)malloc(sizeof(double) * n);
  double* p = (double*)malloc(sizeof(double) * n);
  double* q = (double*)malloc(sizeof(double) * n);
  double* dinv = (double*)malloc(sizeof(double) * n);
  for (int i = 0; i < n; ++i) {
    double d = 0.0;
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k)
      if (col[k] == i) d = vals[k];
    dinv[i] = d != 0.0 ? 1.0 / d : 1.0;
  }
  matvec(n, rp, col, vals, x, q, threads);
  for (int i = 0; i < n; ++i) {
    r[i] = b[i] - q[i];
    z[i] = dinv[i] * r[i];
    p[i] = z[i];
  }
  const double bn = sqrt(dot(n, b, b, threads));
  double rz = dot(n, r, z, threads);
  int it = 0;
  while (it < maxit && sqrt(dot(n, r, r, threads)) > rtol * bn) {
    matvec(n, rp, col, vals, p, q, threads);
    const double alpha = rz / dot(n, p, q, threads);
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) if (threads > 1)
#endif
    for (int i = 0; i < n; ++i) {
      x[i] += alpha * p[i];
      r[i] -= alpha * q[i];
      z[i] = dinv[i] * r[i];
    }
    const double rz2 = dot(n, r, z, threads);
    const double beta = rz2 / rz;
    rz = rz2;
#ifdef _OPENMP
#pragma omp parallel for num_threads(threads) if (threads > 1)
#endif
    for (int i = 0; i < n; ++i) p[i] = z[i] + beta * p[i];
    ++it;
  }
  matvec(n, rp, col, vals, x, q, threads);
  for (int i = 0; i < n; ++i) r[i] = b[i] - q[i];
  if (relres) *relres = bn > 0 ? sqrt(dot(n, r, r, threads)) / bn : 0.0;
  free(r); free(z); free(p); free(q); free(dinv);
  return it;
}

/* Element matrix of 3-D linear elasticity, element e (LINEAR_ELASTICITY_FINITE_ELEMENT_CALCULATE,
 * linear_elasticity_routines.F90:1109-1403): DPHIDX(ns,mi) = sum_ni DXI_DX(mi,ni) dphi_ns/dxi_ni
 * (:1245-1252, as written), JRWG times the tensor of LINEAR_ELASTICITY_TENSOR (:1410-1451) on the
 * diagonal blocks [C11,C66,C44 | C66,C22,C55 | C44,C55,C33] and the off-diagonal pairs
 * (1,2): [C12,C66], (1,3): [C13,C44], (2,3): [C23,C55], upper triangle mirrored (:1373-1380).
 * mat[e] = (E1, E2, E3, v13, v23, v12); Ke[3nn][3nn] component-major. */
static void element_elastic3(int e, int nn, int ng, const double* dN, const double* w, const double* coords,
                             const int32_t* enodes, const double* mat, double* Ke, double* xe, double* dphidx) {
  const int n = 3 * nn;
  memset(Ke, 0, sizeof(double) * n * n);
  for (int a = 0; a < nn; ++a)
    for (int d = 0; d < 3; ++d) xe[a * 3 + d] = coords[(int64_t)enodes[(int64_t)e * nn + a] * 3 + d];
  const double* m = mat + (int64_t)e * 6;
  const double E1 = m[0], E2 = m[1], E3 = m[2], v13 = m[3], v23 = m[4], v12 = m[5];
  const double v31 = v13, v32 = v23, v21 = v12;
  const double gama = 1.0 / (1.0 - v12 * v21 - v23 * v32 - v31 * v13 - 2.0 * v21 * v32 * v13);
  const double C11 = E1 * (1.0 - v23 * v32) * gama, C22 = E2 * (1.0 - v13 * v31) * gama;
  const double C33 = E3 * (1.0 - v12 * v21) * gama, C12 = E1 * (v21 + v31 * v23) * gama;
  const double C13 = E1 * (v31 + v21 * v32) * gama, C23 = E2 * (v32 + v12 * v31) * gama;
  const double C44 = E2 / (2.0 * (1.0 + v23)), C55 = E1 / (2.0 * (1.0 + v13)), C66 = E3 / (2.0 * (1.0 + v12));
  for (int g = 0; g < ng; ++g) {
    double J[3][3] = {{0}};
    for (int a = 0; a < nn; ++a)
      for (int j = 0; j < 3; ++j) {
        const double dn = dN[((int64_t)g * nn + a) * 3 + j];
        for (int d = 0; d < 3; ++d) J[d][j] += xe[a * 3 + d] * dn;
      }
    double GL[3][3], GU[3][3], JI[3][3];
    for (int mi = 0; mi < 3; ++mi)
      for (int ni = 0; ni < 3; ++ni) GL[mi][ni] = J[0][mi] * J[0][ni] + J[1][mi] * J[1][ni] + J[2][mi] * J[2][ni];
    const double jrwg = sqrt(fabs(inv3(GL, GU))) * w[g];
    inv3(J, JI); /* DXI_DX(i,j) = dxi_i/dx_j */
    for (int s = 0; s < nn; ++s)
      for (int mi = 0; mi < 3; ++mi) {
        double v = 0.0;
        for (int ni = 0; ni < 3; ++ni) v += JI[mi][ni] * dN[((int64_t)g * nn + s) * 3 + ni];
        dphidx[s * 3 + mi] = v;
      }
    const double dg[3][3] = {{C11, C66, C44}, {C66, C22, C55}, {C44, C55, C33}};
    for (int xi = 0; xi < 3; ++xi)
      for (int s = 0; s < nn; ++s)
        for (int t = 0; t < nn; ++t) {
          double v = 0.0;
          for (int d = 0; d < 3; ++d) v += dphidx[s * 3 + d] * dphidx[t * 3 + d] * dg[xi][d];
          Ke[(xi * nn + s) * n + xi * nn + t] += v * jrwg;
        }
    const int pr[3][2] = {{0, 1}, {0, 2}, {1, 2}};
    const double ca[3] = {C12, C13, C23}, cb[3] = {C66, C44, C55};
    for (int o = 0; o < 3; ++o) {
      const int r = pr[o][0], c = pr[o][1];
      for (int s = 0; s < nn; ++s)
        for (int t = 0; t < nn; ++t) {
          const double v = (dphidx[s * 3 + r] * dphidx[t * 3 + c] * ca[o] +
                            dphidx[s * 3 + c] * dphidx[t * 3 + r] * cb[o]) * jrwg;
          Ke[(r * nn + s) * n + c * nn + t] += v;
          Ke[(c * nn + t) * n + r * nn + s] += v;
        }
    }
  }
}

/* Linear-elasticity assembly (3 components on one basis): element loop in element order,
 * each entry located by bisection.  edofs [E][3nn]. */
int oracle_assemble_elastic3(int E, int nn, int ng, const double* dN, const double* w, const double* coords,
                             const int32_t* enodes, const int32_t* edofs, const double* mat, const int32_t* rp,
                             const int32_t* col, double* vals, int e_begin, int e_end) {
  const int n = 3 * nn;
  double* Ke = (double*)malloc(sizeof(double) * n * n);
  double* xe = (double*)malloc(sizeof(double) * nn * 3);
  double* dphidx = (double*)malloc(sizeof(double) * nn * 3);
  int bad = 0;
  for (int e = e_begin; e < e_end && e < E; ++e) {
    element_elastic3(e, nn, ng, dN, w, coords, enodes, mat, Ke, xe, dphidx);
    for (int ms = 0; ms < n; ++ms) {
      const int row = edofs[(int64_t)e * n + ms];
      for (int ns = 0; ns < n; ++ns) {
        const int64_t p = location_find(rp, col, row, edofs[(int64_t)e * n + ns]);
        if (p < 0) { bad = 1; continue; }
        vals[p] += Ke[ms * n + ns];
      }
    }
  }
  free(Ke); free(xe); free(dphidx);
  return bad ? -1 : 0;
}

/* As oracle_assemble_elastic3 on `threads` OpenMP threads (atomic adds). */
int oracle_assemble_elastic3_par(int E, int nn, int ng, const double* dN, const double* w, const double* coords,
                                 const int32_t* enodes, const int32_t* edofs, const double* mat, const int32_t* rp,
                                 const int32_t* col, double* vals, int e_begin, int e_end, int threads) {
  const int n = 3 * nn;
  const int e_stop = e_end < E ? e_end : E;
  int bad = 0;
#ifdef _OPENMP
#pragma omp parallel num_threads(threads) reduction(| : bad)
#endif
  {
    double* Ke = (double*)malloc(sizeof(double) * n * n);
    double* xe = (double*)malloc(sizeof(double) * nn * 3);
    double* dphidx = (double*)malloc(sizeof(double) * nn * 3);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 64)
#endif
    for (int e = e_begin; e < e_stop; ++e) {
      element_elastic3(e, nn, ng, dN, w, coords, enodes, mat, Ke, xe, dphidx);
      for (int ms = 0; ms < n; ++ms) {
        const int row = edofs[(int64_t)e * n + ms];
        for (int ns = 0; ns < n; ++ns) {
          const int64_t p = location_find(rp, col, row, edofs[(int64_t)e * n + ns]);
          if (p < 0) { bad = 1; continue; }
#ifdef _OPENMP
#pragma omp atomic
#endif
          vals[p] += Ke[ms * n + ns];
        }
      }
    }
    free(Ke); free(xe); free(dphidx);
  }
  (void)threads;
  return bad ? -1 : 0;
}

/* EQUATIONS_SET_BACKSUBSTITUTE, CSR branch (equations_set_routines.F90:1640-1912), the
 * correct form: R_i = sum_j K_ij u_j - src_i over every equations row. */
void oracle_backsubstitute(int n, const int32_t* rp, const int32_t* col, const double* vals, const double* u,
                           const double* src, double* R, int threads) {
#ifdef _OPENMP
#pragma omp parallel for schedule(static) num_threads(threads) if (threads > 1)
#endif
  for (int i = 0; i < n; ++i) {
    double s = 0.0;
    for (int64_t k = rp[i]; k < rp[i + 1]; ++k) s += vals[k] * u[col[k]];
    R[i] = s - (src ? src[i] : 0.0);
  }
  (void)threads;
}
