/*
 * or_linalg.c -- dense FP64 linear algebra for the oracle (test
 * infrastructure only).  Replaces the MKL 2020u4 BLAS/LAPACK calls the
 * reference makes on this path (blas.cpp:956-1067, 1303-1358;
 * lapack.cpp Potrf/Syevd) with plain loops:
 *   - symmetric eigensolver: Householder tridiagonalisation followed by the
 *     implicit-shift QL iteration of Bowdler, Martin, Reinsch & Wilkinson
 *     (Numer. Math. 11 (1968) 293, "tql2"), eigenvalues sorted ascending;
 *   - Cholesky A = U^T U and X <- X U^{-1};
 *   - EigenSolver::SolveGenEig, sequential branch (eigensolver.cpp:224-297);
 *   - EigenSolver::Orthogonalize, complex-X branch (eigensolver.cpp:386-469).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

void or_gemm_tn(int m, int n1, int n2, const double* A, const double* B, double* C) {
#pragma omp parallel for collapse(2) schedule(static)
  for (int j = 0; j < n2; j++)
    for (int i = 0; i < n1; i++) {
      const double* a = A + (size_t)m * i;
      const double* b = B + (size_t)m * j;
      double s = 0.0;
      for (int r = 0; r < m; r++) s += a[r] * b[r];
      C[i + (size_t)n1 * j] = s;
    }
}

void or_gemm_nn(int m, int n, int k, double alpha, const double* A, int lda, const double* B,
                int ldb, double beta, double* C, int ldc) {
#pragma omp parallel for schedule(static)
  for (int j = 0; j < n; j++) {
    double* c = C + (size_t)ldc * j;
    if (beta == 0.0)
      for (int i = 0; i < m; i++) c[i] = 0.0;
    else if (beta != 1.0)
      for (int i = 0; i < m; i++) c[i] *= beta;
    for (int p = 0; p < k; p++) {
      const double bpj = alpha * B[p + (size_t)ldb * j];
      if (bpj == 0.0) continue;
      const double* a = A + (size_t)lda * p;
      for (int i = 0; i < m; i++) c[i] += a[i] * bpj;
    }
  }
}

int or_potrf_upper(int n, double* A, int lda) {
#define U(i, j) A[(i) + (size_t)lda * (j)]
  for (int j = 0; j < n; j++) {
    double s = U(j, j);
    for (int k = 0; k < j; k++) s -= U(k, j) * U(k, j);
    if (!(s > 0.0)) return j + 1;
    double d = sqrt(s);
    U(j, j) = d;
    for (int i = j + 1; i < n; i++) {
      double t = U(j, i);
      for (int k = 0; k < j; k++) t -= U(k, j) * U(k, i);
      U(j, i) = t / d;
    }
  }
  return 0;
#undef U
}

void or_trsm_right_upper(int m, int n, const double* Um, int ldu, double* X, int ldx) {
  /* solve Y U = X column by column: y_j = (x_j - sum_{k<j} y_k U(k,j)) / U(j,j) */
  for (int j = 0; j < n; j++) {
    double* xj = X + (size_t)ldx * j;
    for (int k = 0; k < j; k++) {
      const double u = Um[k + (size_t)ldu * j];
      if (u == 0.0) continue;
      const double* yk = X + (size_t)ldx * k;
#pragma omp parallel for schedule(static) if (m > 4096)
      for (int i = 0; i < m; i++) xj[i] -= yk[i] * u;
    }
    const double d = 1.0 / Um[j + (size_t)ldu * j];
    for (int i = 0; i < m; i++) xj[i] *= d;
  }
}

/* Householder reduction of a symmetric matrix (lower triangle filled from
 * the upper one) to tridiagonal form; Q accumulated explicitly. */
static void tridiag(int n, double* A /* n x n, ld n, full symmetric */, double* Q, double* d,
                    double* e) {
#define a(i, j) A[(i) + (size_t)n * (j)]
#define q(i, j) Q[(i) + (size_t)n * (j)]
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) q(i, j) = (i == j) ? 1.0 : 0.0;
  double* v = (double*)malloc(sizeof(double) * n);
  double* p = (double*)malloc(sizeof(double) * n);
  double* w = (double*)malloc(sizeof(double) * n);
  for (int k = 0; k < n - 2; k++) {
    double nx = 0.0;
    for (int i = k + 1; i < n; i++) nx += a(i, k) * a(i, k);
    nx = sqrt(nx);
    if (nx == 0.0) continue;
    double alpha = a(k + 1, k) > 0 ? -nx : nx;
    for (int i = k + 1; i < n; i++) v[i] = a(i, k);
    v[k + 1] -= alpha;
    double nv = 0.0;
    for (int i = k + 1; i < n; i++) nv += v[i] * v[i];
    nv = sqrt(nv);
    if (nv == 0.0) continue;
    for (int i = k + 1; i < n; i++) v[i] /= nv;
    /* p = A22 v, K = v^T p, q = p - K v; A22 -= 2 (v q^T + q v^T) */
#pragma omp parallel for schedule(static) if (n > 256)
    for (int i = k + 1; i < n; i++) {
      double s = 0.0; /* A symmetric: row i == column i */
      for (int j = k + 1; j < n; j++) s += a(j, i) * v[j];
      p[i] = s;
    }
    double K = 0.0;
    for (int i = k + 1; i < n; i++) K += v[i] * p[i];
    for (int i = k + 1; i < n; i++) p[i] -= K * v[i];
#pragma omp parallel for schedule(static) if (n > 256)
    for (int j = k + 1; j < n; j++)
      for (int i = k + 1; i < n; i++) a(i, j) -= 2.0 * (v[i] * p[j] + p[i] * v[j]);
    a(k + 1, k) = alpha;
    a(k, k + 1) = alpha;
    for (int i = k + 2; i < n; i++) { a(i, k) = 0.0; a(k, i) = 0.0; }
    /* Q <- Q (I - 2 v v^T) on columns k+1.. */
    for (int i = 0; i < n; i++) w[i] = 0.0;
    for (int j = k + 1; j < n; j++)
      for (int i = 0; i < n; i++) w[i] += q(i, j) * v[j];
#pragma omp parallel for schedule(static) if (n > 256)
    for (int j = k + 1; j < n; j++)
      for (int i = 0; i < n; i++) q(i, j) -= 2.0 * w[i] * v[j];
  }
  for (int i = 0; i < n; i++) d[i] = a(i, i);
  for (int i = 0; i < n - 1; i++) e[i] = a(i + 1, i);
  e[n - 1] = 0.0;
  free(v); free(p); free(w);
#undef a
#undef q
}

/* Implicit QL with shifts on (d, e), e[i] couples i and i+1; Z accumulates. */
static int tql(int n, double* d, double* e, double* Z) {
  const double eps = 2.220446049250313e-16;
  for (int l = 0; l < n; l++) {
    int iter = 0, m;
    do {
      for (m = l; m < n - 1; m++) {
        double dd = fabs(d[m]) + fabs(d[m + 1]);
        if (fabs(e[m]) <= eps * dd) break;
      }
      if (m != l) {
        if (iter++ == 200) return -1;
        double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
        double r = hypot(g, 1.0);
        g = d[m] - d[l] + e[l] / (g + (g >= 0 ? fabs(r) : -fabs(r)));
        double s = 1.0, c = 1.0, p = 0.0;
        int i;
        for (i = m - 1; i >= l; i--) {
          double f = s * e[i], b = c * e[i];
          r = hypot(f, g);
          e[i + 1] = r;
          if (r == 0.0) {
            d[i + 1] -= p;
            e[m] = 0.0;
            break;
          }
          s = f / r;
          c = g / r;
          g = d[i + 1] - p;
          r = (d[i] - g) * s + 2.0 * c * b;
          p = s * r;
          d[i + 1] = g + p;
          g = c * r - b;
          double* zi = Z + (size_t)n * i;
          double* zi1 = Z + (size_t)n * (i + 1);
          for (int k = 0; k < n; k++) {
            double ff = zi1[k];
            zi1[k] = s * zi[k] + c * ff;
            zi[k] = c * zi[k] - s * ff;
          }
        }
        if (r == 0.0 && i >= l) continue;
        d[l] -= p;
        e[l] = g;
        e[m] = 0.0;
      }
    } while (m != l);
  }
  return 0;
}

int or_syev(int n, double* A, int lda, double* w) {
  if (n <= 0) return 0;
  double* M = (double*)malloc(sizeof(double) * (size_t)n * n);
  double* Q = (double*)malloc(sizeof(double) * (size_t)n * n);
  double* e = (double*)malloc(sizeof(double) * n);
  /* full symmetric copy from the upper triangle */
  for (int j = 0; j < n; j++)
    for (int i = 0; i <= j; i++) {
      double v = A[i + (size_t)lda * j];
      M[i + (size_t)n * j] = v;
      M[j + (size_t)n * i] = v;
    }
  tridiag(n, M, Q, w, e);
  int rc = tql(n, w, e, Q);
  /* sort ascending (selection sort keeps it simple and stable enough) */
  int* ord = (int*)malloc(sizeof(int) * n);
  for (int i = 0; i < n; i++) ord[i] = i;
  for (int i = 0; i < n; i++) {
    int k = i;
    for (int j = i + 1; j < n; j++)
      if (w[ord[j]] < w[ord[k]]) k = j;
    int t = ord[i]; ord[i] = ord[k]; ord[k] = t;
  }
  double* ws = (double*)malloc(sizeof(double) * n);
  for (int i = 0; i < n; i++) ws[i] = w[ord[i]];
  for (int j = 0; j < n; j++) {
    const double* src = Q + (size_t)n * ord[j];
    /* deterministic sign: largest-magnitude component positive */
    int im = 0;
    for (int i = 1; i < n; i++)
      if (fabs(src[i]) > fabs(src[im])) im = i;
    double sg = src[im] < 0 ? -1.0 : 1.0;
    for (int i = 0; i < n; i++) A[i + (size_t)lda * j] = sg * src[i];
  }
  memcpy(w, ws, sizeof(double) * n);
  free(ws); free(ord); free(M); free(Q); free(e);
  return rc;
}

int or_solve_gen_eig(int lda, int numCol, int width, double* AMat, double* BMat, double* eigValS) {
#define AM(i, j) AMat[(i) + (size_t)lda * (j)]
#define BM(i, j) BMat[(i) + (size_t)lda * (j)]
  /* symmetrise (eigensolver.cpp:229-235) */
  for (int j = 0; j < numCol; j++)
    for (int i = j + 1; i < numCol; i++) {
      AM(i, j) = AM(j, i);
      BM(i, j) = BM(j, i);
    }
  double* sigma2 = (double*)calloc(lda, sizeof(double));
  double* invsigma = (double*)calloc(lda, sizeof(double));
  int rc = or_syev(numCol, BMat, lda, sigma2);
  if (rc) { free(sigma2); free(invsigma); return -2; }
  int numKeep = 0;
  for (int i = numCol - 1; i >= 0; i--) {
    if (sigma2[i] / sigma2[numCol - 1] > 1e-8) numKeep++;
    else break;
  }
  for (int i = 0; i < numKeep; i++) invsigma[i] = 1.0 / sqrt(sigma2[i + numCol - numKeep]);
  if (numKeep < width) { free(sigma2); free(invsigma); return -3; } /* "there are not enough number of columns." */
  double* T1 = (double*)calloc((size_t)lda * lda, sizeof(double));
  const double* Ukeep = BMat + (size_t)lda * (numCol - numKeep);
  /* AMatT1 = A * U_keep ; A = U_keep^T * AMatT1 */
  or_gemm_nn(numCol, numKeep, numCol, 1.0, AMat, lda, Ukeep, lda, 0.0, T1, lda);
  for (int j = 0; j < numKeep; j++)
    for (int i = 0; i < numKeep; i++) {
      double s = 0.0;
      for (int r = 0; r < numCol; r++) s += Ukeep[r + (size_t)lda * i] * T1[r + (size_t)lda * j];
      AM(i, j) = s;
    }
  for (int j = 0; j < numKeep; j++)
    for (int i = 0; i < numKeep; i++) AM(i, j) *= invsigma[i] * invsigma[j];
  double* eigs = (double*)malloc(sizeof(double) * numKeep);
  rc = or_syev(numKeep, AMat, lda, eigs);
  for (int i = 0; i < width; i++) eigValS[i] = eigs[i];
  for (int j = 0; j < numKeep; j++)
    for (int i = 0; i < numKeep; i++) AM(i, j) *= invsigma[i];
  or_gemm_nn(numCol, numKeep, numKeep, 1.0, Ukeep, lda, AMat, lda, 0.0, T1, lda);
  for (int j = 0; j < numKeep; j++)
    for (int i = 0; i < numCol; i++) AM(i, j) = T1[i + (size_t)lda * j];
  free(eigs); free(T1); free(sigma2); free(invsigma);
  return rc ? -2 : 0;
#undef AM
#undef BM
}

int or_orthogonalize(int npw, int nb, double* X) {
  const int m = 2 * npw;
  double* S = (double*)malloc(sizeof(double) * (size_t)nb * nb);
  or_gemm_tn(m, nb, nb, X, X, S);
  int info = or_potrf_upper(nb, S, nb);
  if (info == 0) or_trsm_right_upper(m, nb, S, nb, X, m);
  free(S);
  return info == 0 ? 0 : -4;
}

/* lapack::Sygvd(1, 'V', 'U', n, A, lda, B, ldb, w) (include/lapack.hpp):
 * A x = lambda B x with B = U^T U (Potrf 'U'), C = U^{-T} A U^{-1} (upper
 * triangles read), C = Y diag(w) Y^T, x = U^{-1} y; A <- X (B-orthonormal
 * columns, each with its largest-magnitude component positive), B <- U.
 * Returns 0, or -4 when B is not positive definite. */
int or_sygvd(int n, double* A, int lda, double* B, int ldb, double* w) {
  if (n <= 0) return 0;
  for (int j = 0; j < n; j++)
    for (int i = j + 1; i < n; i++) B[i + (size_t)ldb * j] = 0.0;
  if (or_potrf_upper(n, B, ldb) != 0) return -4;
  double* C = (double*)malloc(sizeof(double) * (size_t)n * n);
  for (int j = 0; j < n; j++)
    for (int i = 0; i < n; i++) C[i + (size_t)n * j] = (i <= j) ? A[i + (size_t)lda * j] : A[j + (size_t)lda * i];
  /* C <- C U^{-1} (rows), then C <- U^{-T} C */
  or_trsm_right_upper(n, n, B, ldb, C, n);
  for (int j = 0; j < n; j++)      /* solve U^T Z = C column by column (forward substitution) */
    for (int i = 0; i < n; i++) {
      double s = C[i + (size_t)n * j];
      for (int k = 0; k < i; k++) s -= B[k + (size_t)ldb * i] * C[k + (size_t)n * j];
      C[i + (size_t)n * j] = s / B[i + (size_t)ldb * i];
    }
  int rc = or_syev(n, C, n, w);
  /* X = U^{-1} Y: back substitution per column */
  for (int j = 0; j < n; j++)
    for (int i = n - 1; i >= 0; i--) {
      double s = C[i + (size_t)n * j];
      for (int k = i + 1; k < n; k++) s -= B[i + (size_t)ldb * k] * C[k + (size_t)n * j];
      C[i + (size_t)n * j] = s / B[i + (size_t)ldb * i];
    }
  for (int j = 0; j < n; j++) {
    const double* z = C + (size_t)n * j;
    int im = 0;
    for (int i = 1; i < n; i++)
      if (fabs(z[i]) > fabs(z[im])) im = i;
    const double sg = z[im] < 0 ? -1.0 : 1.0;
    for (int i = 0; i < n; i++) A[i + (size_t)lda * j] = sg * z[i];
  }
  free(C);
  return rc ? -2 : 0;
}
