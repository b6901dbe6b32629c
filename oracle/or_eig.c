/*
 * or_eig.c -- FP64 eigensolvers of the oracle (test infrastructure only).
 *
 *   or_lobpcg        EigenSolver::LOBPCGSolveReal, sphere branch
 *                    (lobpcg.cpp:119-576), one rank: the Alltoall transposes
 *                    and MPI reductions are identities.
 *   or_ppcg          EigenSolver::PPCGSolveReal, sphere branch (ppcg.cpp:119-576).
 *   or_lanczos       k-step Lanczos bound of Zhou & Li, Linear Algebra Appl.
 *                    435 (2011) 480: b_up = theta_k + |beta_k|.
 *   or_cheb_filter   scaled Chebyshev filter of Zhou, Saad, Tiago &
 *                    Chelikowsky, J. Comput. Phys. 219 (2006) 172.
 *   or_rayleigh_ritz Cholesky QR (EigenSolver::Orthogonalize,
 *                    eigensolver.cpp:386-469) then Rayleigh-Ritz.
 *   or_chebfsi_sweep one filter + RR cycle.  The reference has no ChebFSI
 *                    (scf.cpp:1306-1316 and 1321-1332 are compiled out), so the
 *                    bound rules below are this build's, stated in oracle.h.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

static double energy(int m, const double* x) {
  double s = 0.0;
  for (int i = 0; i < m; i++) s += x[i] * x[i];
  return s;
}

/* copy a rows x cols block into AMat at (r0, c0) */
static void put_block(double* M, int lda, int r0, int c0, int rows, int cols, const double* B) {
  for (int j = 0; j < cols; j++)
    for (int i = 0; i < rows; i++) M[(r0 + i) + (size_t)lda * (c0 + j)] = B[i + (size_t)rows * j];
}

int or_lobpcg(or_ham* H, int nb, int numEig, int scfIter, int eigMaxIter, double eigMinTolerance,
              double eigTolerance, double* psi, double* eigval, double* resval, int* iters_out) {
  (void)eigMinTolerance;
  extern int or_ham_npw(const or_ham*);
  const int npw = or_ham_npw(H);
  const int m = 2 * npw, width = nb, lda = 3 * width;
  if (numEig > width) return -5; /* lobpcg.cpp:111-117 */
  const size_t blk = (size_t)m * width;
  double* S = (double*)calloc(3 * blk, sizeof(double));
  double* AS = (double*)calloc(3 * blk, sizeof(double));
  double *X = S, *W = S + blk, *P = S + 2 * blk;
  double *AX = AS, *AW = AS + blk, *AP = AS + 2 * blk;
  double* AMat = (double*)calloc((size_t)lda * lda, sizeof(double));
  double* BMat = (double*)calloc((size_t)lda * lda, sizeof(double));
  double* XTX = (double*)calloc((size_t)width * width, sizeof(double));
  double* T = (double*)calloc((size_t)width * width, sizeof(double));
  double* Xtemp = (double*)calloc(blk, sizeof(double));
  double* resNorm = (double*)calloc(width, sizeof(double));
  double* eigValS = (double*)calloc(lda, sizeof(double));
  int rc = 0, iter = 0, numSet = 2;
  double resMax = 0.0;

  memcpy(X, psi, sizeof(double) * blk);
  if (scfIter == 1) {
    rc = or_orthogonalize(npw, width, X);
    if (rc) goto done;
  }
  or_apply_h(H, width, X, AX);

  do {
    iter++;
    numSet = (iter == 1) ? 2 : 3;

    or_gemm_tn(m, width, width, X, AX, XTX);
    put_block(AMat, lda, 0, 0, width, width, XTX);

    /* R = AX - X (X'AX) */
    memcpy(Xtemp, AX, sizeof(double) * blk);
    or_gemm_nn(m, width, width, -1.0, X, m, AMat, lda, 1.0, Xtemp, m);
    for (int k = 0; k < width; k++) {
      double r = sqrt(energy(m, Xtemp + (size_t)m * k));
      double d = fabs(XTX[k + (size_t)width * k]);
      resNorm[k] = r / (d > 1.0 ? d : 1.0);
    }
    resMax = resNorm[0];
    for (int k = 1; k < numEig; k++)
      if (resNorm[k] > resMax) resMax = resNorm[k];

    /* W = T R, normalised (lobpcg.cpp:275-291) */
    memset(W, 0, sizeof(double) * blk);
    or_teter_precond(H, width, Xtemp, W);
    for (int k = 0; k < width; k++) {
      double nrm = sqrt(energy(m, W + (size_t)m * k));
      for (int i = 0; i < m; i++) W[(size_t)m * k + i] *= 1.0 / nrm;
    }
    if (numSet == 3) {
      for (int k = 0; k < width; k++) {
        double nrm = sqrt(energy(m, P + (size_t)m * k));
        for (int i = 0; i < m; i++) {
          P[(size_t)m * k + i] *= 1.0 / nrm;
          AP[(size_t)m * k + i] *= 1.0 / nrm;
        }
      }
    }
    or_apply_h(H, width, W, AW);

    or_gemm_tn(m, width, width, X, AW, T); put_block(AMat, lda, 0, width, width, width, T);
    or_gemm_tn(m, width, width, W, AW, T); put_block(AMat, lda, width, width, width, width, T);
    if (numSet == 3) {
      or_gemm_tn(m, width, width, X, AP, T); put_block(AMat, lda, 0, 2 * width, width, width, T);
      or_gemm_tn(m, width, width, W, AP, T); put_block(AMat, lda, width, 2 * width, width, width, T);
      or_gemm_tn(m, width, width, P, AP, T); put_block(AMat, lda, 2 * width, 2 * width, width, width, T);
    }
    or_gemm_tn(m, width, width, X, X, T); put_block(BMat, lda, 0, 0, width, width, T);
    or_gemm_tn(m, width, width, X, W, T); put_block(BMat, lda, 0, width, width, width, T);
    or_gemm_tn(m, width, width, W, W, T); put_block(BMat, lda, width, width, width, width, T);
    if (numSet == 3) {
      or_gemm_tn(m, width, width, X, P, T); put_block(BMat, lda, 0, 2 * width, width, width, T);
      or_gemm_tn(m, width, width, W, P, T); put_block(BMat, lda, width, 2 * width, width, width, T);
      or_gemm_tn(m, width, width, P, P, T); put_block(BMat, lda, 2 * width, 2 * width, width, width, T);
    }
    const int numCol = (numSet == 3) ? 3 * width : 2 * width;
    rc = or_solve_gen_eig(lda, numCol, width, AMat, BMat, eigValS);
    if (rc) goto done;

    const double* CX = AMat;
    const double* CW = AMat + width;
    const double* CP = AMat + 2 * width;
    if (numSet == 2) {
      or_gemm_nn(m, width, width, 1.0, X, m, CX, lda, 0.0, Xtemp, m);
      or_gemm_nn(m, width, width, 1.0, W, m, CW, lda, 1.0, Xtemp, m);
      memcpy(X, Xtemp, sizeof(double) * blk);
      memcpy(P, W, sizeof(double) * blk);
      or_gemm_nn(m, width, width, 1.0, AX, m, CX, lda, 0.0, Xtemp, m);
      or_gemm_nn(m, width, width, 1.0, AW, m, CW, lda, 1.0, Xtemp, m);
      memcpy(AX, Xtemp, sizeof(double) * blk);
      memcpy(AP, AW, sizeof(double) * blk);
    } else {
      or_gemm_nn(m, width, width, 1.0, W, m, CW, lda, 0.0, Xtemp, m);
      or_gemm_nn(m, width, width, 1.0, P, m, CP, lda, 1.0, Xtemp, m);
      memcpy(P, Xtemp, sizeof(double) * blk);
      or_gemm_nn(m, width, width, 1.0, X, m, CX, lda, 1.0, Xtemp, m);
      memcpy(X, Xtemp, sizeof(double) * blk);
      or_gemm_nn(m, width, width, 1.0, AW, m, CW, lda, 0.0, Xtemp, m);
      or_gemm_nn(m, width, width, 1.0, AP, m, CP, lda, 1.0, Xtemp, m);
      memcpy(AP, Xtemp, sizeof(double) * blk);
      or_gemm_nn(m, width, width, 1.0, AX, m, CX, lda, 1.0, Xtemp, m);
      memcpy(AX, Xtemp, sizeof(double) * blk);
    }
  } while (iter < eigMaxIter && resMax > eigTolerance);

  /* post-processing (lobpcg.cpp:536-548): Syevd of the X'AX formed at the
   * start of the last iteration, then X <- X C. */
  rc = or_syev(width, XTX, width, eigValS);
  if (rc) goto done;
  or_gemm_nn(m, width, width, 1.0, X, m, XTX, width, 0.0, Xtemp, m);
  memcpy(psi, Xtemp, sizeof(double) * blk);
  memcpy(eigval, eigValS, sizeof(double) * width);
  memcpy(resval, resNorm, sizeof(double) * width);
done:
  if (iters_out) *iters_out = iter;
  free(S); free(AS); free(AMat); free(BMat); free(XTX); free(T); free(Xtemp);
  free(resNorm); free(eigValS);
  return rc;
}

int or_lanczos(or_ham* H, int k, const double* v0, double* theta, double* b_up) {
  extern int or_ham_npw(const or_ham*);
  const int m = 2 * or_ham_npw(H);
  double* v = (double*)malloc(sizeof(double) * m);
  double* vp = (double*)calloc(m, sizeof(double));
  double* f = (double*)malloc(sizeof(double) * m);
  double* al = (double*)calloc(k, sizeof(double));
  double* be = (double*)calloc(k, sizeof(double));
  double nrm = sqrt(energy(m, v0));
  for (int i = 0; i < m; i++) v[i] = v0[i] / nrm;
  double beta = 0.0;
  for (int j = 0; j < k; j++) {
    or_apply_h(H, 1, v, f);
    double a = 0.0;
    for (int i = 0; i < m; i++) a += v[i] * f[i];
    for (int i = 0; i < m; i++) f[i] -= a * v[i] + beta * vp[i];
    al[j] = a;
    beta = sqrt(energy(m, f));
    be[j] = beta;
    if (j + 1 < k) {
      for (int i = 0; i < m; i++) {
        vp[i] = v[i];
        v[i] = f[i] / beta;
      }
    }
  }
  /* Ritz values of the k x k tridiagonal T_k */
  double* Tm = (double*)calloc((size_t)k * k, sizeof(double));
  for (int j = 0; j < k; j++) {
    Tm[j + (size_t)k * j] = al[j];
    if (j + 1 < k) Tm[j + (size_t)k * (j + 1)] = be[j];
  }
  int rc = or_syev(k, Tm, k, theta);
  *b_up = theta[k - 1] + fabs(be[k - 1]);
  free(Tm); free(v); free(vp); free(f); free(al); free(be);
  return rc;
}

int or_cheb_filter(or_ham* H, int nb, int deg, double a, double b, double aL, double* X) {
  extern int or_ham_npw(const or_ham*);
  const int m = 2 * or_ham_npw(H);
  const size_t n = (size_t)m * nb;
  if (deg < 1) return 0;
  double* B0 = (double*)malloc(sizeof(double) * n);
  double* B1 = (double*)malloc(sizeof(double) * n);
  double* B2 = (double*)malloc(sizeof(double) * n);
  memcpy(B0, X, sizeof(double) * n);
  double *Xp = B0, *Y = B1, *Yn = B2;
  const double e = (b - a) / 2.0, c = (b + a) / 2.0;
  double sigma = e / (aL - c);
  const double tau = 2.0 / sigma;
  /* Y = (H X - c X) * sigma / e */
  or_apply_h(H, nb, Xp, Y);
  for (size_t i = 0; i < n; i++) Y[i] = (Y[i] - c * Xp[i]) * (sigma / e);
  for (int it = 2; it <= deg; it++) {
    const double sn = 1.0 / (tau - sigma);
    or_apply_h(H, nb, Y, Yn);
    for (size_t i = 0; i < n; i++) Yn[i] = (Yn[i] - c * Y[i]) * (2.0 * sn / e) - (sigma * sn) * Xp[i];
    double* t = Xp; Xp = Y; Y = Yn; Yn = t;
    sigma = sn;
  }
  memcpy(X, Y, sizeof(double) * n);
  free(B0); free(B1); free(B2);
  return 0;
}

int or_rayleigh_ritz(or_ham* H, int nb, double* X, double* eig, double* res) {
  extern int or_ham_npw(const or_ham*);
  const int npw = or_ham_npw(H), m = 2 * npw;
  const size_t n = (size_t)m * nb;
  int rc = or_orthogonalize(npw, nb, X);
  if (rc) return rc;
  double* HX = (double*)malloc(sizeof(double) * n);
  double* T = (double*)malloc(sizeof(double) * n);
  double* Hs = (double*)malloc(sizeof(double) * (size_t)nb * nb);
  or_apply_h(H, nb, X, HX);
  or_gemm_tn(m, nb, nb, X, HX, Hs);
  rc = or_syev(nb, Hs, nb, eig);
  if (rc == 0) {
    or_gemm_nn(m, nb, nb, 1.0, X, m, Hs, nb, 0.0, T, m);
    memcpy(X, T, sizeof(double) * n);
    or_gemm_nn(m, nb, nb, 1.0, HX, m, Hs, nb, 0.0, T, m);
    for (int k = 0; k < nb; k++) {
      double s = 0.0;
      for (int i = 0; i < m; i++) {
        double r = T[(size_t)m * k + i] - eig[k] * X[(size_t)m * k + i];
        s += r * r;
      }
      res[k] = sqrt(s);
    }
  }
  free(HX); free(T); free(Hs);
  return rc;
}

int or_chebfsi_sweep(or_ham* H, int nb, int deg, int k, const double* v0, const double* prev_eig,
                     double* X, double* eig, double* res, double* bounds) {
  double* theta = (double*)malloc(sizeof(double) * (k > 0 ? k : 1));
  double bup = 0.0;
  int rc = or_lanczos(H, k, v0, theta, &bup);
  double a, aL;
  if (prev_eig) {
    a = prev_eig[nb - 1];
    aL = prev_eig[0];
  } else {
    aL = theta[0];
    a = theta[0] + 0.1 * (bup - theta[0]);
  }
  free(theta);
  if (rc) return rc;
  if (bounds) { bounds[0] = a; bounds[1] = bup; bounds[2] = aL; }
  rc = or_cheb_filter(H, nb, deg, a, bup, aL, X);
  if (rc) return rc;
  return or_rayleigh_ritz(H, nb, X, eig, res);
}

/* EigenSolver::PPCGSolveReal, sphere branch (ppcg.cpp:119-576), one rank:
 * no locking; numSet 2 on the first iteration, 3 after; the block of
 * sbSize columns k solves Sygvd on its (x | w | p) triple and updates
 * p <- w cw (+ p cp), x <- x cx + p (and the A-images alike); Cholesky QR
 * (X and AX by U^{-1}) closes every iteration; post-processing Syevd of
 * X^T A X, rotation and residual norms. */
static void pp_block_gram(int m, int sb, const double* a, const double* b, double* M, int ldm, int r0, int c0) {
  double* T = (double*)malloc(sizeof(double) * (size_t)sb * sb);
  or_gemm_tn(m, sb, sb, a, b, T);
  put_block(M, ldm, r0, c0, sb, sb, T);
  free(T);
}

/* The PPCG iteration on real m x nb blocks; hop(ctx, nb, x, y): y = H x;
 * pop(ctx, nb, r, w): w += T r.  The sphere branch runs it on the (2 npw)-row
 * real view, the real-space branch (ppcg.cpp:683-1246, the same steps on
 * DblNumMat) on the ntot-row coarse-grid bands. */
int or_ppcg_op(int m, or_op hop, or_op pop, void* ctx, int nb, int numEig, int scfIter, int eigMaxIter,
               double eigTolerance, int sbSize, double* psi, double* eigval, double* resval, int* iters_out) {
  const int width = nb;
  if (numEig > width) return -5; /* ppcg.cpp:105-112 */
  if (sbSize < 1) return -1;
  const size_t blk = (size_t)m * width;
  double* S = (double*)calloc(3 * blk, sizeof(double));
  double* AS = (double*)calloc(3 * blk, sizeof(double));
  double *X = S, *W = S + blk, *P = S + 2 * blk;
  double *AX = AS, *AW = AS + blk, *AP = AS + 2 * blk;
  double* XTX = (double*)calloc((size_t)width * width, sizeof(double));
  double* Xtemp = (double*)calloc(blk, sizeof(double));
  double* tmp = (double*)calloc(blk, sizeof(double));
  double* resNorm = (double*)calloc(width, sizeof(double));
  double* eigValS = (double*)calloc(3 * width, sizeof(double));
  const int sb3 = 3 * sbSize;
  double* AMat = (double*)calloc((size_t)sb3 * sb3, sizeof(double));
  double* BMat = (double*)calloc((size_t)sb3 * sb3, sizeof(double));
  double* eigs = (double*)calloc(sb3, sizeof(double));
  int rc = 0, iter = 0, numSet = 2;
  double resMax = 0.0;

  memcpy(X, psi, sizeof(double) * blk);
  if (scfIter == 1) {
    rc = or_orthogonalize(m / 2, width, X);  /* Cholesky QR on the m rows (m even) */
    if (rc) goto done;
  }
  hop(ctx, width, X, AX);
  do {
    iter++;
    numSet = (iter == 1) ? 2 : 3;
    or_gemm_tn(m, width, width, X, AX, XTX);
    memcpy(Xtemp, AX, sizeof(double) * blk);
    or_gemm_nn(m, width, width, -1.0, X, m, XTX, width, 1.0, Xtemp, m);
    for (int k = 0; k < width; k++) {
      const double r = sqrt(energy(m, Xtemp + (size_t)m * k));
      const double d = fabs(XTX[k + (size_t)width * k]);
      resNorm[k] = r / (d > 1.0 ? d : 1.0);
    }
    resMax = resNorm[0];
    for (int k = 1; k < numEig; k++)
      if (resNorm[k] > resMax) resMax = resNorm[k];
    /* W = T R, AW = A W */
    memset(W, 0, sizeof(double) * blk);
    pop(ctx, width, Xtemp, W);
    hop(ctx, width, W, AW);
    /* W -= X (X'W), AW -= AX (X'W); normalise */
    or_gemm_tn(m, width, width, X, W, XTX);
    or_gemm_nn(m, width, width, -1.0, X, m, XTX, width, 1.0, W, m);
    or_gemm_nn(m, width, width, -1.0, AX, m, XTX, width, 1.0, AW, m);
    for (int k = 0; k < width; k++) {
      const double s = 1.0 / sqrt(energy(m, W + (size_t)m * k));
      for (int i = 0; i < m; i++) {
        W[(size_t)m * k + i] *= s;
        AW[(size_t)m * k + i] *= s;
      }
    }
    if (numSet == 3) {
      or_gemm_tn(m, width, width, X, P, XTX);
      or_gemm_nn(m, width, width, -1.0, X, m, XTX, width, 1.0, P, m);
      or_gemm_nn(m, width, width, -1.0, AX, m, XTX, width, 1.0, AP, m);
      for (int k = 0; k < width; k++) {
        const double s = 1.0 / sqrt(energy(m, P + (size_t)m * k));
        for (int i = 0; i < m; i++) {
          P[(size_t)m * k + i] *= s;
          AP[(size_t)m * k + i] *= s;
        }
      }
    }
    const int nsb = (width + sbSize - 1) / sbSize;
    for (int k = 0; k < nsb; k++) {
      const int sb = (k == nsb - 1 && width % sbSize) ? width % sbSize : sbSize;
      const int ld = 3 * sb;
      const size_t off = (size_t)m * sbSize * k;
      double *x = X + off, *w = W + off, *p = P + off, *ax = AX + off, *aw = AW + off, *ap = AP + off;
      memset(AMat, 0, sizeof(double) * (size_t)sb3 * sb3);
      memset(BMat, 0, sizeof(double) * (size_t)sb3 * sb3);
      pp_block_gram(m, sb, x, ax, AMat, ld, 0, 0);
      pp_block_gram(m, sb, w, aw, AMat, ld, sb, sb);
      pp_block_gram(m, sb, x, aw, AMat, ld, 0, sb);
      pp_block_gram(m, sb, x, x, BMat, ld, 0, 0);
      pp_block_gram(m, sb, w, w, BMat, ld, sb, sb);
      pp_block_gram(m, sb, x, w, BMat, ld, 0, sb);
      if (numSet == 3) {
        pp_block_gram(m, sb, p, ap, AMat, ld, 2 * sb, 2 * sb);
        pp_block_gram(m, sb, x, ap, AMat, ld, 0, 2 * sb);
        pp_block_gram(m, sb, w, ap, AMat, ld, sb, 2 * sb);
        pp_block_gram(m, sb, p, p, BMat, ld, 2 * sb, 2 * sb);
        pp_block_gram(m, sb, x, p, BMat, ld, 0, 2 * sb);
        pp_block_gram(m, sb, w, p, BMat, ld, sb, 2 * sb);
      }
      const int dim = (numSet == 3) ? 3 * sb : 2 * sb;
      rc = or_sygvd(dim, AMat, ld, BMat, ld, eigs);
      if (rc) goto done;
      const double* cx = AMat;
      const double* cw = AMat + sb;
      const double* cp = AMat + 2 * sb;
      if (numSet == 3) {
        or_gemm_nn(m, sb, sb, 1.0, p, m, cp, ld, 0.0, tmp, m);
        or_gemm_nn(m, sb, sb, 1.0, w, m, cw, ld, 1.0, tmp, m);
        memcpy(p, tmp, sizeof(double) * (size_t)m * sb);
        or_gemm_nn(m, sb, sb, 1.0, ap, m, cp, ld, 0.0, tmp, m);
        or_gemm_nn(m, sb, sb, 1.0, aw, m, cw, ld, 1.0, tmp, m);
        memcpy(ap, tmp, sizeof(double) * (size_t)m * sb);
      } else {
        or_gemm_nn(m, sb, sb, 1.0, w, m, cw, ld, 0.0, p, m);
        or_gemm_nn(m, sb, sb, 1.0, aw, m, cw, ld, 0.0, ap, m);
      }
      memcpy(tmp, p, sizeof(double) * (size_t)m * sb);
      or_gemm_nn(m, sb, sb, 1.0, x, m, cx, ld, 1.0, tmp, m);
      memcpy(x, tmp, sizeof(double) * (size_t)m * sb);
      memcpy(tmp, ap, sizeof(double) * (size_t)m * sb);
      or_gemm_nn(m, sb, sb, 1.0, ax, m, cx, ld, 1.0, tmp, m);
      memcpy(ax, tmp, sizeof(double) * (size_t)m * sb);
    }
    /* Cholesky QR of X; AX transformed by the same U^{-1} */
    or_gemm_tn(m, width, width, X, X, XTX);
    if (or_potrf_upper(width, XTX, width) != 0) { rc = -4; goto done; }
    or_trsm_right_upper(m, width, XTX, width, X, m);
    or_trsm_right_upper(m, width, XTX, width, AX, m);
  } while (iter < eigMaxIter && resMax > eigTolerance);

  /* post-processing (ppcg.cpp:530-576) */
  or_gemm_tn(m, width, width, X, AX, XTX);
  rc = or_syev(width, XTX, width, eigValS);
  if (rc) goto done;
  or_gemm_nn(m, width, width, 1.0, X, m, XTX, width, 0.0, Xtemp, m);
  memcpy(X, Xtemp, sizeof(double) * blk);
  or_gemm_nn(m, width, width, 1.0, AX, m, XTX, width, 0.0, Xtemp, m);
  memcpy(AX, Xtemp, sizeof(double) * blk);
  for (int k = 0; k < width; k++) {
    double s = 0.0;
    for (int i = 0; i < m; i++) {
      const double r = AX[(size_t)m * k + i] - X[(size_t)m * k + i] * eigValS[k];
      s += r * r;
    }
    const double d = fabs(eigValS[k]);
    resNorm[k] = sqrt(s) / (d > 1.0 ? d : 1.0);
  }
  memcpy(psi, X, sizeof(double) * blk);
  memcpy(eigval, eigValS, sizeof(double) * width);
  memcpy(resval, resNorm, sizeof(double) * width);
done:
  if (iters_out) *iters_out = iter;
  free(S); free(AS); free(XTX); free(Xtemp); free(tmp); free(resNorm); free(eigValS);
  free(AMat); free(BMat); free(eigs);
  return rc;
}

static void ham_hop(void* ctx, int nb, const double* x, double* y) { or_apply_h((or_ham*)ctx, nb, x, y); }
static void ham_pop(void* ctx, int nb, const double* r, double* w) { or_teter_precond((or_ham*)ctx, nb, r, w); }

int or_ppcg(or_ham* H, int nb, int numEig, int scfIter, int eigMaxIter, double eigMinTolerance,
            double eigTolerance, int sbSize, double* psi, double* eigval, double* resval, int* iters_out) {
  (void)eigMinTolerance;
  extern int or_ham_npw(const or_ham*);
  return or_ppcg_op(2 * or_ham_npw(H), ham_hop, ham_pop, H, nb, numEig, scfIter, eigMaxIter, eigTolerance,
                    sbSize, psi, eigval, resval, iters_out);
}
