/*
 * or_ham.c -- FP64 restatement of the sphere-cut Gamma-point Hamiltonian
 * (test infrastructure only; see oracle.h).
 *
 *   or_sphere_build      fourier.cpp:97-244 (G grid), 541-612 (half sphere, padding)
 *   or_fock_sphere_build fourier.cpp:642-670
 *   or_ekin_teter        hamiltonian.cpp:2629-2667
 *   or_mult_spinor       hamiltonian.cpp:3270-3452 -> spinor.cpp:527-680
 *   or_nonlocal_multx    eigensolver.cpp:471-515, blas.cpp:997-1067
 *   or_teter_precond     spinor.cpp:348-374
 *   or_ham_set_exx       hamiltonian.cpp:3454-3551 (kernel), 3553-3603 (phi)
 *   or_add_exx           spinor.cpp:783-945
 *   or_ham_build_ace     hamiltonian.cpp:3685-3924 (sphere branch); applied in
 *                        or_mult_spinor as hamiltonian.cpp:3347-3430
 *   or_density           hamiltonian.cpp:2509-2573
 *   or_nonlocal_force    hamiltonian.cpp:2907-3006 (in G-space form)
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "oracle.h"

typedef double _Complex cplx;

void* or_plan3_new(const int* n);
void or_plan3_delete(void* P);
void or_plan3_c2r(void* P, double* work_half, double* out);
void or_plan3_r2c(void* P, const double* in, double* out_half);

static const double PI = 3.14159265358979323846;

double or_volume(const double* s) {
  return fabs(s[0] * (s[4] * s[8] - s[7] * s[5]) - s[3] * (s[1] * s[8] - s[7] * s[2]) +
              s[6] * (s[1] * s[5] - s[4] * s[2]));
}

/* supercell[3*j+i] = component i of lattice vector a_j;
 * recip[3*j+i] = component i of reciprocal vector b_j, a_j . b_k = 2 pi delta. */
void or_recip_cell(const double* s, double* b) {
  /* inverse of S (row j = a_j) */
  double det = s[0] * (s[4] * s[8] - s[5] * s[7]) - s[1] * (s[3] * s[8] - s[5] * s[6]) +
               s[2] * (s[3] * s[7] - s[4] * s[6]);
  double inv[9];
  inv[0] = (s[4] * s[8] - s[5] * s[7]) / det;
  inv[1] = (s[2] * s[7] - s[1] * s[8]) / det;
  inv[2] = (s[1] * s[5] - s[2] * s[4]) / det;
  inv[3] = (s[5] * s[6] - s[3] * s[8]) / det;
  inv[4] = (s[0] * s[8] - s[2] * s[6]) / det;
  inv[5] = (s[2] * s[3] - s[0] * s[5]) / det;
  inv[6] = (s[3] * s[7] - s[4] * s[6]) / det;
  inv[7] = (s[1] * s[6] - s[0] * s[7]) / det;
  inv[8] = (s[0] * s[4] - s[1] * s[3]) / det;
  /* S * B^T = 2 pi I  ->  B^T = 2 pi S^-1  ->  b_j[i] = 2 pi inv[i][j] */
  for (int j = 0; j < 3; j++)
    for (int i = 0; i < 3; i++) b[3 * j + i] = 2.0 * PI * inv[3 * i + j];
}

static inline int kgrid(int i, int n) { return i <= n / 2 ? i : i - n; }

static inline double gkk_of(const double* b, int m0, int m1, int m2) {
  double g0 = b[0] * m0 + b[3] * m1 + b[6] * m2;
  double g1 = b[1] * m0 + b[4] * m1 + b[7] * m2;
  double g2 = b[2] * m0 + b[5] * m1 + b[8] * m2;
  return (g0 * g0 + g1 * g1 + g2 * g2) / 2.0;
}

int or_sphere_build(const int* n, const double* supercell, double ecut, int* idx, double* gkk,
                    int* npad, int* pad_src, int* pad_dst) {
  double b[9];
  or_recip_cell(supercell, b);
  const int n1 = n[0], n2 = n[1], n3 = n[2], h1 = n1 / 2 + 1;
  const double eps = 1e-16;
  int npw = 0, np = 0;
  for (int k = 0; k < n3; k++)
    for (int j = 0; j < n2; j++)
      for (int i = 0; i < h1; i++) {
        int G0 = kgrid(i, n1), G1 = kgrid(j, n2), G2 = kgrid(k, n3);
        int g = i + h1 * (j + n2 * k);
        double e = gkk_of(b, G0, G1, G2);
        int inGrid = 0, doPad = 0, inSphere = e < ecut;
        if (G0 > eps) inGrid = 1;
        else {
          if (G1 > eps) inGrid = 1;
          else if (G1 > -eps) {
            if (G2 >= -eps) inGrid = 1;
          }
        }
        if (G0 < eps) {
          if (G1 < -eps) doPad = 1;
          else if (G1 < eps) {
            if (G2 < -eps) doPad = 1;
          }
        }
        if (inGrid && inSphere) {
          idx[npw] = g;
          gkk[npw] = e;
          npw++;
        }
        if (doPad && inSphere) {
          int m1 = -G1, m2 = -G2;
          if (m1 < 0) m1 += n2;
          if (m2 < 0) m2 += n3;
          int src = G0 + h1 * m1 + h1 * n2 * m2;
          if (src < h1 * n2 * n3) {
            pad_src[np] = src;
            pad_dst[np] = g;
            np++;
          }
        }
      }
  *npad = np;
  return npw;
}

int or_fock_sphere_build(const int* n, const double* supercell, double ecut, int* idx_fock,
                         double* gkk_fock) {
  double b[9];
  or_recip_cell(supercell, b);
  const int n1 = n[0], n2 = n[1], n3 = n[2], h1 = n1 / 2 + 1;
  int cnt = 0;
  /* ascending R2C index == the reference's std::set iteration order */
  for (int k = 0; k < n3; k++)
    for (int j = 0; j < n2; j++)
      for (int i = 0; i < h1; i++) {
        double e = gkk_of(b, kgrid(i, n1), kgrid(j, n2), kgrid(k, n3));
        if (e < ecut * 4.0) {
          idx_fock[cnt] = i + h1 * (j + n2 * k);
          gkk_fock[cnt] = e;
          cnt++;
        }
      }
  return cnt;
}

void or_ekin_teter(int npw, const double* gkk, double* ekin, double* teter) {
  for (int i = 0; i < npw; i++) {
    ekin[i] = gkk[i];
    double a = ekin[i] * 2.0;
    double bb = 27.0 + a * (18.0 + a * (12.0 + a * 8.0));
    teter[i] = bb / (bb + 16.0 * pow(a, 4.0));
  }
}

/* ------------------------------------------------------------------ */

struct or_ham {
  int n[3], h1, N, Nh;
  double vol, supercell[9], recip[9];
  int npw, npad;
  int *idx, *pad_src, *pad_dst;
  double *ekin, *teter, *vtot;
  int nbeta;
  cplx* beta;
  double* wgt;
  /* EXX */
  int nocc, nfock;
  int* fock_idx;
  double* exx_kernel;
  double exx_fraction;
  double* phi_rs; /* N x nocc real-space phi */
  double* occ;
  /* ACE: xi (2npw x nace real view); MultSpinor uses it instead of V_x */
  int nace;
  double* xi;
  int nthreads;
};

or_ham* or_ham_create(const int* n, const double* supercell, int npw, const int* idx, int npad,
                      const int* pad_src, const int* pad_dst, const double* ekin,
                      const double* teter, const double* vtot) {
  or_ham* H = (or_ham*)calloc(1, sizeof(or_ham));
  memcpy(H->n, n, 3 * sizeof(int));
  H->h1 = n[0] / 2 + 1;
  H->N = n[0] * n[1] * n[2];
  H->Nh = H->h1 * n[1] * n[2];
  memcpy(H->supercell, supercell, 9 * sizeof(double));
  or_recip_cell(supercell, H->recip);
  H->vol = or_volume(supercell);
  H->npw = npw;
  H->npad = npad;
  H->idx = (int*)malloc(sizeof(int) * npw);
  memcpy(H->idx, idx, sizeof(int) * npw);
  H->pad_src = (int*)malloc(sizeof(int) * (npad + 1));
  H->pad_dst = (int*)malloc(sizeof(int) * (npad + 1));
  memcpy(H->pad_src, pad_src, sizeof(int) * npad);
  memcpy(H->pad_dst, pad_dst, sizeof(int) * npad);
  H->ekin = (double*)malloc(sizeof(double) * npw);
  H->teter = (double*)malloc(sizeof(double) * npw);
  memcpy(H->ekin, ekin, sizeof(double) * npw);
  memcpy(H->teter, teter, sizeof(double) * npw);
  H->vtot = (double*)malloc(sizeof(double) * H->N);
  memcpy(H->vtot, vtot, sizeof(double) * H->N);
  H->nthreads = 0;
  return H;
}

void or_ham_destroy(or_ham* H) {
  if (!H) return;
  free(H->idx); free(H->pad_src); free(H->pad_dst);
  free(H->ekin); free(H->teter); free(H->vtot);
  free(H->beta); free(H->wgt);
  free(H->fock_idx); free(H->exx_kernel); free(H->phi_rs); free(H->occ);
  free(H->xi);
  free(H);
}

int or_ham_set_threads(or_ham* H, int nthreads) {
  H->nthreads = nthreads;
  return 0;
}

static int nthr(const or_ham* H) {
#ifdef _OPENMP
  return H->nthreads > 0 ? H->nthreads : omp_get_max_threads();
#else
  (void)H;
  return 1;
#endif
}

int or_ham_set_nonlocal(or_ham* H, int nbeta, const double* beta, const double* wgt) {
  free(H->beta); free(H->wgt);
  H->nbeta = nbeta;
  H->beta = (cplx*)malloc(sizeof(cplx) * (size_t)H->npw * (nbeta > 0 ? nbeta : 1));
  H->wgt = (double*)malloc(sizeof(double) * (nbeta > 0 ? nbeta : 1));
  memcpy(H->beta, beta, sizeof(cplx) * (size_t)H->npw * nbeta);
  memcpy(H->wgt, wgt, sizeof(double) * nbeta);
  return 0;
}

/* sphere -> padded R2C half array, with the sqrt(2) convention (spinor.cpp:573-595) */
static void scatter_sphere(const or_ham* H, const cplx* c, cplx* half) {
  memset(half, 0, sizeof(cplx) * (size_t)H->Nh);
  const double is2 = 1.0 / sqrt(2.0);
  half[H->idx[0]] = c[0];
  for (int i = 1; i < H->npw; i++) half[H->idx[i]] = c[i] * is2;
  for (int p = 0; p < H->npad; p++) half[H->pad_dst[p]] = conj(half[H->pad_src[p]]);
}

/* a3 += gather of the R2C half array (spinor.cpp:666-676) */
static void gather_sphere_add(const or_ham* H, const cplx* half, cplx* a3) {
  const double s2 = sqrt(2.0);
  a3[0] += half[H->idx[0]];
  for (int i = 1; i < H->npw; i++) a3[i] += half[H->idx[i]] * s2;
}

int or_mult_spinor(or_ham* H, int nb, const double* psid, double* a3d) {
  const cplx* psi = (const cplx*)psid;
  cplx* a3 = (cplx*)a3d;
  const int npw = H->npw, N = H->N;
  memset(a3, 0, sizeof(cplx) * (size_t)npw * nb); /* hamiltonian.cpp:3300 */
#pragma omp parallel num_threads(nthr(H))
  {
    void* P = or_plan3_new(H->n);
    cplx* half = (cplx*)malloc(sizeof(cplx) * (size_t)H->Nh);
    double* r = (double*)malloc(sizeof(double) * (size_t)N);
#pragma omp for schedule(dynamic, 1)
    for (int k = 0; k < nb; k++) {
      const cplx* c = psi + (size_t)npw * k;
      cplx* out = a3 + (size_t)npw * k;
      scatter_sphere(H, c, half);
      or_plan3_c2r(P, (double*)half, r);
      const double fac = 1.0 / (double)N;
      for (int i = 0; i < N; i++) r[i] *= fac;
      for (int i = 0; i < N; i++) r[i] = r[i] * H->vtot[i];
      for (int i = 0; i < npw; i++) out[i] = c[i] * H->ekin[i];
      or_plan3_r2c(P, r, (double*)half);
      gather_sphere_add(H, half, out);
    }
    free(half);
    free(r);
    or_plan3_delete(P);
  }
  if (H->nocc > 0 && H->nace > 0) {
    /* ACE apply (hamiltonian.cpp:3347-3430): M = Re(xi^H psi) (real
     * nace x nb, blas.cpp:997-1029), a3 -= xi M (blas.cpp:1031-1067) */
    const int m = 2 * H->npw;
    double* M = (double*)malloc(sizeof(double) * (size_t)H->nace * nb);
    or_gemm_tn(m, H->nace, nb, H->xi, psid, M);
    or_gemm_nn(m, nb, H->nace, -1.0, H->xi, m, M, H->nace, 1.0, a3d, m);
    free(M);
    return 0;
  }
  if (H->nocc > 0) return or_add_exx(H, nb, psid, a3d);
  return 0;
}

int or_nonlocal_multx(or_ham* H, int nb, const double* Xd, double* VnlXd) {
  const cplx* X = (const cplx*)Xd;
  cplx* V = (cplx*)VnlXd;
  const int npw = H->npw, nbeta = H->nbeta;
  if (nbeta == 0) {
    memset(V, 0, sizeof(cplx) * (size_t)npw * nb);
    return 0;
  }
  double* W = (double*)calloc((size_t)nbeta * nb, sizeof(double));
  /* W = Re(beta^H X) = beta_r^T X_r + beta_i^T X_i (blas.cpp:997-1029) */
#pragma omp parallel for collapse(2) num_threads(nthr(H))
  for (int b = 0; b < nb; b++)
    for (int p = 0; p < nbeta; p++) {
      const cplx* bp = H->beta + (size_t)npw * p;
      const cplx* xb = X + (size_t)npw * b;
      double s = 0.0;
      for (int i = 0; i < npw; i++) s += creal(bp[i]) * creal(xb[i]) + cimag(bp[i]) * cimag(xb[i]);
      W[p + (size_t)nbeta * b] = s * H->wgt[p];
    }
  /* VnlX = beta * W (blas.cpp:1031-1067) */
#pragma omp parallel for num_threads(nthr(H))
  for (int b = 0; b < nb; b++) {
    cplx* vb = V + (size_t)npw * b;
    for (int i = 0; i < npw; i++) vb[i] = 0;
    for (int p = 0; p < nbeta; p++) {
      const double w = W[p + (size_t)nbeta * b];
      const cplx* bp = H->beta + (size_t)npw * p;
      for (int i = 0; i < npw; i++) vb[i] += bp[i] * w;
    }
  }
  free(W);
  return 0;
}

int or_apply_h(or_ham* H, int nb, const double* X, double* HX) {
  const size_t n2 = (size_t)2 * H->npw * nb;
  double* t = (double*)malloc(sizeof(double) * n2);
  int rc = or_mult_spinor(H, nb, X, HX);
  if (rc == 0) rc = or_nonlocal_multx(H, nb, X, t);
  for (size_t i = 0; i < n2; i++) HX[i] = t[i] + HX[i]; /* AlltoallForwardAdd */
  free(t);
  return rc;
}

int or_teter_precond(or_ham* H, int nb, const double* psid, double* a3d) {
  const cplx* psi = (const cplx*)psid;
  cplx* a3 = (cplx*)a3d;
  for (int k = 0; k < nb; k++)
    for (int i = 0; i < H->npw; i++) a3[(size_t)H->npw * k + i] += psi[(size_t)H->npw * k + i] * H->teter[i];
  return 0;
}

/* ------------------------------ EXX ------------------------------ */

int or_ham_set_exx(or_ham* H, double ecut, double mu, double exx_fraction, int divtype, int nocc,
                   const double* psi_occ, const double* occ, double* exx_div_out) {
  const int n1 = H->n[0], n2 = H->n[1], n3 = H->n[2];
  const double epsDiv = 1e-8;
  free(H->fock_idx); free(H->exx_kernel); free(H->phi_rs); free(H->occ);
  free(H->xi); /* a new exchange operator drops the compressed one */
  H->xi = NULL;
  H->nace = 0;
  H->fock_idx = (int*)malloc(sizeof(int) * H->Nh);
  double* gf = (double*)malloc(sizeof(double) * H->Nh);
  H->nfock = or_fock_sphere_build(H->n, H->supercell, ecut, H->fock_idx, gf);
  const double alpha = 10.0 / (ecut * 2.0);
  double div = 0.0;
  if (divtype == 1) {
    for (int k = 0; k < n3; k++)
      for (int j = 0; j < n2; j++)
        for (int i = 0; i < n1; i++) {
          double g2 = gkk_of(H->recip, kgrid(i, n1), kgrid(j, n2), kgrid(k, n3)) * 2.0;
          if (g2 > epsDiv && g2 < ecut * 4.0) {
            if (mu > 0.0) div += exp(-alpha * g2) / g2 * (1.0 - exp(-g2 / (4.0 * mu * mu)));
            else div += exp(-alpha * g2) / g2;
          }
        }
    if (mu > 0.0) div += 1.0 / (4.0 * mu * mu);
    else div -= alpha;
    div *= 4.0 * PI;
    double aa;
    if (mu > 0.0) aa = 1.0 / sqrt(alpha * PI) - 1.0 / sqrt(PI * (alpha + 1.0 / (4.0 * mu * mu)));
    else aa = 1.0 / sqrt(alpha * PI);
    div -= H->vol * aa;
  }
  if (exx_div_out) *exx_div_out = div;
  H->exx_kernel = (double*)malloc(sizeof(double) * H->nfock);
  for (int ig = 0; ig < H->nfock; ig++) {
    double g2 = gf[ig] * 2.0;
    if (g2 > epsDiv) {
      if (mu > 0) H->exx_kernel[ig] = 4.0 * PI / g2 * (1.0 - exp(-g2 / (4.0 * mu * mu)));
      else H->exx_kernel[ig] = 4.0 * PI / g2;
    } else {
      H->exx_kernel[ig] = -div;
      if (mu > 0) H->exx_kernel[ig] += PI / (mu * mu);
    }
  }
  free(gf);
  H->exx_fraction = exx_fraction;
  H->nocc = nocc;
  H->occ = (double*)malloc(sizeof(double) * (nocc > 0 ? nocc : 1));
  memcpy(H->occ, occ, sizeof(double) * nocc);
  /* phi = psi * sqrt(1/(vol N)) (hamiltonian.cpp:3573), then raw C2R (spinor.cpp:846-864) */
  H->phi_rs = (double*)malloc(sizeof(double) * (size_t)H->N * (nocc > 0 ? nocc : 1));
  const double fac = sqrt(1.0 / H->vol / (double)H->N);
  void* P = or_plan3_new(H->n);
  cplx* half = (cplx*)malloc(sizeof(cplx) * (size_t)H->Nh);
  cplx* c = (cplx*)malloc(sizeof(cplx) * (size_t)H->npw);
  for (int j = 0; j < nocc; j++) {
    const cplx* pj = (const cplx*)psi_occ + (size_t)H->npw * j;
    for (int i = 0; i < H->npw; i++) c[i] = pj[i] * fac;
    scatter_sphere(H, c, half);
    or_plan3_c2r(P, (double*)half, H->phi_rs + (size_t)H->N * j);
  }
  free(c);
  free(half);
  or_plan3_delete(P);
  return 0;
}

int or_add_exx(or_ham* H, int nb, const double* psid, double* a3d) {
  const cplx* psi = (const cplx*)psid;
  cplx* a3 = (cplx*)a3d;
  const int N = H->N, Nh = H->Nh, npw = H->npw;
  const double vol = H->vol;
#pragma omp parallel num_threads(nthr(H))
  {
    void* P = or_plan3_new(H->n);
    cplx* half = (cplx*)malloc(sizeof(cplx) * (size_t)Nh);
    cplx* prod = (cplx*)malloc(sizeof(cplx) * (size_t)Nh);
    double* psr = (double*)malloc(sizeof(double) * (size_t)N);
    double* r = (double*)malloc(sizeof(double) * (size_t)N);
#pragma omp for schedule(dynamic, 1)
    for (int k = 0; k < nb; k++) {
      scatter_sphere(H, psi + (size_t)npw * k, half);
      or_plan3_c2r(P, (double*)half, psr);
      cplx* out = a3 + (size_t)npw * k;
      for (int j = 0; j < H->nocc; j++) {
        const double* ph = H->phi_rs + (size_t)N * j;
        for (int i = 0; i < N; i++) r[i] = psr[i] * ph[i];
        or_plan3_r2c(P, r, (double*)prod);
        const double ff = vol / (double)N; /* FFTWExecute forward */
        memset(half, 0, sizeof(cplx) * (size_t)Nh);
        for (int ig = 0; ig < H->nfock; ig++) {
          int id = H->fock_idx[ig];
          half[id] = prod[id] * ff * H->exx_kernel[ig];
        }
        or_plan3_c2r(P, (double*)half, r);
        const double fb = 1.0 / vol; /* FFTWExecute backward */
        const double fac = -H->exx_fraction * H->occ[j];
        for (int i = 0; i < N; i++) r[i] = (r[i] * fb) * (ph[i] * fac);
        or_plan3_r2c(P, r, (double*)half);
        gather_sphere_add(H, half, out);
      }
    }
    free(half); free(prod); free(psr); free(r);
    or_plan3_delete(P);
  }
  return 0;
}

/* KohnSham::CalculateVexxACE, sphere branch (hamiltonian.cpp:3685-3924):
 * W = V_x psi (AddMultSpinorEXX into a zero block), M = -Re(psi^H W),
 * Potrf('L') M = L L^T, xi = W L^{-T} (Trsm 'R','L','T','N').  With
 * L = U^T this is the upper factor M = U^T U and xi = W U^{-1}.  nb = 0
 * drops the operator.  Returns -4 when M is not positive definite. */
int or_ham_build_ace(or_ham* H, int nb, const double* psi) {
  free(H->xi);
  H->xi = NULL;
  H->nace = 0;
  if (nb == 0) return 0;
  if (H->nocc <= 0) return -3;
  const int m = 2 * H->npw;
  double* W = (double*)calloc((size_t)m * nb, sizeof(double));
  double* M = (double*)malloc(sizeof(double) * (size_t)nb * nb);
  or_add_exx(H, nb, psi, W);
  or_gemm_tn(m, nb, nb, psi, W, M);
  for (size_t i = 0; i < (size_t)nb * nb; i++) M[i] = -M[i];
  int info = or_potrf_upper(nb, M, nb);
  if (info != 0) {
    free(W);
    free(M);
    return -4;
  }
  or_trsm_right_upper(m, nb, M, nb, W, m);
  free(M);
  H->xi = W;
  H->nace = nb;
  return 0;
}

/* ------------------------- density / forces ------------------------- */

int or_density(or_ham* H, int nb, const double* psid, const double* occ, double num_spin,
               double* rho) {
  const cplx* psi = (const cplx*)psid;
  const int N = H->N;
  memset(rho, 0, sizeof(double) * N);
  void* P = or_plan3_new(H->n);
  cplx* half = (cplx*)malloc(sizeof(cplx) * (size_t)H->Nh);
  double* r = (double*)malloc(sizeof(double) * (size_t)N);
  const double fac = 1.0 / sqrt((double)N);
  for (int k = 0; k < nb; k++) {
    scatter_sphere(H, psi + (size_t)H->npw * k, half);
    or_plan3_c2r(P, (double*)half, r);
    const double w = num_spin * occ[k];
    for (int i = 0; i < N; i++) {
      double v = r[i] * fac;
      rho[i] += v * v * w;
    }
  }
  free(half); free(r);
  or_plan3_delete(P);
  return 0;
}

double or_band_energy(or_ham* H, int nb, const double* psi, const double* f) {
  const size_t m = (size_t)2 * H->npw;
  double* hx = (double*)malloc(sizeof(double) * m * nb);
  or_apply_h(H, nb, psi, hx);
  double e = 0.0;
  for (int k = 0; k < nb; k++) {
    double s = 0.0;
    for (size_t i = 0; i < m; i++) s += psi[m * k + i] * hx[m * k + i];
    e += f[k] * s;
  }
  free(hx);
  return e;
}

int or_nonlocal_force(or_ham* H, int nb, const double* psid, const double* f, int natom, int npa,
                      double* force) {
  const cplx* psi = (const cplx*)psid;
  const int npw = H->npw;
  if (natom * npa != H->nbeta) return -1;
  /* G vectors of the sphere */
  double* G = (double*)malloc(sizeof(double) * 3 * npw);
  const int n1 = H->n[0], n2 = H->n[1], n3 = H->n[2], h1 = H->h1;
  for (int i = 0; i < npw; i++) {
    int g = H->idx[i];
    int kx = g % h1, ky = (g / h1) % n2, kz = g / (h1 * n2);
    int m0 = kgrid(kx, n1), m1 = kgrid(ky, n2), m2 = kgrid(kz, n3);
    for (int d = 0; d < 3; d++)
      G[3 * i + d] = H->recip[d] * m0 + H->recip[3 + d] * m1 + H->recip[6 + d] * m2;
  }
  memset(force, 0, sizeof(double) * 3 * natom);
  for (int a = 0; a < natom; a++)
    for (int q = 0; q < npa; q++) {
      int p = a * npa + q;
      const cplx* bp = H->beta + (size_t)npw * p;
      for (int k = 0; k < nb; k++) {
        const cplx* c = psi + (size_t)npw * k;
        double v = 0.0, d[3] = {0, 0, 0};
        for (int i = 0; i < npw; i++) {
          cplx t = conj(bp[i]) * c[i];
          v += creal(t);
          /* <b| d psi/dx_d> with d psi/dx (G) = i G psi(G) */
          for (int dd = 0; dd < 3; dd++) d[dd] += creal(t * (I * G[3 * i + dd]));
        }
        for (int dd = 0; dd < 3; dd++) force[3 * a + dd] += -2.0 * f[k] * H->wgt[p] * v * d[dd];
      }
    }
  free(G);
  return 0;
}

int or_ham_npw(const or_ham* H) { return H->npw; }
