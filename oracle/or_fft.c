/*
 * or_fft.c -- FP64 FFTs for the oracle (test infrastructure only).
 *
 * Replaces the FFTW 3.3.10 calls of the reference (fftw_plan_dft_{r2c,c2r}_3d
 * with dims (n3,n2,n1), fourier.cpp:169-184; executed at spinor.cpp:614-646,
 * 843-930): unnormalised transforms, forward = exp(-i 2 pi k x / n),
 * backward = exp(+i ...).  The 1D engine is a Stockham autosort mixed-radix
 * FFT (radix 4, 2, 3, 5 kernels plus a direct DFT for other primes); roots
 * of unity are computed directly in long double, never by recurrence.
 *
 * c2r follows FFTW's multi-dimensional real-data layout: complex transforms
 * along z and y on the h1 = n1/2+1 half, then a Hermitian-extended complex
 * transform along x whose real part is kept (the imaginary parts of the
 * kx = 0 and kx = n1/2 entries are dropped, as a halfcomplex c2r does).
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

typedef double _Complex cplx;

typedef struct {
  int n, nf;
  int f[64];
  cplx* w; /* w[m] = exp(+2 pi i m / n) */
} plan1;

static void plan1_init(plan1* p, int n) {
  p->n = n;
  p->nf = 0;
  int m = n;
  while (m % 4 == 0) { p->f[p->nf++] = 4; m /= 4; }
  while (m % 2 == 0) { p->f[p->nf++] = 2; m /= 2; }
  while (m % 3 == 0) { p->f[p->nf++] = 3; m /= 3; }
  while (m % 5 == 0) { p->f[p->nf++] = 5; m /= 5; }
  for (int q = 7; m > 1; q += 2)
    while (m % q == 0) { p->f[p->nf++] = q; m /= q; }
  p->w = (cplx*)malloc(sizeof(cplx) * (size_t)n);
  for (int k = 0; k < n; k++) {
    long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)k / (long double)n;
    p->w[k] = (double)cosl(a) + I * (double)sinl(a);
  }
}

static void plan1_free(plan1* p) { free(p->w); p->w = NULL; }

static inline cplx root(const plan1* p, int sign, long m) {
  cplx v = p->w[m % p->n];
  return sign > 0 ? v : conj(v);
}

/* x: in/out (n), y: scratch (n).  sign = -1 forward, +1 backward. */
static void plan1_exec(const plan1* p, cplx* x, cplx* y, int sign) {
  const int n = p->n;
  if (n == 1) return;
  cplx *src = x, *dst = y;
  int Ns = 1;
  cplx v[64], t[64];
  for (int s = 0; s < p->nf; s++) {
    const int R = p->f[s];
    const int m = n / R;
    const int tstep = n / (Ns * R); /* root index step for exp(2 pi i /(Ns R)) */
    for (int j = 0; j < m; j++) {
      const int k = j % Ns;
      for (int r = 0; r < R; r++) v[r] = src[j + r * m];
      for (int r = 1; r < R; r++) v[r] *= root(p, sign, (long)r * k * tstep);
      const double sg = (double)sign;
      if (R == 2) {
        cplx a = v[0], b = v[1];
        v[0] = a + b; v[1] = a - b;
      } else if (R == 4) {
        cplx t0 = v[0] + v[2], t1 = v[0] - v[2];
        cplx t2 = v[1] + v[3], t3 = (v[1] - v[3]) * (sg * I);
        v[0] = t0 + t2; v[1] = t1 + t3; v[2] = t0 - t2; v[3] = t1 - t3;
      } else if (R == 3) {
        const double h = 0.86602540378443864676372317075294; /* sqrt(3)/2 */
        cplx a = v[0], b = v[1], c = v[2];
        cplx tt = a - 0.5 * (b + c);
        cplx u = (sg * h * I) * (b - c);
        v[0] = a + b + c; v[1] = tt + u; v[2] = tt - u;
      } else if (R == 5) {
        const double c1 = 0.30901699437494742410229341718282;  /* cos(2pi/5) */
        const double c2 = -0.80901699437494742410229341718282; /* cos(4pi/5) */
        const double s1 = 0.95105651629515357211643933337938;  /* sin(2pi/5) */
        const double s2 = 0.58778525229247312916870595463907;  /* sin(4pi/5) */
        cplx a = v[0];
        cplx t1 = v[1] + v[4], t2 = v[2] + v[3], t3 = v[1] - v[4], t4 = v[2] - v[3];
        cplx m1 = a + c1 * t1 + c2 * t2, m2 = a + c2 * t1 + c1 * t2;
        cplx n1 = (sg * I) * (s1 * t3 + s2 * t4), n2 = (sg * I) * (s2 * t3 - s1 * t4);
        v[0] = a + t1 + t2; v[1] = m1 + n1; v[4] = m1 - n1; v[2] = m2 + n2; v[3] = m2 - n2;
      } else {
        /* direct DFT of prime length R: exp(sign 2 pi i r q / R) = w[(r q mod R) * n/R] */
        for (int q = 0; q < R; q++) {
          cplx acc = 0;
          for (int r = 0; r < R; r++) acc += v[r] * root(p, sign, (long)((r * q) % R) * (n / R));
          t[q] = acc;
        }
        for (int q = 0; q < R; q++) v[q] = t[q];
      }
      const int base = (j / Ns) * Ns * R + k;
      for (int r = 0; r < R; r++) dst[base + r * Ns] = v[r];
    }
    cplx* tmp = src; src = dst; dst = tmp;
    Ns *= R;
  }
  if (src != x) memcpy(x, src, sizeof(cplx) * (size_t)n);
}

int or_fft1d(int n, int sign, double* xd) {
  if (n <= 0) return -1;
  plan1 p;
  plan1_init(&p, n);
  cplx* y = (cplx*)malloc(sizeof(cplx) * (size_t)n);
  plan1_exec(&p, (cplx*)xd, y, sign);
  free(y);
  plan1_free(&p);
  return 0;
}

/* ---- 3D real transforms, with plans cached per call ---- */

typedef struct {
  int n1, n2, n3, h1;
  plan1 p1, p2, p3;
  cplx *line, *scr;
} plan3;

static void plan3_init(plan3* P, const int* n) {
  P->n1 = n[0]; P->n2 = n[1]; P->n3 = n[2]; P->h1 = n[0] / 2 + 1;
  plan1_init(&P->p1, n[0]);
  plan1_init(&P->p2, n[1]);
  plan1_init(&P->p3, n[2]);
  int mx = n[0] > n[1] ? n[0] : n[1];
  mx = mx > n[2] ? mx : n[2];
  P->line = (cplx*)malloc(sizeof(cplx) * (size_t)mx);
  P->scr = (cplx*)malloc(sizeof(cplx) * (size_t)mx);
}

static void plan3_free(plan3* P) {
  plan1_free(&P->p1); plan1_free(&P->p2); plan1_free(&P->p3);
  free(P->line); free(P->scr);
}

/* half-array work (h1*n2*n3) -> z then y transforms with sign, in place */
static void yz_pass(plan3* P, cplx* w, int sign, int z_first) {
  const int h1 = P->h1, n2 = P->n2, n3 = P->n3;
  for (int pass = 0; pass < 2; pass++) {
    int do_z = (pass == 0) == (z_first != 0);
    if (do_z) {
      for (int ky = 0; ky < n2; ky++)
        for (int kx = 0; kx < h1; kx++) {
          cplx* base = w + kx + (size_t)h1 * ky;
          for (int z = 0; z < n3; z++) P->line[z] = base[(size_t)h1 * n2 * z];
          plan1_exec(&P->p3, P->line, P->scr, sign);
          for (int z = 0; z < n3; z++) base[(size_t)h1 * n2 * z] = P->line[z];
        }
    } else {
      for (int z = 0; z < n3; z++)
        for (int kx = 0; kx < h1; kx++) {
          cplx* base = w + kx + (size_t)h1 * n2 * z;
          for (int y = 0; y < n2; y++) P->line[y] = base[(size_t)h1 * y];
          plan1_exec(&P->p2, P->line, P->scr, sign);
          for (int y = 0; y < n2; y++) base[(size_t)h1 * y] = P->line[y];
        }
    }
  }
}

/* work is h1*n2*n3 complex; it is consumed. */
static void c2r_exec(plan3* P, cplx* work, double* out) {
  const int n1 = P->n1, n2 = P->n2, n3 = P->n3, h1 = P->h1;
  yz_pass(P, work, +1, 1);
  for (int z = 0; z < n3; z++)
    for (int y = 0; y < n2; y++) {
      const cplx* row = work + (size_t)h1 * (y + (size_t)n2 * z);
      for (int k = 0; k < n1; k++) {
        if (k < h1) P->line[k] = row[k];
        else P->line[k] = conj(row[n1 - k]);
      }
      P->line[0] = creal(P->line[0]);
      if (n1 % 2 == 0) P->line[n1 / 2] = creal(P->line[n1 / 2]);
      plan1_exec(&P->p1, P->line, P->scr, +1);
      double* o = out + (size_t)n1 * (y + (size_t)n2 * z);
      for (int x = 0; x < n1; x++) o[x] = creal(P->line[x]);
    }
}

static void r2c_exec(plan3* P, const double* in, cplx* out) {
  const int n1 = P->n1, n2 = P->n2, n3 = P->n3, h1 = P->h1;
  for (int z = 0; z < n3; z++)
    for (int y = 0; y < n2; y++) {
      const double* r = in + (size_t)n1 * (y + (size_t)n2 * z);
      for (int x = 0; x < n1; x++) P->line[x] = r[x];
      plan1_exec(&P->p1, P->line, P->scr, -1);
      cplx* o = out + (size_t)h1 * (y + (size_t)n2 * z);
      for (int k = 0; k < h1; k++) o[k] = P->line[k];
    }
  yz_pass(P, out, -1, 0);
}

int or_fft_c2r(const int* n, const double* in, double* out) {
  plan3 P;
  plan3_init(&P, n);
  size_t nh = (size_t)P.h1 * n[1] * n[2];
  cplx* w = (cplx*)malloc(sizeof(cplx) * nh);
  memcpy(w, in, sizeof(cplx) * nh);
  c2r_exec(&P, w, out);
  free(w);
  plan3_free(&P);
  return 0;
}

int or_fft_r2c(const int* n, const double* in, double* out) {
  plan3 P;
  plan3_init(&P, n);
  r2c_exec(&P, in, (cplx*)out);
  plan3_free(&P);
  return 0;
}

/* ---- internal API used by the Hamiltonian (one plan per thread) ---- */
void* or_plan3_new(const int* n) {
  plan3* P = (plan3*)malloc(sizeof(plan3));
  plan3_init(P, n);
  return P;
}
void or_plan3_delete(void* P) {
  plan3_free((plan3*)P);
  free(P);
}
void or_plan3_c2r(void* P, double* work_half, double* out) { c2r_exec((plan3*)P, (cplx*)work_half, out); }
void or_plan3_r2c(void* P, const double* in, double* out_half) { r2c_exec((plan3*)P, in, (cplx*)out_half); }
