/*
 * or_rs.c -- FP64 CPU restatement of the real-space dual-grid Gamma-point
 * H*psi (Use_Real_Space = 1, FFT_Number_Type "even"/"power").
 *
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 *
 *   or_ip_c2f / or_ip_f2c      IP_c2f / IP_f2c, src/utility.cpp:3150-3212 and
 *                              :3215-3276 (the scatter and the gather, in the
 *                              reference's branch order)
 *   or_rs_mult_spinor          KohnSham::MultSpinor(Spinor&, NumTns<Real>&,
 *                              Fourier&), src/SCF/hamiltonian.cpp:3058-3125,
 *                              with Spinor::AddMultSpinorFineR2C (real a3),
 *                              src/SCF/spinor.cpp:376-525
 *   or_rs_filter               the wavefunction ecut filter of that
 *                              MultSpinor (hamiltonian.cpp:3090-3116,
 *                              3240-3265)
 *   or_rs_density              KohnSham::CalculateDensity, real-space
 *                              branch (hamiltonian.cpp:2453-2627)
 *   or_rs_teter_precond        Spinor::AddTeterPrecond, real version
 *                              (spinor.cpp:309-345)
 *   or_rs_ppcg                 EigenSolver::PPCGSolveReal, real-space branch
 *                              (ppcg.cpp:683-1246): the PPCG iteration of
 *                              or_eig.c over the operators above
 *
 * Scalings are FFTWExecute's (src/SCF/fourier.cpp:695-760): forward
 * transforms are multiplied by vol/N, backward ones by 1/vol.
 *
 * Real-space projectors are grouped per atom, as the reference builds them:
 * every projector of one atom is a SparseVec on the same index list (the
 * sphere of periodtable.cpp:1055-1101, shared by each SparseVec(iv, dv) at
 * :1155-1589); only the VAL column of dv enters H*psi.  Atom a owns
 * npts[a] grid indices idx[...] and nproj[a] projectors; val holds, per atom,
 * an npts x nproj column-major block; wgt one weight per projector.
 */
#include <stdlib.h>
#include <string.h>
#include "oracle.h"

typedef struct { double re, im; } zc;

/* IP_c2f (utility.cpp:3150-3212): coarse half spectrum -> fine half spectrum
 * (tildef must hold zeros where nothing is written, as SetValue leaves it). */
void or_ip_c2f(const int* Nc, const int* Nf, const double* tilde_d, double* tildef_d) {
  const zc* tilde = (const zc*)tilde_d;
  zc* tildef = (zc*)tildef_d;
  const int hc = Nc[0] / 2 + 1, hf = Nf[0] / 2 + 1;
  const int cjF = Nf[1] - Nc[1] / 2, ckF = Nf[2] - Nc[2] / 2;
  for (int k = 0; k < Nc[2]; k++)
    for (int j = 0; j < Nc[1]; j++)
      for (int i = 0; i < hc; i++) {
        const long ptrc = i + (long)hc * (j + (long)Nc[1] * k);
        const int iF = i;
        int jF = (j <= Nc[1] / 2) ? j : Nf[1] - Nc[1] + j;
        int kF = (k <= Nc[2] / 2) ? k : Nf[2] - Nc[2] + k;
        long ptrF = iF + (long)hf * (jF + (long)Nf[1] * kF);
        const zc t = tilde[ptrc];
        const zc h = {t.re / 2.0, t.im / 2.0};
        tildef[ptrF] = t;
        const int ni = (iF == Nc[0] / 2), nj = (jF == Nc[1] / 2), nk = (kF == Nc[2] / 2);
        if (ni && nj && nk) {
          tildef[ptrF] = h;
        } else if (ni && nj) {
          tildef[ptrF] = h;
        } else if (ni && nk) {
          tildef[ptrF] = h;
        } else if (nj && nk) {
          tildef[ptrF] = h;
          ptrF = iF + (long)hf * (cjF + (long)Nf[1] * ckF);
          tildef[ptrF] = h;
        } else if (ni) {
          tildef[ptrF] = h;
        } else if (nj) {
          tildef[ptrF] = h;
          ptrF = iF + (long)hf * (cjF + (long)Nf[1] * kF);
          tildef[ptrF] = h;
        } else if (nk) {
          tildef[ptrF] = h;
          ptrF = iF + (long)hf * (jF + (long)Nf[1] * ckF);
          tildef[ptrF] = h;
        }
      }
}

/* IP_f2c (utility.cpp:3215-3276): tilde += restriction of tildef. */
void or_ip_f2c(const int* Nc, const int* Nf, double* tilde_d, const double* tildef_d) {
  zc* tilde = (zc*)tilde_d;
  const zc* tildef = (const zc*)tildef_d;
  const int hc = Nc[0] / 2 + 1, hf = Nf[0] / 2 + 1;
  const int cjF = Nf[1] - Nc[1] / 2, ckF = Nf[2] - Nc[2] / 2;
#define FIDX(a, b, c) ((a) + (long)hf * ((b) + (long)Nf[1] * (c)))
  for (int k = 0; k < Nc[2]; k++)
    for (int j = 0; j < Nc[1]; j++)
      for (int i = 0; i < hc; i++) {
        const long ptrc = i + (long)hc * (j + (long)Nc[1] * k);
        const int iF = i;
        int jF = (j <= Nc[1] / 2) ? j : Nf[1] - Nc[1] + j;
        int kF = (k <= Nc[2] / 2) ? k : Nf[2] - Nc[2] + k;
        const zc f1 = tildef[FIDX(iF, jF, kF)];
        zc f2, add;
        const int ni = (i == Nc[0] / 2), nj = (j == Nc[1] / 2), nk = (k == Nc[2] / 2);
        if (ni && nj && nk) {
          add.re = f1.re;
          add.im = 0.0;
        } else if (ni && nj) {
          kF = (Nf[2] - kF) % Nf[2];
          f2 = tildef[FIDX(iF, jF, kF)];
          add.re = 0.5 * (f1.re + f2.re);
          add.im = 0.5 * (f1.im - f2.im);
        } else if (ni && nk) {
          jF = (Nf[1] - jF) % Nf[1];
          f2 = tildef[FIDX(iF, jF, kF)];
          add.re = 0.5 * (f1.re + f2.re);
          add.im = 0.5 * (f1.im - f2.im);
        } else if (nj && nk) {
          f2 = tildef[FIDX(iF, cjF, ckF)];
          add.re = 0.5 * (f1.re + f2.re);
          add.im = 0.5 * (f1.im + f2.im);
        } else if (ni) {
          jF = (Nf[1] - jF) % Nf[1];
          kF = (Nf[2] - kF) % Nf[2];
          f2 = tildef[FIDX(iF, jF, kF)];
          add.re = 0.5 * (f1.re + f2.re);
          add.im = 0.5 * (f1.im - f2.im);
        } else if (nj) {
          f2 = tildef[FIDX(iF, cjF, kF)];
          add.re = 0.5 * (f1.re + f2.re);
          add.im = 0.5 * (f1.im + f2.im);
        } else if (nk) {
          f2 = tildef[FIDX(iF, jF, ckF)];
          add.re = 0.5 * (f1.re + f2.re);
          add.im = 0.5 * (f1.im + f2.im);
        } else {
          add = f1;
        }
        tilde[ptrc].re += add.re;
        tilde[ptrc].im += add.im;
      }
#undef FIDX
}

/* a3 = H psi for nb real bands on the coarse grid (ntot = Nc0*Nc1*Nc2 each,
 * band-major).  ekin on the coarse R2C grid, vtot on the fine grid. */
int or_rs_mult_spinor(const int* Nc, const int* Nf, double vol, const double* ekin,
                      const double* vtot, int natom, const int* npts, const int* nproj,
                      const int* idx, const double* val, const double* wgt, int nb,
                      const double* psi, double* a3) {
  const long ntot = (long)Nc[0] * Nc[1] * Nc[2], ntotF = (long)Nf[0] * Nf[1] * Nf[2];
  const long nr2c = (long)(Nc[0] / 2 + 1) * Nc[1] * Nc[2];
  const long nr2cF = (long)(Nf[0] / 2 + 1) * Nf[1] * Nf[2];
  double* outc = malloc(sizeof(double) * 2 * nr2c);
  double* outf = malloc(sizeof(double) * 2 * nr2cF);
  double* psiF = malloc(sizeof(double) * ntotF);
  double* upd = malloc(sizeof(double) * ntotF);
  double* tmp = malloc(sizeof(double) * ntot);
  if (!outc || !outf || !psiF || !upd || !tmp) {
    free(outc); free(outf); free(psiF); free(upd); free(tmp);
    return -1;
  }
  int rc = 0;
  for (int b = 0; b < nb && rc == 0; b++) {
    const double* x = psi + (long)b * ntot;
    /* forwardPlanR2C, scaled by vol/ntot */
    rc |= or_fft_r2c(Nc, x, outc);
    for (long g = 0; g < 2 * nr2c; g++) outc[g] *= vol / (double)ntot;
    /* IP_c2f into a zeroed fine spectrum, backwardPlanR2CFine scaled by 1/vol */
    memset(outf, 0, sizeof(double) * 2 * nr2cF);
    or_ip_c2f(Nc, Nf, outc, outf);
    rc |= or_fft_c2r(Nf, outf, psiF);
    for (long r = 0; r < ntotF; r++) psiF[r] *= 1.0 / vol;
    /* local potential */
    for (long r = 0; r < ntotF; r++) upd[r] = psiF[r] * vtot[r];
    /* nonlocal, atom by atom, projector by projector */
    long po = 0, vo = 0, wo = 0;
    for (int a = 0; a < natom; a++) {
      for (int p = 0; p < nproj[a]; p++) {
        const double* dv = val + vo + (long)p * npts[a];
        double weight = 0.0;
        for (int i = 0; i < npts[a]; i++) weight += dv[i] * psiF[idx[po + i]];
        weight *= vol / (double)ntotF * wgt[wo + p];
        for (int i = 0; i < npts[a]; i++) upd[idx[po + i]] += dv[i] * weight;
      }
      po += npts[a];
      vo += (long)npts[a] * nproj[a];
      wo += nproj[a];
    }
    /* kinetic term on the coarse spectrum */
    for (long g = 0; g < nr2c; g++) {
      outc[2 * g] *= ekin[g];
      outc[2 * g + 1] *= ekin[g];
    }
    /* forwardPlanR2CFine scaled by vol/ntotFine, IP_f2c, backwardPlanR2C */
    rc |= or_fft_r2c(Nf, upd, outf);
    for (long g = 0; g < 2 * nr2cF; g++) outf[g] *= vol / (double)ntotF;
    or_ip_f2c(Nc, Nf, outc, outf);
    rc |= or_fft_c2r(Nc, outc, tmp);
    double* y = a3 + (long)b * ntot;
    for (long r = 0; r < ntot; r++) y[r] = tmp[r] * (1.0 / vol);
  }
  free(outc); free(outf); free(psiF); free(upd); free(tmp);
  return rc;
}

/* Wavefunction ecut filter of the real MultSpinor (hamiltonian.cpp:3090-3116
 * on psi, :3240-3265 on a3): R2C (x vol/ntot), zero every entry with
 * gkkR2C > cutoff, C2R (x 1/vol), in place on nb bands. */
int or_rs_filter(const int* Nc, double vol, const double* gkk, double cutoff, int nb, double* x) {
  const long ntot = (long)Nc[0] * Nc[1] * Nc[2];
  const long nr2c = (long)(Nc[0] / 2 + 1) * Nc[1] * Nc[2];
  double* outc = malloc(sizeof(double) * 2 * nr2c);
  double* tmp = malloc(sizeof(double) * ntot);
  if (!outc || !tmp) {
    free(outc); free(tmp);
    return -1;
  }
  int rc = 0;
  for (int b = 0; b < nb && rc == 0; b++) {
    double* y = x + (long)b * ntot;
    rc |= or_fft_r2c(Nc, y, outc);
    for (long g = 0; g < nr2c; g++) {
      const double f = gkk[g] > cutoff ? 0.0 : vol / (double)ntot;
      outc[2 * g] *= f;
      outc[2 * g + 1] *= f;
    }
    rc |= or_fft_c2r(Nc, outc, tmp);
    for (long r = 0; r < ntot; r++) y[r] = tmp[r] * (1.0 / vol);
  }
  free(outc); free(tmp);
  return rc;
}

/* KohnSham::CalculateDensity, real-space branch, one density component
 * (hamiltonian.cpp:2453-2627): rho = sum_b num_spin occ_b psiF_b^2 on the
 * fine grid, then rho *= num_spin num_occ Nf / (vol sum rho) and *val =
 * sum(rho) vol / Nf.  The reference interpolates with a C2C transform and
 * the idxFineGrid padding (:2526-2534) and squares the real part; that real
 * part is sqrt(Nc/Nf) times the IP_c2f interpolant used here, a factor the
 * renormalisation removes. */
int or_rs_density(const int* Nc, const int* Nf, double vol, int nb, const double* psi, const double* occ,
                  int num_spin, int num_occ, double* rho, double* val) {
  const long ntot = (long)Nc[0] * Nc[1] * Nc[2], ntotF = (long)Nf[0] * Nf[1] * Nf[2];
  const long nr2c = (long)(Nc[0] / 2 + 1) * Nc[1] * Nc[2];
  const long nr2cF = (long)(Nf[0] / 2 + 1) * Nf[1] * Nf[2];
  double* outc = malloc(sizeof(double) * 2 * nr2c);
  double* outf = malloc(sizeof(double) * 2 * nr2cF);
  double* psiF = malloc(sizeof(double) * ntotF);
  if (!outc || !outf || !psiF) {
    free(outc); free(outf); free(psiF);
    return -1;
  }
  int rc = 0;
  memset(rho, 0, sizeof(double) * ntotF);
  for (int b = 0; b < nb && rc == 0; b++) {
    rc |= or_fft_r2c(Nc, psi + (long)b * ntot, outc);
    for (long g = 0; g < 2 * nr2c; g++) outc[g] *= vol / (double)ntot;
    memset(outf, 0, sizeof(double) * 2 * nr2cF);
    or_ip_c2f(Nc, Nf, outc, outf);
    rc |= or_fft_c2r(Nf, outf, psiF);
    const double fac = num_spin * occ[b];
    for (long r = 0; r < ntotF; r++) {
      const double v = psiF[r] * (1.0 / vol);
      rho[r] += v * v * fac;
    }
  }
  double s = 0.0;
  for (long r = 0; r < ntotF; r++) s += rho[r];
  const double scale = (num_spin * (double)num_occ * (double)ntotF) / (vol * s);
  double s2 = 0.0;
  for (long r = 0; r < ntotF; r++) {
    rho[r] *= scale;
    s2 += rho[r] * vol / ntotF;
  }
  if (val) *val = s2;
  free(outc); free(outf); free(psiF);
  return rc;
}

/* Spinor::AddTeterPrecond, real version (spinor.cpp:309-345): per band
 * R2C (x vol/ntot), times teter on the R2C grid, C2R (x 1/vol), a3 += . */
int or_rs_teter_precond(const int* Nc, double vol, const double* teter, int nb, const double* psi, double* a3) {
  const long ntot = (long)Nc[0] * Nc[1] * Nc[2];
  const long nr2c = (long)(Nc[0] / 2 + 1) * Nc[1] * Nc[2];
  double* outc = malloc(sizeof(double) * 2 * nr2c);
  double* tmp = malloc(sizeof(double) * ntot);
  if (!outc || !tmp) {
    free(outc); free(tmp);
    return -1;
  }
  int rc = 0;
  for (int b = 0; b < nb && rc == 0; b++) {
    rc |= or_fft_r2c(Nc, psi + (long)b * ntot, outc);
    for (long g = 0; g < nr2c; g++) {
      const double f = vol / (double)ntot * teter[g];
      outc[2 * g] *= f;
      outc[2 * g + 1] *= f;
    }
    rc |= or_fft_c2r(Nc, outc, tmp);
    double* y = a3 + (long)b * ntot;
    for (long r = 0; r < ntot; r++) y[r] += tmp[r] * (1.0 / vol);
  }
  free(outc); free(tmp);
  return rc;
}

typedef struct {
  const int *Nc, *Nf;
  double vol;
  const double *ekin, *vtot, *teter;
  int natom;
  const int *npts, *nproj, *idx;
  const double *val, *wgt;
} rs_problem;

static void rs_hop(void* ctx, int nb, const double* x, double* y) {
  const rs_problem* p = (const rs_problem*)ctx;
  or_rs_mult_spinor(p->Nc, p->Nf, p->vol, p->ekin, p->vtot, p->natom, p->npts, p->nproj, p->idx, p->val, p->wgt,
                    nb, x, y);
}

static void rs_pop(void* ctx, int nb, const double* r, double* w) {
  const rs_problem* p = (const rs_problem*)ctx;
  or_rs_teter_precond(p->Nc, p->vol, p->teter, nb, r, w);
}

int or_rs_ppcg(const int* Nc, const int* Nf, double vol, const double* ekin, const double* vtot, int natom,
               const int* npts, const int* nproj, const int* idx, const double* val, const double* wgt,
               int nb, int numEig, int scfIter, int eigMaxIter, double eigTolerance, int sbSize, double* psi,
               double* eigval, double* resval, int* iters) {
  const long nr2c = (long)(Nc[0] / 2 + 1) * Nc[1] * Nc[2];
  double* teter = malloc(sizeof(double) * nr2c);
  if (!teter) return -1;
  /* KohnSham::CalculateEkin's teter_ (hamiltonian.cpp:2659-2663) */
  for (long g = 0; g < nr2c; g++) {
    const double a = ekin[g] * 2.0, b = 27.0 + a * (18.0 + a * (12.0 + a * 8.0));
    teter[g] = b / (b + 16.0 * a * a * a * a);
  }
  rs_problem p = {Nc, Nf, vol, ekin, vtot, teter, natom, npts, nproj, idx, val, wgt};
  const long ntot = (long)Nc[0] * Nc[1] * Nc[2];
  const int rc = or_ppcg_op((int)ntot, rs_hop, rs_pop, &p, nb, numEig, scfIter, eigMaxIter, eigTolerance, sbSize,
                            psi, eigval, resval, iters);
  free(teter);
  return rc;
}
