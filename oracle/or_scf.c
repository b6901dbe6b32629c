/* Harris-Foulkes free energy, FP64 CPU oracle (test infrastructure only: used
 * by tests/, never by the product library).
 *
 *   or_harris_energy   SCF::CalculateHarrisEnergy, src/SCF/scf.cpp:1987-2140
 *                      (the real, Gamma-point branch :2018-2025 and :2111-2133)
 *
 * Arrays follow the reference's DblNumMat(ntot, ncomp) storage, column-major:
 * density(i, RHO = 0 ..), spindensity(i, UP = 0 / DN = 1), vxc(i, is).
 * terms = {Exc, Eself, EIonSR, EVdw, Eext, fermi, Tbeta, temperature}.
 */
#include <math.h>

#include "oracle.h"

int or_harris_energy(int ntot, int nspin, double vol, const double* density, const double* spindensity,
                     const double* vxc, const double* vhart, const double* pseudo, int nstate, const double* eig,
                     const double* occ, int num_spin, int num_occ, const double* terms, double* efree) {
  if (ntot <= 0 || nstate < 0 || !(nspin == 1 || nspin == 2 || nspin == 4)) return 1;
  const double Exc = terms[0], Eself = terms[1], EIonSR = terms[2], EVdw = terms[3], Eext = terms[4];
  const double fermi = terms[5], Tbeta = terms[6], temperature = terms[7];
  /* band-energy sum (scf.cpp:2022-2025) */
  double Ekin = 0.0;
  for (int i = 0; i < nstate; i++) Ekin += num_spin * eig[i] * occ[i];
  /* double-counting terms of the input density (scf.cpp:2052-2077) */
  double EVxc = 0.0, Ehart = 0.0;
  if (nspin == 1 || nspin == 4) {
    for (int is = 0; is < nspin; is++)
      for (int i = 0; i < ntot; i++) EVxc += vxc[i + (long)ntot * is] * density[i + (long)ntot * is];
  } else {
    for (int i = 0; i < ntot; i++) {
      EVxc += vxc[i] * spindensity[i];
      EVxc += vxc[i + (long)ntot] * spindensity[i + (long)ntot];
    }
  }
  for (int i = 0; i < ntot; i++) Ehart += 0.5 * vhart[i] * (density[i] + pseudo[i]);
  Ehart *= vol / (double)ntot;
  EVxc *= vol / (double)ntot;
  const double Ecor = (Exc - EVxc) - Ehart - Eself + EIonSR + EVdw + Eext;
  double E;
  if (num_occ == nstate || temperature == 0.0) {
    E = Ekin + Ecor; /* zero temperature (scf.cpp:2081-2086) */
  } else {
    /* Fermi-Dirac free energy, split by the sign of eig - fermi so that the
     * exponent is never positive (scf.cpp:2121-2133) */
    E = 0.0;
    for (int k = 0; k < nstate; k++) {
      const double x = eig[k] - fermi;
      if (x >= 0)
        E += -num_spin / Tbeta * log(1.0 + exp(-Tbeta * x));
      else
        E += num_spin * x - num_spin / Tbeta * log(1.0 + exp(Tbeta * x));
    }
    E += Ecor + fermi * num_occ * num_spin;
  }
  *efree = E;
  return 0;
}
