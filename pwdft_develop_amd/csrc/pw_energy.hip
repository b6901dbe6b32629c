// Harris-Foulkes free energy of the eigensolver's output against the input
// density: SCF::CalculateHarrisEnergy (src/SCF/scf.cpp:1987-2140), the real
// Gamma-point branch.  The two grid integrals (the XC double-counting term
// and the Hartree energy of the input density) are FP64 dot products over
// the fine grid on the device; the band-energy and Fermi-Dirac sums over
// the eigenvalues and the scalar terms are evaluated on the host.
#include <cmath>

#include "pw_internal.h"

namespace pw {

constexpr int kHarrisBlocks = 512;

// Block b sums a fixed slice of the two integrands in a fixed order:
//   p[2 b]     = sum_e vxc[e] x[e],          e < nx (x = density components,
//                or spin densities for nspin = 2)
//   p[2 b + 1] = sum_i vhart[i] (rho[i] + pseudo[i]),  i < ntot
__global__ void __launch_bounds__(256)
k_harris_partial(const double* __restrict__ vxc, const double* __restrict__ x, long nx,
                 const double* __restrict__ vhart, const double* __restrict__ rho,
                 const double* __restrict__ pseudo, long ntot, double* __restrict__ p) {
  __shared__ double sred[2][4];
  const long stride = (long)gridDim.x * blockDim.x;
  double a = 0.0, h = 0.0;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < nx; e += stride) a += vxc[e] * x[e];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < ntot; i += stride)
    h += vhart[i] * (rho[i] + pseudo[i]);
  a = wave_sum(a);
  h = wave_sum(h);
  const int w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sred[0][w] = a;
    sred[1][w] = h;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    p[2 * blockIdx.x] = sred[0][0] + sred[0][1] + sred[0][2] + sred[0][3];
    p[2 * blockIdx.x + 1] = sred[1][0] + sred[1][1] + sred[1][2] + sred[1][3];
  }
}

// out[0..1] = sums of the nb block partials, in block order.
__global__ void k_harris_final(const double* __restrict__ p, int nb, double* __restrict__ out) {
  if (threadIdx.x < 2) {
    double s = 0.0;
    for (int b = 0; b < nb; b++) s += p[2 * b + threadIdx.x];
    out[threadIdx.x] = s;
  }
}

}  // namespace pw

using namespace pw;

extern "C" int pw_harris_energy(pw_ctx* c, int ntot, int nspin, double vol, const double* density,
                                const double* spindensity, const double* vxc, const double* vhart,
                                const double* pseudo_charge, int nstate, const double* eig, const double* occ,
                                int num_spin, int num_occupied, const double* terms, double* efree) {
  if (!c || !efree || !terms) return fail(PW_ERR_ARG, "pw_harris_energy: null argument");
  if (ntot <= 0 || nstate < 0 || !(nspin == 1 || nspin == 2 || nspin == 4))
    return fail(PW_ERR_ARG, "pw_harris_energy: ntot %d, nspin %d, nstate %d", ntot, nspin, nstate);
  if (!density || !vxc || !vhart || !pseudo_charge || (nspin == 2 && !spindensity) || (nstate > 0 && (!eig || !occ)))
    return fail(PW_ERR_ARG, "pw_harris_energy: null array");
  const double Exc = terms[0], Eself = terms[1], EIonSR = terms[2], EVdw = terms[3], Eext = terms[4];
  const double fermi = terms[5], Tbeta = terms[6], temperature = terms[7];
  // band-energy sum (scf.cpp:2022-2025)
  double Ekin = 0.0;
  for (int i = 0; i < nstate; i++) Ekin += num_spin * eig[i] * occ[i];
  // grid integrals of the input density (scf.cpp:2052-2077): nspin 1 or 4
  // pairs vxc with the density components, nspin 2 with the spin densities
  const double* x = nspin == 2 ? spindensity : density;
  double* d = nullptr;
  PW_HIP(hipMalloc((void**)&d, sizeof(double) * (2 * kHarrisBlocks + 2)));
  hipLaunchKernelGGL(k_harris_partial, dim3(kHarrisBlocks), dim3(256), 0, c->stream, vxc, x,
                     (long)ntot * nspin, vhart, density, pseudo_charge, (long)ntot, d);
  hipLaunchKernelGGL(k_harris_final, dim3(1), dim3(64), 0, c->stream, (const double*)d, kHarrisBlocks,
                     d + 2 * kHarrisBlocks);
  double sums[2] = {0.0, 0.0};
  hipError_t e = hipGetLastError();
  if (e == hipSuccess) e = hipMemcpyAsync(sums, d + 2 * kHarrisBlocks, sizeof(sums), hipMemcpyDeviceToHost, c->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(PW_ERR_HIP, "pw_harris_energy: %s", hipGetErrorString(e));
  const double EVxc = sums[0] * vol / (double)ntot;
  const double Ehart = 0.5 * sums[1] * vol / (double)ntot;
  const double Ecor = (Exc - EVxc) - Ehart - Eself + EIonSR + EVdw + Eext;
  if (num_occupied == nstate || temperature == 0.0) {
    *efree = Ekin + Ecor;  // zero temperature (scf.cpp:2081-2086)
    return PW_OK;
  }
  // Fermi-Dirac free energy, each term in the form whose exponent is not
  // positive (scf.cpp:2121-2133)
  double E = 0.0;
  for (int k = 0; k < nstate; k++) {
    const double t = eig[k] - fermi;
    if (t >= 0)
      E += -num_spin / Tbeta * std::log(1.0 + std::exp(-Tbeta * t));
    else
      E += num_spin * t - num_spin / Tbeta * std::log(1.0 + std::exp(Tbeta * t));
  }
  *efree = E + Ecor + fermi * num_occupied * num_spin;
  return PW_OK;
}
