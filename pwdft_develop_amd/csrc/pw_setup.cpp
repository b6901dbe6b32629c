// Host-side setup of the sphere-cut Gamma-point problem: G grid and half
// sphere (src/SCF/fourier.cpp:97-244, 541-672), kinetic/Teter arrays
// (src/SCF/hamiltonian.cpp:2629-2667) and the Gaussian projector block
// (src/SCF/hamiltonian.cpp:693-797 with the radial integral of
// src/SCF/periodtable.cpp:1599-1849 in closed form).  No device code.
#include <cmath>
#include <complex>
#include <cstring>
#include <map>
#include <mutex>

#include "pw_internal.h"

namespace pw {

static thread_local std::string g_err = "no error";

void set_error(const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
}

int fail(int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

double cell_volume(const double* s) {
  return std::fabs(s[0] * (s[4] * s[8] - s[5] * s[7]) - s[1] * (s[3] * s[8] - s[5] * s[6]) +
                   s[2] * (s[3] * s[7] - s[4] * s[6]));
}

void recip_cell(const double* s, double* b) {
  // b_j . a_k = 2 pi delta_jk: B = 2 pi (S^-1)^T with S rows = lattice vectors
  const double det = s[0] * (s[4] * s[8] - s[5] * s[7]) - s[1] * (s[3] * s[8] - s[5] * s[6]) +
                     s[2] * (s[3] * s[7] - s[4] * s[6]);
  const double c[9] = {s[4] * s[8] - s[5] * s[7], s[5] * s[6] - s[3] * s[8], s[3] * s[7] - s[4] * s[6],
                       s[2] * s[7] - s[1] * s[8], s[0] * s[8] - s[2] * s[6], s[1] * s[6] - s[0] * s[7],
                       s[1] * s[5] - s[2] * s[4], s[2] * s[3] - s[0] * s[5], s[0] * s[4] - s[1] * s[3]};
  // c is the cofactor matrix laid out so that c[3*j+i] = cof(j,i); (S^-1)^T = cof / det
  const double twopi = 2.0 * M_PI;
  for (int k = 0; k < 9; k++) b[k] = twopi * c[k] / det;
}

bool smooth235(int n) {
  if (n <= 0) return false;
  for (int p : {2, 3, 5})
    while (n % p == 0) n /= p;
  return n == 1;
}

int ensure_lds(const void* kern, size_t bytes) {
  static std::mutex mu;
  static std::map<const void*, size_t> done;
  if (bytes <= 64 * 1024) return PW_OK;
  if (bytes > 160 * 1024) return fail(PW_ERR_UNSUPPORTED, "kernel needs %zu bytes of LDS (> 160 KiB)", bytes);
  std::lock_guard<std::mutex> lock(mu);
  size_t& have = done[kern];
  if (have >= bytes) return PW_OK;
  PW_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  have = bytes;
  return PW_OK;
}

std::vector<int> radix_plan(int n) {
  std::vector<int> f;
  while (n % 4 == 0) { f.push_back(4); n /= 4; }
  while (n % 2 == 0) { f.push_back(2); n /= 2; }
  while (n % 3 == 0) { f.push_back(3); n /= 3; }
  while (n % 5 == 0) { f.push_back(5); n /= 5; }
  return f;
}

static inline void gvec(const double* b, int m0, int m1, int m2, double* g) {
  for (int d = 0; d < 3; d++) g[d] = b[d] * m0 + b[3 + d] * m1 + b[6 + d] * m2;
}

}  // namespace pw

using namespace pw;

extern "C" const char* pw_last_error(void) { return g_err.c_str(); }

extern "C" int pw_sphere_build(const int grid[3], const double supercell[9], double ecut, int* npw,
                               int* npad, int* idx, double* gkk, int* pad_src, int* pad_dst) {
  if (!grid || !supercell || !npw || !npad) return fail(PW_ERR_ARG, "pw_sphere_build: null argument");
  const int n1 = grid[0], n2 = grid[1], n3 = grid[2], h1 = n1 / 2 + 1;
  if (n1 <= 0 || n2 <= 0 || n3 <= 0) return fail(PW_ERR_ARG, "pw_sphere_build: bad grid");
  double b[9];
  recip_cell(supercell, b);
  int cnt = 0, np = 0;
  for (int k = 0; k < n3; k++)
    for (int j = 0; j < n2; j++)
      for (int i = 0; i < h1; i++) {
        const int G0 = kgrid(i, n1), G1 = kgrid(j, n2), G2 = kgrid(k, n3);
        double g[3];
        gvec(b, G0, G1, G2, g);
        const double e = (g[0] * g[0] + g[1] * g[1] + g[2] * g[2]) / 2.0;
        if (!(e < ecut)) continue;
        // upper half space of the kx = 0 plane, and all of kx > 0
        const bool upper = G0 > 0 || (G0 == 0 && (G1 > 0 || (G1 == 0 && G2 >= 0)));
        const int r2c = i + h1 * (j + n2 * k);
        if (upper) {
          if (idx) { idx[cnt] = r2c; gkk[cnt] = e; }
          cnt++;
        } else {  // G0 == 0 lower half: Hermitian padding from (0, -G1, -G2)
          int m1 = -G1, m2 = -G2;
          if (m1 < 0) m1 += n2;
          if (m2 < 0) m2 += n3;
          if (pad_src) { pad_src[np] = h1 * (m1 + n2 * m2); pad_dst[np] = r2c; }
          np++;
        }
      }
  *npw = cnt;
  *npad = np;
  return PW_OK;
}

extern "C" int pw_fock_sphere_build(const int grid[3], const double supercell[9], double ecut,
                                    int* nfock, int* idx, double* gkk) {
  if (!grid || !supercell || !nfock) return fail(PW_ERR_ARG, "pw_fock_sphere_build: null argument");
  const int n1 = grid[0], n2 = grid[1], n3 = grid[2], h1 = n1 / 2 + 1;
  double b[9];
  recip_cell(supercell, b);
  int cnt = 0;
  for (int k = 0; k < n3; k++)
    for (int j = 0; j < n2; j++)
      for (int i = 0; i < h1; i++) {
        double g[3];
        gvec(b, kgrid(i, n1), kgrid(j, n2), kgrid(k, n3), g);
        const double e = (g[0] * g[0] + g[1] * g[1] + g[2] * g[2]) / 2.0;
        if (e < ecut * 4.0) {
          if (idx) { idx[cnt] = i + h1 * (j + n2 * k); gkk[cnt] = e; }
          cnt++;
        }
      }
  *nfock = cnt;
  return PW_OK;
}

extern "C" int pw_ekin_teter(int npw, const double* gkk, double* ekin, double* teter) {
  if (npw < 0 || !gkk || !ekin || !teter) return fail(PW_ERR_ARG, "pw_ekin_teter: bad argument");
  for (int i = 0; i < npw; i++) {
    const double e = gkk[i];
    ekin[i] = e;
    const double a = 2.0 * e;  // |G|^2
    const double bb = 27.0 + a * (18.0 + a * (12.0 + a * 8.0));
    teter[i] = bb / (bb + 16.0 * (a * a) * (a * a));
  }
  return PW_OK;
}

extern "C" int pw_gaussian_projectors(const int grid[3], const double supercell[9], int npw,
                                      const int* idx, int natom, const double* pos, double sigma,
                                      int lmax, const double* wgt_channel, double* beta_out,
                                      double* w_out) {
  if (!grid || !supercell || !idx || !pos || !wgt_channel || !beta_out || !w_out)
    return fail(PW_ERR_ARG, "pw_gaussian_projectors: null argument");
  if (lmax < 0 || lmax > 1) return fail(PW_ERR_UNSUPPORTED, "pw_gaussian_projectors: lmax must be 0 or 1");
  const int n1 = grid[0], n2 = grid[1], h1 = n1 / 2 + 1;
  const int npa = (lmax + 1) * (lmax + 1);
  double b[9];
  recip_cell(supercell, b);
  const double vol = cell_volume(supercell);
  std::vector<double> G(3 * (size_t)npw), q(npw);
  for (int i = 0; i < npw; i++) {
    const int g = idx[i];
    const int kx = g % h1, ky = (g / h1) % n2, kz = g / (h1 * n2);
    gvec(b, kgrid(kx, n1), kgrid(ky, n2), kgrid(kz, grid[2]), &G[3 * i]);
    q[i] = std::sqrt(G[3 * i] * G[3 * i] + G[3 * i + 1] * G[3 * i + 1] + G[3 * i + 2] * G[3 * i + 2]);
  }
  typedef std::complex<double> cd;
  cd* beta = reinterpret_cast<cd*>(beta_out);
  const double s2 = std::sqrt(2.0);
  const double c0 = std::sqrt(1.0 / (4.0 * M_PI));
  const double c1 = std::sqrt(3.0 / (4.0 * M_PI));
  const double rad = std::sqrt(M_PI / 2.0);
  for (int a = 0; a < natom; a++) {
    const double* R = pos + 3 * a;
    for (int p = 0; p < npa; p++) {
      const int col = a * npa + p;
      const int l = (p == 0) ? 0 : 1;
      w_out[col] = wgt_channel[a * (lmax + 1) + l] * (4.0 * M_PI) * (4.0 * M_PI) / vol;
      cd* bc = beta + (size_t)npw * col;
      for (int i = 0; i < npw; i++) {
        const double* g = &G[3 * i];
        const double gauss = std::exp(-q[i] * q[i] * sigma * sigma / 2.0);
        cd vkb;
        if (l == 0) {
          vkb = c0 * rad * std::pow(sigma, 3) * gauss;
        } else {
          // (-i) sqrt(3/4pi) (G_c/|G|) * sqrt(pi/2) sigma^5 |G| e^{..}; c = z, x, y
          const double gc = (p == 1) ? g[2] : (p == 2) ? g[0] : g[1];
          vkb = cd(0.0, -1.0) * c1 * rad * std::pow(sigma, 5) * gc * gauss;
        }
        const double ph = -(g[0] * R[0] + g[1] * R[1] + g[2] * R[2]);
        cd v = vkb * cd(std::cos(ph), std::sin(ph));
        bc[i] = (i == 0) ? v : v * s2;
      }
    }
  }
  return PW_OK;
}
