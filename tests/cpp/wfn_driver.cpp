// Cross-check of host/wfn_io.hpp against the Python mirror: "write <dir>"
// writes a fixed restart (index 3); "read <dir>" reads the files the Python
// side wrote (index 5) and prints their sizes and checksums.
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#include "../../pwdft_develop_amd/host/wfn_io.hpp"

using namespace pwdft_amd::wfn;

int main(int argc, char** argv) {
  if (argc != 3) return 2;
  const std::string mode = argv[1], dir = argv[2];
  try {
    if (mode == "write") {
      const Int npw = 7, nband = 3;
      std::vector<Complex> psi((size_t)npw * nband);
      for (Int k = 0; k < nband; k++)
        for (Int i = 0; i < npw; i++) psi[i + (size_t)npw * k] = Complex(i + 0.25 * k, -0.5 * i + k);
      std::vector<double> occ = {2.0, 2.0, 0.5};
      std::vector<Int> idx = {0, 1, 5, 6, 33, 40, 41};
      write_restart(dir + "/WFN", dir + "/IDX", 3, psi.data(), npw, 1, nband, occ.data(), 3, idx.data());
      return 0;
    }
    Restart r = read_restart(dir + "/WFN", dir + "/IDX", 5);
    double s = 0.0, so = 0.0;
    long si = 0;
    for (size_t i = 0; i < r.psi.size(); i++) s += r.psi[i].real() * (i + 1) + r.psi[i].imag() * (2 * i + 1);
    for (double o : r.occ) so += o;
    for (Int v : r.idx) si += v;
    std::printf("%d %d %d %.17g %.17g %ld\n", r.npw, r.ncom, r.nband, s, so, si);
    return 0;
  } catch (const std::exception& e) {
    std::printf("error: %s\n", e.what());
    return 1;
  }
}
