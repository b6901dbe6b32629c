// Drives the C++ adapter (pwdft_develop_amd/host/pwdft_adapter.hpp) the way
// the reference's SCF driver calls its classes: Fourier::Initialize,
// KohnSham::CalculateEkin / MultSpinor, EigenSolver::NonlocalMultX,
// LOBPCGSolveReal and GeneralChebyStep.
//
//   adapter_driver <input.bin> <output.bin>
//
// input: int32 n1 n2 n3 nb nbeta npw; float64 supercell[9], ecut; vtot[N];
//        beta[2 npw nbeta]; wgt[nbeta]; psi[2 npw nb]; v0[2 npw]
// output: float64 Hpsi[2 npw nb] (MultSpinor + NonlocalMultX), LOBPCG
//         eigenvalues[nb] (3 iterations), ChebFSI eigenvalues[nb] (one
//         GeneralChebyStep after the LOBPCG, degree 10, 8 Lanczos steps)
#include <cstdio>
#include <vector>

#include "../../pwdft_develop_amd/host/pwdft_adapter.hpp"

using namespace pwdft;

template <class T>
static void rd(FILE* f, T* p, size_t n) {
  if (fread(p, sizeof(T), n, f) != n) ErrorHandling("short read");
}

int main(int argc, char** argv) {
  if (argc != 3) {
    fprintf(stderr, "usage: %s in.bin out.bin\n", argv[0]);
    return 2;
  }
  try {
    FILE* f = fopen(argv[1], "rb");
    if (!f) ErrorHandling("cannot open input");
    int hdr[6];
    rd(f, hdr, 6);
    const int nb = hdr[3], nbeta = hdr[4], npw = hdr[5];
    Domain dm;
    dm.numGrid[0] = hdr[0]; dm.numGrid[1] = hdr[1]; dm.numGrid[2] = hdr[2];
    double sc[9], ecut;
    rd(f, sc, 9);
    rd(f, &ecut, 1);
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < 3; i++) dm.supercell(j, i) = sc[3 * j + i];
    const int N = dm.NumGridTotal();
    DblNumVec vtot(N);
    rd(f, vtot.Data(), N);
    CpxNumMat beta(npw, nbeta);
    rd(f, reinterpret_cast<double*>(beta.Data()), 2 * (size_t)npw * nbeta);
    DblNumVec wgt(nbeta);
    rd(f, wgt.Data(), nbeta);
    std::vector<Complex> psi((size_t)npw * nb), v0(npw);
    rd(f, reinterpret_cast<double*>(psi.data()), 2 * psi.size());
    rd(f, reinterpret_cast<double*>(v0.data()), 2 * v0.size());
    fclose(f);

    Fourier fft;
    fft.Initialize(dm, ecut, nb);
    if (fft.domain.numGridSphere != npw) ErrorHandling("sphere size differs from the input's");
    KohnSham ham;
    ham.CalculateEkin(fft);
    ham.SetVtot(vtot);
    ham.SetVnlc(beta, wgt);

    // H psi as LOBPCG assembles it: MultSpinor + NonlocalMultX
    Spinor spn(fft.domain, 1, nb, false, psi.data());
    NumTns<Complex> a3(npw, 1, nb);
    ham.MultSpinor(spn, a3, fft);
    CpxNumMat X(npw, nb, false, psi.data()), VnlX(npw, nb);
    EigenSolver eig;
    eig.Setup(ham, spn, fft);
    eig.NonlocalMultX(npw, nb, X, VnlX);
    std::vector<double> out;
    for (size_t i = 0; i < (size_t)npw * nb; i++) {
      const Complex h = a3.Data()[i] + VnlX.Data()[i];
      out.push_back(h.real());
      out.push_back(h.imag());
    }
    eig.LOBPCGSolveReal(nb, 1, 3, 0.0, 1e-20);
    for (int i = 0; i < nb; i++) out.push_back(eig.EigVal()[i]);
    eig.SetLanczosStart(v0, 8);
    eig.GeneralChebyStep(nb, 10);
    for (int i = 0; i < nb; i++) out.push_back(eig.EigVal()[i]);
    FILE* g = fopen(argv[2], "wb");
    if (!g || fwrite(out.data(), sizeof(double), out.size(), g) != out.size()) ErrorHandling("cannot write output");
    fclose(g);
  } catch (const std::exception& e) {
    fprintf(stderr, "ERROR: %s\n", e.what());
    return 1;
  }
  return 0;
}
