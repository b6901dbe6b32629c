// C++ adapter with the reference's interfaces for the H*psi / eigensolver
// path, implemented over the C-ABI of libpwdft_hip.so (include/pwdft_hip.h).
//
// The classes keep the names, signatures, data layouts and error behaviour of
// the reference so that its SCF driver can call them unchanged:
//   Domain            include/domain.hpp:60-144 (orthorhombic or general cell)
//   NumVec/NumMat/NumTns  include/num{vec,mat,tns}_*.hpp (column-major,
//                     owndata = false wraps caller memory)
//   Fourier           include/fourier.hpp:69-176, src/SCF/fourier.cpp:97-672
//   Spinor            include/spinor.hpp:69-372, src/SCF/spinor.cpp
//   KohnSham          include/hamiltonian.hpp:383-501 (MultSpinor, CalculateEkin,
//                     Teter, Vnlc, EigVal, density)
//   EigenSolver       include/eigensolver.hpp:59-219 (Setup, LOBPCGSolveReal,
//                     NonlocalMultX, FirstChebyStep, GeneralChebyStep, EigVal,
//                     ResVal)
//   ErrorHandling     src/global.cpp:71-91: throws std::runtime_error
//
// Host arrays stay host arrays (as in the reference); each call moves the
// band block to the device, runs the kernels there and copies results back.
// A Fourier owns one device context, so (like the reference's Fourier with its
// shared scratch) it is not re-entrant.
#pragma once
#include <hip/hip_runtime.h>

#include <complex>
#include <cstring>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../include/pwdft_hip.h"

namespace pwdft {

typedef int Int;
typedef double Real;
typedef std::complex<double> Complex;

inline void ErrorHandling(const std::string& msg) { throw std::runtime_error(msg); }
inline void Check(int rc) {
  if (rc != PW_OK) ErrorHandling(pw_last_error());
}

// ---- containers: column-major, optionally wrapping foreign memory ----
template <class F>
class NumTns {
 public:
  NumTns() = default;
  NumTns(Int m, Int n, Int p, bool owndata = true, F* data = nullptr) { Resize(m, n, p, owndata, data); }
  void Resize(Int m, Int n, Int p, bool owndata = true, F* data = nullptr) {
    m_ = m; n_ = n; p_ = p; owndata_ = owndata;
    if (owndata) { store_.assign((size_t)m * n * p, F(0)); data_ = store_.data(); if (data) std::memcpy(data_, data, sizeof(F) * store_.size()); }
    else data_ = data;
  }
  Int m() const { return m_; }
  Int n() const { return n_; }
  Int p() const { return p_; }
  F* Data() const { return data_; }
  F* VecData(Int j, Int k) const { return data_ + (size_t)m_ * (j + (size_t)n_ * k); }
  F* MatData(Int k) const { return data_ + (size_t)m_ * n_ * k; }
  F& operator()(Int i, Int j, Int k) const { return data_[i + (size_t)m_ * (j + (size_t)n_ * k)]; }

 private:
  Int m_ = 0, n_ = 0, p_ = 0;
  bool owndata_ = true;
  F* data_ = nullptr;
  std::vector<F> store_;
};

template <class F>
class NumMat {
 public:
  NumMat() = default;
  NumMat(Int m, Int n, bool owndata = true, F* data = nullptr) { Resize(m, n, owndata, data); }
  void Resize(Int m, Int n, bool owndata = true, F* data = nullptr) {
    m_ = m; n_ = n;
    if (owndata) { store_.assign((size_t)m * n, F(0)); data_ = store_.data(); if (data) std::memcpy(data_, data, sizeof(F) * store_.size()); }
    else data_ = data;
  }
  Int m() const { return m_; }
  Int n() const { return n_; }
  F* Data() const { return data_; }
  F* VecData(Int j) const { return data_ + (size_t)m_ * j; }
  F& operator()(Int i, Int j) const { return data_[i + (size_t)m_ * j]; }

 private:
  Int m_ = 0, n_ = 0;
  F* data_ = nullptr;
  std::vector<F> store_;
};

template <class F>
class NumVec {
 public:
  NumVec() = default;
  explicit NumVec(Int m) { Resize(m); }
  void Resize(Int m) { v_.assign(m, F(0)); }
  Int m() const { return (Int)v_.size(); }
  Int Size() const { return (Int)v_.size(); }
  F* Data() { return v_.data(); }
  const F* Data() const { return v_.data(); }
  F& operator[](Int i) { return v_[i]; }
  const F& operator[](Int i) const { return v_[i]; }
  F& operator()(Int i) { return v_[i]; }
  const F& operator()(Int i) const { return v_[i]; }

 private:
  std::vector<F> v_;
};

typedef NumVec<Int> IntNumVec;
typedef NumVec<Real> DblNumVec;
typedef NumMat<Real> DblNumMat;
typedef NumMat<Complex> CpxNumMat;

// ---- Domain (include/domain.hpp) ----
struct Domain {
  DblNumMat supercell{3, 3};  // supercell(i, j): component j of lattice vector i
  Int numGrid[3] = {0, 0, 0};
  Int numGridSphere = 0;
  Real Volume() const {
    const Real* s = supercell.Data();
    return std::abs(s[0] * (s[4] * s[8] - s[5] * s[7]) - s[3] * (s[1] * s[8] - s[2] * s[7]) +
                    s[6] * (s[1] * s[5] - s[2] * s[4]));
  }
  Int NumGridTotal() const { return numGrid[0] * numGrid[1] * numGrid[2]; }
  // row-major lattice (vector j in row j) for the C-ABI
  void Lattice(double* out) const {
    for (int j = 0; j < 3; j++)
      for (int i = 0; i < 3; i++) out[3 * j + i] = supercell(j, i);
  }
};

// A device block mirror of a host band block (complex double).
class DeviceBlock {
 public:
  DeviceBlock() = default;
  explicit DeviceBlock(size_t n) { Alloc(n); }
  ~DeviceBlock() { if (p_) (void)hipFree(p_); }
  DeviceBlock(const DeviceBlock&) = delete;
  DeviceBlock& operator=(const DeviceBlock&) = delete;
  void Alloc(size_t n) {
    if (n <= n_) return;
    if (p_) (void)hipFree(p_);
    if (hipMalloc(&p_, sizeof(Complex) * (n ? n : 1)) != hipSuccess) ErrorHandling("hipMalloc failed");
    n_ = n;
  }
  void Upload(const Complex* h, size_t n) {
    Alloc(n);
    if (hipMemcpy(p_, h, sizeof(Complex) * n, hipMemcpyHostToDevice) != hipSuccess) ErrorHandling("H2D copy failed");
  }
  void Download(Complex* h, size_t n) const {
    if (hipMemcpy(h, p_, sizeof(Complex) * n, hipMemcpyDeviceToHost) != hipSuccess) ErrorHandling("D2H copy failed");
  }
  void* Ptr() const { return p_; }

 private:
  void* p_ = nullptr;
  size_t n_ = 0;
};

// ---- Fourier (include/fourier.hpp, src/SCF/fourier.cpp) ----
class Fourier {
 public:
  Domain domain;
  bool isInitialized = false;
  Int numGridTotal = 0, numGridTotalR2C = 0;
  IntNumVec idxCoarseCut;
  std::pair<IntNumVec, IntNumVec> idxCoarsePadding;
  DblNumVec gkkSphere;  // |G|^2/2 on the sphere
  Real ecutWavefunction = 0;

  ~Fourier() { if (ctx_) pw_ctx_destroy(ctx_); }

  // Initialize + InitializeSphere (fourier.cpp:97-244, 541-672)
  void Initialize(const Domain& dm, Real ecut, Int maxBands) {
    if (isInitialized) ErrorHandling("Fourier has been initialized.");
    domain = dm;
    ecutWavefunction = ecut;
    numGridTotal = dm.NumGridTotal();
    numGridTotalR2C = (dm.numGrid[0] / 2 + 1) * dm.numGrid[1] * dm.numGrid[2];
    double lat[9];
    dm.Lattice(lat);
    int npw = 0, npad = 0;
    Check(pw_sphere_build(dm.numGrid, lat, ecut, &npw, &npad, nullptr, nullptr, nullptr, nullptr));
    idxCoarseCut.Resize(npw);
    gkkSphere.Resize(npw);
    idxCoarsePadding.first.Resize(npad);
    idxCoarsePadding.second.Resize(npad);
    Check(pw_sphere_build(dm.numGrid, lat, ecut, &npw, &npad, idxCoarseCut.Data(), gkkSphere.Data(),
                          idxCoarsePadding.first.Data(), idxCoarsePadding.second.Data()));
    domain.numGridSphere = npw;
    maxBands_ = maxBands;
    isInitialized = true;
  }

  // The device context, created on first use with the kinetic/Teter arrays.
  pw_ctx* Ctx(const DblNumVec& ekin, const DblNumVec& teter) {
    if (!isInitialized) ErrorHandling("Fourier is not prepared.");
    if (!ctx_) {
      double lat[9];
      domain.Lattice(lat);
      Check(pw_ctx_create(&ctx_, domain.numGrid, lat, domain.numGridSphere, idxCoarseCut.Data(),
                          idxCoarsePadding.first.m(), idxCoarsePadding.first.Data(), idxCoarsePadding.second.Data(),
                          ekin.Data(), teter.Data(), maxBands_, nullptr));
    }
    return ctx_;
  }
  pw_ctx* Ctx() const {
    if (!ctx_) ErrorHandling("Fourier is not prepared.");
    return ctx_;
  }

 private:
  pw_ctx* ctx_ = nullptr;
  Int maxBands_ = 0;
};

// ---- Spinor (include/spinor.hpp) ----
class Spinor {
 public:
  Spinor() = default;
  // spinor.cpp:247-286: wraps caller memory when owndata == false
  Spinor(const Domain& dm, Int numComponent, Int numStateTotal, bool owndata, Complex* data) {
    Setup(dm, numComponent, numStateTotal, owndata, data);
  }
  void Setup(const Domain& dm, Int numComponent, Int numStateTotal, bool owndata, Complex* data) {
    domain_ = dm;
    numComponent_ = numComponent;
    numStateTotal_ = numStateTotal;
    wavefunG_.Resize(dm.numGridSphere, numComponent, numStateTotal, owndata, data);
  }
  NumTns<Complex>& WavefunG() { return wavefunG_; }
  Int NumGridTotal() const { return wavefunG_.m(); }
  Int NumComponent() const { return numComponent_; }
  Int NumState() const { return wavefunG_.p(); }
  Int NumStateTotal() const { return numStateTotal_; }

  // spinor.cpp:348-374: a3 += teter .* psi
  void AddTeterPrecond(Fourier* fftPtr, DblNumVec& teter, NumTns<Complex>& a3) {
    if (!fftPtr->isInitialized) ErrorHandling("Fourier is not prepared.");
    const Int ntot = NumGridTotal(), nocc = NumState();
    if (fftPtr->domain.numGridSphere != ntot) ErrorHandling("Domain size does not match.");
    (void)teter;
    pw_ctx* ctx = fftPtr->Ctx();
    DeviceBlock x, y;
    x.Upload(wavefunG_.Data(), (size_t)ntot * nocc);
    y.Upload(a3.Data(), (size_t)ntot * nocc);
    Check(pw_teter_precond(ctx, PW_F64, nocc, x.Ptr(), y.Ptr()));
    y.Download(a3.Data(), (size_t)ntot * nocc);
  }

 private:
  Domain domain_;
  Int numComponent_ = 1, numStateTotal_ = 0;
  NumTns<Complex> wavefunG_;
};

// ---- Hamiltonian / KohnSham (include/hamiltonian.hpp) ----
class Hamiltonian {
 public:
  virtual ~Hamiltonian() = default;
  virtual void MultSpinor(Spinor& psi, NumTns<Complex>& a3, Fourier& fft) = 0;
};

class KohnSham : public Hamiltonian {
 public:
  // CalculateEkin (hamiltonian.cpp:2629-2667)
  void CalculateEkin(Fourier& fft) {
    const Int npw = fft.domain.numGridSphere;
    ekin_.Resize(npw);
    teter_.Resize(npw);
    Check(pw_ekin_teter(npw, fft.gkkSphere.Data(), ekin_.Data(), teter_.Data()));
    ctx_ = fft.Ctx(ekin_, teter_);
  }
  // vtot_[:, spinswitch_] (hamiltonian.cpp:3333-3334)
  void SetVtot(const DblNumVec& vtot) {
    vtot_ = vtot;
    Check(pw_set_vtot(ctx_, vtot_.Data()));
  }
  // vnlc_[0] = (beta, weights) (hamiltonian.cpp:693-797)
  void SetVnlc(const CpxNumMat& beta, const DblNumVec& wgt) {
    vnlc_.first.Resize(beta.m(), beta.n(), true, beta.Data());
    vnlc_.second = wgt;
    Check(pw_set_nonlocal(ctx_, beta.n(), reinterpret_cast<const double*>(beta.Data()), wgt.Data()));
  }
  // MultSpinor(Spinor&, NumTns<Complex>&, Fourier&) (hamiltonian.cpp:3270-3452):
  // a3 = (T + V_loc [+ V_x]) psi; the nonlocal term is NonlocalMultX's.
  void MultSpinor(Spinor& psi, NumTns<Complex>& a3, Fourier& fft) override {
    if (!fft.isInitialized) ErrorHandling("Fourier is not prepared.");
    const Int ntot = psi.NumGridTotal(), nb = psi.NumState();
    if (fft.domain.numGridSphere != ntot) ErrorHandling("Domain size does not match3.");
    if (psi.NumComponent() != 1)
      ErrorHandling("Component number of spinor should be one in spin-restricted and spin-unrestricted calculations.");
    DeviceBlock x, y;
    x.Upload(psi.WavefunG().Data(), (size_t)ntot * nb);
    y.Alloc((size_t)ntot * nb);
    Check(pw_mult_spinor(ctx_, PW_F64, nb, x.Ptr(), y.Ptr()));
    y.Download(a3.Data(), (size_t)ntot * nb);
  }
  DblNumVec& Teter() { return teter_; }
  DblNumVec& Ekin() { return ekin_; }
  std::pair<CpxNumMat, DblNumVec>& Vnlc() { return vnlc_; }
  DblNumVec& EigVal() { return eigVal_; }
  pw_ctx* Ctx() const { return ctx_; }

 private:
  pw_ctx* ctx_ = nullptr;
  DblNumVec ekin_, teter_, vtot_, eigVal_;
  std::pair<CpxNumMat, DblNumVec> vnlc_;
};

// ---- EigenSolver (include/eigensolver.hpp) ----
class EigenSolver {
 public:
  void Setup(KohnSham& ham, Spinor& psi, Fourier& fft) {
    hamPtr_ = &ham;
    psiPtr_ = &psi;
    fftPtr_ = &fft;
    eigVal_.Resize(psi.NumStateTotal());
    resVal_.Resize(psi.NumStateTotal());
  }
  DblNumVec& EigVal() { return eigVal_; }
  DblNumVec& ResVal() { return resVal_; }
  Spinor& Psi() { return *psiPtr_; }
  KohnSham& Ham() { return *hamPtr_; }
  Fourier& FFT() { return *fftPtr_; }

  // NonlocalMultX (eigensolver.cpp:471-515): VnlX = beta (w .* Re(beta^H X))
  void NonlocalMultX(Int heightLocal, Int width, CpxNumMat& X, CpxNumMat& VnlX) {
    DeviceBlock x, y;
    x.Upload(X.Data(), (size_t)heightLocal * width);
    y.Alloc((size_t)heightLocal * width);
    Check(pw_nonlocal_multx(hamPtr_->Ctx(), PW_F64, width, x.Ptr(), y.Ptr()));
    y.Download(VnlX.Data(), (size_t)heightLocal * width);
  }

  // LOBPCGSolveReal (lobpcg.cpp:59-576)
  void LOBPCGSolveReal(Int numEig, Int scfIter, Int eigMaxIter, Real eigMinTolerance, Real eigTolerance) {
    Spinor& psi = *psiPtr_;
    const Int ntot = psi.NumGridTotal(), nb = psi.NumStateTotal();
    DeviceBlock x;
    x.Upload(psi.WavefunG().Data(), (size_t)ntot * nb);
    int iters = 0;
    Check(pw_lobpcg(hamPtr_->Ctx(), nb, numEig, scfIter, eigMaxIter, eigMinTolerance, eigTolerance, x.Ptr(),
                    eigVal_.Data(), resVal_.Data(), &iters));
    x.Download(psi.WavefunG().Data(), (size_t)ntot * nb);
    hamPtr_->EigVal() = eigVal_;
  }

  // The ChebFSI steps the reference calls at scf.cpp:1306-1316 and 1321-1332.
  // FirstChebyStep: Rayleigh-Ritz on the input block seeds the filter bounds,
  // then cycleNum filter + RR cycles; GeneralChebyStep: one cycle with the
  // bounds from the previous Ritz values.  mixed = FP32 filter.
  void FirstChebyStep(Int numEig, Int cycleNum, Int filterOrder, bool mixed = false) {
    (void)numEig;
    Spinor& psi = *psiPtr_;
    const Int ntot = psi.NumGridTotal(), nb = psi.NumStateTotal();
    DeviceBlock x;
    x.Upload(psi.WavefunG().Data(), (size_t)ntot * nb);
    Check(pw_rayleigh_ritz(hamPtr_->Ctx(), PW_F64, nb, x.Ptr(), eigVal_.Data(), resVal_.Data()));
    for (Int c = 0; c < cycleNum; c++) Sweep(x, filterOrder, mixed);
    x.Download(psi.WavefunG().Data(), (size_t)ntot * nb);
    hamPtr_->EigVal() = eigVal_;
  }
  void GeneralChebyStep(Int numEig, Int filterOrder, bool mixed = false) {
    (void)numEig;
    Spinor& psi = *psiPtr_;
    const Int ntot = psi.NumGridTotal(), nb = psi.NumStateTotal();
    DeviceBlock x;
    x.Upload(psi.WavefunG().Data(), (size_t)ntot * nb);
    Sweep(x, filterOrder, mixed);
    x.Download(psi.WavefunG().Data(), (size_t)ntot * nb);
    hamPtr_->EigVal() = eigVal_;
  }
  // Lanczos start vector of the spectral-bound estimate (one band).
  void SetLanczosStart(const std::vector<Complex>& v0, Int steps) {
    lanczosStart_ = v0;
    lanczosSteps_ = steps;
  }

 private:
  void Sweep(DeviceBlock& x, Int filterOrder, bool mixed) {
    const Int nb = psiPtr_->NumStateTotal();
    if (lanczosStart_.empty()) ErrorHandling("EigenSolver: no Lanczos start vector (SetLanczosStart)");
    DeviceBlock v0;
    v0.Upload(lanczosStart_.data(), lanczosStart_.size());
    std::vector<double> prev(eigVal_.Data(), eigVal_.Data() + nb);
    const bool havePrev = !(prev.empty() || (prev.front() == 0.0 && prev.back() == 0.0));
    double bounds[3];
    Check(pw_chebfsi_sweep(hamPtr_->Ctx(), mixed ? PW_F32 : PW_F64, nb, filterOrder, lanczosSteps_, v0.Ptr(),
                           havePrev ? prev.data() : nullptr, x.Ptr(), eigVal_.Data(), resVal_.Data(), bounds));
  }

  KohnSham* hamPtr_ = nullptr;
  Spinor* psiPtr_ = nullptr;
  Fourier* fftPtr_ = nullptr;
  DblNumVec eigVal_, resVal_;
  std::vector<Complex> lanczosStart_;
  Int lanczosSteps_ = 8;
};

}  // namespace pwdft
