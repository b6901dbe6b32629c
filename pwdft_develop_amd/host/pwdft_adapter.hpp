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
//                     PPCGSolveReal, NonlocalMultX, FirstChebyStep,
//                     GeneralChebyStep, EigVal, ResVal)
//   esdfParam         include/esdf.hpp: the two keys the solvers read
//                     (PPCGsbSize, isUseRealSpace)
//   ErrorHandling     src/global.cpp:71-91: throws std::runtime_error
//   SparseVec, NonlocalPP, PseudoPot   include/utility.hpp:120-152 (the
//                     real-space projectors, vnlList only)
// The real-space dual-grid path (Use_Real_Space = 1) adds Domain::numGridFine,
// Spinor::Wavefun (real coarse-grid values), KohnSham::CalculateEkin over the
// whole R2C grid, the fine-grid vtot and pseudo_, and the NumTns<Real>
// MultSpinor overload (hamiltonian.cpp:3058-3125).
//
// Host arrays stay host arrays (as in the reference); each call moves the
// band block to the device, runs the kernels there and copies results back
// through staging blocks that the Hamiltonian and the EigenSolver keep and
// only grow (no device allocation per call once the largest block was seen).
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

// The input keys of esdfParam (include/esdf.hpp:473 and its neighbours)
// that EigenSolver reads: PPCG's sub-block size (src/esdf.cpp:2722, default
// 1) and the real-space switch (ppcg.cpp:75).
struct ESDFInputParam {
  Int PPCGsbSize = 1;
  bool isUseRealSpace = false;
};
inline ESDFInputParam esdfParam;
inline void Check(int rc) {
  if (rc != PW_OK) ErrorHandling(pw_last_error());
}

// ---- containers: column-major, optionally wrapping foreign memory ----
template <class F>
class NumTns {
 public:
  NumTns() = default;
  NumTns(Int m, Int n, Int p, bool owndata = true, F* data = nullptr) { Resize(m, n, p, owndata, data); }
  // copies own their data when the source does (a deep copy), else share
  // the wrapped memory, as the reference's containers do
  NumTns(const NumTns& o) { *this = o; }
  NumTns(NumTns&&) = default;
  NumTns& operator=(NumTns&&) = default;
  NumTns& operator=(const NumTns& o) {
    if (this == &o) return *this;
    m_ = o.m_; n_ = o.n_; p_ = o.p_; owndata_ = o.owndata_;
    store_ = o.store_;
    data_ = o.owndata_ ? store_.data() : o.data_;
    return *this;
  }
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
  NumMat(const NumMat& o) { *this = o; }
  NumMat(NumMat&&) = default;
  NumMat& operator=(NumMat&&) = default;
  NumMat& operator=(const NumMat& o) {
    if (this == &o) return *this;
    m_ = o.m_; n_ = o.n_; owndata_ = o.owndata_;
    store_ = o.store_;
    data_ = o.owndata_ ? store_.data() : o.data_;
    return *this;
  }
  void Resize(Int m, Int n, bool owndata = true, F* data = nullptr) {
    m_ = m; n_ = n; owndata_ = owndata;
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
  bool owndata_ = true;
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

// include/utility.hpp:120-152 (real-space projectors; column VAL = 0 of dv)
typedef std::pair<IntNumVec, DblNumMat> SparseVec;
typedef std::pair<SparseVec, Real> NonlocalPP;
struct PseudoPot {
  std::vector<NonlocalPP> vnlList;
};
enum { VAL = 0, DX = 1, DY = 2, DZ = 3 };

// ---- Domain (include/domain.hpp) ----
struct Domain {
  DblNumMat supercell{3, 3};  // supercell(i, j): component j of lattice vector i
  Int numGrid[3] = {0, 0, 0};
  Int numGridFine[3] = {0, 0, 0};
  Int numGridSphere = 0;
  Real Volume() const {
    const Real* s = supercell.Data();
    return std::abs(s[0] * (s[4] * s[8] - s[5] * s[7]) - s[3] * (s[1] * s[8] - s[2] * s[7]) +
                    s[6] * (s[1] * s[5] - s[2] * s[4]));
  }
  Int NumGridTotal() const { return numGrid[0] * numGrid[1] * numGrid[2]; }
  Int NumGridTotalFine() const { return numGridFine[0] * numGridFine[1] * numGridFine[2]; }
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
  // real arrays: n doubles (the block is sized in complex slots)
  void UploadReal(const Real* h, size_t n) {
    Alloc((n + 1) / 2);
    if (hipMemcpy(p_, h, sizeof(Real) * n, hipMemcpyHostToDevice) != hipSuccess) ErrorHandling("H2D copy failed");
  }
  void DownloadReal(Real* h, size_t n) const {
    if (hipMemcpy(h, p_, sizeof(Real) * n, hipMemcpyDeviceToHost) != hipSuccess) ErrorHandling("D2H copy failed");
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

  ~Fourier() {
    if (ctx_) pw_ctx_destroy(ctx_);
    if (rs_) pw_rs_destroy(rs_);
  }

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
  // Real-space dual-grid handle (the R2C / R2CFine plans, fourier.cpp:
  // 166-183, 246-330), created on first use with ekin on the R2C grid.
  pw_rs* RsCtx(const DblNumVec& ekinR2C) {
    if (!isInitialized) ErrorHandling("Fourier is not prepared.");
    if (!rs_) {
      if (domain.NumGridTotalFine() <= 0) ErrorHandling("numGridFine is not set.");
      if (ekinR2C.m() != numGridTotalR2C) ErrorHandling("Domain size does not match.");
      Check(pw_rs_create(&rs_, domain.numGrid, domain.numGridFine, domain.Volume(), ekinR2C.Data(), maxBands_,
                         nullptr));
    }
    return rs_;
  }
  pw_rs* RsCtx() const {
    if (!rs_) ErrorHandling("Fourier is not prepared.");
    return rs_;
  }

 private:
  pw_ctx* ctx_ = nullptr;
  pw_rs* rs_ = nullptr;
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
  // Real-space Gamma wavefunctions on the coarse grid (spinor.cpp:85-112):
  // ntot x numComponent x numState, wrapping caller memory when owndata ==
  // false.
  void SetupReal(const Domain& dm, Int numComponent, Int numStateTotal, bool owndata, Real* data) {
    domain_ = dm;
    numComponent_ = numComponent;
    numStateTotal_ = numStateTotal;
    wavefun_.Resize(dm.NumGridTotal(), numComponent, numStateTotal, owndata, data);
  }
  NumTns<Real>& Wavefun() { return wavefun_; }
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
    if (!stx_) {
      stx_ = std::make_shared<DeviceBlock>();
      sty_ = std::make_shared<DeviceBlock>();
    }
    stx_->Upload(wavefunG_.Data(), (size_t)ntot * nocc);
    sty_->Upload(a3.Data(), (size_t)ntot * nocc);
    Check(pw_teter_precond(ctx, PW_F64, nocc, stx_->Ptr(), sty_->Ptr()));
    sty_->Download(a3.Data(), (size_t)ntot * nocc);
  }

 private:
  Domain domain_;
  Int numComponent_ = 1, numStateTotal_ = 0;
  NumTns<Complex> wavefunG_;
  NumTns<Real> wavefun_;
  // device staging of AddTeterPrecond, grown on demand (shared by copies,
  // which keeps Spinor copyable; calls run in order on one stream)
  std::shared_ptr<DeviceBlock> stx_, sty_;
};

// ---- Hamiltonian / KohnSham (include/hamiltonian.hpp) ----
// The base class holds what the solvers read (hamiltonian.hpp:71-376: the
// device context in place of the host operator data, ekin_/teter_, vnlc_,
// eigVal_) so that EigenSolver::Setup takes a Hamiltonian&, as the
// reference's does.
class Hamiltonian {
 public:
  virtual ~Hamiltonian() = default;
  virtual void MultSpinor(Spinor& psi, NumTns<Complex>& a3, Fourier& fft) = 0;
  virtual void MultSpinor(Spinor& psi, NumTns<Real>& a3, Fourier& fft) = 0;
  DblNumVec& Teter() { return teter_; }
  DblNumVec& Ekin() { return ekin_; }
  std::pair<CpxNumMat, DblNumVec>& Vnlc() { return vnlc_; }
  DblNumVec& EigVal() { return eigVal_; }
  pw_ctx* Ctx() const { return ctx_; }
  pw_rs* RsCtx() const { return rs_; }

 protected:
  pw_ctx* ctx_ = nullptr;
  pw_rs* rs_ = nullptr;
  DblNumVec ekin_, teter_, vtot_, eigVal_;
  std::pair<CpxNumMat, DblNumVec> vnlc_;
  DeviceBlock dx_, dy_;  // staging of MultSpinor's blocks, grown on demand
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
    dx_.Upload(psi.WavefunG().Data(), (size_t)ntot * nb);
    dy_.Alloc((size_t)ntot * nb);
    Check(pw_mult_spinor(ctx_, PW_F64, nb, dx_.Ptr(), dy_.Ptr()));
    dy_.Download(a3.Data(), (size_t)ntot * nb);
  }
  // CalculateEkin without sphere cut (hamiltonian.cpp:2629-2667 with
  // isUseSphereCut false): |G|^2/2 and Teter over the coarse R2C grid, G from
  // the k-grid of fourier.cpp:125-131 and the reciprocal lattice.
  void CalculateEkinR2C(Fourier& fft) {
    const Domain& dm = fft.domain;
    const Int n1 = dm.numGrid[0], n2 = dm.numGrid[1], n3 = dm.numGrid[2], h1 = n1 / 2 + 1;
    double A[9], B[9];
    dm.Lattice(A);
    const double det = A[0] * (A[4] * A[8] - A[5] * A[7]) - A[1] * (A[3] * A[8] - A[5] * A[6]) +
                       A[2] * (A[3] * A[7] - A[4] * A[6]);
    if (det == 0.0) ErrorHandling("singular supercell");
    // rows of B = 2 pi (A^{-1})^T: b_j . a_i = 2 pi delta_ij
    const double c = 2.0 * 3.14159265358979323846 / det;
    B[0] = c * (A[4] * A[8] - A[5] * A[7]); B[1] = c * (A[5] * A[6] - A[3] * A[8]); B[2] = c * (A[3] * A[7] - A[4] * A[6]);
    B[3] = c * (A[2] * A[7] - A[1] * A[8]); B[4] = c * (A[0] * A[8] - A[2] * A[6]); B[5] = c * (A[1] * A[6] - A[0] * A[7]);
    B[6] = c * (A[1] * A[5] - A[2] * A[4]); B[7] = c * (A[2] * A[3] - A[0] * A[5]); B[8] = c * (A[0] * A[4] - A[1] * A[3]);
    ekinR2C_.Resize(h1 * n2 * n3);
    teterR2C_.Resize(h1 * n2 * n3);
    Int g = 0;
    for (Int k = 0; k < n3; k++)
      for (Int j = 0; j < n2; j++)
        for (Int i = 0; i < h1; i++, g++) {
          const double m[3] = {(double)i, (double)(j <= n2 / 2 ? j : j - n2), (double)(k <= n3 / 2 ? k : k - n3)};
          double G[3] = {0, 0, 0};
          for (int d = 0; d < 3; d++) G[d] = m[0] * B[d] + m[1] * B[3 + d] + m[2] * B[6 + d];
          ekinR2C_[g] = (G[0] * G[0] + G[1] * G[1] + G[2] * G[2]) / 2;
          const double a = ekinR2C_[g] * 2.0, b = 27.0 + a * (18.0 + a * (12.0 + a * 8.0));
          teterR2C_[g] = b / (b + 16.0 * a * a * a * a);
        }
    rs_ = fft.RsCtx(ekinR2C_);
  }
  // vtot_[:, spinswitch_] on the fine grid (hamiltonian.cpp:3120-3122)
  void SetVtotFine(const DblNumVec& vtotFine) {
    if (!rs_) ErrorHandling("Fourier is not prepared.");
    Check(pw_rs_set_vtot(rs_, vtotFine.Data()));
  }
  // pseudo_ (hamiltonian.cpp:325-470): every projector of an atom is a
  // SparseVec on the atom's index list (periodtable.cpp:1063-1589)
  void SetPseudo(const std::vector<PseudoPot>& pseudo) {
    if (!rs_) ErrorHandling("Fourier is not prepared.");
    std::vector<int> npts, nproj, idx;
    std::vector<double> val, wgt;
    for (const PseudoPot& pp : pseudo) {
      const Int np = (Int)pp.vnlList.size();
      const Int m = np ? pp.vnlList[0].first.first.m() : 0;
      for (const NonlocalPP& v : pp.vnlList) {
        const IntNumVec& iv = v.first.first;
        if (iv.m() != m || std::memcmp(iv.Data(), pp.vnlList[0].first.first.Data(), sizeof(Int) * m) != 0)
          ErrorHandling("projectors of one atom must share their grid points");
        if (v.first.second.m() != m) ErrorHandling("Domain size does not match.");
      }
      npts.push_back(m);
      nproj.push_back(np);
      if (np) idx.insert(idx.end(), pp.vnlList[0].first.first.Data(), pp.vnlList[0].first.first.Data() + m);
      for (const NonlocalPP& v : pp.vnlList) {
        const DblNumMat& dv = v.first.second;
        for (Int i = 0; i < m; i++) val.push_back(dv(i, VAL));
        wgt.push_back(v.second);
      }
    }
    Check(pw_rs_set_nonlocal(rs_, (int)pseudo.size(), npts.data(), nproj.data(), idx.data(), val.data(),
                             wgt.data()));
  }
  // set_wfn_filter (include/hamiltonian.hpp:347-356)
  void set_wfn_filter(int apply_filter, int apply_first, Real wfn_cutoff) {
    apply_filter_ = apply_filter;
    apply_first_ = apply_first;
    wfn_cutoff_ = wfn_cutoff;
  }
  void set_wfn_filter(int apply_first) { apply_first_ = apply_first; }
  // MultSpinor(Spinor&, NumTns<Real>&, Fourier&) (hamiltonian.cpp:3058-3125,
  // 3240-3265) with AddMultSpinorFineR2C (spinor.cpp:376-525): a3 = H psi;
  // with the wavefunction filter on, psi is filtered in place on the first
  // call and a3 on every call.
  void MultSpinor(Spinor& psi, NumTns<Real>& a3, Fourier& fft) override {
    if (!fft.isInitialized) ErrorHandling("Fourier is not prepared.");
    const Int ntot = psi.Wavefun().m(), nb = psi.Wavefun().p();
    if (fft.domain.NumGridTotal() != ntot) ErrorHandling("Domain size does not match3.");
    if (psi.Wavefun().n() != 1)
      ErrorHandling("Component number of spinor should be one in spin-restricted and spin-unrestricted calculations.");
    pw_rs* rs = fft.RsCtx();
    dx_.UploadReal(psi.Wavefun().Data(), (size_t)ntot * nb);
    if (apply_filter_ && apply_first_) {
      apply_first_ = 0;
      Check(pw_rs_filter(rs, wfn_cutoff_, nb, static_cast<double*>(dx_.Ptr())));
      dx_.DownloadReal(psi.Wavefun().Data(), (size_t)ntot * nb);
    }
    dy_.Alloc(((size_t)ntot * nb + 1) / 2);
    Check(pw_rs_mult_spinor(rs, nb, static_cast<const double*>(dx_.Ptr()), static_cast<double*>(dy_.Ptr())));
    if (apply_filter_) Check(pw_rs_filter(rs, wfn_cutoff_, nb, static_cast<double*>(dy_.Ptr())));
    dy_.DownloadReal(a3.Data(), (size_t)ntot * nb);
  }
  DblNumVec& EkinR2C() { return ekinR2C_; }

 private:
  bool apply_filter_ = false, apply_first_ = false;
  Real wfn_cutoff_ = 0.0;
  DblNumVec ekinR2C_, teterR2C_;
};

// ---- EigenSolver (include/eigensolver.hpp) ----
class EigenSolver {
 public:
  void Setup(Hamiltonian& ham, Spinor& psi, Fourier& fft) {
    hamPtr_ = &ham;
    psiPtr_ = &psi;
    fftPtr_ = &fft;
    eigVal_.Resize(psi.NumStateTotal());
    resVal_.Resize(psi.NumStateTotal());
  }
  DblNumVec& EigVal() { return eigVal_; }
  DblNumVec& ResVal() { return resVal_; }
  Spinor& Psi() { return *psiPtr_; }
  Hamiltonian& Ham() { return *hamPtr_; }
  Fourier& FFT() { return *fftPtr_; }

  // NonlocalMultX (eigensolver.cpp:471-515): VnlX = beta (w .* Re(beta^H X))
  void NonlocalMultX(Int heightLocal, Int width, CpxNumMat& X, CpxNumMat& VnlX) {
    dx_.Upload(X.Data(), (size_t)heightLocal * width);
    dy_.Alloc((size_t)heightLocal * width);
    Check(pw_nonlocal_multx(hamPtr_->Ctx(), PW_F64, width, dx_.Ptr(), dy_.Ptr()));
    dy_.Download(VnlX.Data(), (size_t)heightLocal * width);
  }

  // LOBPCGSolveReal (lobpcg.cpp:59-576)
  void LOBPCGSolveReal(Int numEig, Int scfIter, Int eigMaxIter, Real eigMinTolerance, Real eigTolerance) {
    Spinor& psi = *psiPtr_;
    const Int ntot = psi.NumGridTotal(), nb = psi.NumStateTotal();
    dx_.Upload(psi.WavefunG().Data(), (size_t)ntot * nb);
    int iters = 0;
    Check(pw_lobpcg(hamPtr_->Ctx(), nb, numEig, scfIter, eigMaxIter, eigMinTolerance, eigTolerance, dx_.Ptr(),
                    eigVal_.Data(), resVal_.Data(), &iters));
    dx_.Download(psi.WavefunG().Data(), (size_t)ntot * nb);
    hamPtr_->EigVal() = eigVal_;
  }

  // PPCGSolveReal (ppcg.cpp:59-1246): the sphere branch on WavefunG or, with
  // esdfParam.isUseRealSpace, the real-space branch on Wavefun (ppcg.cpp:75,
  // 683); sub-blocks of esdfParam.PPCGsbSize (ppcg.cpp:335).
  void PPCGSolveReal(Int numEig, Int scfIter, Int eigMaxIter, Real eigMinTolerance, Real eigTolerance) {
    Spinor& psi = *psiPtr_;
    int iters = 0;
    if (esdfParam.isUseRealSpace) {
      const Int ntot = psi.Wavefun().m(), nb = psi.Wavefun().p();
      pw_rs* rs = hamPtr_->RsCtx();
      if (!rs) ErrorHandling("Fourier is not prepared.");
      dx_.UploadReal(psi.Wavefun().Data(), (size_t)ntot * nb);
      Check(pw_rs_ppcg(rs, nb, numEig, scfIter, eigMaxIter, eigMinTolerance, eigTolerance, esdfParam.PPCGsbSize,
                       static_cast<double*>(dx_.Ptr()), eigVal_.Data(), resVal_.Data(), &iters));
      dx_.DownloadReal(psi.Wavefun().Data(), (size_t)ntot * nb);
    } else {
      const Int ntot = psi.NumGridTotal(), nb = psi.NumStateTotal();
      dx_.Upload(psi.WavefunG().Data(), (size_t)ntot * nb);
      Check(pw_ppcg(hamPtr_->Ctx(), nb, numEig, scfIter, eigMaxIter, eigMinTolerance, eigTolerance,
                    esdfParam.PPCGsbSize, dx_.Ptr(), eigVal_.Data(), resVal_.Data(), &iters));
      dx_.Download(psi.WavefunG().Data(), (size_t)ntot * nb);
    }
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
    dx_.Upload(psi.WavefunG().Data(), (size_t)ntot * nb);
    Check(pw_rayleigh_ritz(hamPtr_->Ctx(), PW_F64, nb, dx_.Ptr(), eigVal_.Data(), resVal_.Data()));
    for (Int c = 0; c < cycleNum; c++) Sweep(dx_, filterOrder, mixed);
    dx_.Download(psi.WavefunG().Data(), (size_t)ntot * nb);
    hamPtr_->EigVal() = eigVal_;
  }
  void GeneralChebyStep(Int numEig, Int filterOrder, bool mixed = false) {
    (void)numEig;
    Spinor& psi = *psiPtr_;
    const Int ntot = psi.NumGridTotal(), nb = psi.NumStateTotal();
    dx_.Upload(psi.WavefunG().Data(), (size_t)ntot * nb);
    Sweep(dx_, filterOrder, mixed);
    dx_.Download(psi.WavefunG().Data(), (size_t)ntot * nb);
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
    DeviceBlock& v0 = dv0_;
    v0.Upload(lanczosStart_.data(), lanczosStart_.size());
    std::vector<double> prev(eigVal_.Data(), eigVal_.Data() + nb);
    const bool havePrev = !(prev.empty() || (prev.front() == 0.0 && prev.back() == 0.0));
    double bounds[3];
    Check(pw_chebfsi_sweep(hamPtr_->Ctx(), mixed ? PW_F32 : PW_F64, nb, filterOrder, lanczosSteps_, v0.Ptr(),
                           havePrev ? prev.data() : nullptr, x.Ptr(), eigVal_.Data(), resVal_.Data(), bounds));
  }

  Hamiltonian* hamPtr_ = nullptr;
  Spinor* psiPtr_ = nullptr;
  DeviceBlock dx_, dy_, dv0_;  // staging blocks, grown on demand
  Fourier* fftPtr_ = nullptr;
  DblNumVec eigVal_, resVal_;
  std::vector<Complex> lanczosStart_;
  Int lanczosSteps_ = 8;
};

}  // namespace pwdft
