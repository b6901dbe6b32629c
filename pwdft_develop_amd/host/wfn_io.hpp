// Restart files of the wavefunctions in the reference's serialize layout
// (include/utility.hpp:761-830 scalars, :1088-1094 IntNumVec, :1373-1381
// NumVec, :1439-1452 NumTns; written by SCF::Iterate, src/SCF/scf.cpp:
// 1112-1151, read at :294-330; file names from SeparateWrite/SeparateRead
// with an output index, src/utility.cpp:465-510: "<name>_<index>").
//
//   Int     4 bytes, native order          Real   8 bytes
//   Complex 16 bytes (real, imaginary)
//   NumVec<T> / IntNumVec : Int m, then m values
//   NumTns<T>             : Int m, n, p, then the values with i fastest,
//                           then j, then k (column-major)
//
// Records written here follow the reference's complex-wavefunction branch
// (scf.cpp:1115-1142), which is the one that carries sphere-cut
// coefficients: WFN_<index> = [CpxNumTns npw x ncom x nband][DblNumVec
// occupation]; IDX_<index> = [IntNumVec idxCoarseCut].  (The reference's
// real Gamma build writes Psi().Wavefun(), its real-space block, which is
// left empty when the sphere-cut coefficients live in WavefunG(); the
// coefficients are what a restart of this path needs.)
//
// Header-only host code (no GPU); errors throw std::runtime_error like the
// reference's ErrorHandling (src/global.cpp:71-91).
#pragma once
#include <complex>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iterator>
#include <stdexcept>
#include <string>
#include <vector>

namespace pwdft_amd {
namespace wfn {

typedef int32_t Int;
typedef std::complex<double> Complex;

inline std::string file_name(const std::string& name, int index) { return name + "_" + std::to_string(index); }

class Writer {
 public:
  template <typename T>
  void scalar(const T& v) {
    const char* p = reinterpret_cast<const char*>(&v);
    buf_.insert(buf_.end(), p, p + sizeof(T));
  }
  // NumVec<T> / IntNumVec
  template <typename T>
  void vec(const T* v, Int m) {
    scalar(m);
    const char* p = reinterpret_cast<const char*>(v);
    buf_.insert(buf_.end(), p, p + sizeof(T) * (size_t)m);
  }
  // NumTns<T>, column-major m x n x p
  template <typename T>
  void tns(const T* v, Int m, Int n, Int p) {
    scalar(m);
    scalar(n);
    scalar(p);
    const char* q = reinterpret_cast<const char*>(v);
    buf_.insert(buf_.end(), q, q + sizeof(T) * (size_t)m * n * p);
  }
  const std::vector<char>& bytes() const { return buf_; }
  void save(const std::string& filename) const {
    std::ofstream f(filename, std::ios::binary | std::ios::trunc);
    if (!f.good()) throw std::runtime_error("File cannot be open!");
    f.write(buf_.data(), (std::streamsize)buf_.size());
  }

 private:
  std::vector<char> buf_;
};

class Reader {
 public:
  explicit Reader(const std::string& filename) {
    std::ifstream f(filename, std::ios::binary);
    if (!f.good()) throw std::runtime_error("File cannot be open!");
    buf_.assign(std::istreambuf_iterator<char>(f), std::istreambuf_iterator<char>());
  }
  template <typename T>
  T scalar() {
    T v;
    take(&v, sizeof(T));
    return v;
  }
  template <typename T>
  std::vector<T> vec() {
    const Int m = scalar<Int>();
    if (m < 0) throw std::runtime_error("wfn: negative vector length");
    std::vector<T> v((size_t)m);
    take(v.data(), sizeof(T) * (size_t)m);
    return v;
  }
  template <typename T>
  std::vector<T> tns(Int* m, Int* n, Int* p) {
    *m = scalar<Int>();
    *n = scalar<Int>();
    *p = scalar<Int>();
    if (*m < 0 || *n < 0 || *p < 0) throw std::runtime_error("wfn: negative tensor dimension");
    std::vector<T> v((size_t)*m * *n * *p);
    take(v.data(), sizeof(T) * v.size());
    return v;
  }
  bool done() const { return pos_ == buf_.size(); }

 private:
  void take(void* dst, size_t nbytes) {
    if (pos_ + nbytes > buf_.size()) throw std::runtime_error("wfn: file shorter than its records");
    std::memcpy(dst, buf_.data() + pos_, nbytes);
    pos_ += nbytes;
  }
  std::vector<char> buf_;
  size_t pos_ = 0;
};

// WFN_<index>: coefficients (npw x ncom x nband, column-major) and the
// occupation rates; IDX_<index>: the sphere's grid indices.
inline void write_restart(const std::string& wfn_name, const std::string& idx_name, int index, const Complex* psi,
                          Int npw, Int ncom, Int nband, const double* occ, Int nocc, const Int* idx) {
  Writer w;
  w.tns(psi, npw, ncom, nband);
  w.vec(occ, nocc);
  w.save(file_name(wfn_name, index));
  Writer wi;
  wi.vec(idx, npw);
  wi.save(file_name(idx_name, index));
}

struct Restart {
  Int npw = 0, ncom = 0, nband = 0;
  std::vector<Complex> psi;
  std::vector<double> occ;
  std::vector<Int> idx;
};

inline Restart read_restart(const std::string& wfn_name, const std::string& idx_name, int index) {
  Restart r;
  Reader rd(file_name(wfn_name, index));
  r.psi = rd.tns<Complex>(&r.npw, &r.ncom, &r.nband);
  r.occ = rd.vec<double>();
  if (!rd.done()) throw std::runtime_error("wfn: trailing bytes after the occupation rates");
  Reader ri(file_name(idx_name, index));
  r.idx = ri.vec<Int>();
  if ((Int)r.idx.size() != r.npw) throw std::runtime_error("wfn: index file does not match the coefficients");
  return r;
}

}  // namespace wfn
}  // namespace pwdft_amd
