// H*psi on gfx950: sphere-cut Gamma-point local + kinetic operator as three
// kernels over a stick (z-column) decomposition of the half sphere, plus the
// nonlocal projector GEMMs.
//
// Replaces Spinor::AddMultSpinorFineR2C (src/SCF/spinor.cpp:527-680): per
// band, scatter into the padded R2C array, C2R, x 1/N, x V, R2C, gather with
// sqrt(2) and add ekin .* c.  The reference does two dense 3D FFTs per band
// (FFTW plan built per band).  Here the transform is pruned to where the data
// lives: only the nstick sticks of the half sphere are transformed along z
// (kernels 1 and 3) and only kx <= kxmax columns along y (kernel 2); the x
// transform pairs two real rows into one complex Stockham FFT.  Many bands go
// per launch (grid.y = band).  Results equal the dense transform's up to
// rounding: the skipped entries are exact zeros.
//
//   kernel 1  sticks_inverse_z : gather sphere (+ conj mirror of the (0,0)
//             column) -> LDS pencils, inverse z FFT -> T1[band][z][stick]
//   kernel 2  slab_xy          : per (z, band): scatter sticks (+ conj
//             mirrors of the kx = 0 plane), inverse y FFT, paired-row x c2r,
//             x V/N, x r2c, forward y FFT, gather sticks -> T2
//   kernel 3  sticks_forward_z : T2 -> forward z FFT -> sphere gather with
//             sqrt(2), + ekin .* c -> a3
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <vector>

#include "pw_fft.h"
#include "pw_internal.h"
#include "pw_slab_reg.h"

namespace pw {

// ---------------------------------------------------------------- kernel 1
template <typename T, int N3>
__global__ void __launch_bounds__(256)
k_sticks_inverse_z(const typename cx<T>::t* __restrict__ psi, int ldpsi,
                   typename cx<T>::t* __restrict__ t1, int n3_, int nstick, int sc, int nchunk,
                   const int* __restrict__ chunk_off, const int* __restrict__ pt_idx,
                   const int* __restrict__ pt_sz, const typename cx<T>::t* __restrict__ tw3g) {
  typedef typename cx<T>::t C;
  const int n3 = N3 > 0 ? N3 : n3_;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* tw = reinterpret_cast<C*>(smem_raw);
  C* A = tw + n3;
  C* B = A + sc * n3;
  // one row of nchunk x nb blocks; the chunks of a band in order on one XCD
  // (neighbouring chunks share the cache lines at their ends)
  const int item = xcd_item(blockIdx.x, gridDim.x);
  const int band = item / nchunk, c = item - band * nchunk;
  const int s0 = c * sc;
  const C* ps = psi + (size_t)band * ldpsi;
  for (int t = threadIdx.x; t < n3; t += blockDim.x) tw[t] = tw3g[t];
  for (int t = threadIdx.x; t < sc * n3; t += blockDim.x) A[t] = C{0, 0};
  __syncthreads();
  const T is2 = (T)0.70710678118654752440084436210485;
  const int e = chunk_off[c + 1];
  for (int t = chunk_off[c] + threadIdx.x; t < e; t += blockDim.x) {
    const int i = pt_idx[t];
    const int sz = pt_sz[t];
    C v;
    if (i >= 0) {
      v = ps[i];
      if (i != 0) v = cscale(v, is2);
    } else {
      C u = ps[-i - 1];
      v = C{u.x * is2, -u.y * is2};
    }
    A[(sz & 0xffff) * n3 + (sz >> 16)] = v;
  }
  __syncthreads();
  C* R = lds_fft_n<T, N3>(A, B, n3, sc, n3, tw, (T)1);
  const int nsl = min(sc, nstick - s0);
  for (int t = threadIdx.x; t < n3 * nsl; t += blockDim.x) {
    const int z = t / nsl, sl = t - z * nsl;
    t1[((size_t)band * n3 + z) * nstick + s0 + sl] = R[sl * n3 + z];
  }
}

// ---------------------------------------------------------------- kernel 2
template <typename T, int N1, int N2>
__global__ void __launch_bounds__(1024)  // 256 threads, or up to 1024 for planes that hold one workgroup per CU
k_slab_xy(const typename cx<T>::t* __restrict__ t1, typename cx<T>::t* __restrict__ t2, int n1_,
          int n2_, int n3, int nstick, int kxmax, const int* __restrict__ pos,
          const int* __restrict__ mpos, const T* __restrict__ vtot,
          const typename cx<T>::t* __restrict__ tw1g, const typename cx<T>::t* __restrict__ tw2g,
          int rp, int cg) {
  typedef typename cx<T>::t C;
  const int n1 = N1 > 0 ? N1 : n1_;  // compile-time when the grid is one of the dispatched sizes
  const int n2 = N2 > 0 ? N2 : n2_;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int nk = kxmax + 1;
  C* tw1 = reinterpret_cast<C*>(smem_raw);
  C* tw2 = tw1 + n1;
  C* Y0 = tw2 + n2;
  // cg == 0: two y buffers (ping-pong over all nk columns) and two x-row
  // buffers.  cg > 0 (planes whose double y-space does not fit in LDS): one
  // y buffer, transformed cg columns at a time through a scratch that the
  // x pass reuses for its two row buffers.
  C* Y1 = cg == 0 ? Y0 + nk * n2 : nullptr;
  C* Z0 = cg == 0 ? Y1 + nk * n2 : Y0 + nk * n2;
  C* Z1 = Z0 + rp * n1;
  const int z = blockIdx.x, band = blockIdx.y;
  for (int t = threadIdx.x; t < n1; t += blockDim.x) tw1[t] = tw1g[t];
  for (int t = threadIdx.x; t < n2; t += blockDim.x) tw2[t] = tw2g[t];
  for (int t = threadIdx.x; t < nk * n2; t += blockDim.x) Y0[t] = C{0, 0};
  __syncthreads();
  const C* src = t1 + ((size_t)band * n3 + z) * nstick;
  for (int s = threadIdx.x; s < nstick; s += blockDim.x) {
    const C v = src[s];
    Y0[pos[s]] = v;
    const int mp = mpos[s];
    if (mp >= 0) Y0[mp] = C{v.x, -v.y};
  }
  __syncthreads();
  // inverse y on the kx <= kxmax columns
  C* Y = slab_y_fft<T, N2>(Y0, Y1, Z0, n2, nk, cg, tw2, (T)1);
  const int npair = (n2 + 1) / 2;
  const T* vz = vtot + (size_t)z * n1 * n2;
  const int h = n1 / 2;
  for (int q0 = 0; q0 < npair; q0 += rp) {
    const int nq = min(rp, npair - q0);
    // Hermitian-extended spectra of rows y1 = 2q, y2 = 2q+1 packed as X1 + i X2
    for (int t = threadIdx.x; t < nq * n1; t += blockDim.x) {
      const int ql = t / n1, k = t - ql * n1;
      const int y1 = 2 * (q0 + ql), y2 = y1 + 1;
      C a1 = C{0, 0}, a2 = C{0, 0};
      int kk = k;
      bool cj = false;
      if (k > h) { kk = n1 - k; cj = true; }
      if (kk <= kxmax) {
        a1 = Y[kk * n2 + y1];
        if (y2 < n2) a2 = Y[kk * n2 + y2];
        if (cj) { a1.y = -a1.y; a2.y = -a2.y; }
        if (k == 0 || (2 * k == n1)) { a1.y = 0; a2.y = 0; }
      }
      Z0[ql * n1 + k] = C{a1.x - a2.y, a1.y + a2.x};
    }
    __syncthreads();
    C* Zr = lds_fft_n<T, N1>(Z0, Z1, n1, nq, n1, tw1, (T)1);
    // real rows -> x V/N (V already carries the 1/N of spinor.cpp:620-621)
    for (int t = threadIdx.x; t < nq * n1; t += blockDim.x) {
      const int ql = t / n1, x = t - ql * n1;
      const int y1 = 2 * (q0 + ql), y2 = y1 + 1;
      C u = Zr[ql * n1 + x];
      const T v1 = vz[(size_t)y1 * n1 + x];
      const T v2 = (y2 < n2) ? vz[(size_t)y2 * n1 + x] : (T)0;
      Zr[ql * n1 + x] = C{u.x * v1, u.y * v2};
    }
    __syncthreads();
    C* Zo = (Zr == Z0) ? Z1 : Z0;
    C* Zf = lds_fft_n<T, N1>(Zr, Zo, n1, nq, n1, tw1, (T)-1);
    // unpack the two real-row spectra for kx <= kxmax back into Y
    for (int t = threadIdx.x; t < nq * nk; t += blockDim.x) {
      const int ql = t / nk, k = t - ql * nk;
      const int y1 = 2 * (q0 + ql), y2 = y1 + 1;
      const C w = Zf[ql * n1 + k];
      const C wc = Zf[ql * n1 + (k == 0 ? 0 : n1 - k)];
      // U1 = (w + conj(wc))/2, U2 = (w - conj(wc))/(2i)
      Y[k * n2 + y1] = C{(T)0.5 * (w.x + wc.x), (T)0.5 * (w.y - wc.y)};
      if (y2 < n2) Y[k * n2 + y2] = C{(T)0.5 * (w.y + wc.y), (T)0.5 * (wc.x - w.x)};
    }
    __syncthreads();
  }
  C* Yf = slab_y_fft<T, N2>(Y, (Y == Y0) ? Y1 : Y0, Z0, n2, nk, cg, tw2, (T)-1);
  C* dst = t2 + ((size_t)band * n3 + z) * nstick;
  for (int s = threadIdx.x; s < nstick; s += blockDim.x) dst[s] = Yf[pos[s]];
}

// ------------------------------------------------ kernel 2, register variant
// The same slab transform for compile-time N1 = N2 with every 1D FFT held by
// one thread in registers (reg_fft: constant twiddles).  A workgroup takes one z plane of SB
// bands: V's plane is staged once in LDS as (row 2q, row 2q + 1) pairs, one
// 64-bit word per x, pair rows padded to N1 + 1 words so that neighbouring
// threads hit distinct banks; each band's y-space lives in LDS, and the five
// phases (scatter, y IFFT per column, x c2r / V / r2c per row pair, y FFT per
// column, gather) are separated by one barrier each.  In the row-pair phase a
// thread holds N1 complex values; its LDS reads are issued in chunks fenced
// by scheduling barriers so that the compiler does not hoist all of them next
// to the spectrum (that spills past the 256 VGPRs of two waves per SIMD).
// Same arithmetic, in the same order, as k_slab_xy.
// LY lanes per y column transform (1, 2 or 4); NT threads per workgroup.
template <typename T, int N1, int N2, int SB, int LB, int LY, int NT = 256>
__global__ void __launch_bounds__(NT, LB)  // LB waves per SIMD: <= 512 / LB registers
k_slab_xy_reg(const typename cx<T>::t* __restrict__ t1, typename cx<T>::t* __restrict__ t2, int n3, int nstick,
              int kxmax, int nb, const int* __restrict__ pos, const int* __restrict__ mpos,
              const T* __restrict__ vtot) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  slab_xy_reg_body<T, N1, N2, SB, LY, NT>(blockIdx.x, blockIdx.y * SB, t1, t2, n3, nstick, kxmax, nb, pos, mpos,
                                          vtot, smem_raw);
}

// The same transform as a persistent grid over (plane, band group) items
// with the next item prefetched (slab_xy_reg_persistent).
template <typename T, int N1, int N2, int SB, int LB, int LY>
__global__ void __launch_bounds__(256, LB)
k_slab_xy_reg_p(const typename cx<T>::t* __restrict__ t1, typename cx<T>::t* __restrict__ t2, int n3, int nstick,
                int kxmax, int nb, const int* __restrict__ pos, const int* __restrict__ mpos,
                const T* __restrict__ vtot) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  slab_xy_reg_persistent<T, N1, N2, SB, LY, 256>(t1, t2, n3, nstick, kxmax, nb, pos, mpos, vtot, smem_raw);
}

// ---------------------------------------------------------------- kernel 3
template <typename T, int N3>
__global__ void __launch_bounds__(256)
k_sticks_forward_z(const typename cx<T>::t* __restrict__ t2, const typename cx<T>::t* __restrict__ psi,
                   typename cx<T>::t* __restrict__ out, int ld, int n3_, int nstick, int sc, int nchunk,
                   const int* __restrict__ gchunk_off, const int* __restrict__ gpt_idx,
                   const int* __restrict__ gpt_sz, const T* __restrict__ ekin,
                   const typename cx<T>::t* __restrict__ tw3g) {
  typedef typename cx<T>::t C;
  const int n3 = N3 > 0 ? N3 : n3_;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* tw = reinterpret_cast<C*>(smem_raw);
  C* A = tw + n3;
  C* B = A + sc * n3;
  const int item = xcd_item(blockIdx.x, gridDim.x);  // as in k_sticks_inverse_z
  const int band = item / nchunk, c = item - band * nchunk;
  const int s0 = c * sc;
  const int nsl = min(sc, nstick - s0);
  for (int t = threadIdx.x; t < n3; t += blockDim.x) tw[t] = tw3g[t];
  for (int t = threadIdx.x; t < n3 * sc; t += blockDim.x) {
    const int z = t / sc, sl = t - z * sc;
    A[sl * n3 + z] = (sl < nsl) ? t2[((size_t)band * n3 + z) * nstick + s0 + sl] : C{0, 0};
  }
  __syncthreads();
  C* R = lds_fft_n<T, N3>(A, B, n3, sc, n3, tw, (T)-1);
  const C* ps = psi + (size_t)band * ld;
  C* o = out + (size_t)band * ld;
  const T s2 = (T)1.4142135623730950488016887242097;
  const int e = gchunk_off[c + 1];
  for (int t = gchunk_off[c] + threadIdx.x; t < e; t += blockDim.x) {
    const int i = gpt_idx[t];
    const int sz = gpt_sz[t];
    C y = R[(sz & 0xffff) * n3 + (sz >> 16)];
    if (i != 0) y = cscale(y, s2);
    if (ekin) {
      const C cc = ps[i];
      const T ek = ekin[i];
      o[i] = C{ek * cc.x + y.x, ek * cc.y + y.y};
    } else {  // accumulate: a3 += gather (spinor.cpp:933-935)
      const C a = o[i];
      o[i] = C{a.x + y.x, a.y + y.y};
    }
  }
}

// Register forms of kernels 1 and 3 for compile-time N3: each stick's z
// column is split over L adjacent lanes (lanes_fft: L = 4 where 4 divides
// N3, else 2), lane h holding z = m L + h on input and the block q M ..
// q M + M - 1 on output; the sphere side is read or written through the
// (z, stick) table StickPlan::zs.  Thread e = L (band nstick + stick) + h,
// so that at each step neighbouring lane groups take neighbouring sticks
// (adjacent t1 / t2 entries; adjacent points of the sphere's (kz, ky, kx)
// order).  No LDS, no barrier.
template <int N3>
struct ZLanes {
  static constexpr int L = N3 % 4 == 0 ? 4 : 2, M = N3 / L;
};

template <typename T, int N3>
__global__ void __launch_bounds__(256)
k_sticks_inverse_z_reg(const typename cx<T>::t* __restrict__ psi, int ldpsi, typename cx<T>::t* __restrict__ t1,
                       int nstick, long nwork, const int* __restrict__ zs) {
  typedef typename cx<T>::t C;
  constexpr int L = ZLanes<N3>::L, M = ZLanes<N3>::M;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long w = e / L;  // (band, stick); the L lanes of one stick exit together
  if (w >= nwork) return;
  const int h = (int)(e - w * L);
  const int band = (int)(w / nstick), s = (int)(w - (long)band * nstick);
  const C* ps = psi + (size_t)band * ldpsi;
  const T is2 = (T)0.70710678118654752440084436210485;
  C a[M];
#pragma unroll
  for (int m = 0; m < M; m++) {
    const int i = zs[(size_t)(m * L + h) * nstick + s];
    C v = C{0, 0};
    if (i >= 0) {
      v = ps[i];
      if (i != 0) v = cscale(v, is2);
    } else if (i != kZsAbsent) {
      const C u = ps[-i - 1];
      v = C{u.x * is2, -u.y * is2};
    }
    a[m] = v;
  }
  const int q = lanes_fft<T, N3, 1, L>(a, h);
  C* o = t1 + ((size_t)band * N3 + q * M) * nstick + s;
#pragma unroll
  for (int k = 0; k < M; k++) o[(size_t)k * nstick] = a[k];
}

template <typename T, int N3>
__global__ void __launch_bounds__(256)
k_sticks_forward_z_reg(const typename cx<T>::t* __restrict__ t2, const typename cx<T>::t* __restrict__ psi,
                       typename cx<T>::t* __restrict__ out, int ld, int nstick, long nwork,
                       const int* __restrict__ zs, const T* __restrict__ ekin) {
  typedef typename cx<T>::t C;
  constexpr int L = ZLanes<N3>::L, M = ZLanes<N3>::M;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  const long w = e / L;
  if (w >= nwork) return;
  const int h = (int)(e - w * L);
  const int band = (int)(w / nstick), s = (int)(w - (long)band * nstick);
  const C* src = t2 + (size_t)band * N3 * nstick + s;
  C a[M];
#pragma unroll
  for (int m = 0; m < M; m++) a[m] = src[(size_t)(m * L + h) * nstick];
  const int q = lanes_fft<T, N3, -1, L>(a, h);
  const C* ps = psi + (size_t)band * ld;
  C* o = out + (size_t)band * ld;
  const T s2 = (T)1.4142135623730950488016887242097;
#pragma unroll
  for (int k = 0; k < M; k++) {
    const int i = zs[(size_t)(q * M + k) * nstick + s];
    if (i >= 0) {  // direct points only (mirrors and absent entries are negative)
      C y = a[k];
      if (i != 0) y = cscale(y, s2);
      if (ekin) {
        const C cc = ps[i];
        const T ek = ekin[i];
        o[i] = C{ek * cc.x + y.x, ek * cc.y + y.y};
      } else {  // accumulate: a3 += gather (spinor.cpp:933-935)
        const C acc = o[i];
        o[i] = C{acc.x + y.x, acc.y + y.y};
      }
    }
  }
}

// Grid sizes whose FFT lengths are compiled in (all lengths of the built
// workloads); other 2/3/5-smooth sizes use the runtime-length engine.
#define PW_FFT_SIZES(X) X(24) X(30) X(32) X(36) X(40) X(48) X(60) X(72) X(96) X(180)

template <typename T>
auto slab_kernel(int n1, int n2) -> decltype(&k_slab_xy<T, 0, 0>) {
  if (n1 == n2) {
    switch (n1) {
#define PW_CASE(n) \
  case n:          \
    return &k_slab_xy<T, n, n>;
      PW_FFT_SIZES(PW_CASE)
#undef PW_CASE
      default:
        break;
    }
  }
  return &k_slab_xy<T, 0, 0>;
}

// Register slab kernel: FP32 square planes of these sizes (a row of N1
// complex floats per thread stays within the register budget up to 72).
// bands per z plane and waves per SIMD of the register slab kernel: FP32
// 8 bands, two waves (4 bands at three waves measured slower: 116 against
// 98 us at 60^3); FP64 4 bands (the y-space of 8 would not leave LDS for a
// second workgroup either), two waves up to 48 points, one beyond (the
// 60-point double FFT pair needs about 350 registers)
template <typename T> struct SlabCfg { static constexpr int SB = 8; };
constexpr int kSlabWideSB = 16;  // the wide FP32 variant: 16 bands per plane, 512 threads
template <> struct SlabCfg<double> { static constexpr int SB = 4; };
template <typename T, int N> constexpr int slab_lb() { return (sizeof(T) == 4 || N <= 48) ? 2 : 1; }
// lanes per y column transform when no plan chooses
template <typename T> static int default_slab_ly(int n) {
  (void)n;
  return 1;
}
#define PW_REG_SIZES(X) X(24) X(30) X(32) X(36) X(40) X(48) X(60) X(64) X(72)
template <typename T>
using slab_reg_fn = decltype(&k_slab_xy_reg<T, 24, 24, SlabCfg<T>::SB, 2, 1>);
template <typename T, int N>
static slab_reg_fn<T> slab_reg_ly(int ly, int sb, bool persist) {
  constexpr int SB = SlabCfg<T>::SB, LB = slab_lb<T, N>();
  if (persist) {
    // persistent variant: FP32 only (the FP64 prefetch registers spill),
    // the default band count, one lane per column
    if constexpr (sizeof(T) == 4) {
      if (ly == 1 && sb == SB) return &k_slab_xy_reg_p<T, N, N, SB, LB, 1>;
    }
    return nullptr;
  }
  if (sb == kSlabWideSB) {
    // FP32 planes up to 64 points: 16 bands per z plane, 512 threads
    if constexpr (sizeof(T) == 4 && N <= 64)
      if (ly == 1) return &k_slab_xy_reg<T, N, N, kSlabWideSB, 2, 1, 512>;
    return nullptr;
  }
  if (ly == 1) return &k_slab_xy_reg<T, N, N, SB, LB, 1>;
  if (ly == 2) return &k_slab_xy_reg<T, N, N, SB, LB, 2>;
  if constexpr (N % 4 == 0)
    if (ly == 4) return &k_slab_xy_reg<T, N, N, SB, LB, 4>;
  return nullptr;
}
// the register slab kernel of an n1 x n2 plane with ly lanes per y column
// transform and sb bands per z plane (<= 0: the defaults for the precision),
// or null if there is none
template <typename T>
static slab_reg_fn<T> slab_reg_kernel(int n1, int n2, int ly = 0, int sb = 0, bool persist = false) {
  if (n1 != n2) return nullptr;
  if (ly <= 0) ly = default_slab_ly<T>(n1);
  if (sb <= 0) sb = SlabCfg<T>::SB;
  switch (n1) {
#define PW_CASE(n) \
  case n:          \
    return slab_reg_ly<T, n>(ly, sb, persist);
    PW_REG_SIZES(PW_CASE)
#undef PW_CASE
    default:
      return nullptr;
  }
}
// LDS bytes of the register slab kernel at sb bands per plane
static size_t slab_reg_lds(const pw_ctx* c, size_t csz, int sb) {
  return csz * ((size_t)sb * (c->sp.kxmax + 1) * c->n2) + csz * (size_t)(c->n2 / 2) * (c->n1 + 1) +
         sizeof(int) * 2 * (size_t)c->sp.nstick;
}

template <typename T>
auto inv_z_kernel(int n3) -> decltype(&k_sticks_inverse_z<T, 0>) {
  switch (n3) {
#define PW_CASE(n) \
  case n:          \
    return &k_sticks_inverse_z<T, n>;
    PW_FFT_SIZES(PW_CASE)
#undef PW_CASE
    default:
      return &k_sticks_inverse_z<T, 0>;
  }
}

// register z kernels (even z lengths of the compiled list), else null
template <typename T>
auto inv_z_reg_kernel(int n3) -> decltype(&k_sticks_inverse_z_reg<T, 24>) {
  switch (n3) {
#define PW_CASE(n)                                                        \
  case n:                                                                 \
    if constexpr (n % 2 == 0) return &k_sticks_inverse_z_reg<T, n>; \
    return nullptr;
    PW_FFT_SIZES(PW_CASE)
#undef PW_CASE
    default:
      return nullptr;
  }
}
template <typename T>
auto fwd_z_reg_kernel(int n3) -> decltype(&k_sticks_forward_z_reg<T, 24>) {
  switch (n3) {
#define PW_CASE(n)                                                        \
  case n:                                                                 \
    if constexpr (n % 2 == 0) return &k_sticks_forward_z_reg<T, n>; \
    return nullptr;
    PW_FFT_SIZES(PW_CASE)
#undef PW_CASE
    default:
      return nullptr;
  }
}

template <typename T>
int launch_inv_z(pw_ctx* c, int nb, const typename cx<T>::t* psi, typename cx<T>::t* t1,
                 const typename cx<T>::t* tw3);
template <typename T>
int launch_fwd_z(pw_ctx* c, int nb, const typename cx<T>::t* t2, const typename cx<T>::t* psi,
                 typename cx<T>::t* out, const T* ek, const typename cx<T>::t* tw3);

template <typename T>
auto fwd_z_kernel(int n3) -> decltype(&k_sticks_forward_z<T, 0>) {
  switch (n3) {
#define PW_CASE(n) \
  case n:          \
    return &k_sticks_forward_z<T, n>;
    PW_FFT_SIZES(PW_CASE)
#undef PW_CASE
    default:
      return &k_sticks_forward_z<T, 0>;
  }
}

template <typename T>
__global__ void k_teter(const typename cx<T>::t* __restrict__ psi, typename cx<T>::t* __restrict__ a3,
                        const T* __restrict__ teter, int npw, long total) {
  typedef typename cx<T>::t C;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total; t += (long)gridDim.x * blockDim.x) {
    const int i = (int)(t % npw);
    const C v = psi[t];
    C a = a3[t];
    a.x += v.x * teter[i];
    a.y += v.y * teter[i];
    a3[t] = a;
  }
}

}  // namespace pw

using namespace pw;

// ----------------------------------------------------------------- host side

// ------------------------------------------------ allocation-free projector product
// Scratch of pw_nonlocal_multx_noalloc: the projections W (nbeta x nb, then
// padded to 256 bytes) followed by the split-K partial tiles.  A file-scope
// device array: each GPU the library runs on holds its own copy until the
// process ends.  Host code never passes the array's name to a kernel (on the
// host it is only a stand-in for the symbol); the device address is looked
// up once per device, with that device current, when a context is created,
// and that pointer is what the kernels get.
constexpr size_t kNoAllocScratchBytes = (size_t)32 << 20;
constexpr int kMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static void* g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mu;

static int scratch_lookup(int device) {
  if (device < 0 || device >= kMaxDevices) return fail(PW_ERR_UNSUPPORTED, "device %d beyond %d", device, kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_ptr[device]) return PW_OK;
  void* p = nullptr;
  const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) return fail(PW_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
  g_scratch_ptr[device] = p;
  return PW_OK;
}

template <typename T>
__global__ void k_zero_block(T* __restrict__ p, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = (T)0;
}

template <typename T>
static int nonlocal_noalloc(pw_ctx* c, int nb, const void* X, void* V, void* sp) {
  const int m = 2 * c->npw;
  if (c->nbeta == 0) {
    const long n = (long)m * nb;
    hipLaunchKernelGGL(k_zero_block<T>, dim3((int)std::min<long>((n + 255) / 256, 8192)), dim3(256), 0, c->stream,
                       (T*)V, n);
    PW_HIP(hipGetLastError());
    return PW_OK;
  }
  const size_t wbytes = (sizeof(T) * (size_t)c->nbeta * nb + 255) / 256 * 256;
  if (wbytes >= kNoAllocScratchBytes)
    return fail(PW_ERR_UNSUPPORTED,
                "pw_nonlocal_multx_noalloc: %d projectors x %d bands need more than the %zu-byte scratch", c->nbeta,
                nb, kNoAllocScratchBytes);
  T* W = (T*)sp;
  void* ws = (unsigned char*)sp + wbytes;
  const T* beta = (const T*)(sizeof(T) == 8 ? (const void*)c->d_beta : (const void*)c->d_beta_f);
  const T* betaw = (const T*)(sizeof(T) == 8 ? (const void*)c->d_betaw : (const void*)c->d_betaw_f);
  PW_TRY(gemm_tn_ws<T>(c, m, c->nbeta, nb, beta, m, (const T*)X, m, W, c->nbeta, nullptr, (T)0, false, ws,
                       kNoAllocScratchBytes - wbytes));
  return gemm_nn<T>(c, m, nb, c->nbeta, (T)1, betaw, m, W, c->nbeta, (T)0, (T*)V, m);
}

static void build_stick_plan(pw_ctx* c, const int* idx) {
  StickPlan& sp = c->sp;
  const int h1 = c->h1, n2 = c->n2, n3 = c->n3, npw = c->npw;
  // stick id per (kx, ky): sticks ordered by (ky storage index, kx)
  std::vector<int> sid((size_t)h1 * n2, -1);
  std::vector<char> used((size_t)h1 * n2, 0);
  for (int i = 0; i < npw; i++) {
    const int g = idx[i];
    used[g % (h1 * n2)] = 1;
  }
  sp.kx.clear();
  sp.ky.clear();
  sp.kxmax = 0;
  for (int ky = 0; ky < n2; ky++)
    for (int kx = 0; kx < h1; kx++)
      if (used[kx + (size_t)h1 * ky]) {
        sid[kx + (size_t)h1 * ky] = (int)sp.kx.size();
        sp.kx.push_back(kx);
        sp.ky.push_back(ky);
        if (kx > sp.kxmax) sp.kxmax = kx;
      }
  sp.nstick = (int)sp.kx.size();
  sp.nchunk = (sp.nstick + sp.sc - 1) / sp.sc;
  sp.pos.resize(sp.nstick);
  sp.mpos.resize(sp.nstick);
  for (int s = 0; s < sp.nstick; s++) {
    sp.pos[s] = sp.kx[s] * n2 + sp.ky[s];
    sp.mpos[s] = (sp.kx[s] == 0 && sp.ky[s] > 0) ? (n2 - sp.ky[s]) % n2 : -1;
  }
  // point lists per chunk, (z, stick)-ordered
  struct P { int z, sl, i; };
  std::vector<std::vector<P>> fw(sp.nchunk), inv(sp.nchunk);
  for (int i = 0; i < npw; i++) {
    const int g = idx[i];
    const int col = g % (h1 * n2), kz = g / (h1 * n2);
    const int s = sid[col];
    const int ch = s / sp.sc, sl = s % sp.sc;
    fw[ch].push_back({kz, sl, i});
    inv[ch].push_back({kz, sl, i});
    // the (0,0) column holds kz >= 0 only; its conj mirror fills -kz
    if (sp.kx[s] == 0 && sp.ky[s] == 0 && kz != 0) inv[ch].push_back({(n3 - kz) % n3, sl, -(i + 1)});
  }
  auto flatten = [](std::vector<std::vector<P>>& L, std::vector<int>& off, std::vector<int>& pi,
                    std::vector<int>& psz) {
    off.assign(L.size() + 1, 0);
    pi.clear();
    psz.clear();
    for (size_t c = 0; c < L.size(); c++) {
      std::sort(L[c].begin(), L[c].end(), [](const P& a, const P& b) {
        return a.z != b.z ? a.z < b.z : a.sl < b.sl;
      });
      for (const P& p : L[c]) {
        pi.push_back(p.i);
        psz.push_back(p.sl | (p.z << 16));
      }
      off[c + 1] = (int)pi.size();
    }
  };
  flatten(inv, sp.chunk_off, sp.pt_idx, sp.pt_sz);
  flatten(fw, sp.gchunk_off, sp.gpt_idx, sp.gpt_sz);
  // (z, stick) table: direct points first, so that a mirror never replaces
  // one (they could meet only on a kz = n3/2 plane)
  sp.zs.assign((size_t)n3 * sp.nstick, kZsAbsent);
  for (int i = 0; i < npw; i++) {
    const int g = idx[i];
    sp.zs[(size_t)(g / (h1 * n2)) * sp.nstick + sid[g % (h1 * n2)]] = i;
  }
  for (int i = 0; i < npw; i++) {
    const int g = idx[i], col = g % (h1 * n2), kz = g / (h1 * n2), s = sid[col];
    if (sp.kx[s] == 0 && sp.ky[s] == 0 && kz != 0) {
      int& e = sp.zs[(size_t)((n3 - kz) % n3) * sp.nstick + s];
      if (e == kZsAbsent) e = -(i + 1);
    }
  }
}

template <typename X>
static int upload(X** d, const X* h, size_t n) {
  PW_HIP(hipMalloc((void**)d, sizeof(X) * (n > 0 ? n : 1)));
  if (n) PW_HIP(hipMemcpy(*d, h, sizeof(X) * n, hipMemcpyHostToDevice));
  return PW_OK;
}

static void roots(int n, std::vector<double>& out) {
  for (int m = 0; m < n; m++) {
    long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)n;
    out.push_back((double)cosl(a));
    out.push_back((double)sinl(a));
  }
}

extern "C" int pw_ctx_create(pw_ctx** out, const int grid[3], const double supercell[9], int npw,
                             const int* idx, int npad, const int* pad_src, const int* pad_dst,
                             const double* ekin, const double* teter, int max_bands, void* stream) {
  if (!out || !grid || !supercell || !idx || !ekin || !teter || npw <= 0 || max_bands <= 0)
    return fail(PW_ERR_ARG, "pw_ctx_create: bad argument");
  if (npad > 0 && (!pad_src || !pad_dst)) return fail(PW_ERR_ARG, "pw_ctx_create: null padding arrays");
  for (int d = 0; d < 3; d++)
    if (!smooth235(grid[d]))
      return fail(PW_ERR_UNSUPPORTED, "pw_ctx_create: grid size %d is not 2/3/5-smooth", grid[d]);
  if (grid[2] >= 32768) return fail(PW_ERR_UNSUPPORTED, "pw_ctx_create: n3 too large");
  pw_ctx* c = new pw_ctx();
  {
    // register z kernels on request only: at si64_60 (256 bands) they take
    // 22.8 / 25.9 us (FP32 inverse / forward) against the LDS kernels'
    // 23.6 / 23.5 us, FP64 35.9 / 43.3 against 37.9 / 36.4 us
    const char* zr = getenv("PW_ZREG");
    if (zr && atoi(zr) == 1) c->zreg = true;
    const char* bf = getenv("PW_BF16");
    if (bf) c->bf16_split = (atoi(bf) == 2) ? 2 : (atoi(bf) == 0 ? 0 : 3);
    const char* sl = getenv("PW_SLAB");
    for (auto& pl : c->plan) {
      pl.reg = !(sl && std::string(sl).compare(0, 3, "lds") == 0);
      // "reg:<ly>[:<sb>][:p]": the register kernel with ly lanes per y
      // column transform (sb bands per z plane; 16: the wide FP32 variant;
      // p: on a persistent grid)
      if (sl && std::string(sl).compare(0, 4, "reg:") == 0) {
        pl.ly = std::max(1, atoi(sl + 4));
        const char* b = strchr(sl + 4, ':');
        if (b) pl.sb = std::max(0, atoi(b + 1));
        pl.persist = strstr(sl + 4, ":p") != nullptr;  // ":p": the persistent grid
      }
      // "lds:<g>[:<r>]": the LDS kernel with its large-plane layout, g y
      // columns per group (and r x row pairs per step)
      if (sl && std::string(sl).compare(0, 4, "lds:") == 0) {
        pl.cg = std::max(1, atoi(sl + 4));
        const char* r = strchr(sl + 4, ':');
        if (r) pl.rp = std::max(1, std::min(64, atoi(r + 1)));
        const char* th = r ? strchr(r + 1, ':') : nullptr;  // ":<threads>" (256, 512 or 1024)
        if (th) {
          const int t = atoi(th + 1);
          pl.threads = (t == 512 || t == 1024) ? t : 256;
        }
      }
    }
  }
  c->n1 = grid[0]; c->n2 = grid[1]; c->n3 = grid[2];
  c->h1 = grid[0] / 2 + 1;
  c->N = (long)grid[0] * grid[1] * grid[2];
  c->Nh = (long)c->h1 * grid[1] * grid[2];
  memcpy(c->supercell, supercell, sizeof(double) * 9);
  recip_cell(supercell, c->recip);
  c->vol = cell_volume(supercell);
  c->npw = npw;
  c->npad = npad;
  c->max_bands = max_bands;
  c->stream = (hipStream_t)stream;
  // validate the sphere: G = 0 first, all entries in range, and the padding is
  // exactly the conj mirror of the kx = 0 plane (fourier.cpp:580-612)
  if (idx[0] != 0) { delete c; return fail(PW_ERR_DOMAIN, "pw_ctx_create: idx[0] must be G = 0"); }
  for (int i = 0; i < npw; i++)
    if (idx[i] < 0 || idx[i] >= c->Nh) { delete c; return fail(PW_ERR_DOMAIN, "pw_ctx_create: idx out of range"); }
  {
    std::vector<int> where((size_t)c->Nh, -1);
    for (int i = 0; i < npw; i++) where[idx[i]] = i;
    int expect = 0;
    for (int i = 0; i < npw; i++) {
      const int g = idx[i];
      const int kx = g % c->h1, ky = (g / c->h1) % c->n2, kz = g / (c->h1 * c->n2);
      if (kx == 0 && !(ky == 0 && kz == 0)) expect++;
    }
    if (expect != npad) { delete c; return fail(PW_ERR_DOMAIN, "pw_ctx_create: padding count %d != %d", npad, expect); }
    for (int p = 0; p < npad; p++) {
      const int s = pad_src[p], d = pad_dst[p];
      if (s < 0 || s >= c->Nh || d < 0 || d >= c->Nh || where[s] < 0) {
        delete c;
        return fail(PW_ERR_DOMAIN, "pw_ctx_create: padding pair %d out of the sphere", p);
      }
      const int ky = (s / c->h1) % c->n2, kz = s / (c->h1 * c->n2);
      const int mky = (c->n2 - ky) % c->n2, mkz = (c->n3 - kz) % c->n3;
      if (s % c->h1 != 0 || d != c->h1 * (mky + c->n2 * mkz)) {
        delete c;
        return fail(PW_ERR_DOMAIN, "pw_ctx_create: padding pair %d is not a Hermitian mirror", p);
      }
    }
  }
  build_stick_plan(c, idx);
  c->h_idx.assign(idx, idx + npw);
  c->f1 = radix_plan(c->n1);
  c->f2 = radix_plan(c->n2);
  c->f3 = radix_plan(c->n3);
  int rc;
  PW_HIP(hipGetDevice(&c->device));
  if ((rc = scratch_lookup(c->device)) != PW_OK) {
    pw_ctx_destroy(c);
    return rc;
  }
#define UP(d, h, n)                                       \
  if ((rc = upload(&(d), (h), (n))) != PW_OK) {           \
    pw_ctx_destroy(c);                                    \
    return rc;                                            \
  }
  UP(c->d_pt_idx, c->sp.pt_idx.data(), c->sp.pt_idx.size());
  UP(c->d_pt_sz, c->sp.pt_sz.data(), c->sp.pt_sz.size());
  UP(c->d_chunk_off, c->sp.chunk_off.data(), c->sp.chunk_off.size());
  UP(c->d_gpt_idx, c->sp.gpt_idx.data(), c->sp.gpt_idx.size());
  UP(c->d_gpt_sz, c->sp.gpt_sz.data(), c->sp.gpt_sz.size());
  UP(c->d_gchunk_off, c->sp.gchunk_off.data(), c->sp.gchunk_off.size());
  UP(c->d_zs, c->sp.zs.data(), c->sp.zs.size());
  UP(c->d_pos, c->sp.pos.data(), c->sp.pos.size());
  UP(c->d_mpos, c->sp.mpos.data(), c->sp.mpos.size());
  UP(c->d_ekin, ekin, (size_t)npw);
  UP(c->d_teter, teter, (size_t)npw);
  {
    std::vector<float> ef(ekin, ekin + npw), tf(teter, teter + npw);
    UP(c->d_ekin_f, ef.data(), ef.size());
    UP(c->d_teter_f, tf.data(), tf.size());
    std::vector<double> tw;
    roots(c->n1, tw);
    roots(c->n2, tw);
    roots(c->n3, tw);
    std::vector<float> twf(tw.begin(), tw.end());
    UP(c->d_tw64, tw.data(), tw.size());
    UP(c->d_tw32, twf.data(), twf.size());
  }
#undef UP
  c->tbytes = sizeof(double2) * (size_t)max_bands * c->n3 * c->sp.nstick;
  if (hipMalloc(&c->d_t1, c->tbytes) != hipSuccess || hipMalloc(&c->d_t2, c->tbytes) != hipSuccess) {
    pw_ctx_destroy(c);
    return fail(PW_ERR_NOMEM, "pw_ctx_create: cannot allocate %zu bytes of stick buffers", 2 * c->tbytes);
  }
  if (rocblas_create_handle(&c->blas) != rocblas_status_success) {
    pw_ctx_destroy(c);
    return fail(PW_ERR_HIP, "pw_ctx_create: rocblas_create_handle failed");
  }
  rocblas_set_stream(c->blas, c->stream);
  *out = c;
  return PW_OK;
}

extern "C" int pw_ctx_destroy(pw_ctx* c) {
  if (!c) return PW_OK;
  pw_free_workspace(c);
  exx_free(c);
  if (c->d_gemm_ws) (void)hipFree(c->d_gemm_ws);
  void* ptrs[] = {c->d_pt_idx, c->d_pt_sz, c->d_chunk_off, c->d_gpt_idx, c->d_gpt_sz, c->d_gchunk_off, c->d_zs,
                  c->d_pos, c->d_mpos, c->d_ekin, c->d_teter, c->d_ekin_f, c->d_teter_f, c->d_tw64,
                  c->d_tw32, c->d_vtot, c->d_vtot_f, c->d_beta, c->d_betaw, c->d_W, c->d_beta_f,
                  c->d_betaw_f, c->d_W_f, c->d_t1, c->d_t2, c->d_betaw_bf, c->d_W_bf};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (c->blas) rocblas_destroy_handle(c->blas);
  if (c->ev_fork) (void)hipEventDestroy(c->ev_fork);
  if (c->ev_join) (void)hipEventDestroy(c->ev_join);
  if (c->side) (void)hipStreamDestroy(c->side);
  for (auto& e : c->ev_pool) (void)hipEventDestroy(e);
  delete c;
  return PW_OK;
}

extern "C" int pw_ctx_set_stream(pw_ctx* c, void* stream) {
  if (!c) return fail(PW_ERR_ARG, "null context");
  c->stream = (hipStream_t)stream;
  rocblas_set_stream(c->blas, c->stream);
  return PW_OK;
}

extern "C" int pw_ctx_info(const pw_ctx* c, int* npw, int* nstick, int* kxmax, long* ngrid) {
  if (!c) return fail(PW_ERR_ARG, "null context");
  if (npw) *npw = c->npw;
  if (nstick) *nstick = c->sp.nstick;
  if (kxmax) *kxmax = c->sp.kxmax;
  if (ngrid) *ngrid = c->N;
  return PW_OK;
}

extern "C" int pw_set_vtot(pw_ctx* c, const double* vtot) {
  if (!c || !vtot) return fail(PW_ERR_ARG, "pw_set_vtot: null argument");
  std::vector<double> v(vtot, vtot + c->N);
  const double inv = 1.0 / (double)c->N;
  for (auto& x : v) x *= inv;  // fold the C2R normalisation of spinor.cpp:620-621
  std::vector<float> vf(v.begin(), v.end());
  if (!c->d_vtot) {
    invalidate_graph(c);
    PW_HIP(hipMalloc((void**)&c->d_vtot, sizeof(double) * c->N));
    PW_HIP(hipMalloc((void**)&c->d_vtot_f, sizeof(float) * c->N));
  }
  PW_HIP(hipMemcpy(c->d_vtot, v.data(), sizeof(double) * c->N, hipMemcpyHostToDevice));
  PW_HIP(hipMemcpy(c->d_vtot_f, vf.data(), sizeof(float) * c->N, hipMemcpyHostToDevice));
  c->have_vtot = true;
  return PW_OK;
}

extern "C" int pw_set_nonlocal(pw_ctx* c, int nbeta, const double* beta, const double* w) {
  if (!c || nbeta < 0 || (nbeta > 0 && (!beta || !w))) return fail(PW_ERR_ARG, "pw_set_nonlocal: bad argument");
  invalidate_graph(c);
  void* ptrs[] = {c->d_beta, c->d_betaw, c->d_W, c->d_beta_f, c->d_betaw_f, c->d_W_f, c->d_betaw_bf, c->d_W_bf};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  c->d_beta = c->d_betaw = c->d_W = nullptr;
  c->d_beta_f = c->d_betaw_f = c->d_W_f = nullptr;
  c->d_betaw_bf = c->d_W_bf = nullptr;
  c->nbeta = nbeta;
  c->h_wgt.assign(w ? w : (const double*)nullptr, w ? w + nbeta : (const double*)nullptr);
  if (nbeta == 0) return PW_OK;
  const size_t m = 2 * (size_t)c->npw;
  std::vector<double> bw(beta, beta + m * nbeta);
  for (int p = 0; p < nbeta; p++)
    for (size_t r = 0; r < m; r++) bw[m * p + r] *= w[p];
  std::vector<float> bf(beta, beta + m * nbeta), bwf(bw.begin(), bw.end());
  int rc;
  if ((rc = upload(&c->d_beta, beta, m * nbeta)) != PW_OK) return rc;
  if ((rc = upload(&c->d_betaw, bw.data(), m * nbeta)) != PW_OK) return rc;
  if ((rc = upload(&c->d_beta_f, bf.data(), m * nbeta)) != PW_OK) return rc;
  if ((rc = upload(&c->d_betaw_f, bwf.data(), m * nbeta)) != PW_OK) return rc;
  PW_HIP(hipMalloc((void**)&c->d_W, sizeof(double) * (size_t)nbeta * c->max_bands));
  PW_HIP(hipMalloc((void**)&c->d_W_f, sizeof(float) * (size_t)nbeta * c->max_bands));
  // the FP32 filter's projector apply on split BF16 products: betaw split once
  PW_TRY(split_a_operand(c, c->d_betaw_f, (int)m, (int)m, nbeta, &c->d_betaw_bf));
  PW_HIP(hipMalloc((void**)&c->d_W_bf, split_b_bytes(c->max_bands, nbeta)));
  PW_HIP(hipStreamSynchronize(c->stream));
  return PW_OK;
}

extern "C" int pw_set_timing(pw_ctx* c, int enable) {
  if (!c) return fail(PW_ERR_ARG, "null context");
  if (enable && c->ev_pool.empty()) {
    c->ev_pool.resize(5 * pw_ctx::kTimedCalls);
    for (auto& e : c->ev_pool) PW_HIP(hipEventCreate(&e));
  }
  c->timing = enable != 0;
  c->ev_calls = 0;
  c->ev_total = 0;  // a new timing session counts its own calls
  for (int i = 0; i < 4; i++) c->stage_ms[i] = 0.0;
  return PW_OK;
}

extern "C" int pw_get_timing(pw_ctx* c, double* ms, int* ncalls) {
  if (!c || !ms) return fail(PW_ERR_ARG, "null argument");
  const int n = std::min(c->ev_calls, pw_ctx::kTimedCalls);
  if (n > 0) PW_HIP(hipEventSynchronize(c->ev_pool[5 * (n - 1) + 4]));
  for (int k = 0; k < n; k++)
    for (int i = 0; i < 4; i++) {
      float t = 0.f;
      PW_HIP(hipEventElapsedTime(&t, c->ev_pool[5 * k + i], c->ev_pool[5 * k + i + 1]));
      c->stage_ms[i] += t;
    }
  for (int i = 0; i < 4; i++) ms[i] = c->stage_ms[i];
  if (ncalls) *ncalls = c->ev_total + n;
  c->ev_total += n;
  c->ev_calls = 0;
  return PW_OK;
}

namespace pw {

// Event k of the current timed call (k = 0..4); no-op when timing is off or
// the pool is full (the caller then sees fewer timed calls).
static int mark(pw_ctx* c, int k) {
  if (!c->timing || c->ev_calls >= pw_ctx::kTimedCalls) return PW_OK;
  PW_HIP(hipEventRecord(c->ev_pool[5 * c->ev_calls + k], c->stream));
  if (k == 4) c->ev_calls++;
  return PW_OK;
}

// LDS bytes of k_slab_xy: twiddles, y-space (two copies when cg == 0, one
// plus a cg-column scratch shared with the x rows otherwise), x rows.
static size_t slab_lds(const pw_ctx* c, int esz, int rp, int cg) {
  const size_t nk = c->sp.kxmax + 1, n1 = c->n1, n2 = c->n2;
  if (cg == 0) return esz * (n1 + n2 + 2 * nk * n2 + 2 * (size_t)rp * n1);
  return esz * (n1 + n2 + nk * n2 + std::max((size_t)cg * n2, 2 * (size_t)rp * n1));
}

template <typename T>
static int local_kinetic(pw_ctx* c, int nb, const void* psi, void* a3) {
  typedef typename cx<T>::t C;
  const int esz = sizeof(C);
  const T* tw = (const T*)(sizeof(T) == 8 ? (const void*)c->d_tw64 : (const void*)c->d_tw32);
  const C* tw1 = (const C*)tw;
  const C* tw2 = tw1 + c->n1;
  const C* tw3 = tw2 + c->n2;
  const T* vt = (const T*)(sizeof(T) == 8 ? (const void*)c->d_vtot : (const void*)c->d_vtot_f);
  const T* ek = (const T*)(sizeof(T) == 8 ? (const void*)c->d_ekin : (const void*)c->d_ekin_f);
  const int sc = c->sp.sc;
  C* t1 = (C*)c->d_t1;
  C* t2 = (C*)c->d_t2;
  // x-pass row-pair chunk: as many pairs as fit next to the y buffers in 64 KiB
  const pw_ctx::Plan& pl = c->plan[sizeof(T) == 8 ? PW_F64 : PW_F32];
  int rp = 8, cg = 0;
  while (rp > 1 && slab_lds(c, esz, rp, 0) > 64 * 1024) rp /= 2;
  if (pl.cg == 0 && pl.rp > 0) {
    // double-buffered layout with a chosen row-pair count (the autotuner's)
    rp = std::min(pl.rp, (c->n2 + 1) / 2);
  } else if (pl.cg > 0) {
    cg = std::min(c->sp.kxmax + 1, pl.cg);
    if (pl.rp > 0) rp = pl.rp;
  } else if (slab_lds(c, esz, rp, 0) > 64 * 1024) {
    // large planes: the double-buffered layout would hold one workgroup per
    // CU with one x row pair per step.  One y buffer and a column-group
    // scratch (the x rows' space) instead: one group of all columns when it
    // fits, else the most row pairs (<= 8) that leave room for two
    // workgroups per CU, else for one.
    const int nk = c->sp.kxmax + 1;
    bool found = false;
    // first choice: all nk columns in one group (the scratch is a second y
    // buffer) and as many x row pairs as that scratch holds (FP32 180 x 180,
    // 22 row pairs: 13.1 ms; 16: 13.3; 8-column groups with 4 row pairs: 14.8)
    const int r = std::max(1, std::min((c->n2 + 1) / 2, nk * c->n2 / (2 * c->n1)));
    if (slab_lds(c, esz, r, nk) <= 160 * 1024) {
      rp = r;
      cg = nk;
      found = true;
    }
    for (size_t budget : {(size_t)80 * 1024, (size_t)160 * 1024}) {
      if (found) break;
      for (rp = 8; rp >= 1 && !found; rp /= 2) {
        cg = std::min(nk, std::max(1, 2 * rp * c->n1 / c->n2));
        found = slab_lds(c, esz, rp, cg) <= budget;
        if (found) break;
      }
      if (found) break;
    }
    if (!found) { rp = 1; cg = 1; }
  }
  // one or two bands (the Lanczos steps): the LDS kernel with all y columns
  // in one group and all x row pairs in one step at 512 threads spreads each
  // plane's transforms over the workgroup, where the register kernel leaves
  // most of its threads idle (si64_60, one band: FP32 13.6 us against 24.0,
  // FP64 15.7 against 29.3)
  const int nkc = c->sp.kxmax + 1, npr = (c->n2 + 1) / 2;
  const bool few = nb <= 2 && slab_lds(c, esz, npr, nkc) <= 160 * 1024;
  if (few) {
    cg = nkc;
    rp = npr;
  }
  const size_t lds2 = slab_lds(c, esz, rp, cg);
  if (lds2 > 160 * 1024)
    return fail(PW_ERR_UNSUPPORTED, "H*psi slab needs %zu bytes of LDS (> 160 KiB)", lds2);
  const size_t lds1 = (size_t)esz * (c->n3 + 2 * (size_t)sc * c->n3);
  PW_TRY(mark(c, 0));
  PW_TRY(launch_inv_z<T>(c, nb, (const C*)psi, t1, tw3));
  PW_TRY(mark(c, 1));
  // the wide variant only where its y-space fits the LDS
  const bool wide = pl.sb == kSlabWideSB && slab_reg_lds(c, sizeof(C), kSlabWideSB) <= 160 * 1024;
  // the persistent grid where that variant exists, else the one-item grid
  slab_reg_fn<T> kreg =
      pl.reg && !few && pl.persist && !wide ? slab_reg_kernel<T>(c->n1, c->n2, pl.ly, 0, true) : nullptr;
  const bool persist = kreg != nullptr;
  if (pl.reg && !few && !kreg) kreg = slab_reg_kernel<T>(c->n1, c->n2, wide ? 1 : pl.ly, wide ? kSlabWideSB : 0);
  if (kreg) {
    // one z plane of SB bands per workgroup (or per item of the persistent
    // grid), V plane + y-space in LDS
    const int SB = wide ? kSlabWideSB : SlabCfg<T>::SB;
    const size_t ldsr = slab_reg_lds(c, sizeof(C), SB);
    PW_TRY(ensure_lds((const void*)kreg, ldsr));
    dim3 grid(c->n3, (nb + SB - 1) / SB);
    if (persist) {
      // as many workgroups as the device holds at once
      pw_ctx::Plan& pm = c->plan[sizeof(T) == 8 ? PW_F64 : PW_F32];
      if (pm.pkern != (const void*)kreg || pm.plds != ldsr) {
        int per_cu = 0, ncu = 0;
        PW_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)kreg, 256, ldsr));
        PW_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
        pm.pkern = (const void*)kreg;
        pm.plds = ldsr;
        pm.pgrid = std::max(1, per_cu) * std::max(1, ncu);
      }
      grid = dim3((unsigned)std::min<long>((long)grid.x * grid.y, pm.pgrid));
    }
    hipLaunchKernelGGL(kreg, grid, dim3(wide ? 512 : 256), ldsr, c->stream, (const C*)t1, (C*)t2, c->n3,
                       c->sp.nstick, c->sp.kxmax, nb, c->d_pos, c->d_mpos, (const T*)vt);
  } else {
    auto kslab = slab_kernel<T>(c->n1, c->n2);
    PW_TRY(ensure_lds((const void*)kslab, lds2));
    // large planes (one or two workgroups per CU): more waves per workgroup
    // to cover the LDS latency of the transform passes
    // (180 x 180 planes, 1024 bands: 1024 threads 15.0 / 33.3 ms FP32 / FP64,
    // 256 threads 18.9 / 43.2 ms)
    int threads = cg > 0 ? (lds2 > 80 * 1024 ? 1024 : 512) : 256;
    if (pl.threads > 0) threads = pl.threads;
    if (few) threads = 512;
    hipLaunchKernelGGL(kslab, dim3(c->n3, nb), dim3(threads), lds2, c->stream, (const C*)t1, t2, c->n1,
                       c->n2, c->n3, c->sp.nstick, c->sp.kxmax, c->d_pos, c->d_mpos, vt, tw1, tw2, rp, cg);
  }
  PW_HIP(hipGetLastError());
  PW_TRY(mark(c, 2));
  PW_TRY(launch_fwd_z<T>(c, nb, (const C*)t2, (const C*)psi, (C*)a3, ek, tw3));
  PW_TRY(mark(c, 3));
  return PW_OK;
}

bool has_reg_slab(const pw_ctx* c, int dtype, int ly, int sb, bool persist) {
  if (sb == kSlabWideSB && slab_reg_lds(c, dtype == PW_F64 ? 16 : 8, sb) > 160 * 1024) return false;
  return dtype == PW_F64 ? slab_reg_kernel<double>(c->n1, c->n2, ly, sb, persist) != nullptr
                         : slab_reg_kernel<float>(c->n1, c->n2, ly, sb, persist) != nullptr;
}
size_t slab_lds_bytes(const pw_ctx* c, int esz, int rp, int cg) { return slab_lds(c, esz, rp, cg); }

// Kernel 1 over nb bands (psi: npw-strided bands; t1: [band][z][stick]):
// the register form where the z length has one, else the LDS form.
template <typename T>
int launch_inv_z(pw_ctx* c, int nb, const typename cx<T>::t* psi, typename cx<T>::t* t1,
                 const typename cx<T>::t* tw3) {
  typedef typename cx<T>::t C;
  const auto kr = c->zreg ? inv_z_reg_kernel<T>(c->n3) : nullptr;
  if (kr) {
    const long nwork = (long)nb * c->sp.nstick, nthr = nwork * (c->n3 % 4 == 0 ? 4 : 2);  // ZLanes<n3>::L
    hipLaunchKernelGGL(kr, dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, c->stream, psi, c->npw, t1,
                       c->sp.nstick, nwork, c->d_zs);
  } else {
    const size_t lds1 = sizeof(C) * (c->n3 + 2 * (size_t)c->sp.sc * c->n3);
    hipLaunchKernelGGL(inv_z_kernel<T>(c->n3), dim3(c->sp.nchunk * nb), dim3(256), lds1, c->stream, psi, c->npw,
                       t1, c->n3, c->sp.nstick, c->sp.sc, c->sp.nchunk, c->d_chunk_off, c->d_pt_idx, c->d_pt_sz,
                       tw3);
  }
  PW_HIP(hipGetLastError());
  return PW_OK;
}

// Kernel 3 over nb bands: out = ekin .* psi + gather (ekin non-null), or
// out += gather.
template <typename T>
int launch_fwd_z(pw_ctx* c, int nb, const typename cx<T>::t* t2, const typename cx<T>::t* psi,
                 typename cx<T>::t* out, const T* ek, const typename cx<T>::t* tw3) {
  typedef typename cx<T>::t C;
  const auto kr = c->zreg ? fwd_z_reg_kernel<T>(c->n3) : nullptr;
  if (kr) {
    const long nwork = (long)nb * c->sp.nstick, nthr = nwork * (c->n3 % 4 == 0 ? 4 : 2);  // ZLanes<n3>::L
    hipLaunchKernelGGL(kr, dim3((unsigned)((nthr + 255) / 256)), dim3(256), 0, c->stream, t2, psi, out, c->npw,
                       c->sp.nstick, nwork, c->d_zs, ek);
  } else {
    const size_t lds1 = sizeof(C) * (c->n3 + 2 * (size_t)c->sp.sc * c->n3);
    hipLaunchKernelGGL(fwd_z_kernel<T>(c->n3), dim3(c->sp.nchunk * nb), dim3(256), lds1, c->stream, t2, psi, out,
                       c->npw, c->n3, c->sp.nstick, c->sp.sc, c->sp.nchunk, c->d_gchunk_off, c->d_gpt_idx,
                       c->d_gpt_sz, ek, tw3);
  }
  PW_HIP(hipGetLastError());
  return PW_OK;
}

// Sphere coefficients -> z-transformed sticks T[band][z][stick] (kernel 1).
template <typename T>
int sphere_to_sticks(pw_ctx* c, int nb, const void* psi, void* t1) {
  typedef typename cx<T>::t C;
  const C* tw3 = (const C*)(sizeof(T) == 8 ? (const void*)c->d_tw64 : (const void*)c->d_tw32) + c->n1 + c->n2;
  const size_t lds1 = sizeof(C) * (c->n3 + 2 * (size_t)c->sp.sc * c->n3);
  return launch_inv_z<T>(c, nb, (const C*)psi, (C*)t1, tw3);
}

// Sticks -> forward z transform -> out += sqrt(2)-scaled sphere gather.
template <typename T>
int sticks_to_sphere_add(pw_ctx* c, int nb, const void* t2, void* out) {
  typedef typename cx<T>::t C;
  const C* tw3 = (const C*)(sizeof(T) == 8 ? (const void*)c->d_tw64 : (const void*)c->d_tw32) + c->n1 + c->n2;
  const size_t lds1 = sizeof(C) * (c->n3 + 2 * (size_t)c->sp.sc * c->n3);
  return launch_fwd_z<T>(c, nb, (const C*)t2, (const C*)nullptr, (C*)out, (const T*)nullptr, tw3);
}

template int sphere_to_sticks<double>(pw_ctx*, int, const void*, void*);
template int sphere_to_sticks<float>(pw_ctx*, int, const void*, void*);
template int sticks_to_sphere_add<double>(pw_ctx*, int, const void*, void*);
template int sticks_to_sphere_add<float>(pw_ctx*, int, const void*, void*);

// W = Re(beta^H X) on the real (2 npw) view, then Y (+)= (beta .* w) W.
template <typename T>
static int nonlocal(pw_ctx* c, int nb, const void* X, void* Y, bool accumulate) {
  const int m = 2 * c->npw;
  const T one = 1, zero = 0, b2 = accumulate ? (T)1 : (T)0;
  if (c->nbeta == 0) {
    if (!accumulate) PW_HIP(hipMemsetAsync(Y, 0, sizeof(T) * (size_t)m * nb, c->stream));
    return PW_OK;
  }
  const T* beta = (const T*)(sizeof(T) == 8 ? (const void*)c->d_beta : (const void*)c->d_beta_f);
  const T* betaw = (const T*)(sizeof(T) == 8 ? (const void*)c->d_betaw : (const void*)c->d_betaw_f);
  T* W = (T*)(sizeof(T) == 8 ? (void*)c->d_W : (void*)c->d_W_f);
  PW_TRY(gemm_tn<T>(c, m, c->nbeta, nb, beta, m, (const T*)X, m, W, c->nbeta, nullptr, zero));
  if (!accumulate) PW_HIP(hipMemsetAsync(Y, 0, sizeof(T) * (size_t)m * nb, c->stream));
  // Y += (beta w) W on the MFMA kernel (FP64 / FP32)
  (void)one;
  (void)b2;
  PW_TRY(gemm_nn_ep<T>(c, m, nb, c->nbeta, betaw, m, W, c->nbeta, (const T*)Y, nullptr, nullptr, (T*)Y, m, nullptr,
                       0, false));
  return PW_OK;
}

static bool aligned8(const void* p) { return ((uintptr_t)p & 7) == 0; }
static int finish_step(pw_ctx* c, int rc) {
  if (rc != PW_OK) return rc;
  return mark(c, 4);
}

// One scaled Chebyshev step around H: Ynew = s1 (H Y - c Y) - s2 Yp with
// (c, s1, s2) = coef[3 step ..] on the device (Yp == null: s2 = 0).  The
// local, kinetic and exchange terms go into HY; the projector product's
// epilogue adds its term and forms the recurrence, so H Y is never
// completed in memory.
template <typename T>
static int cheb_step(pw_ctx* c, int nb, const void* Y, void* HY, const void* Yp, void* Ynew, const T* coef,
                     int step) {
  PW_TRY(local_kinetic<T>(c, nb, Y, HY));
  if (c->exx) PW_TRY(exchange_add<T>(c, nb, Y, HY));
  const int m = 2 * c->npw;
  const T* betaw = (const T*)(sizeof(T) == 8 ? (const void*)c->d_betaw : (const void*)c->d_betaw_f);
  if (c->nbeta > 0) {
    const T* beta = (const T*)(sizeof(T) == 8 ? (const void*)c->d_beta : (const void*)c->d_beta_f);
    T* W = (T*)(sizeof(T) == 8 ? (void*)c->d_W : (void*)c->d_W_f);
    const int split = c->bf16_split;
    PW_TRY(gemm_tn<T>(c, m, c->nbeta, nb, beta, m, (const T*)Y, m, W, c->nbeta, nullptr, (T)0, split));
    if constexpr (sizeof(T) == 4) {
      if ((split == 2 || split == 3) && c->d_betaw_bf && aligned8(HY) && aligned8(Y) && aligned8(Yp) &&
          aligned8(Ynew))
        return finish_step(c, gemm_nn_pre(c, m, nb, c->nbeta, c->d_betaw_bf, W, c->nbeta, c->d_W_bf,
                                          (const float*)HY, (const float*)Y, (const float*)Yp, (float*)Ynew, m,
                                          coef, step, true, split));
    }
    PW_TRY(gemm_nn_ep<T>(c, m, nb, c->nbeta, betaw, m, W, c->nbeta, (const T*)HY, (const T*)Y, (const T*)Yp,
                         (T*)Ynew, m, coef, step, true, split));
  } else {
    PW_TRY(gemm_nn_ep<T>(c, m, nb, 0, nullptr, m, nullptr, 0, (const T*)HY, (const T*)Y, (const T*)Yp, (T*)Ynew, m,
                         coef, step, true));
  }
  PW_TRY(mark(c, 4));
  return PW_OK;
}

static int check_call(pw_ctx* c, int dtype, int nb, const void* a, const void* b) {
  if (!c) return fail(PW_ERR_ARG, "null context");
  if (dtype != PW_F64 && dtype != PW_F32) return fail(PW_ERR_ARG, "unknown dtype %d", dtype);
  if (nb < 0 || nb > c->max_bands) return fail(PW_ERR_ARG, "nb = %d outside [0, %d]", nb, c->max_bands);
  if (nb > 0 && (!a || !b)) return fail(PW_ERR_ARG, "null block pointer");
  if (!c->have_vtot) return fail(PW_ERR_STATE, "vtot has not been set");
  return PW_OK;
}

}  // namespace pw

int pw::cheb_step_apply(pw_ctx* c, int dtype, int nb, const void* Y, void* HY, const void* Yp, void* Ynew,
                        int step) {
  PW_TRY(check_call(c, dtype, nb, Y, HY));
  if (dtype == PW_F64) return cheb_step<double>(c, nb, Y, HY, Yp, Ynew, c->d_coef, step);
  return cheb_step<float>(c, nb, Y, HY, Yp, Ynew, c->d_coef_f, step);
}

extern "C" int pw_mult_spinor(pw_ctx* c, int dtype, int nb, const void* psi, void* a3) {
  PW_TRY(check_call(c, dtype, nb, psi, a3));
  if (nb == 0) return PW_OK;
  if (psi == a3) return fail(PW_ERR_ARG, "pw_mult_spinor: psi and a3 must not alias");
  int rc = dtype == PW_F64 ? local_kinetic<double>(c, nb, psi, a3) : local_kinetic<float>(c, nb, psi, a3);
  // hybrid functional with EXX active: + V_x psi (hamiltonian.cpp:3431-3448)
  if (rc == PW_OK && c->exx) rc = dtype == PW_F64 ? exchange_add<double>(c, nb, psi, a3) : exchange_add<float>(c, nb, psi, a3);
  if (rc == PW_OK) PW_TRY(mark(c, 4));
  return rc;
}

extern "C" int pw_nonlocal_multx(pw_ctx* c, int dtype, int nb, const void* X, void* V) {
  if (!c) return fail(PW_ERR_ARG, "null context");
  if (dtype != PW_F64 && dtype != PW_F32) return fail(PW_ERR_ARG, "unknown dtype %d", dtype);
  if (nb < 0 || nb > c->max_bands) return fail(PW_ERR_ARG, "nb = %d outside [0, %d]", nb, c->max_bands);
  if (nb == 0) return PW_OK;
  if (!X || !V || X == V) return fail(PW_ERR_ARG, "pw_nonlocal_multx: bad block pointers");
  return dtype == PW_F64 ? nonlocal<double>(c, nb, X, V, false) : nonlocal<float>(c, nb, X, V, false);
}

// EigenSolver::NonlocalMultX without any allocation: the projections and the
// split-K partials in the per-device scratch array; calls that overlap on one
// GPU share (and race on) that scratch.
extern "C" int pw_nonlocal_multx_noalloc(pw_ctx* c, int dtype, int nb, const void* X, void* V) {
  if (!c) return fail(PW_ERR_ARG, "null context");
  if (dtype != PW_F64 && dtype != PW_F32) return fail(PW_ERR_ARG, "unknown dtype %d", dtype);
  if (nb < 0 || nb > c->max_bands) return fail(PW_ERR_ARG, "nb = %d outside [0, %d]", nb, c->max_bands);
  if (nb == 0) return PW_OK;
  if (!X || !V || X == V) return fail(PW_ERR_ARG, "pw_nonlocal_multx_noalloc: bad block pointers");
  void* sp = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch_ptr[c->device] : nullptr;
  if (!sp) return fail(PW_ERR_STATE, "pw_nonlocal_multx_noalloc: no scratch address for device %d", c->device);
  return dtype == PW_F64 ? nonlocal_noalloc<double>(c, nb, X, V, sp) : nonlocal_noalloc<float>(c, nb, X, V, sp);
}

namespace pw {
int side_stream(pw_ctx* c) {
  if (!c->side) PW_HIP(hipStreamCreateWithFlags(&c->side, hipStreamNonBlocking));
  if (!c->ev_fork) PW_HIP(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
  if (!c->ev_join) PW_HIP(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
  return PW_OK;
}
}  // namespace pw

extern "C" int pw_apply_h(pw_ctx* c, int dtype, int nb, const void* X, void* HX) {
  PW_TRY(check_call(c, dtype, nb, X, HX));
  if (nb == 0) return PW_OK;
  if (X == HX) return fail(PW_ERR_ARG, "pw_apply_h: X and HX must not alias");
  int rc = dtype == PW_F64 ? local_kinetic<double>(c, nb, X, HX) : local_kinetic<float>(c, nb, X, HX);
  if (rc == PW_OK && c->exx) rc = dtype == PW_F64 ? exchange_add<double>(c, nb, X, HX) : exchange_add<float>(c, nb, X, HX);
  if (rc != PW_OK) return rc;
  rc = dtype == PW_F64 ? nonlocal<double>(c, nb, X, HX, true) : nonlocal<float>(c, nb, X, HX, true);
  if (rc != PW_OK) return rc;
  PW_TRY(mark(c, 4));
  return PW_OK;
}

extern "C" int pw_teter_precond(pw_ctx* c, int dtype, int nb, const void* psi, void* a3) {
  if (!c) return fail(PW_ERR_ARG, "null context");
  if (nb < 0 || nb > c->max_bands) return fail(PW_ERR_ARG, "nb out of range");
  if (nb == 0) return PW_OK;
  const long total = (long)c->npw * nb;
  const int blocks = (int)std::min<long>((total + 255) / 256, 4096);
  if (dtype == PW_F64)
    hipLaunchKernelGGL(k_teter<double>, dim3(blocks), dim3(256), 0, c->stream, (const double2*)psi,
                       (double2*)a3, c->d_teter, c->npw, total);
  else if (dtype == PW_F32)
    hipLaunchKernelGGL(k_teter<float>, dim3(blocks), dim3(256), 0, c->stream, (const float2*)psi,
                       (float2*)a3, c->d_teter_f, c->npw, total);
  else
    return fail(PW_ERR_ARG, "unknown dtype %d", dtype);
  PW_HIP(hipGetLastError());
  return PW_OK;
}
