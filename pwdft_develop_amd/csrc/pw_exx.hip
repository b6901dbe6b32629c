// Screened exact exchange on gfx950: Spinor::AddMultSpinorEXX, sphere-cut
// branch (src/SCF/spinor.cpp:783-945), with the kernel of
// KohnSham::InitializeEXX (src/SCF/hamiltonian.cpp:3454-3551) and the phi
// scaling of KohnSham::SetPhiEXX (hamiltonian.cpp:3569-3603).
//
// The reference runs 3 FFTs per (band k, occupied j) pair: R2C of the pair
// density psi_k phi_j, C2R of the kernel-weighted spectrum, and an R2C of
// phi_j times the result that is gathered into a3_k.  The last transform is
// linear, so it is taken once per band on the accumulated real-space sum:
// 2 FFTs per pair + 1 per band, equal up to rounding.  All transforms are
// batched over pairs; the Fock sphere (|G|^2/2 < 4 Ecut) gets its own stick
// plan, so its z transforms touch only the Fock columns.
//
//   k_exx_slab_inv   wavefunction sticks -> y IFFT -> x c2r -> real grid (psi_k, phi_j)
//   k_exx_pair_fwd   psi_k(r) phi_j(r) -> x r2c -> y FFT -> Fock sticks
//                    (k_exx_pair_fwd_reg / _acc_reg: the register-FFT forms
//                    for square planes of compile-time length)
//   k_exx_fock_z     Fock sticks: z FFT, x K(G)/N, z IFFT (in place);
//                    k_exx_fock_z_reg for compile-time z lengths
//   k_exx_pair_acc   Fock sticks -> y IFFT -> x c2r -> acc_k += f_j phi_j u_kj, over j
//   k_exx_slab_fwd   acc_k(r) -> x r2c -> y FFT -> wavefunction sticks -> a3 += gather
#include <algorithm>
#include <cstdio>
#include <cmath>
#include <string>
#include <vector>

#include <rocsolver/rocsolver.h>

#include "pw_fft.h"
#include "pw_internal.h"

namespace pw {

struct ExxState {
  double ecut = 0, mu = 0, frac = 0, div = 0;
  int divtype = 1, nocc = 0;
  int nfock = 0;
  // Fock stick plan
  int nstick = 0, kxmax = 0, sc = 16, nchunk = 0;
  int* d_pos = nullptr;          // kx*n2 + ky per Fock stick
  double* d_K64 = nullptr;       // K(G)/N per (stick, z)
  float* d_K32 = nullptr;
  double* d_K64t = nullptr;      // the same per (z, stick)
  float* d_K32t = nullptr;
  double* d_phi64 = nullptr;     // phi_j(r), nocc x N
  float* d_phi32 = nullptr;
  double* d_fac64 = nullptr;     // -frac * occ_j
  float* d_fac32 = nullptr;
  std::vector<double> occ;
  // scratch
  void* d_psir = nullptr;        // max_bands x N (double-sized)
  void* d_acc = nullptr;         // max_bands x N
  void* d_fst = nullptr;         // pair Fock sticks: npair_max x n3 x nstick complex double
  int npair_max = 0;
  // adaptively compressed exchange: xi (2 npw x nace real view) and the
  // projection scratch M = Re(xi^H psi) (nace x max_bands)
  int nace = 0;
  double* d_xi64 = nullptr;
  float* d_xi32 = nullptr;
  double* d_M64 = nullptr;
  float* d_M32 = nullptr;
};

static void ace_free(ExxState* E) {
  void* p[] = {E->d_xi64, E->d_xi32, E->d_M64, E->d_M32};
  for (void* q : p)
    if (q) (void)hipFree(q);
  E->d_xi64 = nullptr;
  E->d_xi32 = nullptr;
  E->d_M64 = nullptr;
  E->d_M32 = nullptr;
  E->nace = 0;
}

template <typename S, typename D>
__global__ void k_exx_convert(D* __restrict__ out, const S* __restrict__ in, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    out[i] = (D)in[i];
}

template <typename T>
__global__ void k_exx_negate(T* __restrict__ p, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = -p[i];
}

static inline int grid_of(long n) { return (int)std::max<long>(1, std::min<long>((n + 255) / 256, 8192)); }

// cg == 0: two y buffers and two x-row buffers; cg > 0: one y buffer whose
// transforms go cg columns at a time through the x rows' scratch
struct SlabGeom {
  int n1, n2, n3, nstick, kxmax, rp, cg;
};

#define PW_EXX_PLANES(X) X(24) X(30) X(32) X(36) X(40) X(48) X(60) X(64) X(72) X(96)
#define PW_EXX_Z(X) X(24) X(30) X(32) X(36) X(40) X(48) X(60) X(72) X(96)

template <typename T>
__device__ void load_tw(typename cx<T>::t* tw1, typename cx<T>::t* tw2, const typename cx<T>::t* g1,
                        const typename cx<T>::t* g2, int n1, int n2) {
  for (int t = threadIdx.x; t < n1; t += blockDim.x) tw1[t] = g1[t];
  for (int t = threadIdx.x; t < n2; t += blockDim.x) tw2[t] = g2[t];
}

// sticks (with optional conj mirrors) -> y IFFT; returns the y-space buffer
template <typename T, int N2 = 0>
__device__ typename cx<T>::t* sticks_to_yspace(const typename cx<T>::t* src, const int* pos, const int* mpos,
                                               const SlabGeom& g, typename cx<T>::t* Y0, typename cx<T>::t* Y1,
                                               typename cx<T>::t* S, const typename cx<T>::t* tw2) {
  typedef typename cx<T>::t C;
  const int nk = g.kxmax + 1;
  for (int t = threadIdx.x; t < nk * g.n2; t += blockDim.x) Y0[t] = C{0, 0};
  __syncthreads();
  for (int s = threadIdx.x; s < g.nstick; s += blockDim.x) {
    const C v = src[s];
    Y0[pos[s]] = v;
    if (mpos) {
      const int mp = mpos[s];
      if (mp >= 0) Y0[mp] = C{v.x, -v.y};
    }
  }
  __syncthreads();
  return slab_y_fft<T, N2>(Y0, Y1, S, g.n2, nk, g.cg, tw2, (T)1);
}

#define SLAB_LDS_SETUP                                      \
  typedef typename cx<T>::t C;                              \
  extern __shared__ __align__(16) unsigned char smem_raw[]; \
  const int nk = g.kxmax + 1;                               \
  C* tw1 = reinterpret_cast<C*>(smem_raw);                  \
  C* tw2 = tw1 + g.n1;                                      \
  C* Y0 = tw2 + g.n2;                                       \
  C* Y1 = g.cg == 0 ? Y0 + nk * g.n2 : nullptr;            \
  C* Z0 = g.cg == 0 ? Y1 + nk * g.n2 : Y0 + nk * g.n2;      \
  C* Z1 = Z0 + g.rp * g.n1;                                 \
  const int n1 = N1 > 0 ? N1 : g.n1;                        \
  (void)Z1;                                                 \
  (void)n1;

// wavefunction sticks -> real grid, scaled
template <typename T, int N1 = 0, int N2 = 0>
__global__ void __launch_bounds__(1024)
k_exx_slab_inv(const typename cx<T>::t* __restrict__ st, T* __restrict__ out, const int* __restrict__ pos,
               const int* __restrict__ mpos, SlabGeom g, T scale, const typename cx<T>::t* tw1g,
               const typename cx<T>::t* tw2g) {
  SLAB_LDS_SETUP
  const int z = blockIdx.x, b = blockIdx.y;
  load_tw<T>(tw1, tw2, tw1g, tw2g, g.n1, g.n2);
  C* Y = sticks_to_yspace<T, N2>(st + ((size_t)b * g.n3 + z) * g.nstick, pos, mpos, g, Y0, Y1, Z0, tw2);
  T* o = out + ((size_t)b * g.n3 + z) * g.n1 * g.n2;
  const int npair = (g.n2 + 1) / 2;
  for (int q0 = 0; q0 < npair; q0 += g.rp) {
    const int nq = min(g.rp, npair - q0);
    C* Zr = x_c2r_rows<T, N1>(Y, g.kxmax, g.n1, g.n2, q0, nq, Z0, Z1, tw1);
    for (int t = threadIdx.x; t < nq * n1; t += blockDim.x) {
      const int ql = t / n1, x = t - ql * n1;
      const int y1 = 2 * (q0 + ql), y2 = y1 + 1;
      const C u = Zr[ql * n1 + x];
      o[(size_t)y1 * g.n1 + x] = u.x * scale;
      if (y2 < g.n2) o[(size_t)y2 * g.n1 + x] = u.y * scale;
    }
    __syncthreads();
  }
}

// pair densities psi_k phi_j -> Fock sticks; pair p = k + nb * jl
template <typename T, int N1 = 0, int N2 = 0>
__global__ void __launch_bounds__(1024)
k_exx_pair_fwd(const T* __restrict__ psir, const T* __restrict__ phir, int nb, int j0,
               typename cx<T>::t* __restrict__ fst, const int* __restrict__ pos, SlabGeom g,
               const typename cx<T>::t* tw1g, const typename cx<T>::t* tw2g) {
  SLAB_LDS_SETUP
  const int z = blockIdx.x, p = blockIdx.y;
  const int k = p % nb, j = j0 + p / nb;
  const size_t plane = (size_t)g.n1 * g.n2;
  const T* a = psir + ((size_t)k * g.n3 + z) * plane;
  const T* bph = phir + ((size_t)j * g.n3 + z) * plane;
  load_tw<T>(tw1, tw2, tw1g, tw2g, g.n1, g.n2);
  __syncthreads();
  const int npair = (g.n2 + 1) / 2;
  for (int q0 = 0; q0 < npair; q0 += g.rp) {
    const int nq = min(g.rp, npair - q0);
    for (int t = threadIdx.x; t < nq * n1; t += blockDim.x) {
      const int ql = t / n1, x = t - ql * n1;
      const int y1 = 2 * (q0 + ql), y2 = y1 + 1;
      const T u1 = a[(size_t)y1 * n1 + x] * bph[(size_t)y1 * n1 + x];
      const T u2 = (y2 < g.n2) ? a[(size_t)y2 * n1 + x] * bph[(size_t)y2 * n1 + x] : (T)0;
      Z0[ql * n1 + x] = C{u1, u2};
    }
    __syncthreads();
    x_r2c_rows<T, N1>(Z0, Z1, Y0, g.kxmax, g.n1, g.n2, q0, nq, tw1);
  }
  C* Yf = slab_y_fft<T, N2>(Y0, Y1, Z0, g.n2, nk, g.cg, tw2, (T)-1);
  C* dst = fst + ((size_t)p * g.n3 + z) * g.nstick;
  for (int s = threadIdx.x; s < g.nstick; s += blockDim.x) dst[s] = Yf[pos[s]];
}

// Fock sticks: forward z, x K/N, inverse z (in place); grid (chunk, pair)
template <typename T, int N3 = 0>
__global__ void __launch_bounds__(256)
k_exx_fock_z(typename cx<T>::t* __restrict__ fst, const T* __restrict__ K, int n3, int nstick, int sc,
             const typename cx<T>::t* __restrict__ tw3g) {
  typedef typename cx<T>::t C;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* tw = reinterpret_cast<C*>(smem_raw);
  C* A = tw + n3;
  C* B = A + sc * n3;
  const int c = blockIdx.x, p = blockIdx.y;
  const int s0 = c * sc;
  const int nsl = min(sc, nstick - s0);
  for (int t = threadIdx.x; t < n3; t += blockDim.x) tw[t] = tw3g[t];
  C* base = fst + (size_t)p * n3 * nstick;
  for (int t = threadIdx.x; t < n3 * sc; t += blockDim.x) {
    const int z = t / sc, sl = t - z * sc;
    A[sl * n3 + z] = (sl < nsl) ? base[(size_t)z * nstick + s0 + sl] : C{0, 0};
  }
  __syncthreads();
  C* R = lds_fft_n<T, N3>(A, B, n3, sc, n3, tw, (T)-1);
  for (int t = threadIdx.x; t < n3 * nsl; t += blockDim.x) {
    const int sl = t / n3, z = t - sl * n3;
    R[sl * n3 + z] = cscale(R[sl * n3 + z], K[(size_t)(s0 + sl) * n3 + z]);
  }
  __syncthreads();
  C* O = (R == A) ? B : A;
  C* Q = lds_fft_n<T, N3>(R, O, n3, sc, n3, tw, (T)1);
  for (int t = threadIdx.x; t < n3 * nsl; t += blockDim.x) {
    const int z = t / nsl, sl = t - z * nsl;
    base[(size_t)z * nstick + s0 + sl] = Q[sl * n3 + z];
  }
}

// The same for compile-time N3 with one thread per (pair, stick): the z
// column in registers (reg_fft), no LDS and no barrier; loads and stores of
// a z row are coalesced over the sticks, K read per (z, stick).
template <typename T, int N3>
__global__ void __launch_bounds__(256)
k_exx_fock_z_reg(typename cx<T>::t* __restrict__ fst, const T* __restrict__ Kt, int nstick, long nwork) {
  typedef typename cx<T>::t C;
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nwork) return;
  const long p = e / nstick;
  const int s = (int)(e - p * nstick);
  C* base = fst + (size_t)p * N3 * nstick + s;
  C a[N3];
#pragma unroll
  for (int z = 0; z < N3; z++) a[z] = base[(size_t)z * nstick];
  pin_regs(a);  // phase fences: keep K's loads after the first transform
  reg_fft<T, N3, -1>(a);
  pin_regs(a);
#pragma unroll
  for (int z = 0; z < N3; z++) a[z] = cscale(a[z], Kt[(size_t)z * nstick + s]);
  pin_regs(a);
  reg_fft<T, N3, 1>(a);
  pin_regs(a);
#pragma unroll
  for (int z = 0; z < N3; z++) base[(size_t)z * nstick] = a[z];
}

// acc_k(r) += sum_jl fac_j phi_j(r) c2r(Fock sticks of pair (k, jl))(r)
template <typename T, int N1 = 0, int N2 = 0>
__global__ void __launch_bounds__(1024)
k_exx_pair_acc(const typename cx<T>::t* __restrict__ fst, const T* __restrict__ phir, const T* __restrict__ fac,
               int nb, int j0, int nj, T* __restrict__ acc, const int* __restrict__ pos, SlabGeom g,
               const typename cx<T>::t* tw1g, const typename cx<T>::t* tw2g) {
  SLAB_LDS_SETUP
  const int z = blockIdx.x, k = blockIdx.y;
  const size_t plane = (size_t)g.n1 * g.n2;
  T* ak = acc + ((size_t)k * g.n3 + z) * plane;
  load_tw<T>(tw1, tw2, tw1g, tw2g, g.n1, g.n2);
  const int npair = (g.n2 + 1) / 2;
  for (int jl = 0; jl < nj; jl++) {
    const int j = j0 + jl;
    const T fj = fac[j];
    const T* ph = phir + ((size_t)j * g.n3 + z) * plane;
    const int p = k + nb * jl;
    C* Y = sticks_to_yspace<T, N2>(fst + ((size_t)p * g.n3 + z) * g.nstick, pos, nullptr, g, Y0, Y1, Z0, tw2);
    for (int q0 = 0; q0 < npair; q0 += g.rp) {
      const int nq = min(g.rp, npair - q0);
      C* Zr = x_c2r_rows<T, N1>(Y, g.kxmax, g.n1, g.n2, q0, nq, Z0, Z1, tw1);
      for (int t = threadIdx.x; t < nq * n1; t += blockDim.x) {
        const int ql = t / n1, x = t - ql * n1;
        const int y1 = 2 * (q0 + ql), y2 = y1 + 1;
        const C u = Zr[ql * n1 + x];
        ak[(size_t)y1 * n1 + x] += fj * ph[(size_t)y1 * n1 + x] * u.x;
        if (y2 < g.n2) ak[(size_t)y2 * n1 + x] += fj * ph[(size_t)y2 * n1 + x] * u.y;
      }
      __syncthreads();
    }
  }
}

// Register variants of the pair kernels for square N x N planes: every 1D
// transform is split over L adjacent lanes (lanes_fft / col_fft_lanes: L =
// 4 where 4 divides N, else 2), so that each thread holds N / L values and
// the passes need no LDS round trips or barriers of their own.  LDS holds
// one y-space (kxmax + 1 <= N / 2 + 1 columns of N), which the pair
// densities' row pairs share before the x transform; ExxReg<T, N> carries
// the shape.  One workgroup per (z plane, pair) as in the LDS kernels;
// the row-pair and column phases run on N / 2 L and (kxmax + 1) L threads.
template <typename T, int N>
struct ExxReg {
  static constexpr int L = N % 4 == 0 ? 4 : 2, M = N / L, NP = N / 2;
  static constexpr int NT = ((NP * L + 63) / 64) * 64;  // threads per workgroup
  static constexpr int LDS = (NP + 1) * N * (int)sizeof(typename cx<T>::t);
  static constexpr int UMAX = (NP * (NP + 1) + NT - 1) / NT;  // unpack outputs per thread
};

// pair densities psi_k phi_j -> Fock sticks (as k_exx_pair_fwd)
template <typename T, int N>
__global__ void __launch_bounds__((ExxReg<T, N>::NT))
k_exx_pair_fwd_reg(const T* __restrict__ psir, const T* __restrict__ phir, int nb, int j0,
                   typename cx<T>::t* __restrict__ fst, const int* __restrict__ pos, int n3, int nstick, int kxmax) {
  typedef typename cx<T>::t C;
  typedef ExxReg<T, N> R;
  constexpr int L = R::L, NP = R::NP, NT = R::NT;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* Z = reinterpret_cast<C*>(smem_raw);  // row pairs Z[q][x], then y-space Y[k][y]
  T* Zr = reinterpret_cast<T*>(smem_raw);
  const int t = threadIdx.x, z = blockIdx.x, p = blockIdx.y;
  const int k = p % nb, j = j0 + p / nb;
  const size_t plane = (size_t)N * N;
  const T* a = psir + ((size_t)k * n3 + z) * plane;
  const T* b = phir + ((size_t)j * n3 + z) * plane;
  // products, rows 2q and 2q + 1 as the real and imaginary parts of Z[q]
  for (int e0 = 0; e0 < N * N; e0 += 4 * NT) {
    T u[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int e = e0 + t + NT * i;
      u[i] = e < N * N ? a[e] * b[e] : (T)0;
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int e = e0 + t + NT * i;
      if (e < N * N) {
        const int y = e / N, x = e - y * N;
        Zr[2 * ((y >> 1) * N + x) + (y & 1)] = u[i];
      }
    }
  }
  __syncthreads();
  if (t < NP * L) col_fft_lanes<T, N, -1, L>(Z + (t / L) * N, t % L);
  __syncthreads();
  // unpack the two rows' spectra for kx <= kxmax (reads, barrier, writes:
  // the y-space overlays the row pairs)
  const int nk = kxmax + 1;
  C o1[R::UMAX], o2[R::UMAX];
#pragma unroll
  for (int u = 0; u < R::UMAX; u++) {
    const int e = t + NT * u;
    if (e < NP * nk) {
      const int q = e / nk, kk = e - q * nk;
      const C w = Z[q * N + kk], wc = Z[q * N + (kk == 0 ? 0 : N - kk)];
      o1[u] = C{(T)0.5 * (w.x + wc.x), (T)0.5 * (w.y - wc.y)};
      o2[u] = C{(T)0.5 * (w.y + wc.y), (T)0.5 * (wc.x - w.x)};
    }
  }
  __syncthreads();
#pragma unroll
  for (int u = 0; u < R::UMAX; u++) {
    const int e = t + NT * u;
    if (e < NP * nk) {
      const int q = e / nk, kk = e - q * nk;
      Z[kk * N + 2 * q] = o1[u];
      Z[kk * N + 2 * q + 1] = o2[u];
    }
  }
  __syncthreads();
  for (int c = t / L; c < nk; c += NT / L) col_fft_lanes<T, N, -1, L>(Z + c * N, t % L);
  __syncthreads();
  C* dst = fst + ((size_t)p * n3 + z) * nstick;
  for (int s0 = 0; s0 < nstick; s0 += 4 * NT) {
    C g[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int s1 = s0 + t + NT * i;
      g[i] = s1 < nstick ? Z[pos[s1]] : C{0, 0};
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int s1 = s0 + t + NT * i;
      if (s1 < nstick) dst[s1] = g[i];
    }
  }
}

// acc_k(r) += sum_jl fac_j phi_j(r) c2r(Fock sticks of pair (k, jl))(r) (as
// k_exx_pair_acc); thread (q, h) of the row phase owns the same x samples
// of rows 2q, 2q + 1 for every jl, so acc stays in its registers.
template <typename T, int N>
__global__ void __launch_bounds__((ExxReg<T, N>::NT))
k_exx_pair_acc_reg(const typename cx<T>::t* __restrict__ fst, const T* __restrict__ phir,
                   const T* __restrict__ fac, int nb, int j0, int nj, T* __restrict__ acc,
                   const int* __restrict__ pos, int n3, int nstick, int kxmax) {
  typedef typename cx<T>::t C;
  typedef ExxReg<T, N> R;
  constexpr int L = R::L, M = R::M, NP = R::NP, NT = R::NT;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  C* Y = reinterpret_cast<C*>(smem_raw);
  const int t = threadIdx.x, z = blockIdx.x, k = blockIdx.y;
  const int nk = kxmax + 1;
  const size_t plane = (size_t)N * N;
  T* ak = acc + ((size_t)k * n3 + z) * plane;
  const bool rowt = t < NP * L;
  const int q = t / L, h = t % L, y1 = 2 * q, y2 = y1 + 1;
  // output block of this lane after lanes_fft (the same for every call)
  const int qo = L == 2 ? (h & 1) : ((h >> 1) + 2 * (h & 1));
  T r1[M], r2[M];
#pragma unroll
  for (int x = 0; x < M; x++) {
    r1[x] = rowt ? ak[(size_t)y1 * N + qo * M + x] : (T)0;
    r2[x] = rowt ? ak[(size_t)y2 * N + qo * M + x] : (T)0;
  }
  for (int jl = 0; jl < nj; jl++) {
    const int j = j0 + jl;
    const T fj = fac[j];
    const int p = k + nb * jl;
    for (int e = t; e < nk * N; e += NT) Y[e] = C{0, 0};
    __syncthreads();
    const C* src = fst + ((size_t)p * n3 + z) * nstick;
    for (int s0 = 0; s0 < nstick; s0 += 4 * NT) {
      C g[4];
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int s1 = s0 + t + NT * i;
        g[i] = s1 < nstick ? src[s1] : C{0, 0};
      }
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int s1 = s0 + t + NT * i;
        if (s1 < nstick) Y[pos[s1]] = g[i];
      }
    }
    __syncthreads();
    for (int c = t / L; c < nk; c += NT / L) col_fft_lanes<T, N, 1, L>(Y + c * N, t % L);
    __syncthreads();
    if (rowt) {
      // Hermitian-extended spectra of rows y1, y2 at x frequencies m L + h,
      // packed as X1 + i X2 (x_c2r_rows' packing)
      C a[M];
#pragma unroll
      for (int m = 0; m < M; m++) {
        const int kf = m * L + h;
        int kk = kf;
        bool cj = false;
        if (kf > N / 2) { kk = N - kf; cj = true; }
        C a1 = C{0, 0}, a2 = C{0, 0};
        if (kk <= kxmax) {
          a1 = Y[kk * N + y1];
          a2 = Y[kk * N + y2];
          if (cj) { a1.y = -a1.y; a2.y = -a2.y; }
          if (kf == 0 || 2 * kf == N) { a1.y = 0; a2.y = 0; }
        }
        a[m] = C{a1.x - a2.y, a1.y + a2.x};
      }
      lanes_fft<T, N, 1, L>(a, h);
      const T* ph = phir + ((size_t)j * n3 + z) * plane;
#pragma unroll
      for (int x = 0; x < M; x++) {
        r1[x] += fj * ph[(size_t)y1 * N + qo * M + x] * a[x].x;
        r2[x] += fj * ph[(size_t)y2 * N + qo * M + x] * a[x].y;
      }
    }
    __syncthreads();  // the next pair's zero fill overwrites Y
  }
  if (rowt) {
#pragma unroll
    for (int x = 0; x < M; x++) {
      ak[(size_t)y1 * N + qo * M + x] = r1[x];
      ak[(size_t)y2 * N + qo * M + x] = r2[x];
    }
  }
}

// real grid acc_k -> x r2c -> y FFT -> wavefunction sticks
template <typename T, int N1 = 0, int N2 = 0>
__global__ void __launch_bounds__(1024)
k_exx_slab_fwd(const T* __restrict__ in, typename cx<T>::t* __restrict__ st, const int* __restrict__ pos,
               SlabGeom g, const typename cx<T>::t* tw1g, const typename cx<T>::t* tw2g) {
  SLAB_LDS_SETUP
  const int z = blockIdx.x, b = blockIdx.y;
  const T* a = in + ((size_t)b * g.n3 + z) * g.n1 * g.n2;
  load_tw<T>(tw1, tw2, tw1g, tw2g, g.n1, g.n2);
  __syncthreads();
  const int npair = (g.n2 + 1) / 2;
  for (int q0 = 0; q0 < npair; q0 += g.rp) {
    const int nq = min(g.rp, npair - q0);
    for (int t = threadIdx.x; t < nq * n1; t += blockDim.x) {
      const int ql = t / n1, x = t - ql * n1;
      const int y1 = 2 * (q0 + ql), y2 = y1 + 1;
      Z0[ql * n1 + x] = C{a[(size_t)y1 * n1 + x], (y2 < g.n2) ? a[(size_t)y2 * n1 + x] : (T)0};
    }
    __syncthreads();
    x_r2c_rows<T, N1>(Z0, Z1, Y0, g.kxmax, g.n1, g.n2, q0, nq, tw1);
  }
  C* Yf = slab_y_fft<T, N2>(Y0, Y1, Z0, g.n2, nk, g.cg, tw2, (T)-1);
  C* dst = st + ((size_t)b * g.n3 + z) * g.nstick;
  for (int s = threadIdx.x; s < g.nstick; s += blockDim.x) dst[s] = Yf[pos[s]];
}

template <typename T>
__global__ void k_fill_zero(T* __restrict__ p, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = (T)0;
}

static size_t slab_bytes(const pw_ctx* c, int esz, int kxmax, int rp, int cg = 0) {
  const size_t nk = kxmax + 1, n1 = c->n1, n2 = c->n2;
  if (cg == 0) return esz * (n1 + n2 + 2 * nk * n2 + 2 * (size_t)rp * n1);
  return esz * (n1 + n2 + nk * n2 + std::max((size_t)cg * n2, 2 * (size_t)rp * n1));
}

// Layout of the pair kernels: planes whose double-buffered y-space exceeds
// 64 KiB take the grouped layout (one group of all columns when it fits,
// else the most x row pairs that leave room for two workgroups per CU when
// that is at least 4, else the most that fit one workgroup per CU).
static int pick_pair_layout(const pw_ctx* c, int esz, SlabGeom& g) {
  // PW_EXX_LAYOUT=<cg>:<rp>:<threads> forces the grouped layout's shape
  if (const char* fl = getenv("PW_EXX_LAYOUT")) {
    int cg = 0, rp = 0, th = 0;
    if (sscanf(fl, "%d:%d:%d", &cg, &rp, &th) == 3 && cg > 0 && rp > 0 && (th == 256 || th == 512 || th == 1024)) {
      g.cg = std::min(cg, g.kxmax + 1);
      g.rp = std::min(rp, 64);
      return th;
    }
  }
  g.cg = 0;
  g.rp = 8;
  while (g.rp > 1 && slab_bytes(c, esz, g.kxmax, g.rp) > 96 * 1024) g.rp /= 2;
  if (slab_bytes(c, esz, g.kxmax, g.rp) <= 64 * 1024) return 256;
  const int nk = g.kxmax + 1;
  {
    // first choice: all nk columns in one group (the scratch is a second y
    // buffer) and as many x row pairs as that scratch holds; 1024 threads
    // when one workgroup fills a CU's LDS, else 512 (96 x 96 x 40, 432 x 64
    // pairs, 24 row pairs: FP32 185 ms, FP64 294 ms; 16 row pairs 192 / 311;
    // 16 columns per group, 8 row pairs, 1024 threads 339 / 399)
    const int rp = std::max(1, std::min((c->n2 + 1) / 2, nk * c->n2 / (2 * c->n1)));
    const size_t b = slab_bytes(c, esz, g.kxmax, rp, nk);
    if (b <= 160 * 1024) {
      g.rp = rp;
      g.cg = nk;
      return b > 80 * 1024 ? 1024 : 512;
    }
  }
  auto cg_of = [&](int rp) { return std::min(nk, std::max(1, 2 * rp * c->n1 / c->n2)); };
  for (size_t budget : {(size_t)80 * 1024, (size_t)160 * 1024})
    for (int rp = 8; rp >= (budget == 80 * 1024 ? 4 : 1); rp /= 2)
      if (slab_bytes(c, esz, g.kxmax, rp, cg_of(rp)) <= budget) {
        g.rp = rp;
        g.cg = cg_of(rp);
        return 1024;
      }
  return 256;  // keeps the double-buffered layout (set_lds reports it when too large)
}

static int pick_rp(const pw_ctx* c, int esz, int kxmax) {
  int rp = 8;
  while (rp > 1 && slab_bytes(c, esz, kxmax, rp) > 96 * 1024) rp /= 2;
  return rp;
}

// Layout of the wavefunction-side slab kernels (k_exx_slab_inv,
// k_exx_slab_fwd, k_density_acc): the double-buffered layout at 256 threads
// whenever it fits the LDS; planes beyond that (180 x 180 in FP64: 265 KiB)
// take the grouped layout of the pair kernels, picked the same way.
// PW_WFN_LAYOUT=<cg>:<rp>:<threads> forces a grouped shape.  Returns the
// threads per workgroup.
static int pick_wfn_layout(const pw_ctx* c, int esz, SlabGeom& g) {
  if (const char* fl = getenv("PW_WFN_LAYOUT")) {
    int cg = 0, rp = 0, th = 0;
    if (sscanf(fl, "%d:%d:%d", &cg, &rp, &th) == 3 && cg > 0 && rp > 0 && (th == 256 || th == 512 || th == 1024)) {
      g.cg = std::min(cg, g.kxmax + 1);
      g.rp = std::min(rp, 64);
      return th;
    }
  }
  g.cg = 0;
  g.rp = pick_rp(c, esz, g.kxmax);
  if (slab_bytes(c, esz, g.kxmax, g.rp) <= 160 * 1024) return 256;
  const int nk = g.kxmax + 1;
  const int rp = std::max(1, std::min((c->n2 + 1) / 2, nk * c->n2 / (2 * c->n1)));
  const size_t b = slab_bytes(c, esz, g.kxmax, rp, nk);
  if (b <= 160 * 1024) {
    g.rp = rp;
    g.cg = nk;
    return b > 80 * 1024 ? 1024 : 512;
  }
  auto cg_of = [&](int r) { return std::min(nk, std::max(1, 2 * r * c->n1 / c->n2)); };
  for (size_t budget : {(size_t)80 * 1024, (size_t)160 * 1024})
    for (int r = 8; r >= (budget == 80 * 1024 ? 4 : 1); r /= 2)
      if (slab_bytes(c, esz, g.kxmax, r, cg_of(r)) <= budget) {
        g.rp = r;
        g.cg = cg_of(r);
        return 1024;
      }
  g.cg = 0;
  g.rp = 1;
  return 256;  // set_lds reports the plane as too large
}

template <typename K>
static int set_lds(K kern, size_t bytes) {
  if (bytes > 160 * 1024) return fail(PW_ERR_UNSUPPORTED, "exchange slab needs %zu bytes of LDS (> 160 KiB)", bytes);
  return ensure_lds((const void*)kern, bytes);
}

// psi (sphere, dtype T) -> real grid (scale) for nb bands, via the wavefunction sticks
// The exchange kernels with compile-time transform lengths for the square
// planes (and z lengths) of the register-FFT size list, else the
// runtime-length ones.
template <typename T>
struct ExxKernels {
  decltype(&k_exx_slab_inv<T>) slab_inv;
  decltype(&k_exx_pair_fwd<T>) pair_fwd;
  decltype(&k_exx_pair_acc<T>) pair_acc;
  decltype(&k_exx_slab_fwd<T>) slab_fwd;
  decltype(&k_exx_fock_z<T>) fock_z;
  decltype(&k_exx_fock_z_reg<T, 24>) fock_z_reg;  // null: the LDS kernel
};
template <typename T>
static ExxKernels<T> exx_kernels(int n1, int n2, int n3) {
  ExxKernels<T> k{&k_exx_slab_inv<T>, &k_exx_pair_fwd<T>, &k_exx_pair_acc<T>, &k_exx_slab_fwd<T>, &k_exx_fock_z<T>,
                  nullptr};
  if (n1 == n2) {
    switch (n1) {
#define PW_CASE(n)                                                                                        \
  case n:                                                                                                 \
    k.slab_inv = &k_exx_slab_inv<T, n, n>;                                                                \
    k.pair_fwd = &k_exx_pair_fwd<T, n, n>;                                                                \
    k.pair_acc = &k_exx_pair_acc<T, n, n>;                                                                \
    k.slab_fwd = &k_exx_slab_fwd<T, n, n>;                                                                \
    break;
      PW_EXX_PLANES(PW_CASE)
#undef PW_CASE
      default:
        break;
    }
  }
  switch (n3) {
#define PW_CASE(n)                                                   \
  case n:                                                            \
    k.fock_z = &k_exx_fock_z<T, n>;                                  \
    if constexpr (sizeof(T) * n <= 480) /* no spills: FP64 to 60 */  \
      k.fock_z_reg = &k_exx_fock_z_reg<T, n>;                        \
    break;
    PW_EXX_Z(PW_CASE)
#undef PW_CASE
    default:
      break;
  }
  return k;
}

// the register pair kernels of an n1 x n2 plane (null: none; the LDS
// kernels then), their block size and LDS bytes
template <typename T>
struct ExxRegKernels {
  decltype(&k_exx_pair_fwd_reg<T, 24>) fwd = nullptr;
  decltype(&k_exx_pair_acc_reg<T, 24>) acc = nullptr;
  int threads = 0;
  size_t lds = 0;
};
template <typename T>
static ExxRegKernels<T> exx_reg_kernels(int n1, int n2) {
  ExxRegKernels<T> k;
  // FP32 takes the register kernels, FP64 the LDS ones (config-4 exchange,
  // 1024 x 432 pairs on 96 x 96 x 40: FP32 1.14 s register against 1.89 s
  // LDS; FP64 3.34 against 3.32 s, the FP64 register kernels at one wave
  // per SIMD).  PW_EXX_PAIR (for comparisons): "lds" both LDS kernels,
  // "reg" both register kernels in either precision, "fwd" the register
  // forward kernel only.
  const char* f = getenv("PW_EXX_PAIR");
  const std::string mode = f ? f : "";
  if (n1 != n2 || mode == "lds" || (sizeof(T) == 8 && mode != "reg")) return k;
  switch (n1) {
#define PW_CASE(n)                              \
  case n:                                       \
    k.fwd = &k_exx_pair_fwd_reg<T, n>;          \
    k.acc = &k_exx_pair_acc_reg<T, n>;          \
    k.threads = ExxReg<T, n>::NT;               \
    k.lds = (size_t)ExxReg<T, n>::LDS;          \
    break;
    PW_EXX_PLANES(PW_CASE)
#undef PW_CASE
    default:
      break;
  }
  if (mode == "fwd") k.acc = nullptr;
  return k;
}

template <typename T>
static int sphere_to_real(pw_ctx* c, int nb, const void* psi, T* out, T scale) {
  typedef typename cx<T>::t C;
  PW_TRY(sphere_to_sticks<T>(c, nb, psi, c->d_t1));
  const C* tw = (const C*)(sizeof(T) == 8 ? (const void*)c->d_tw64 : (const void*)c->d_tw32);
  SlabGeom g{c->n1, c->n2, c->n3, c->sp.nstick, c->sp.kxmax, 0, 0};
  const int thr = pick_wfn_layout(c, sizeof(C), g);
  const size_t lds = slab_bytes(c, sizeof(C), g.kxmax, g.rp, g.cg);
  const auto kinv = exx_kernels<T>(c->n1, c->n2, c->n3).slab_inv;
  PW_TRY(set_lds(kinv, lds));
  hipLaunchKernelGGL(kinv, dim3(c->n3, nb), dim3(thr), lds, c->stream, (const C*)c->d_t1, out,
                     c->d_pos, c->d_mpos, g, scale, tw, tw + c->n1);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

template <typename T>
int exx_add(pw_ctx* c, int nb, const void* psi, void* a3) {
  typedef typename cx<T>::t C;
  ExxState* E = c->exx;
  if (!E || E->nocc == 0 || nb == 0) return PW_OK;
  const long N = c->N;
  T* psir = (T*)E->d_psir;
  T* acc = (T*)E->d_acc;
  const T* phir = (const T*)(sizeof(T) == 8 ? (const void*)E->d_phi64 : (const void*)E->d_phi32);
  const T* fac = (const T*)(sizeof(T) == 8 ? (const void*)E->d_fac64 : (const void*)E->d_fac32);
  const T* K = (const T*)(sizeof(T) == 8 ? (const void*)E->d_K64 : (const void*)E->d_K32);
  const T* Kt = (const T*)(sizeof(T) == 8 ? (const void*)E->d_K64t : (const void*)E->d_K32t);
  const C* tw = (const C*)(sizeof(T) == 8 ? (const void*)c->d_tw64 : (const void*)c->d_tw32);
  PW_TRY(sphere_to_real<T>(c, nb, psi, psir, (T)1));
  // zero fill by a kernel, not a memset node: this runs inside captured
  // filter graphs
  hipLaunchKernelGGL(k_fill_zero<T>, dim3((int)std::min<long>((N * nb + 255) / 256, 8192)), dim3(256), 0, c->stream,
                     acc, N * (long)nb);
  SlabGeom gf{c->n1, c->n2, c->n3, E->nstick, E->kxmax, 0, 0};
  const int thr = pick_pair_layout(c, sizeof(C), gf);
  const size_t ldsf = slab_bytes(c, sizeof(C), gf.kxmax, gf.rp, gf.cg);
  const ExxKernels<T> kx = exx_kernels<T>(c->n1, c->n2, c->n3);
  PW_TRY(set_lds(kx.pair_fwd, ldsf));
  PW_TRY(set_lds(kx.pair_acc, ldsf));
  const ExxRegKernels<T> kr = exx_reg_kernels<T>(c->n1, c->n2);
  if (kr.fwd) PW_TRY(set_lds(kr.fwd, kr.lds));
  if (kr.acc) PW_TRY(set_lds(kr.acc, kr.lds));
  // pairs per launch: kc bands x jc occupied bands (kc jc <= npair_max);
  // up to 64 bands per chunk so that the accumulation kernel's read and
  // write of acc is shared by jc pairs, and 40 planes x 64 bands still fill
  // the device
  const int jc = std::max(1, std::min(E->nocc, E->npair_max / std::min(nb, 64)));
  const int kc = std::max(1, std::min(nb, E->npair_max / jc));
  const size_t ldsz = sizeof(C) * (c->n3 + 2 * (size_t)E->sc * c->n3);
  const size_t bstride = (size_t)N;  // one band of psir / acc
  for (int k0 = 0; k0 < nb; k0 += kc) {
    const int nk = std::min(kc, nb - k0);
    const T* pk = psir + (size_t)k0 * bstride;
    T* ak = acc + (size_t)k0 * bstride;
    for (int j0 = 0; j0 < E->nocc; j0 += jc) {
      const int nj = std::min(jc, E->nocc - j0);
      if (kr.fwd)
        hipLaunchKernelGGL(kr.fwd, dim3(c->n3, nk * nj), dim3(kr.threads), kr.lds, c->stream, pk, phir, nk, j0,
                           (C*)E->d_fst, E->d_pos, c->n3, E->nstick, E->kxmax);
      else
        hipLaunchKernelGGL(kx.pair_fwd, dim3(c->n3, nk * nj), dim3(thr), ldsf, c->stream, pk, phir, nk, j0,
                           (C*)E->d_fst, E->d_pos, gf, tw, tw + c->n1);
      PW_HIP(hipGetLastError());
      if (kx.fock_z_reg) {
        const long nwork = (long)nk * nj * E->nstick;
        hipLaunchKernelGGL(kx.fock_z_reg, dim3((unsigned)((nwork + 255) / 256)), dim3(256), 0, c->stream,
                           (C*)E->d_fst, Kt, E->nstick, nwork);
      } else {
        hipLaunchKernelGGL(kx.fock_z, dim3(E->nchunk, nk * nj), dim3(256), ldsz, c->stream, (C*)E->d_fst, K,
                           c->n3, E->nstick, E->sc, tw + c->n1 + c->n2);
      }
      PW_HIP(hipGetLastError());
      if (kr.acc)
        hipLaunchKernelGGL(kr.acc, dim3(c->n3, nk), dim3(kr.threads), kr.lds, c->stream, (const C*)E->d_fst, phir,
                           fac, nk, j0, nj, ak, E->d_pos, c->n3, E->nstick, E->kxmax);
      else
        hipLaunchKernelGGL(kx.pair_acc, dim3(c->n3, nk), dim3(thr), ldsf, c->stream, (const C*)E->d_fst, phir,
                           fac, nk, j0, nj, ak, E->d_pos, gf, tw, tw + c->n1);
      PW_HIP(hipGetLastError());
    }
  }
  SlabGeom g{c->n1, c->n2, c->n3, c->sp.nstick, c->sp.kxmax, 0, 0};
  const int thw = pick_wfn_layout(c, sizeof(C), g);
  const size_t lds = slab_bytes(c, sizeof(C), g.kxmax, g.rp, g.cg);
  PW_TRY(set_lds(kx.slab_fwd, lds));
  hipLaunchKernelGGL(kx.slab_fwd, dim3(c->n3, nb), dim3(thw), lds, c->stream, (const T*)acc, (C*)c->d_t2,
                     c->d_pos, g, tw, tw + c->n1);
  PW_HIP(hipGetLastError());
  return sticks_to_sphere_add<T>(c, nb, c->d_t2, a3);
}

template int exx_add<double>(pw_ctx*, int, const void*, void*);
template int exx_add<float>(pw_ctx*, int, const void*, void*);

// ------------------------------------------------------------------ adjoint
// dE/dV(r) of E = sum_i f_i <psi_i|H|psi_i>: sum_i f_i psi_i(r)^2 / N with
// psi_i(r) = C2R(c_i) -- the accumulation of KohnSham::CalculateDensity
// (src/SCF/hamiltonian.cpp:2509-2573) with f_i = numSpin occ_i.  Bands are
// recomputed from their sticks chunk by chunk; real-space psi is never
// stored.  Block (z, split) sums the bands b = split, split + nsplit, ...
// of the chunk into its own partial plane; partials are reduced in order.
template <typename T, int N1 = 0>
__global__ void __launch_bounds__(1024)
k_density_acc(const typename cx<T>::t* __restrict__ st, const double* __restrict__ f, int b0, int nbc, int nsplit,
              double* __restrict__ part, const int* __restrict__ pos, const int* __restrict__ mpos, SlabGeom g,
              double invN, const typename cx<T>::t* tw1g, const typename cx<T>::t* tw2g) {
  SLAB_LDS_SETUP
  const int z = blockIdx.x, sp = blockIdx.y;
  const size_t plane = (size_t)g.n1 * g.n2;
  double* acc = part + ((size_t)sp * g.n3 + z) * plane;
  load_tw<T>(tw1, tw2, tw1g, tw2g, g.n1, g.n2);
  const int npair = (g.n2 + 1) / 2;
  for (int bl = sp; bl < nbc; bl += nsplit) {
    const double fb = f[b0 + bl] * invN;
    C* Y = sticks_to_yspace<T>(st + ((size_t)bl * g.n3 + z) * g.nstick, pos, mpos, g, Y0, Y1, Z0, tw2);
    for (int q0 = 0; q0 < npair; q0 += g.rp) {
      const int nq = min(g.rp, npair - q0);
      C* Zr = x_c2r_rows<T>(Y, g.kxmax, g.n1, g.n2, q0, nq, Z0, Z1, tw1);
      for (int t = threadIdx.x; t < nq * g.n1; t += blockDim.x) {
        const int ql = t / g.n1, x = t - ql * g.n1;
        const int y1 = 2 * (q0 + ql), y2 = y1 + 1;
        const C u = Zr[ql * g.n1 + x];
        acc[(size_t)y1 * g.n1 + x] += fb * (double)u.x * (double)u.x;
        if (y2 < g.n2) acc[(size_t)y2 * g.n1 + x] += fb * (double)u.y * (double)u.y;
      }
      __syncthreads();
    }
  }
}

__global__ void k_sum_partials(const double* __restrict__ part, int nsplit, long n, double* __restrict__ out,
                               int accumulate) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
    double s = accumulate ? out[i] : 0.0;
    for (int q = 0; q < nsplit; q++) s += part[(size_t)q * n + i];
    out[i] = s;
  }
}

// Per-block partial sums of x (fixed order: the host adds the partials in
// block order, so the total is reproducible run to run).
__global__ void k_block_sums(const double* __restrict__ x, long n, double* __restrict__ part) {
  __shared__ double red[256];
  double s = 0.0;
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) s += x[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = blockDim.x / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

__global__ void k_scale(double* __restrict__ x, long n, double s) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) x[i] *= s;
}

// F[a][d] = -2 sum_i f_i sum_{p in a} w_p W[p,i] D_d[p,i]; one block per
// (a, d), summed in FP64 whatever the precision of W and D
template <typename T>
__global__ void k_force_reduce(const T* __restrict__ W, const T* __restrict__ D, const double* __restrict__ w,
                               const double* __restrict__ f, int nbeta, int nb, int npa, double* __restrict__ F) {
  const int a = blockIdx.x, d = blockIdx.y;
  const T* Dd = D + (size_t)d * nbeta * nb;
  double s = 0.0;
  for (int t = threadIdx.x; t < npa * nb; t += blockDim.x) {
    const int q = t % npa, i = t / npa;
    const int p = a * npa + q;
    s += f[i] * w[p] * (double)W[p + (size_t)nbeta * i] * (double)Dd[p + (size_t)nbeta * i];
  }
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = blockDim.x / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) F[3 * a + d] = -2.0 * red[0];
}

// E = sum_i f_i Re<x_i| h_i>: per-band dots (FP64 sums) reduced on the host
template <typename T>
__global__ void k_band_dots(const T* __restrict__ X, const T* __restrict__ HX, long m, double* __restrict__ out) {
  const int j = blockIdx.x;
  const T* x = X + (size_t)j * m;
  const T* h = HX + (size_t)j * m;
  double s = 0.0;
  for (long i = threadIdx.x; i < m; i += blockDim.x) s += (double)x[i] * (double)h[i];
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = blockDim.x / 2; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = red[0];
}

void exx_free(pw_ctx* c) {
  ExxState* E = c->exx;
  if (!E) return;
  invalidate_graph(c);
  void* p[] = {E->d_pos, E->d_K64, E->d_K32, E->d_K64t, E->d_K32t, E->d_phi64, E->d_phi32, E->d_fac64, E->d_fac32,
               E->d_psir, E->d_acc, E->d_fst};
  for (void* q : p)
    if (q) (void)hipFree(q);
  ace_free(E);
  delete E;
  c->exx = nullptr;
}

template <typename X>
static int up(X** d, const X* h, size_t n) {
  PW_HIP(hipMalloc((void**)d, sizeof(X) * (n ? n : 1)));
  if (n) PW_HIP(hipMemcpy(*d, h, sizeof(X) * n, hipMemcpyHostToDevice));
  return PW_OK;
}

}  // namespace pw

using namespace pw;

extern "C" int pw_exx_divergence(const int grid[3], const double supercell[9], double ecut, double mu,
                                 int divtype, double* div) {
  if (!grid || !supercell || !div) return fail(PW_ERR_ARG, "pw_exx_divergence: null argument");
  const int n1 = grid[0], n2 = grid[1], n3 = grid[2];
  double b[9];
  recip_cell(supercell, b);
  const double epsDiv = 1e-8, alpha = 10.0 / (ecut * 2.0);
  double d = 0.0;
  if (divtype == 1) {
    for (int k = 0; k < n3; k++)
      for (int j = 0; j < n2; j++)
        for (int i = 0; i < n1; i++) {
          const int m0 = kgrid(i, n1), m1 = kgrid(j, n2), m2 = kgrid(k, n3);
          double g[3];
          for (int q = 0; q < 3; q++) g[q] = b[q] * m0 + b[3 + q] * m1 + b[6 + q] * m2;
          const double g2 = g[0] * g[0] + g[1] * g[1] + g[2] * g[2];
          if (g2 > epsDiv && g2 < ecut * 4.0)
            d += mu > 0.0 ? std::exp(-alpha * g2) / g2 * (1.0 - std::exp(-g2 / (4.0 * mu * mu)))
                          : std::exp(-alpha * g2) / g2;
        }
    d += mu > 0.0 ? 1.0 / (4.0 * mu * mu) : -alpha;
    d *= 4.0 * M_PI;
    const double aa = mu > 0.0 ? 1.0 / std::sqrt(alpha * M_PI) - 1.0 / std::sqrt(M_PI * (alpha + 1.0 / (4.0 * mu * mu)))
                               : 1.0 / std::sqrt(alpha * M_PI);
    d -= cell_volume(supercell) * aa;
  } else if (divtype != 0) {
    return fail(PW_ERR_UNSUPPORTED, "EXX divergence type %d", divtype);
  }
  *div = d;
  return PW_OK;
}

extern "C" int pw_set_exx(pw_ctx* c, double ecut, double mu, double exx_fraction, int divtype, int nocc,
                          const void* psi_occ, const double* occ) {
  if (!c || nocc < 0 || (nocc > 0 && (!psi_occ || !occ))) return fail(PW_ERR_ARG, "pw_set_exx: bad argument");
  if (nocc > c->max_bands) return fail(PW_ERR_ARG, "pw_set_exx: nocc > max_bands");
  exx_free(c);
  invalidate_graph(c);
  if (nocc == 0) return PW_OK;
  ExxState* E = new ExxState();
  c->exx = E;
  E->ecut = ecut; E->mu = mu; E->frac = exx_fraction; E->divtype = divtype; E->nocc = nocc;
  E->occ.assign(occ, occ + nocc);
  PW_TRY(pw_exx_divergence(&c->n1, c->supercell, ecut, mu, divtype, &E->div));
  // Fock sphere and its sticks; kernel K(G)/N (InitializeEXX folded with the
  // vol/N and 1/vol factors of FFTWExecute)
  const int n1 = c->n1, n2 = c->n2, n3 = c->n3, h1 = c->h1;
  int nf = 0;
  PW_TRY(pw_fock_sphere_build(&c->n1, c->supercell, ecut, &nf, nullptr, nullptr));
  std::vector<int> fidx(nf);
  std::vector<double> fg(nf);
  PW_TRY(pw_fock_sphere_build(&c->n1, c->supercell, ecut, &nf, fidx.data(), fg.data()));
  E->nfock = nf;
  std::vector<int> sid((size_t)h1 * n2, -1);
  std::vector<int> pos;
  int kxmax = 0;
  for (int i = 0; i < nf; i++) sid[fidx[i] % (h1 * n2)] = 1;
  for (int ky = 0; ky < n2; ky++)
    for (int kx = 0; kx < h1; kx++)
      if (sid[kx + (size_t)h1 * ky] > 0) {
        sid[kx + (size_t)h1 * ky] = (int)pos.size();
        pos.push_back(kx * n2 + ky);
        kxmax = std::max(kxmax, kx);
      }
  E->nstick = (int)pos.size();
  E->kxmax = kxmax;
  E->nchunk = (E->nstick + E->sc - 1) / E->sc;
  std::vector<double> K((size_t)E->nstick * n3, 0.0);
  const double inv = 1.0 / (double)c->N;
  for (int i = 0; i < nf; i++) {
    const double g2 = fg[i] * 2.0;
    double kv;
    if (g2 > 1e-8) kv = mu > 0 ? 4.0 * M_PI / g2 * (1.0 - std::exp(-g2 / (4.0 * mu * mu))) : 4.0 * M_PI / g2;
    else kv = -E->div + (mu > 0 ? M_PI / (mu * mu) : 0.0);
    const int col = fidx[i] % (h1 * n2), kz = fidx[i] / (h1 * n2);
    K[(size_t)sid[col] * n3 + kz] = kv * inv;
  }
  std::vector<float> Kf(K.begin(), K.end());
  std::vector<double> Kt(K.size());
  for (int st = 0; st < E->nstick; st++)
    for (int z = 0; z < n3; z++) Kt[(size_t)z * E->nstick + st] = K[(size_t)st * n3 + z];
  std::vector<float> Ktf(Kt.begin(), Kt.end());
  PW_TRY(up(&E->d_pos, pos.data(), pos.size()));
  PW_TRY(up(&E->d_K64, K.data(), K.size()));
  PW_TRY(up(&E->d_K32, Kf.data(), Kf.size()));
  PW_TRY(up(&E->d_K64t, Kt.data(), Kt.size()));
  PW_TRY(up(&E->d_K32t, Ktf.data(), Ktf.size()));
  std::vector<double> fac(nocc);
  for (int j = 0; j < nocc; j++) fac[j] = -exx_fraction * occ[j];
  std::vector<float> facf(fac.begin(), fac.end());
  PW_TRY(up(&E->d_fac64, fac.data(), fac.size()));
  PW_TRY(up(&E->d_fac32, facf.data(), facf.size()));
  const size_t N = (size_t)c->N;
  PW_HIP(hipMalloc(&E->d_psir, sizeof(double) * N * c->max_bands));
  PW_HIP(hipMalloc(&E->d_acc, sizeof(double) * N * c->max_bands));
  PW_HIP(hipMalloc((void**)&E->d_phi64, sizeof(double) * N * nocc));
  PW_HIP(hipMalloc((void**)&E->d_phi32, sizeof(float) * N * nocc));
  // pair stick buffer: up to ~4 GiB, at least one pair per band
  const size_t per_pair = sizeof(double) * 2 * (size_t)n3 * E->nstick;
  size_t npm = std::max<size_t>((size_t)c->max_bands, (size_t)(4ull << 30) / per_pair);
  npm = std::min<size_t>(npm, (size_t)c->max_bands * nocc);
  E->npair_max = (int)npm;
  PW_HIP(hipMalloc(&E->d_fst, per_pair * npm));
  // phi_j(r) = C2R(psi_occ_j * sqrt(1/(vol N))) in both precisions
  const double fphi = std::sqrt(1.0 / c->vol / (double)c->N);
  PW_TRY(sphere_to_real<double>(c, nocc, psi_occ, E->d_phi64, fphi));
  {
    // FP32 copy of the FP64 real-space phi
    const long n = (long)N * nocc;
    hipLaunchKernelGGL((k_exx_convert<double, float>), dim3(grid_of(n)), dim3(256), 0, c->stream, E->d_phi32,
                       (const double*)E->d_phi64, n);
    PW_HIP(hipGetLastError());
  }
  return PW_OK;
}

// ---------------------------------------------- adaptively compressed exchange

namespace pw {

// MultSpinor's exchange term (hamiltonian.cpp:3343-3449): the compressed
// operator a3 -= xi Re(xi^H psi) when one is built (:3347-3430), else the
// full pair-density apply.
template <typename T>
int exchange_add(pw_ctx* c, int nb, const void* psi, void* a3) {
  ExxState* E = c->exx;
  if (E->nace == 0) return exx_add<T>(c, nb, psi, a3);
  const int m = 2 * c->npw;
  const T* xi = sizeof(T) == 8 ? (const T*)(const void*)E->d_xi64 : (const T*)(const void*)E->d_xi32;
  T* M = sizeof(T) == 8 ? (T*)(void*)E->d_M64 : (T*)(void*)E->d_M32;
  PW_TRY(gemm_tn<T>(c, m, E->nace, nb, xi, m, (const T*)psi, m, M, E->nace, nullptr, (T)0));
  PW_TRY(gemm_nn<T>(c, m, nb, E->nace, (T)-1, xi, m, M, E->nace, (T)1, (T*)a3, m));
  return PW_OK;
}
template int exchange_add<double>(pw_ctx*, int, const void*, void*);
template int exchange_add<float>(pw_ctx*, int, const void*, void*);

}  // namespace pw

// KohnSham::CalculateVexxACE, sphere branch (hamiltonian.cpp:3685-3924):
// W = V_x psi, M = -Re(psi^H W), M = U^T U (the reference's Potrf('L') with
// L = U^T), xi = W U^{-1} (its Trsm('R','L','T','N')).  FP64 throughout; an
// FP32 copy of xi serves FP32 applies.
extern "C" int pw_build_ace(pw_ctx* c, int nb, const void* psi) {
  if (!c) return fail(PW_ERR_ARG, "null context");
  if (!c->exx) return fail(PW_ERR_STATE, "pw_build_ace: no exchange operator set (pw_set_exx)");
  if (nb < 0 || nb > c->max_bands) return fail(PW_ERR_ARG, "pw_build_ace: nb = %d outside [0, %d]", nb, c->max_bands);
  ExxState* E = c->exx;
  invalidate_graph(c);
  PW_HIP(hipStreamSynchronize(c->stream));  // earlier applies may still read xi
  ace_free(E);
  if (nb == 0) return PW_OK;
  if (!psi) return fail(PW_ERR_ARG, "pw_build_ace: null block");
  const int m = 2 * c->npw;
  const long nw = (long)m * nb;
  double* S = nullptr;
  int* info = nullptr;
  if (hipMalloc((void**)&E->d_xi64, sizeof(double) * nw) != hipSuccess ||
      hipMalloc((void**)&E->d_xi32, sizeof(float) * nw) != hipSuccess ||
      hipMalloc((void**)&E->d_M64, sizeof(double) * (size_t)nb * c->max_bands) != hipSuccess ||
      hipMalloc((void**)&E->d_M32, sizeof(float) * (size_t)nb * c->max_bands) != hipSuccess ||
      hipMalloc((void**)&S, sizeof(double) * (size_t)nb * nb) != hipSuccess ||
      hipMalloc((void**)&info, sizeof(int)) != hipSuccess) {
    if (S) (void)hipFree(S);
    if (info) (void)hipFree(info);
    ace_free(E);
    return fail(PW_ERR_NOMEM, "pw_build_ace: cannot allocate the compressed operator (%d bands)", nb);
  }
  auto cleanup = [&](int rc) {
    (void)hipStreamSynchronize(c->stream);
    (void)hipFree(S);
    (void)hipFree(info);
    if (rc != PW_OK) ace_free(E);
    return rc;
  };
  hipStream_t s = c->stream;
  const double one = 1.0;
  int rc = PW_OK;
  if (hipMemsetAsync(E->d_xi64, 0, sizeof(double) * nw, s) != hipSuccess)
    return cleanup(fail(PW_ERR_HIP, "pw_build_ace: memset failed"));
  if ((rc = exx_add<double>(c, nb, psi, E->d_xi64)) != PW_OK) return cleanup(rc);
  if ((rc = gemm_tn<double>(c, m, nb, nb, (const double*)psi, m, E->d_xi64, m, S, nb, nullptr, 0.0)) != PW_OK)
    return cleanup(rc);
  hipLaunchKernelGGL(k_exx_negate<double>, dim3(grid_of((long)nb * nb)), dim3(256), 0, s, S, (long)nb * nb);
  if (rocsolver_dpotrf(c->blas, rocblas_fill_upper, nb, S, nb, info) != rocblas_status_success)
    return cleanup(fail(PW_ERR_HIP, "pw_build_ace: rocsolver_dpotrf failed"));
  int hinfo = 0;
  if (hipMemcpyAsync(&hinfo, info, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return cleanup(fail(PW_ERR_HIP, "pw_build_ace: reading the Cholesky status failed"));
  if (hinfo != 0)
    return cleanup(fail(PW_ERR_LINALG, "Potrf: -psi^T V_x psi is not positive definite (info = %d)", hinfo));
  if (rocblas_dtrsm(c->blas, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                    rocblas_diagonal_non_unit, m, nb, &one, S, nb, E->d_xi64, m) != rocblas_status_success)
    return cleanup(fail(PW_ERR_HIP, "pw_build_ace: rocblas_dtrsm failed"));
  hipLaunchKernelGGL((k_exx_convert<double, float>), dim3(grid_of(nw)), dim3(256), 0, s, E->d_xi32,
                     (const double*)E->d_xi64, nw);
  if (hipGetLastError() != hipSuccess) return cleanup(fail(PW_ERR_HIP, "pw_build_ace: kernel launch failed"));
  E->nace = nb;
  return cleanup(PW_OK);
}

extern "C" int pw_add_exx(pw_ctx* c, int dtype, int nb, const void* psi, void* a3) {
  if (!c) return fail(PW_ERR_ARG, "null context");
  if (!c->exx) return fail(PW_ERR_STATE, "pw_add_exx: no exchange operator set (pw_set_exx)");
  if (nb < 0 || nb > c->max_bands) return fail(PW_ERR_ARG, "nb out of range");
  if (nb == 0) return PW_OK;
  if (!psi || !a3 || psi == a3) return fail(PW_ERR_ARG, "pw_add_exx: bad block pointers");
  if (dtype == PW_F64) return exx_add<double>(c, nb, psi, a3);
  if (dtype == PW_F32) return exx_add<float>(c, nb, psi, a3);
  return fail(PW_ERR_ARG, "unknown dtype %d", dtype);
}

// ------------------------------------------------------------ adjoint entries

namespace pw {

template <typename T>
static int density_impl(pw_ctx* c, int nb, const void* psi, const double* f, double* rho, int accumulate) {
  typedef typename cx<T>::t C;
  const long N = c->N;
  const int nsplit = 16;
  double *df = nullptr, *part = nullptr;
  auto done = [&](int rc) {
    if (df) (void)hipFree(df);
    if (part) (void)hipFree(part);
    return rc;
  };
  if (hipMalloc((void**)&df, sizeof(double) * nb) != hipSuccess ||
      hipMalloc((void**)&part, sizeof(double) * N * nsplit) != hipSuccess)
    return done(fail(PW_ERR_NOMEM, "pw_density: scratch allocation failed"));
  if (hipMemcpy(df, f, sizeof(double) * nb, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemsetAsync(part, 0, sizeof(double) * N * nsplit, c->stream) != hipSuccess)
    return done(fail(PW_ERR_HIP, "pw_density: copy failed"));
  int rc = sphere_to_sticks<T>(c, nb, psi, c->d_t1);
  if (rc) return done(rc);
  const C* tw = (const C*)(sizeof(T) == 8 ? (const void*)c->d_tw64 : (const void*)c->d_tw32);
  SlabGeom g{c->n1, c->n2, c->n3, c->sp.nstick, c->sp.kxmax, 0, 0};
  const int thr = pick_wfn_layout(c, sizeof(C), g);
  const size_t lds = slab_bytes(c, sizeof(C), g.kxmax, g.rp, g.cg);
  if ((rc = set_lds(k_density_acc<T>, lds)) != PW_OK) return done(rc);
  hipLaunchKernelGGL(k_density_acc<T>, dim3(c->n3, nsplit), dim3(thr), lds, c->stream, (const C*)c->d_t1,
                     (const double*)df, 0, nb, nsplit, part, c->d_pos, c->d_mpos, g, 1.0 / (double)N, tw,
                     tw + c->n1);
  hipLaunchKernelGGL(k_sum_partials, dim3((int)std::min<long>((N + 255) / 256, 4096)), dim3(256), 0, c->stream,
                     (const double*)part, nsplit, N, rho, accumulate);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(c->stream) != hipSuccess)
    return done(fail(PW_ERR_HIP, "pw_density: kernel failed"));
  return done(PW_OK);
}

template <typename C>
__global__ void k_make_dbeta(const C* __restrict__ beta, const double* __restrict__ G, int npw, int nbeta,
                             C* __restrict__ out) {
  const long total = (long)npw * nbeta;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e % npw);
    const C b = beta[e];
    for (int d = 0; d < 3; d++) {
      const double gd = G[3 * i + d];
      out[(size_t)d * total + e] = C{(decltype(b.x))(gd * b.y), (decltype(b.x))(-gd * b.x)};  // -i G_d beta
    }
  }
}

}  // namespace pw

extern "C" int pw_density(pw_ctx* c, int dtype, int nb, const void* psi, const double* f, double* rho,
                          int accumulate) {
  if (!c || !psi || !f || !rho || nb < 1 || nb > c->max_bands) return fail(PW_ERR_ARG, "pw_density: bad argument");
  if (dtype == PW_F64) return density_impl<double>(c, nb, psi, f, rho, accumulate);
  if (dtype == PW_F32) return density_impl<float>(c, nb, psi, f, rho, accumulate);
  return fail(PW_ERR_ARG, "unknown dtype %d", dtype);
}

namespace pw {
static int device_sum(pw_ctx* c, const double* x, long n, double* out) {
  const int nblk = 256;
  double* part = nullptr;
  if (hipMalloc((void**)&part, sizeof(double) * nblk) != hipSuccess)
    return fail(PW_ERR_NOMEM, "device_sum: scratch allocation failed");
  hipLaunchKernelGGL(k_block_sums, dim3(nblk), dim3(256), 0, c->stream, x, n, part);
  double h[nblk];
  const bool ok = hipGetLastError() == hipSuccess &&
                  hipMemcpyAsync(h, part, sizeof(h), hipMemcpyDeviceToHost, c->stream) == hipSuccess &&
                  hipStreamSynchronize(c->stream) == hipSuccess;
  (void)hipFree(part);
  if (!ok) return fail(PW_ERR_HIP, "device_sum: reduction failed");
  double s = 0.0;
  for (int i = 0; i < nblk; i++) s += h[i];
  *out = s;
  return PW_OK;
}
}  // namespace pw

// KohnSham::CalculateDensity, sphere branch with one density component
// (src/SCF/hamiltonian.cpp:2453-2627): rho = sum_i numSpin occ_i
// (C2R(c_i)/sqrt(N))^2, then rho *= numSpin nocc N / (vol sum(rho))
// (:2599-2600); *val = sum(rho) vol / N after the scaling (:2604-2607).
extern "C" int pw_calculate_density(pw_ctx* c, int dtype, int nb, const void* psi, const double* occ, int num_spin,
                                    int num_occupied, double* rho, double* val) {
  if (!c || !psi || !occ || !rho || nb < 1 || nb > c->max_bands || num_spin < 1 || num_occupied < 0)
    return fail(PW_ERR_ARG, "pw_calculate_density: bad argument");
  std::vector<double> f(occ, occ + nb);
  for (auto& x : f) x *= num_spin;
  PW_TRY(pw_density(c, dtype, nb, psi, f.data(), rho, 0));
  double s = 0.0;
  PW_TRY(device_sum(c, rho, c->N, &s));
  if (!(s > 0.0)) return fail(PW_ERR_DOMAIN, "CalculateDensity: the density sums to %g", s);
  const double scale = (double)num_spin * (double)num_occupied * (double)c->N / (c->vol * s);
  hipLaunchKernelGGL(k_scale, dim3((int)std::min<long>((c->N + 255) / 256, 4096)), dim3(256), 0, c->stream, rho,
                     c->N, scale);
  PW_HIP(hipGetLastError());
  PW_TRY(device_sum(c, rho, c->N, &s));
  if (val) *val = s * c->vol / (double)c->N;
  return PW_OK;
}

namespace pw {
template <typename T>
static int band_energy_impl(pw_ctx* c, int dtype, int nb, const void* X, const double* f, double* E) {
  const long m = 2L * c->npw;
  T* HX = nullptr;
  double* dots = nullptr;
  if (hipMalloc((void**)&HX, sizeof(T) * m * nb) != hipSuccess ||
      hipMalloc((void**)&dots, sizeof(double) * nb) != hipSuccess) {
    if (HX) (void)hipFree(HX);
    return fail(PW_ERR_NOMEM, "pw_band_energy: scratch allocation failed");
  }
  int rc = pw_apply_h(c, dtype, nb, X, HX);
  std::vector<double> h(nb);
  if (rc == PW_OK) {
    hipLaunchKernelGGL(k_band_dots<T>, dim3(nb), dim3(256), 0, c->stream, (const T*)X, (const T*)HX, m, dots);
    if (hipMemcpyAsync(h.data(), dots, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
      rc = fail(PW_ERR_HIP, "pw_band_energy: reduction failed");
  }
  (void)hipFree(HX);
  (void)hipFree(dots);
  if (rc != PW_OK) return rc;
  double e = 0.0;
  for (int i = 0; i < nb; i++) e += f[i] * h[i];
  *E = e;
  return PW_OK;
}

// Nonlocal force in the precision of the block (T): the projections
// W = Re(beta^H X) and D_d = Re((-i G_d beta)^H X) are T GEMMs on MFMA, the
// force sums are FP64.
template <typename T>
static int nonlocal_force_impl(pw_ctx* c, int nb, const void* X, const double* f, int natom, int npa,
                               double* force) {
  typedef typename cx<T>::t Cx;
  const int npw = c->npw, nbeta = c->nbeta, m = 2 * npw;
  const T* beta = sizeof(T) == 8 ? (const T*)(const void*)c->d_beta : (const T*)(const void*)c->d_beta_f;
  if (!beta) return fail(PW_ERR_STATE, "pw_nonlocal_force: no projectors set");
  // G vectors of the sphere
  std::vector<double> G(3 * (size_t)npw);
  for (int i = 0; i < npw; i++) {
    const int g = c->h_idx[i];
    const int kx = g % c->h1, ky = (g / c->h1) % c->n2, kz = g / (c->h1 * c->n2);
    const int m0 = kgrid(kx, c->n1), m1 = kgrid(ky, c->n2), m2 = kgrid(kz, c->n3);
    for (int d = 0; d < 3; d++) G[3 * i + d] = c->recip[d] * m0 + c->recip[3 + d] * m1 + c->recip[6 + d] * m2;
  }
  double *dG = nullptr, *dw = nullptr, *dfv = nullptr, *dF = nullptr;
  T *dB = nullptr, *dW = nullptr, *dD = nullptr;
  auto done = [&](int rc) {
    void* p[] = {dG, dB, dW, dD, dw, dfv, dF};
    for (void* q : p)
      if (q) (void)hipFree(q);
    return rc;
  };
  const size_t blk = (size_t)m * nbeta;
  if (hipMalloc((void**)&dG, sizeof(double) * 3 * npw) != hipSuccess ||
      hipMalloc((void**)&dB, sizeof(T) * 3 * blk) != hipSuccess ||
      hipMalloc((void**)&dW, sizeof(T) * (size_t)nbeta * nb) != hipSuccess ||
      hipMalloc((void**)&dD, sizeof(T) * 3 * (size_t)nbeta * nb) != hipSuccess ||
      hipMalloc((void**)&dw, sizeof(double) * nbeta) != hipSuccess ||
      hipMalloc((void**)&dfv, sizeof(double) * nb) != hipSuccess ||
      hipMalloc((void**)&dF, sizeof(double) * 3 * natom) != hipSuccess)
    return done(fail(PW_ERR_NOMEM, "pw_nonlocal_force: scratch allocation failed"));
  if (hipMemcpy(dG, G.data(), sizeof(double) * 3 * npw, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dw, c->h_wgt.data(), sizeof(double) * nbeta, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(dfv, f, sizeof(double) * nb, hipMemcpyHostToDevice) != hipSuccess)
    return done(fail(PW_ERR_HIP, "pw_nonlocal_force: upload failed"));
  hipLaunchKernelGGL(k_make_dbeta<Cx>, dim3(4096), dim3(256), 0, c->stream, (const Cx*)beta, (const double*)dG, npw,
                     nbeta, (Cx*)dB);
  int rc = gemm_tn<T>(c, m, nbeta, nb, beta, m, (const T*)X, m, dW, nbeta, nullptr, (T)0);
  for (int d = 0; d < 3 && rc == PW_OK; d++)
    rc = gemm_tn<T>(c, m, nbeta, nb, dB + d * blk, m, (const T*)X, m, dD + (size_t)d * nbeta * nb, nbeta, nullptr,
                    (T)0);
  if (rc != PW_OK) return done(rc);
  hipLaunchKernelGGL(k_force_reduce<T>, dim3(natom, 3), dim3(256), 0, c->stream, (const T*)dW, (const T*)dD,
                     (const double*)dw, (const double*)dfv, nbeta, nb, npa, dF);
  if (hipMemcpyAsync(force, dF, sizeof(double) * 3 * natom, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    return done(fail(PW_ERR_HIP, "pw_nonlocal_force: reduction failed"));
  return done(PW_OK);
}
}  // namespace pw

extern "C" int pw_band_energy(pw_ctx* c, int dtype, int nb, const void* X, const double* f, double* E) {
  if (!c || !X || !f || !E || nb < 1 || nb > c->max_bands) return fail(PW_ERR_ARG, "pw_band_energy: bad argument");
  if (dtype == PW_F64) return band_energy_impl<double>(c, dtype, nb, X, f, E);
  if (dtype == PW_F32) return band_energy_impl<float>(c, dtype, nb, X, f, E);
  return fail(PW_ERR_ARG, "unknown dtype %d", dtype);
}

extern "C" int pw_nonlocal_force(pw_ctx* c, int dtype, int nb, const void* X, const double* f, int natom, int npa,
                                 double* force) {
  if (!c || !X || !f || !force || nb < 1 || nb > c->max_bands) return fail(PW_ERR_ARG, "pw_nonlocal_force: bad argument");
  if (natom * npa != c->nbeta) return fail(PW_ERR_ARG, "pw_nonlocal_force: natom*npa = %d != nbeta = %d", natom * npa, c->nbeta);
  if (dtype == PW_F64) return nonlocal_force_impl<double>(c, nb, X, f, natom, npa, force);
  if (dtype == PW_F32) return nonlocal_force_impl<float>(c, nb, X, f, natom, npa, force);
  return fail(PW_ERR_ARG, "unknown dtype %d", dtype);
}
