// Many independent cells of uneven grids and band counts in one persistent
// launch (FP32 H*psi and the scaled Chebyshev filter around it).
//
// The reference runs one cell per MPI job and applies H one band at a time
// (Spinor::AddMultSpinorFineR2C, src/SCF/spinor.cpp:562-677; the
// NonlocalMultX GEMMs, src/EigenSolver/eigensolver.cpp:471-515).  Here a
// batch holds many cells, each a device context of its own (sphere, stick
// plan, potential, projectors).  The filter acts on every band column by
// itself, so a (cell, band chunk) pair is an independent work item: one
// workgroup takes an item from a queue (an atomic counter over the items,
// largest first) and runs the whole degree-m filter on it --
//
//   per step:  sphere -> z-inverse sticks (LDS Stockham, several 16-stick
//              chunks per pass) -> per z plane: y IFFT, paired-row x c2r,
//              x V/N, x r2c, y FFT -> z-forward sticks -> sphere with
//              sqrt(2) and ekin c  (the three kernels of pw_hpsi.hip as
//              workgroup-level device functions);
//              projectors: W = Re(beta^H Y) and (beta w) W on FP32 MFMA
//              (v_mfma_f32_16x16x4f32), fused with the three-term
//              recurrence Y_k = s1 (H Y_{k-1} - c Y_{k-1}) - s2 Y_{k-2}
//
// -- with its sticks and blocks in a private slice of a global scratch, then
// takes the next item.  A workgroup never waits for another one: the grid
// drains when the queue is empty.  Transform lengths are compile-time for
// the cubic grids listed in PW_BATCH_SIZES (one switch per item) and
// runtime otherwise.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "pw_fft.h"
#include "pw_internal.h"

namespace pw {

typedef float f4v_b __attribute__((ext_vector_type(4)));

// Device view of one cell (pointers into its context).
struct BatchCell {
  int n1, n2, n3, npw, nstick, kxmax, sc, nchunk, nbeta, nb;
  int chunk;  // bands per work item
  int rp;     // x row pairs per slab step
  int zc;     // 16-stick chunks per z pass
  long xoff;  // complex offset of the cell's npw x nb block in the batch buffer
  const int *chunk_off, *pt_idx, *pt_sz, *gchunk_off, *gpt_idx, *gpt_sz, *pos, *mpos;
  const float *vtot, *ekin, *beta, *betaw;
  const float2* tw;  // roots of n1 | n2 | n3
  int coef_off;      // offset of this cell's (c, s1, s2) table in the coefficient array
};

constexpr int kBatchThreads = 256;
constexpr int kBatchMaxChunk = 16;  // bands per item: one MFMA column tile
constexpr int kNlK = 32;            // real rows per staged projector tile
constexpr int kNlTiles = 4;         // 16-projector tiles per wavefront (nbeta <= 256)
constexpr int kMaxDegree = 64;

// ------------------------------------------------------------ z transforms
// Sphere coefficients (src, ld npw) -> sticks T[b][z][s] for nbi bands.
template <int N3>
__device__ void bt_inverse_z(const BatchCell& c, const float2* __restrict__ src, int nbi, float2* __restrict__ T,
                             const float2* tw3, float2* A, float2* Bz) {
  const int n3 = N3 > 0 ? N3 : c.n3;
  const int sc = c.sc, nstick = c.nstick;
  const float is2 = 0.70710678118654752440f;
  const int ngrp = (c.nchunk + c.zc - 1) / c.zc;
  for (int job = 0; job < ngrp * nbi; job++) {
    const int g = job / nbi, b = job - g * nbi;
    const int ch0 = g * c.zc, nch = min(c.zc, c.nchunk - ch0);
    const int P = nch * sc;
    const float2* ps = src + (size_t)b * c.npw;
    for (int t = threadIdx.x; t < P * n3; t += blockDim.x) A[t] = float2{0.f, 0.f};
    __syncthreads();
    for (int j = 0; j < nch; j++) {
      const int e = c.chunk_off[ch0 + j + 1];
      for (int t = c.chunk_off[ch0 + j] + threadIdx.x; t < e; t += blockDim.x) {
        const int i = c.pt_idx[t];
        const int sz = c.pt_sz[t];
        float2 v;
        if (i >= 0) {
          v = ps[i];
          if (i != 0) v = cscale(v, is2);
        } else {
          const float2 u = ps[-i - 1];
          v = float2{u.x * is2, -u.y * is2};
        }
        A[(j * sc + (sz & 0xffff)) * n3 + (sz >> 16)] = v;
      }
    }
    __syncthreads();
    float2* R = lds_fft_n<float, N3>(A, Bz, n3, P, n3, tw3, 1.f);
    const int s0 = ch0 * sc, nsl = min(P, nstick - s0);
    for (int t = threadIdx.x; t < n3 * nsl; t += blockDim.x) {
      const int z = t / nsl, sl = t - z * nsl;
      T[((size_t)b * n3 + z) * nstick + s0 + sl] = R[sl * n3 + z];
    }
    __syncthreads();
  }
}

// Sticks T -> forward z -> sphere: out[b][i] = ekin_i src[b][i] + (sqrt 2) y.
template <int N3>
__device__ void bt_forward_z(const BatchCell& c, const float2* __restrict__ src, int nbi,
                             const float2* __restrict__ T, float2* __restrict__ out, const float2* tw3, float2* A,
                             float2* Bz) {
  const int n3 = N3 > 0 ? N3 : c.n3;
  const int sc = c.sc, nstick = c.nstick;
  const float s2 = 1.4142135623730950488f;
  const int ngrp = (c.nchunk + c.zc - 1) / c.zc;
  for (int job = 0; job < ngrp * nbi; job++) {
    const int g = job / nbi, b = job - g * nbi;
    const int ch0 = g * c.zc, nch = min(c.zc, c.nchunk - ch0);
    const int P = nch * sc, s0 = ch0 * sc, nsl = min(P, nstick - s0);
    for (int t = threadIdx.x; t < n3 * P; t += blockDim.x) {
      const int z = t / P, sl = t - z * P;
      A[sl * n3 + z] = (sl < nsl) ? T[((size_t)b * n3 + z) * nstick + s0 + sl] : float2{0.f, 0.f};
    }
    __syncthreads();
    const float2* R = lds_fft_n<float, N3>(A, Bz, n3, P, n3, tw3, -1.f);
    const float2* ps = src + (size_t)b * c.npw;
    float2* o = out + (size_t)b * c.npw;
    for (int j = 0; j < nch; j++) {
      const int e = c.gchunk_off[ch0 + j + 1];
      for (int t = c.gchunk_off[ch0 + j] + threadIdx.x; t < e; t += blockDim.x) {
        const int i = c.gpt_idx[t];
        const int sz = c.gpt_sz[t];
        float2 y = R[(j * sc + (sz & 0xffff)) * n3 + (sz >> 16)];
        if (i != 0) y = cscale(y, s2);
        const float2 cc = ps[i];
        const float ek = c.ekin[i];
        o[i] = float2{ek * cc.x + y.x, ek * cc.y + y.y};
      }
    }
    __syncthreads();
  }
}

// ------------------------------------------------------------ xy slab
// Per (z, band): scatter sticks (+ conj mirrors of the kx = 0 plane), inverse
// y on the kx <= kxmax columns, paired-row x c2r, x V/N, x r2c, forward y,
// gather the sticks back in place (the arithmetic of k_slab_xy).
template <int N>
__device__ void bt_slab(const BatchCell& c, int nbi, float2* __restrict__ T, const float2* tw1, const float2* tw2,
                        float2* lds) {
  const int n1 = N > 0 ? N : c.n1, n2 = N > 0 ? N : c.n2, n3 = c.n3;
  const int nk = c.kxmax + 1, rp = c.rp, nstick = c.nstick;
  float2* Y0 = lds;
  float2* Y1 = Y0 + nk * n2;
  float2* Z0 = Y1 + nk * n2;
  float2* Z1 = Z0 + rp * n1;
  const int npair = (n2 + 1) / 2;
  for (int job = 0; job < n3 * nbi; job++) {
    const int z = job / nbi, b = job - z * nbi;
    float2* st = T + ((size_t)b * n3 + z) * nstick;
    for (int t = threadIdx.x; t < nk * n2; t += blockDim.x) Y0[t] = float2{0.f, 0.f};
    __syncthreads();
    for (int s = threadIdx.x; s < nstick; s += blockDim.x) {
      const float2 v = st[s];
      Y0[c.pos[s]] = v;
      const int mp = c.mpos[s];
      if (mp >= 0) Y0[mp] = float2{v.x, -v.y};
    }
    __syncthreads();
    float2* Y = lds_fft_n<float, N>(Y0, Y1, n2, nk, n2, tw2, 1.f);
    const float* vz = c.vtot + (size_t)z * n1 * n2;
    for (int q0 = 0; q0 < npair; q0 += rp) {
      const int nq = min(rp, npair - q0);
      float2* Zr = x_c2r_rows<float>(Y, c.kxmax, n1, n2, q0, nq, Z0, Z1, tw1);
      for (int t = threadIdx.x; t < nq * n1; t += blockDim.x) {
        const int ql = t / n1, x = t - ql * n1;
        const int y1 = 2 * (q0 + ql), y2 = y1 + 1;
        const float2 u = Zr[ql * n1 + x];
        const float v1 = vz[(size_t)y1 * n1 + x];
        const float v2 = (y2 < n2) ? vz[(size_t)y2 * n1 + x] : 0.f;
        Zr[ql * n1 + x] = float2{u.x * v1, u.y * v2};
      }
      __syncthreads();
      x_r2c_rows<float>(Zr, (Zr == Z0) ? Z1 : Z0, Y, c.kxmax, n1, n2, q0, nq, tw1);
    }
    const float2* Yf = lds_fft_n<float, N>(Y, (Y == Y0) ? Y1 : Y0, n2, nk, n2, tw2, -1.f);
    for (int s = threadIdx.x; s < nstick; s += blockDim.x) st[s] = Yf[c.pos[s]];
    __syncthreads();
  }
}

// ------------------------------------------------------------ projectors
// Ws (16-row tiles of projectors x 16 bands) = Re(beta^H Y) over the m =
// 2 npw real rows; bands b >= nbi and projectors p >= nbeta are zero.
__device__ void bt_project(const BatchCell& c, const float* __restrict__ Y, int nbi, float* lds) {
  const int m = 2 * c.npw, nbeta = c.nbeta;
  const int np16 = (nbeta + 15) / 16;
  float* Bs = lds;
  float* Ys = Bs + np16 * 16 * (kNlK + 1);
  float* Ws = Ys + 16 * (kNlK + 1);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  f4v_b acc[kNlTiles];
#pragma unroll
  for (int j = 0; j < kNlTiles; j++) acc[j] = f4v_b{0.f, 0.f, 0.f, 0.f};
  for (int r0 = 0; r0 < m; r0 += kNlK) {
    for (int t = threadIdx.x; t < np16 * 16 * kNlK; t += blockDim.x) {
      const int p = t / kNlK, k = t - p * kNlK, r = r0 + k;
      Bs[p * (kNlK + 1) + k] = (p < nbeta && r < m) ? c.beta[(size_t)p * m + r] : 0.f;
    }
    for (int t = threadIdx.x; t < 16 * kNlK; t += blockDim.x) {
      const int b = t / kNlK, k = t - b * kNlK, r = r0 + k;
      Ys[b * (kNlK + 1) + k] = (b < nbi && r < m) ? Y[(size_t)b * m + r] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kNlK; kk += 4) {
      const float bv = Ys[(lane & 15) * (kNlK + 1) + kk + (lane >> 4)];
#pragma unroll
      for (int j = 0; j < kNlTiles; j++) {
        const int pt = w + 4 * j;
        if (pt < np16) {
          const float av = Bs[(pt * 16 + (lane & 15)) * (kNlK + 1) + kk + (lane >> 4)];
          acc[j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc[j], 0, 0, 0);
        }
      }
    }
    __syncthreads();
  }
  // FP32 accumulator layout: register i of lane l = (row 4 (l/16) + i, column l % 16)
#pragma unroll
  for (int j = 0; j < kNlTiles; j++) {
    const int pt = w + 4 * j;
    if (pt < np16)
#pragma unroll
      for (int i = 0; i < 4; i++) Ws[(pt * 16 + 4 * (lane >> 4) + i) * 16 + (lane & 15)] = acc[j][i];
  }
  __syncthreads();
}

// h = HY + (beta w) Ws on the real rows; mode 0: HY <- h; mode 1:
// out <- s1 (h - cc Y) - s2 Yp (Yp may be null; out may alias Yp).
__device__ void bt_apply_combine(const BatchCell& c, const float* __restrict__ Y, float* HY, float* out,
                                 const float* Yp, int nbi, float cc, float s1, float s2, int mode, const float* lds) {
  const int m = 2 * c.npw, nbeta = c.nbeta;
  const int np16 = (nbeta + 15) / 16;
  const float* Ws = lds + np16 * 16 * (kNlK + 1) + 16 * (kNlK + 1);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nrt = (m + 15) / 16;
  const int b = lane & 15;
  for (int rt = w; rt < nrt; rt += 4) {
    f4v_b acc = f4v_b{0.f, 0.f, 0.f, 0.f};
    const int r = rt * 16 + (lane & 15);
    for (int k0 = 0; k0 < np16 * 16; k0 += 4) {
      const int p = k0 + (lane >> 4);
      const float av = (r < m && p < nbeta) ? c.betaw[(size_t)p * m + r] : 0.f;
      const float bv = Ws[p * 16 + b];
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc, 0, 0, 0);
    }
    if (b < nbi) {
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int rr = rt * 16 + 4 * (lane >> 4) + i;
        if (rr < m) {
          const size_t e = (size_t)b * m + rr;
          const float h = HY[e] + acc[i];
          if (mode == 0) {
            HY[e] = h;
          } else {
            float v = (h - cc * Y[e]) * s1;
            if (Yp) v -= s2 * Yp[e];
            out[e] = v;
          }
        }
      }
    }
  }
  __syncthreads();
}

// ------------------------------------------------------------ one item
// H Y (local + kinetic into HY, then the projector term fused with the
// recurrence step or added).
template <int N>
__device__ void bt_hpsi_step(const BatchCell& c, const float2* src, int nbi, float2* T, float2* HY, float2* out,
                             const float2* Yp, float cc, float s1, float s2, int mode, float2* lds) {
  const float2* tw1 = lds;
  const float2* tw2 = tw1 + c.n1;
  const float2* tw3 = tw2 + c.n2;
  float2* work = lds + c.n1 + c.n2 + c.n3;
  const int zlen = c.zc * c.sc * c.n3;
  bt_inverse_z<N>(c, src, nbi, T, tw3, work, work + zlen);
  bt_slab<N>(c, nbi, T, tw1, tw2, work);
  bt_forward_z<N>(c, src, nbi, T, HY, tw3, work, work + zlen);
  float* lf = reinterpret_cast<float*>(work);
  if (c.nbeta > 0) bt_project(c, reinterpret_cast<const float*>(src), nbi, lf);
  bt_apply_combine(c, reinterpret_cast<const float*>(src), reinterpret_cast<float*>(HY),
                   reinterpret_cast<float*>(out), reinterpret_cast<const float*>(Yp), nbi, cc, s1, s2, mode, lf);
}

template <int N>
__device__ void bt_item(const BatchCell& c, int b0, int nbi, float2* __restrict__ X, float2* __restrict__ HXout,
                        const float* __restrict__ coefs, int degree, float2* ws, float2* lds) {
  for (int t = threadIdx.x; t < c.n1 + c.n2 + c.n3; t += blockDim.x) lds[t] = c.tw[t];
  __syncthreads();
  const size_t tsz = (size_t)c.chunk * c.n3 * c.nstick, bsz = (size_t)c.chunk * c.npw;
  float2* T = ws;
  float2* A = T + tsz;
  float2* B = A + bsz;
  float2* HY = B + bsz;
  float2* Xc = X + c.xoff + (size_t)b0 * c.npw;
  if (degree == 0) {
    // H X into HXout (HY holds the local part, then the projector term is added)
    float2* Hc = HXout + c.xoff + (size_t)b0 * c.npw;
    bt_hpsi_step<N>(c, Xc, nbi, T, Hc, nullptr, nullptr, 0.f, 0.f, 0.f, 0, lds);
    return;
  }
  const float* cf = coefs + c.coef_off;
  const float2* cur = Xc;
  const float2* prev = nullptr;
  for (int step = 1; step <= degree; step++) {
    float2* dst = step == 1 ? A : (step == 2 ? B : const_cast<float2*>(prev));
    const float cc = cf[3 * (step - 1)], s1 = cf[3 * (step - 1) + 1], s2 = cf[3 * (step - 1) + 2];
    bt_hpsi_step<N>(c, cur, nbi, T, HY, dst, step == 1 ? nullptr : prev, cc, s1, s2, 1, lds);
    prev = cur;
    cur = dst;
  }
  for (int t = threadIdx.x; t < nbi * c.npw; t += blockDim.x) Xc[t] = cur[t];
  __syncthreads();
}

#define PW_BATCH_SIZES(X) X(24) X(30) X(32) X(36) X(40) X(48) X(60) X(72)

// Persistent kernel: every workgroup takes items until the queue is empty.
// degree 0: HX = H X; degree m > 0: X <- p_m(H) X in place.
__global__ void __launch_bounds__(kBatchThreads, 2)
k_batch(const BatchCell* __restrict__ cells, const int4* __restrict__ items, int nitems, int* __restrict__ counter,
        float2* __restrict__ X, float2* __restrict__ HX, const float* __restrict__ coefs, int degree,
        float2* __restrict__ scratch, long wg_scratch) {
  extern __shared__ __align__(16) unsigned char bsmem[];
  float2* lds = reinterpret_cast<float2*>(bsmem);
  __shared__ int s_item;
  float2* ws = scratch + (size_t)blockIdx.x * wg_scratch;
  for (;;) {
    __syncthreads();
    if (threadIdx.x == 0) s_item = atomicAdd(counter, 1);
    __syncthreads();
    // uniform across the workgroup: scalar loads of the item and its cell
    const int it = __builtin_amdgcn_readfirstlane(s_item);
    if (it >= nitems) break;
    const int ci = __builtin_amdgcn_readfirstlane(items[it].x);
    const int ib0 = __builtin_amdgcn_readfirstlane(items[it].y);
    const int inb = __builtin_amdgcn_readfirstlane(items[it].z);
    const int4 item = int4{ci, ib0, inb, 0};
    const BatchCell c = cells[ci];
    const bool cubic = c.n1 == c.n2 && c.n2 == c.n3;
    switch (cubic ? c.n1 : 0) {
#define PW_CASE(n)                                                        \
  case n:                                                                 \
    bt_item<n>(c, item.y, item.z, X, HX, coefs, degree, ws, lds);         \
    break;
      PW_BATCH_SIZES(PW_CASE)
#undef PW_CASE
      default:
        bt_item<0>(c, item.y, item.z, X, HX, coefs, degree, ws, lds);
        break;
    }
  }
}

__global__ void k_batch_reset(int* counter) {
  if (threadIdx.x == 0) counter[0] = 0;
}

__global__ void k_tune_fill_b(float* __restrict__ x, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    x[i] = __sinf(0.7548776662f * (float)(i % 100003) + 0.1f);
}

}  // namespace pw

using namespace pw;

struct pw_batch {
  int device = 0;
  hipStream_t stream = nullptr;
  int ncell = 0;
  std::vector<pw_ctx*> ctx;
  std::vector<int> nb;
  std::vector<long> xoff;
  long xtotal = 0;
  std::vector<BatchCell> h_cells;
  BatchCell* d_cells = nullptr;
  std::vector<int4> h_items;
  int4* d_items = nullptr;
  int* d_counter = nullptr;
  float2* d_scratch = nullptr;
  long wg_scratch = 0;  // complex floats per workgroup
  int nwg = 0;
  size_t lds = 0;
  long target = 0;  // work per item (band x grid points), 0 = default
  int rp_force = 0, zc_force = 0, nwg_force = 0;
  float* d_coef = nullptr;
  std::vector<float> h_coef;
  bool use_graphs = true;
  hipStream_t gstream = nullptr;
  hipEvent_t gev[2] = {nullptr, nullptr};
  hipGraphExec_t gexec = nullptr;
  long gkey = -1, gwarm = -1;
  const void *gX = nullptr, *gHX = nullptr;
};

namespace pw {

// LDS bytes an item of cell c needs: twiddles + the largest stage.
static size_t batch_cell_lds(const BatchCell& c) {
  const size_t tw = (size_t)(c.n1 + c.n2 + c.n3) * sizeof(float2);
  const size_t z = 2 * (size_t)c.zc * c.sc * c.n3 * sizeof(float2);
  const size_t nk = c.kxmax + 1;
  const size_t slab = (2 * nk * c.n2 + 2 * (size_t)c.rp * c.n1) * sizeof(float2);
  const size_t np16 = (c.nbeta + 15) / 16;
  const size_t nl = c.nbeta > 0 ? (np16 * 16 * (kNlK + 1) + 16 * (kNlK + 1) + np16 * 16 * 16) * sizeof(float) : 0;
  return tw + std::max(z, std::max(slab, nl));
}

// Items, their band chunks, the per-workgroup scratch, the LDS and the
// persistent grid from the current tuning (target work per item, x row
// pairs, z chunks per pass).
static int batch_plan(pw_batch* b) {
  const long target = b->target > 0 ? b->target : 16L * 24 * 24 * 24;
  b->h_items.clear();
  b->wg_scratch = 0;
  b->lds = 0;
  for (int i = 0; i < b->ncell; i++) {
    BatchCell& c = b->h_cells[i];
    const long N = (long)c.n1 * c.n2 * c.n3;
    c.chunk = (int)std::max(1L, std::min<long>(kBatchMaxChunk, (target + N / 2) / N));
    c.chunk = std::min(c.chunk, c.nb);
    // x row pairs: all of them when the slab's LDS stays within 64 KiB
    const int npair = (c.n2 + 1) / 2;
    c.rp = b->rp_force > 0 ? std::min(b->rp_force, npair) : npair;
    while (c.rp > 1 && (2 * (size_t)(c.kxmax + 1) * c.n2 + 2 * (size_t)c.rp * c.n1) * sizeof(float2) > 64 * 1024)
      c.rp = (c.rp + 1) / 2;
    // z passes: as many 16-stick chunks as keep two pencil buffers within 48 KiB
    c.zc = b->zc_force > 0 ? b->zc_force
                           : (int)std::max<size_t>(1, (48 * 1024) / (2 * (size_t)c.sc * c.n3 * sizeof(float2)));
    c.zc = std::max(1, std::min(c.zc, c.nchunk));
    b->lds = std::max(b->lds, batch_cell_lds(c));
    const long need = (long)c.chunk * ((long)c.n3 * c.nstick + 3L * c.npw);
    b->wg_scratch = std::max(b->wg_scratch, need);
    for (int b0 = 0; b0 < c.nb; b0 += c.chunk) b->h_items.push_back(int4{i, b0, std::min(c.chunk, c.nb - b0), 0});
  }
  if (b->lds > 160 * 1024) return fail(PW_ERR_UNSUPPORTED, "batch: an item needs %zu bytes of LDS", b->lds);
  // largest items first (longest-processing-time order)
  auto cost = [&](const int4& it) {
    const BatchCell& c = b->h_cells[it.x];
    const double N = (double)c.n1 * c.n2 * c.n3;
    return (double)it.z * (N * std::log2(N) + 8.0 * c.npw * std::max(c.nbeta, 16));
  };
  std::stable_sort(b->h_items.begin(), b->h_items.end(),
                   [&](const int4& x, const int4& y) { return cost(x) > cost(y); });
  // persistent grid: as many workgroups as fit the CUs at this LDS size
  int ncu = 0;
  PW_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, b->device));
  const int per_cu = std::max(1, std::min(8, (int)((160 * 1024) / std::max<size_t>(b->lds, 1))));
  b->nwg = b->nwg_force > 0 ? b->nwg_force : ncu * per_cu;
  b->nwg = std::max(1, std::min(b->nwg, (int)b->h_items.size()));
  return PW_OK;
}

static void batch_free_device(pw_batch* b) {
  void* p[] = {b->d_cells, b->d_items, b->d_counter, b->d_scratch, b->d_coef};
  for (void* q : p)
    if (q) (void)hipFree(q);
  b->d_cells = nullptr;
  b->d_items = nullptr;
  b->d_counter = nullptr;
  b->d_scratch = nullptr;
  b->d_coef = nullptr;
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  b->gexec = nullptr;
  b->gkey = b->gwarm = -1;
}

static int batch_upload(pw_batch* b) {
  batch_free_device(b);
  PW_TRY(batch_plan(b));
  PW_HIP(hipMalloc((void**)&b->d_cells, sizeof(BatchCell) * b->ncell));
  PW_HIP(hipMemcpy(b->d_cells, b->h_cells.data(), sizeof(BatchCell) * b->ncell, hipMemcpyHostToDevice));
  PW_HIP(hipMalloc((void**)&b->d_items, sizeof(int4) * b->h_items.size()));
  PW_HIP(hipMemcpy(b->d_items, b->h_items.data(), sizeof(int4) * b->h_items.size(), hipMemcpyHostToDevice));
  PW_HIP(hipMalloc((void**)&b->d_counter, sizeof(int) * 4));
  PW_HIP(hipMemset(b->d_counter, 0, sizeof(int) * 4));
  const size_t sb = sizeof(float2) * (size_t)b->wg_scratch * b->nwg;
  if (hipMalloc((void**)&b->d_scratch, sb) != hipSuccess)
    return fail(PW_ERR_NOMEM, "batch: cannot allocate %zu bytes of workgroup scratch", sb);
  PW_HIP(hipMalloc((void**)&b->d_coef, sizeof(float) * 3 * kMaxDegree * b->ncell));
  PW_TRY(ensure_lds((const void*)k_batch, b->lds));
  return PW_OK;
}

static int batch_launch(pw_batch* b, const void* X, void* HX, int degree) {
  hipStream_t s = b->stream;
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, b->d_counter);
  hipLaunchKernelGGL(k_batch, dim3(b->nwg), dim3(kBatchThreads), b->lds, s, (const BatchCell*)b->d_cells,
                     (const int4*)b->d_items, (int)b->h_items.size(), b->d_counter, (float2*)X, (float2*)HX,
                     (const float*)b->d_coef, degree, b->d_scratch, b->wg_scratch);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

// The reset and the persistent launch, captured into a graph on the second
// call of a shape (degree, X, HX) and replayed from the third on.
static int batch_run(pw_batch* b, const void* X, void* HX, int degree) {
  const long key = degree;
  if (!b->use_graphs) return batch_launch(b, X, HX, degree);
  if (b->gexec && b->gkey == key && b->gX == X && b->gHX == HX) {
    PW_HIP(hipEventRecord(b->gev[0], b->stream));
    PW_HIP(hipStreamWaitEvent(b->gstream, b->gev[0], 0));
    PW_HIP(hipGraphLaunch(b->gexec, b->gstream));
    PW_HIP(hipEventRecord(b->gev[1], b->gstream));
    PW_HIP(hipStreamWaitEvent(b->stream, b->gev[1], 0));
    return PW_OK;
  }
  if (b->gwarm != key || b->gX != X || b->gHX != HX) {
    b->gwarm = key;
    b->gX = X;
    b->gHX = HX;
    return batch_launch(b, X, HX, degree);
  }
  if (!b->gstream) PW_HIP(hipStreamCreateWithFlags(&b->gstream, hipStreamNonBlocking));
  if (!b->gev[0]) {
    PW_HIP(hipEventCreateWithFlags(&b->gev[0], hipEventDisableTiming));
    PW_HIP(hipEventCreateWithFlags(&b->gev[1], hipEventDisableTiming));
  }
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  b->gexec = nullptr;
  hipStream_t user = b->stream;
  b->stream = b->gstream;
  hipGraph_t g = nullptr;
  int rc = PW_OK;
  if (hipStreamBeginCapture(b->gstream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    rc = fail(PW_ERR_HIP, "batch: hipStreamBeginCapture failed");
  } else {
    rc = batch_launch(b, X, HX, degree);
    const hipError_t e = hipStreamEndCapture(b->gstream, &g);
    if (rc == PW_OK && e != hipSuccess) rc = fail(PW_ERR_HIP, "batch: hipStreamEndCapture: %s", hipGetErrorString(e));
    if (rc == PW_OK && hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0) != hipSuccess)
      rc = fail(PW_ERR_HIP, "batch: hipGraphInstantiate failed");
  }
  if (g) (void)hipGraphDestroy(g);
  b->stream = user;
  if (rc != PW_OK) return rc;
  b->gkey = key;
  PW_HIP(hipEventRecord(b->gev[0], b->stream));
  PW_HIP(hipStreamWaitEvent(b->gstream, b->gev[0], 0));
  PW_HIP(hipGraphLaunch(b->gexec, b->gstream));
  PW_HIP(hipEventRecord(b->gev[1], b->gstream));
  PW_HIP(hipStreamWaitEvent(b->stream, b->gev[1], 0));
  return PW_OK;
}

}  // namespace pw

extern "C" int pw_batch_destroy(pw_batch* b) {
  if (!b) return PW_OK;
  batch_free_device(b);
  for (pw_ctx* c : b->ctx)
    if (c) pw_ctx_destroy(c);
  if (b->gstream) (void)hipStreamDestroy(b->gstream);
  for (auto& e : b->gev)
    if (e) (void)hipEventDestroy(e);
  delete b;
  return PW_OK;
}

extern "C" int pw_batch_create(pw_batch** out, int ncell, const int* grid, const double* supercell, const int* npw,
                               const int* idx, const int* npad, const int* pad_src, const int* pad_dst,
                               const double* ekin, const double* teter, const double* vtot, const int* nbeta,
                               const double* beta, const double* wgt, const int* nb, void* stream) {
  if (!out || ncell < 1 || !grid || !supercell || !npw || !idx || !npad || !ekin || !teter || !vtot || !nbeta ||
      !nb)
    return fail(PW_ERR_ARG, "pw_batch_create: bad argument");
  pw_batch* b = new pw_batch();
  b->stream = (hipStream_t)stream;
  b->ncell = ncell;
  if (hipGetDevice(&b->device) != hipSuccess) {
    delete b;
    return fail(PW_ERR_HIP, "pw_batch_create: no device");
  }
  long io = 0, ip = 0, iv = 0, ib = 0, iw = 0;
  for (int i = 0; i < ncell; i++) {
    if (nb[i] < 1 || npw[i] < 1 || nbeta[i] < 0 || nbeta[i] > 16 * kNlTiles * 4) {
      pw_batch_destroy(b);
      return fail(PW_ERR_ARG, "pw_batch_create: cell %d: nb %d, npw %d, nbeta %d out of range", i, nb[i], npw[i],
                  nbeta[i]);
    }
    pw_ctx* c = nullptr;
    int rc = pw_ctx_create(&c, grid + 3 * i, supercell + 9 * i, npw[i], idx + io, npad[i], pad_src + ip, pad_dst + ip,
                           ekin + io, teter + io, nb[i], stream);
    if (rc == PW_OK) rc = pw_set_vtot(c, vtot + iv);
    if (rc == PW_OK && nbeta[i] > 0) rc = pw_set_nonlocal(c, nbeta[i], beta + ib, wgt + iw);
    if (rc != PW_OK) {
      if (c) pw_ctx_destroy(c);
      pw_batch_destroy(b);
      return rc;
    }
    b->ctx.push_back(c);
    b->nb.push_back(nb[i]);
    b->xoff.push_back(b->xtotal);
    BatchCell bc{};
    bc.n1 = c->n1; bc.n2 = c->n2; bc.n3 = c->n3;
    bc.npw = c->npw; bc.nstick = c->sp.nstick; bc.kxmax = c->sp.kxmax; bc.sc = c->sp.sc; bc.nchunk = c->sp.nchunk;
    bc.nbeta = nbeta[i]; bc.nb = nb[i];
    bc.xoff = b->xtotal;
    bc.chunk_off = c->d_chunk_off; bc.pt_idx = c->d_pt_idx; bc.pt_sz = c->d_pt_sz;
    bc.gchunk_off = c->d_gchunk_off; bc.gpt_idx = c->d_gpt_idx; bc.gpt_sz = c->d_gpt_sz;
    bc.pos = c->d_pos; bc.mpos = c->d_mpos;
    bc.vtot = c->d_vtot_f; bc.ekin = c->d_ekin_f; bc.beta = c->d_beta_f; bc.betaw = c->d_betaw_f;
    bc.tw = (const float2*)c->d_tw32;
    bc.coef_off = 3 * kMaxDegree * i;
    b->h_cells.push_back(bc);
    b->xtotal += (long)npw[i] * nb[i];
    io += npw[i];
    ip += npad[i];
    iv += (long)grid[3 * i] * grid[3 * i + 1] * grid[3 * i + 2];
    ib += 2L * npw[i] * nbeta[i];
    iw += nbeta[i];
  }
  if (const char* e = getenv("PW_BATCH")) {
    // PW_BATCH=<target>:<rp>:<zc>:<nwg> (0 = default) for measurement
    long t = 0;
    int r = 0, z = 0, n = 0;
    if (sscanf(e, "%ld:%d:%d:%d", &t, &r, &z, &n) >= 1) {
      b->target = t;
      b->rp_force = r;
      b->zc_force = z;
      b->nwg_force = n;
    }
  }
  const int rc = batch_upload(b);
  if (rc != PW_OK) {
    pw_batch_destroy(b);
    return rc;
  }
  *out = b;
  return PW_OK;
}

extern "C" int pw_batch_info(const pw_batch* b, long* xtotal, int* nitems, int* nwg, long* target) {
  if (!b) return fail(PW_ERR_ARG, "null batch");
  if (xtotal) *xtotal = b->xtotal;
  if (nitems) *nitems = (int)b->h_items.size();
  if (nwg) *nwg = b->nwg;
  if (target) *target = b->target > 0 ? b->target : 16L * 24 * 24 * 24;
  return PW_OK;
}

extern "C" int pw_batch_tune(pw_batch* b, long target, int rp, int zc, int nwg) {
  if (!b || target < 0 || rp < 0 || zc < 0 || nwg < 0) return fail(PW_ERR_ARG, "pw_batch_tune: bad argument");
  b->target = target;
  b->rp_force = rp;
  b->zc_force = zc;
  b->nwg_force = nwg;
  return batch_upload(b);
}

extern "C" int pw_batch_set_graphs(pw_batch* b, int enable) {
  if (!b) return fail(PW_ERR_ARG, "null batch");
  b->use_graphs = enable != 0;
  return PW_OK;
}

extern "C" int pw_batch_apply_h(pw_batch* b, const void* X, void* HX) {
  if (!b || !X || !HX || X == HX) return fail(PW_ERR_ARG, "pw_batch_apply_h: bad argument");
  return batch_run(b, X, HX, 0);
}

extern "C" int pw_batch_cheb_filter(pw_batch* b, int m, const double* bounds, void* X) {
  if (!b || !X || !bounds || m < 1 || m > kMaxDegree) return fail(PW_ERR_ARG, "pw_batch_cheb_filter: bad argument");
  b->h_coef.assign((size_t)3 * kMaxDegree * b->ncell, 0.f);
  std::vector<double> cf(3 * kMaxDegree);
  for (int i = 0; i < b->ncell; i++) {
    const double a = bounds[3 * i], bu = bounds[3 * i + 1], aL = bounds[3 * i + 2];
    if (!(bu > a) || !(aL < a)) return fail(PW_ERR_ARG, "pw_batch_cheb_filter: cell %d needs aL < a < b", i);
    cheb_coef(a, bu, aL, m, cf.data());
    for (int k = 0; k < 3 * m; k++) b->h_coef[3 * kMaxDegree * i + k] = (float)cf[k];
  }
  PW_HIP(hipMemcpyAsync(b->d_coef, b->h_coef.data(), sizeof(float) * b->h_coef.size(), hipMemcpyHostToDevice,
                        b->stream));
  return batch_run(b, X, nullptr, m);
}

// The same filter cell by cell on each cell's own context (pw_cheb_filter,
// FP32): the one-by-one reference point of the batched launch.
extern "C" int pw_batch_cheb_filter_serial(pw_batch* b, int m, const double* bounds, void* X) {
  if (!b || !X || !bounds || m < 1 || m > kMaxDegree) return fail(PW_ERR_ARG, "pw_batch_cheb_filter_serial: bad argument");
  for (int i = 0; i < b->ncell; i++) {
    pw_ctx* c = b->ctx[i];
    PW_TRY(pw_ctx_set_stream(c, b->stream));
    PW_TRY(pw_cheb_filter(c, PW_F32, b->nb[i], m, bounds[3 * i], bounds[3 * i + 1], bounds[3 * i + 2],
                          (float2*)X + b->xoff[i]));
  }
  return PW_OK;
}

extern "C" int pw_batch_set_stream(pw_batch* b, void* stream) {
  if (!b) return fail(PW_ERR_ARG, "null batch");
  b->stream = (hipStream_t)stream;
  for (pw_ctx* c : b->ctx) PW_TRY(pw_ctx_set_stream(c, stream));
  return PW_OK;
}

// Autotuner of the batched launch's work split (see pw_tune.cpp for the
// single-cell plans): work per item, x row pairs per slab step and stick
// chunks per z pass, timed on the batch's own H X over every cell; the
// fastest is kept and recorded in cache_path under a key of the device and
// the batch's shapes.
extern "C" int pw_batch_autotune(pw_batch* b, const char* cache_path, char* summary, int len) {
  if (!b) return fail(PW_ERR_ARG, "null batch");
  std::string key;
  {
    hipDeviceProp_t p;
    std::string dev = "dev";
    if (hipGetDeviceProperties(&p, b->device) == hipSuccess) {
      dev = p.gcnArchName;
      const size_t colon = dev.find(':');
      if (colon != std::string::npos) dev = dev.substr(0, colon);
      dev += "-" + std::to_string(p.multiProcessorCount) + "cu";
    }
    unsigned long h = 1469598103934665603UL;  // FNV-1a over the cells' shapes
    for (const BatchCell& c : b->h_cells)
      for (int v : {c.n1, c.n2, c.n3, c.npw, c.nb, c.nbeta}) h = (h ^ (unsigned long)v) * 1099511628211UL;
    char hb[32];
    snprintf(hb, sizeof(hb), "%016lx", h);
    key = dev + "|batch" + std::to_string(b->ncell) + "|" + hb;
  }
  std::string out;
  if (cache_path) {
    FILE* f = fopen(cache_path, "r");
    if (f) {
      char line[512];
      long t = 0;
      int r = 0, z = 0, found = 0;
      while (fgets(line, sizeof(line), f)) {
        char k[400];
        long t2;
        int r2, z2;
        if (sscanf(line, "%399s %ld %d %d", k, &t2, &r2, &z2) == 4 && key == k) {
          t = t2; r = r2; z = z2; found = 1;
        }
      }
      fclose(f);
      if (found) {
        b->target = t;
        b->rp_force = r;
        b->zc_force = z;
        PW_TRY(batch_upload(b));
        char s[160];
        snprintf(s, sizeof(s), "batch target %ld rp %d zc %d (cached)", t, r, z);
        if (summary && len > 0) snprintf(summary, (size_t)len, "%s", s);
        return PW_OK;
      }
    }
  }
  float2 *X = nullptr, *HX = nullptr;
  auto done = [&](int rc) {
    if (X) (void)hipFree(X);
    if (HX) (void)hipFree(HX);
    return rc;
  };
  if (hipMalloc((void**)&X, sizeof(float2) * b->xtotal) != hipSuccess ||
      hipMalloc((void**)&HX, sizeof(float2) * b->xtotal) != hipSuccess)
    return done(fail(PW_ERR_NOMEM, "pw_batch_autotune: cannot allocate the timing blocks"));
  const long n = 2 * b->xtotal;
  hipLaunchKernelGGL(k_tune_fill_b, dim3((int)std::min<long>((n + 255) / 256, 8192)), dim3(256), 0, b->stream,
                     (float*)X, n);
  const bool graphs = b->use_graphs;
  b->use_graphs = false;
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return done(fail(PW_ERR_HIP, "pw_batch_autotune: event creation failed"));
  const long unit = 24L * 24 * 24;
  double tbest = 1e30;
  long bt = 0;
  int br = 0, bz = 0, tried = 0;
  int rc = PW_OK;
  for (long tg : {2 * unit, 4 * unit, 8 * unit, 16 * unit, 48 * unit})
    for (int r : {0, 8})
      for (int z : {0, 1}) {
        b->target = tg;
        b->rp_force = r;
        b->zc_force = z;
        if ((rc = batch_upload(b)) != PW_OK) break;
        if ((rc = batch_launch(b, X, HX, 0)) != PW_OK) break;
        float best = 1e30f;
        for (int rep = 0; rep < 2 && rc == PW_OK; rep++) {
          (void)hipEventRecord(e0, b->stream);
          rc = batch_launch(b, X, HX, 0);
          (void)hipEventRecord(e1, b->stream);
          if (hipEventSynchronize(e1) != hipSuccess) rc = fail(PW_ERR_HIP, "pw_batch_autotune: device error");
          float t = 0.f;
          (void)hipEventElapsedTime(&t, e0, e1);
          best = std::min(best, t);
        }
        if (rc != PW_OK) break;
        tried++;
        if (best < tbest) {
          tbest = best;
          bt = tg;
          br = r;
          bz = z;
        }
      }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  b->use_graphs = graphs;
  if (rc != PW_OK) return done(rc);
  b->target = bt;
  b->rp_force = br;
  b->zc_force = bz;
  PW_TRY(batch_upload(b));
  if (cache_path) {
    FILE* f = fopen(cache_path, "a");
    if (f) {
      fprintf(f, "%s %ld %d %d\n", key.c_str(), bt, br, bz);
      fclose(f);
    }
  }
  char s[200];
  snprintf(s, sizeof(s), "batch target %ld rp %d zc %d (%.3f ms H X over %d splits)", bt, br, bz, tbest, tried);
  out = s;
  if (summary && len > 0) snprintf(summary, (size_t)len, "%s", out.c_str());
  return done(PW_OK);
}
