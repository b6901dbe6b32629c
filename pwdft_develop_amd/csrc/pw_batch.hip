// Many independent cells of uneven grids and band counts in one persistent
// launch (FP32 H*psi and the scaled Chebyshev filter around it).
//
// The reference runs one cell per MPI job and applies H one band at a time
// (Spinor::AddMultSpinorFineR2C, src/SCF/spinor.cpp:562-677; the
// NonlocalMultX GEMMs, src/EigenSolver/eigensolver.cpp:471-515).  Here a
// batch holds many cells, each a device context of its own (sphere, stick
// plan, potential, projectors), and one launch of a persistent grid runs the
// whole degree-m filter (or one H X) for all of them.  Every filter step is
// four stages over work queues that mix all cells' (cell, band, ...) items:
//
//   1. sphere -> z-inverse sticks of one band (several 16-stick chunks per
//      LDS Stockham pass)  +  W = Re(beta^H Y) tiles of 16 projectors x 16
//      bands (FP32 MFMA, the K rows split over the four wavefronts)
//   2. one z plane of sb bands: y IFFT, paired-row x c2r, x V/N, x r2c, y FFT
//      (each 1D FFT in one thread's registers: the body of k_slab_xy_reg)
//   3. z-forward sticks of one band -> sphere: HY = ekin c + (sqrt 2) y
//   4. 64 real rows x 16 bands: (beta w) W on FP32 MFMA, fused with the
//      three-term recurrence Y_k = s1 (HY + (beta w) W - c Y_{k-1}) - s2 Y_{k-2}
//
// The workgroups take the items of a stage in a fixed stride (queues sorted
// by cost, largest first) and meet at a device-wide barrier between stages:
// an atomic arrival counter with agent-scope release / acquire fences, whose
// wait is bounded (a barrier that does not fill within about four seconds
// sets an error word and lets every workgroup run to its end).  The grid is
// sized to what the device holds at once (occupancy x CUs), so every
// workgroup is resident and each barrier fills.  Transform lengths are
// compile-time for the cubic grids of PW_BATCH_SIZES, runtime otherwise.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "pw_fft.h"
#include "pw_internal.h"
#include "pw_slab_reg.h"

namespace pw {

typedef float f4v_b __attribute__((ext_vector_type(4)));

// Device view of one cell (pointers into its context and the batch buffers).
struct BatchCell {
  int n1, n2, n3, npw, nstick, kxmax, sc, nchunk, nbeta, nb;
  int rp;     // x row pairs per slab step (LDS slab of runtime lengths)
  int zc;     // 16-stick chunks per z job
  int sb;     // bands per slab job (register-FFT slab: 8 up to 48 points, 4 beyond; 1 otherwise)
  long xoff;  // complex offset of the cell's npw x nb block in the batch buffers
  const int *chunk_off, *pt_idx, *pt_sz, *gchunk_off, *gpt_idx, *gpt_sz, *pos, *mpos;
  const float *vtot, *ekin, *beta, *betaw;
  const float2* tw;  // roots of n1 | n2 | n3
  float2* T;         // sticks, nb x n3 x nstick
  float* W;          // projections, nbeta x nb
  int coef_off;      // offset of this cell's (c, s1, s2) table in the coefficient array
};

constexpr int kBatchThreads = 256;
constexpr int kMaxDegree = 64;
constexpr long kBarrierSpins = 1L << 26;

enum : int { kJobInvZ = 0, kJobProj = 1, kJobSlab = 2, kJobFwdZ = 3, kJobApply = 4 };

// ------------------------------------------------------------ device barrier
// Arrival counter bar[0] (zeroed before the launch), error word bar[1].
// Barrier k of the launch waits for k * gridDim.x arrivals.  Every wavefront
// fences at agent scope before the arrival (its stores complete and are
// written back past its XCD's L2) and after the release (its caches hold no
// stale lines), since the workgroups of a stage run on all XCDs.
__device__ __forceinline__ void grid_barrier(unsigned* bar, unsigned k) {
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    atomicAdd(bar, 1u);
    const unsigned target = k * gridDim.x;
    long spins = 0;
    while (__hip_atomic_load(bar, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < target) {
      if (__hip_atomic_load(bar + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) break;
      __builtin_amdgcn_s_sleep(2);
      if (++spins > kBarrierSpins) {
        __hip_atomic_store(bar + 1, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
    }
  }
  __syncthreads();
  __threadfence();
}

// ------------------------------------------------------------ stage 1/3 jobs
// z-inverse of zc 16-stick chunks (group g) of band b: sphere (src, ld npw)
// -> sticks T[b][z][s].
template <int N3>
__device__ void bt_invz_job(const BatchCell& c, const float2* __restrict__ src, int b, int g, const float2* tw3,
                            float2* A, float2* Bz) {
  const int n3 = N3 > 0 ? N3 : c.n3;
  const int sc = c.sc, nstick = c.nstick;
  const float is2 = 0.70710678118654752440f;
  const int ch0 = g * c.zc, nch = min(c.zc, c.nchunk - ch0);
  const int P = nch * sc;
  const float2* ps = src + (size_t)b * c.npw;
  for (int t = threadIdx.x; t < P * n3; t += blockDim.x) A[t] = float2{0.f, 0.f};
  __syncthreads();
  // the chunks' points are contiguous in pt_idx / pt_sz: gather them in
  // batches of 8 per thread, every index load and then every coefficient
  // load in flight before the LDS stores
  {
    const int t0 = c.chunk_off[ch0], t1 = c.chunk_off[ch0 + nch];
    for (int tb = t0 + threadIdx.x; tb < t1; tb += 8 * blockDim.x) {
      int ii[8], zz[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int t = tb + u * blockDim.x;
        ii[u] = t < t1 ? c.pt_idx[t] : 0;
        zz[u] = t < t1 ? c.pt_sz[t] : -1;
      }
      float2 vv[8];
#pragma unroll
      for (int u = 0; u < 8; u++) vv[u] = ps[ii[u] >= 0 ? ii[u] : -ii[u] - 1];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int t = tb + u * blockDim.x;
        if (t >= t1) continue;
        // chunk j of the group: the first j with chunk_off[ch0 + j + 1] > t
        int j = 0;
        while (c.chunk_off[ch0 + j + 1] <= t) j++;
        float2 v = vv[u];
        if (ii[u] < 0) v = float2{v.x * is2, -v.y * is2};
        else if (ii[u] != 0) v = cscale(v, is2);
        A[(j * sc + (zz[u] & 0xffff)) * n3 + (zz[u] >> 16)] = v;
      }
    }
  }
  __syncthreads();
  float2* R = lds_fft_n<float, N3>(A, Bz, n3, P, n3, tw3, 1.f);
  const int s0 = ch0 * sc, nsl = min(P, nstick - s0);
  float2* T = c.T;
  for (int t = threadIdx.x; t < n3 * nsl; t += blockDim.x) {
    const int z = t / nsl, sl = t - z * nsl;
    T[((size_t)b * n3 + z) * nstick + s0 + sl] = R[sl * n3 + z];
  }
  __syncthreads();
}

// Sticks -> z-forward (group g of band b) -> sphere: HY[b][i] = ekin_i
// src[b][i] + (sqrt 2) y.
template <int N3>
__device__ void bt_fwdz_job(const BatchCell& c, const float2* __restrict__ src, float2* __restrict__ HY, int b, int g,
                            const float2* tw3, float2* A, float2* Bz) {
  const int n3 = N3 > 0 ? N3 : c.n3;
  const int sc = c.sc, nstick = c.nstick;
  const float s2 = 1.4142135623730950488f;
  const int ch0 = g * c.zc, nch = min(c.zc, c.nchunk - ch0);
  const int P = nch * sc, s0 = ch0 * sc, nsl = min(P, nstick - s0);
  const float2* T = c.T;
  for (int tb = threadIdx.x; tb < n3 * P; tb += 8 * blockDim.x) {
    float2 v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int t = tb + u * blockDim.x, z = t / P, sl = t - z * P;
      v[u] = (t < n3 * P && sl < nsl) ? T[((size_t)b * n3 + z) * nstick + s0 + sl] : float2{0.f, 0.f};
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int t = tb + u * blockDim.x, z = t / P, sl = t - z * P;
      if (t < n3 * P) A[sl * n3 + z] = v[u];
    }
  }
  __syncthreads();
  const float2* R = lds_fft_n<float, N3>(A, Bz, n3, P, n3, tw3, -1.f);
  const float2* ps = src + (size_t)b * c.npw;
  float2* o = HY + (size_t)b * c.npw;
  // batches of 8 points per thread: index loads, then coefficient and
  // ekin loads in flight together, then the sums
  {
    const int t0 = c.gchunk_off[ch0], t1 = c.gchunk_off[ch0 + nch];
    for (int tb = t0 + threadIdx.x; tb < t1; tb += 8 * blockDim.x) {
      int ii[8], zz[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int t = tb + u * blockDim.x;
        ii[u] = t < t1 ? c.gpt_idx[t] : 0;
        zz[u] = t < t1 ? c.gpt_sz[t] : 0;
      }
      float2 cc[8];
      float ek[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        cc[u] = ps[ii[u]];
        ek[u] = c.ekin[ii[u]];
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int t = tb + u * blockDim.x;
        if (t >= t1) continue;
        int j = 0;
        while (c.gchunk_off[ch0 + j + 1] <= t) j++;
        float2 y = R[(j * sc + (zz[u] & 0xffff)) * n3 + (zz[u] >> 16)];
        if (ii[u] != 0) y = cscale(y, s2);
        o[ii[u]] = float2{ek[u] * cc[u].x + y.x, ek[u] * cc[u].y + y.y};
      }
    }
  }
  __syncthreads();
}

// ------------------------------------------------------------ stage 2 jobs
// One z plane of band b: scatter sticks (+ conj mirrors of the kx = 0 plane),
// inverse y on the kx <= kxmax columns, paired-row x c2r, x V/N, x r2c,
// forward y, gather the sticks back in place (the arithmetic of k_slab_xy).
template <int N>
__device__ void bt_slab_job(const BatchCell& c, int z, int b, const float2* tw1, const float2* tw2, float2* lds) {
  const int n1 = N > 0 ? N : c.n1, n2 = N > 0 ? N : c.n2, n3 = c.n3;
  const int nk = c.kxmax + 1, rp = c.rp, nstick = c.nstick;
  float2* Y0 = lds;
  float2* Y1 = Y0 + nk * n2;
  float2* Z0 = Y1 + nk * n2;
  float2* Z1 = Z0 + rp * n1;
  const int npair = (n2 + 1) / 2;
  float2* st = c.T + ((size_t)b * n3 + z) * nstick;
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

// ------------------------------------------------------------ projector jobs
// W[p0 .. p0+15][b0 .. b0+15] = Re(beta^H Y) over the m = 2 npw real rows:
// wavefront w sums one quarter of the rows on v_mfma_f32_16x16x4f32 with
// operands straight from memory (A lane l = (projector p0 + l % 16, row
// k + l / 16), B lane l = (row k + l / 16, band b0 + l % 16)); the four
// partial tiles are added in LDS in wavefront order.
__device__ void bt_proj_job(const BatchCell& c, const float* __restrict__ Y, int p0, int b0, float* lds) {
  const int m = 2 * c.npw;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int p = p0 + (lane & 15), b = b0 + (lane & 15), kl = lane >> 4;
  const bool pv = p < c.nbeta, bv = b < c.nb;
  const float* bp = c.beta + (size_t)(pv ? p : 0) * m;
  const float* yp = Y + (size_t)(bv ? b : 0) * m;
  const int qr = ((m + 3) / 4 + 3) & ~3;  // rows per wavefront, a multiple of 4
  const int kb = w * qr, ke = min(m, kb + qr);
  f4v_b acc = f4v_b{0.f, 0.f, 0.f, 0.f};
  int k = kb;
  for (; k + 16 <= ke; k += 16) {
    float a[4], y[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int r = k + 4 * u + kl;
      a[u] = pv ? bp[r] : 0.f;
      y[u] = bv ? yp[r] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a[u], y[u], acc, 0, 0, 0);
  }
  for (; k < ke; k += 4) {
    const int r = k + kl;
    const float a = (pv && r < ke) ? bp[r] : 0.f;
    const float y = (bv && r < ke) ? yp[r] : 0.f;
    acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, y, acc, 0, 0, 0);
  }
  // accumulator register i of lane l: (projector 4 (l / 16) + i, band l % 16)
#pragma unroll
  for (int i = 0; i < 4; i++) lds[w * 256 + (4 * (lane >> 4) + i) * 16 + (lane & 15)] = acc[i];
  __syncthreads();
  if (threadIdx.x < 256) {
    const int pi = threadIdx.x >> 4, bi = threadIdx.x & 15;
    const float s = ((lds[threadIdx.x] + lds[256 + threadIdx.x]) + lds[512 + threadIdx.x]) + lds[768 + threadIdx.x];
    if (p0 + pi < c.nbeta && b0 + bi < c.nb) c.W[(size_t)(b0 + bi) * c.nbeta + p0 + pi] = s;
  }
  __syncthreads();
}

// 64 real rows (r0 ..) x 16 bands (b0 ..): h = HY + (beta w) W; mode 0:
// out = h; mode 1: out = s1 (h - cc Yc) - s2 Yp (Yp may be null, out may
// alias Yp).  Wavefront w takes rows r0 + 16 w .. + 15.
__device__ void bt_apply_job(const BatchCell& c, const float* __restrict__ HY, const float* __restrict__ Yc,
                             const float* Yp, float* out, int r0, int b0, float cc, float s1, float s2, int mode) {
  const int m = 2 * c.npw, nbeta = c.nbeta;
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int r = r0 + 16 * w + (lane & 15), b = b0 + (lane & 15), kl = lane >> 4;
  f4v_b acc = f4v_b{0.f, 0.f, 0.f, 0.f};
  if (nbeta > 0 && r0 + 16 * w < m) {
    const bool rv = r < m, bv = b < c.nb;
    for (int p = 0; p < nbeta; p += 4) {
      const int pp = p + kl;
      const float a = (rv && pp < nbeta) ? c.betaw[(size_t)pp * m + r] : 0.f;
      const float y = (bv && pp < nbeta) ? c.W[(size_t)b * nbeta + pp] : 0.f;
      acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, y, acc, 0, 0, 0);
    }
  }
  if (b < c.nb) {
#pragma unroll
    for (int i = 0; i < 4; i++) {
      const int rr = r0 + 16 * w + 4 * kl + i;
      if (rr < m) {
        const size_t e = (size_t)b * m + rr;
        const float h = HY[e] + acc[i];
        if (mode == 0) {
          out[e] = h;
        } else {
          float v = (h - cc * Yc[e]) * s1;
          if (Yp) v -= s2 * Yp[e];
          out[e] = v;
        }
      }
    }
  }
}

#define PW_BATCH_SIZES(X) X(24) X(30) X(32) X(36) X(40) X(48) X(60) X(72)

// One job of a transform stage on the cell's compile-time size.
template <int N>
__device__ void bt_transform_job(const BatchCell& c, int type, int a, int b, const float2* src, float2* HY,
                                 float2* lds) {
  const float2* tw1 = lds;
  const float2* tw2 = tw1 + c.n1;
  const float2* tw3 = tw2 + c.n2;
  float2* work = lds + c.n1 + c.n2 + c.n3;
  for (int t = threadIdx.x; t < c.n1 + c.n2 + c.n3; t += blockDim.x) lds[t] = c.tw[t];
  __syncthreads();
  const int zlen = c.zc * c.sc * c.n3;
  if (type == kJobInvZ) {
    bt_invz_job<N>(c, src, a, b, tw3, work, work + zlen);
  } else if (type == kJobSlab) {
    if constexpr (N > 0 && N % 2 == 0) {
      // plane a of bands b .. b + sb - 1, register FFTs
      if constexpr (N <= 48)
        slab_xy_reg_body<float, N, N, 8>(a, b, c.T, c.T, c.n3, c.nstick, c.kxmax, c.nb, c.pos, c.mpos, c.vtot,
                                         reinterpret_cast<unsigned char*>(lds));
      else
        slab_xy_reg_body<float, N, N, 4>(a, b, c.T, c.T, c.n3, c.nstick, c.kxmax, c.nb, c.pos, c.mpos, c.vtot,
                                         reinterpret_cast<unsigned char*>(lds));
    } else {
      bt_slab_job<N>(c, a, b, tw1, tw2, work);
    }
  } else {
    bt_fwdz_job<N>(c, src, HY, a, b, tw3, work, work + zlen);
  }
}

__device__ void bt_dispatch(const BatchCell& c, int type, int a, int b, const float2* src, float2* HY, float2* lds) {
  const bool cubic = c.n1 == c.n2 && c.n2 == c.n3;
  switch (cubic ? c.n1 : 0) {
#define PW_CASE(n)                                    \
  case n:                                             \
    bt_transform_job<n>(c, type, a, b, src, HY, lds); \
    break;
    PW_BATCH_SIZES(PW_CASE)
#undef PW_CASE
    default:
      bt_transform_job<0>(c, type, a, b, src, HY, lds);
      break;
  }
}

// The persistent launch.  jobs: int4 (type, cell, a, b), stage s's jobs at
// [joff[s], joff[s + 1]) for the four stages 0..3.  Buffers in the batch
// layout (complex offset xoff per cell): X (input, and the filter's result),
// YA, YB (recurrence), HY (local terms); degree 0 writes H X to HX.
struct BatchArgs {
  const BatchCell* cells;
  const int4* jobs;
  int joff[5];
  unsigned* bar;
  float2 *X, *YA, *YB, *HY, *HX;
  const float* coefs;
  int degree;
  long xtotal;
  unsigned long long* stamps;  // null, or one wall-clock stamp per stage end (+ the start)
};

__global__ void __launch_bounds__(kBatchThreads, 2) k_batch(BatchArgs g) {
  extern __shared__ __align__(16) unsigned char bsmem[];
  float2* lds = reinterpret_cast<float2*>(bsmem);
  unsigned nbar = 0;
  const int steps = g.degree > 0 ? g.degree : 1;
  if (g.stamps && blockIdx.x == 0 && threadIdx.x == 0) g.stamps[0] = wall_clock64();
  for (int step = 1; step <= steps; step++) {
    // recurrence buffers: step 1 reads X and writes YA; step 2 reads YA and
    // X and writes YB; later steps write over Y_{k-2}
    float2* cur = step == 1 ? g.X : ((step & 1) ? g.YB : g.YA);
    float2* prev = step == 1 ? nullptr : (step == 2 ? g.X : ((step & 1) ? g.YA : g.YB));
    float2* dst = g.degree == 0 ? g.HX : ((step & 1) ? g.YA : g.YB);
    for (int s = 0; s < 4; s++) {
      // round r of a stage gives job joff + r G + xcd_item(b, G) to
      // workgroup b (G workgroups): each XCD a contiguous run of the
      // round's jobs, so that neighbouring stick chunks of a band (which
      // share cache lines) meet in one L2
      for (int j = g.joff[s] + xcd_item(blockIdx.x, gridDim.x); j < g.joff[s + 1]; j += gridDim.x) {
        const int4 jb = g.jobs[j];
        const int type = __builtin_amdgcn_readfirstlane(jb.x), ci = __builtin_amdgcn_readfirstlane(jb.y);
        const int a = __builtin_amdgcn_readfirstlane(jb.z), b = __builtin_amdgcn_readfirstlane(jb.w);
        const BatchCell& c = g.cells[ci];
        const float2* src = cur + c.xoff;
        if (type == kJobProj) {
          bt_proj_job(c, reinterpret_cast<const float*>(src), a, b, reinterpret_cast<float*>(lds));
        } else if (type == kJobApply) {
          const float* cf = g.coefs + c.coef_off + 3 * (step - 1);
          const float cc = g.degree > 0 ? cf[0] : 0.f, s1 = g.degree > 0 ? cf[1] : 0.f;
          const float s2 = g.degree > 0 ? cf[2] : 0.f;
          bt_apply_job(c, reinterpret_cast<const float*>(g.HY + c.xoff), reinterpret_cast<const float*>(src),
                       prev ? reinterpret_cast<const float*>(prev + c.xoff) : nullptr,
                       reinterpret_cast<float*>(dst + c.xoff), a, b, cc, s1, s2, g.degree > 0 ? 1 : 0);
        } else {
          bt_dispatch(c, type, a, b, src, g.HY + c.xoff, lds);
        }
      }
      grid_barrier(g.bar, ++nbar);
      if (g.stamps && blockIdx.x == 0 && threadIdx.x == 0) g.stamps[nbar] = wall_clock64();
    }
  }
  if (g.degree > 0) {
    // the result (Y_m: YA for odd m, YB for even m) back into X
    const float2* res = (g.degree & 1) ? g.YA : g.YB;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < g.xtotal; i += (long)gridDim.x * blockDim.x)
      g.X[i] = res[i];
  }
}

__global__ void k_batch_reset(unsigned* bar) {
  if (threadIdx.x < 2) bar[threadIdx.x] = 0u;
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
  std::vector<int4> h_jobs;
  int joff[5] = {0, 0, 0, 0, 0};
  int4* d_jobs = nullptr;
  unsigned* d_bar = nullptr;
  float2 *d_YA = nullptr, *d_YB = nullptr, *d_HY = nullptr;
  float2* d_T = nullptr;  // all cells' stick buffers
  float* d_W = nullptr;   // all cells' projections
  int nwg = 0;
  size_t lds = 0;
  int rp_force = 0, zc_force = 0, nwg_force = 0;
  float* d_coef = nullptr;
  std::vector<float> h_coef;
  bool use_graphs = true;
  hipStream_t gstream = nullptr;
  hipEvent_t gev[2] = {nullptr, nullptr};
  hipGraphExec_t gexec = nullptr;
  long gkey = -1, gwarm = -1;
  const void *gX = nullptr, *gHX = nullptr;
  unsigned long long* d_stamps = nullptr;  // stage timing (pw_batch_set_timing)
  int stamp_steps = 0;                      // steps of the last timed launch
};

namespace pw {

// LDS bytes of a transform job of cell c: twiddles + the larger of the z
// pencils and the slab buffers (a projector job needs 4 KiB).
static bool batch_reg_slab(const BatchCell& c) {
  if (c.n1 != c.n2 || c.n2 != c.n3 || c.n1 % 2) return false;
  switch (c.n1) {
#define PW_CASE(n) case n:
    PW_BATCH_SIZES(PW_CASE)
#undef PW_CASE
    return true;
    default:
      return false;
  }
}

static size_t batch_cell_lds(const BatchCell& c) {
  const size_t tw = (size_t)(c.n1 + c.n2 + c.n3) * sizeof(float2);
  const size_t z = 2 * (size_t)c.zc * c.sc * c.n3 * sizeof(float2);
  const size_t nk = c.kxmax + 1;
  size_t slab = tw + (2 * nk * c.n2 + 2 * (size_t)c.rp * c.n1) * sizeof(float2);
  if (c.sb > 1)  // register-FFT slab: SB y-spaces, V row pairs, stick positions (no twiddles)
    slab = (size_t)c.sb * nk * c.n2 * sizeof(float2) + (size_t)(c.n2 / 2) * (c.n1 + 1) * sizeof(float2) +
           2 * sizeof(int) * (size_t)c.nstick;
  return std::max(std::max(tw + z, slab), (size_t)4096);
}

// Layouts (row pairs, z chunks), the job queues of the four stages (each
// sorted by cost, largest first) and the grid.
static int batch_plan(pw_batch* b) {
  b->lds = 0;
  for (BatchCell& c : b->h_cells) {
    const int npair = (c.n2 + 1) / 2;
    c.rp = b->rp_force > 0 ? std::min(b->rp_force, npair) : npair;
    while (c.rp > 1 && (2 * (size_t)(c.kxmax + 1) * c.n2 + 2 * (size_t)c.rp * c.n1) * sizeof(float2) > 64 * 1024)
      c.rp = (c.rp + 1) / 2;
    c.zc = b->zc_force > 0 ? b->zc_force
                           : (int)std::max<size_t>(1, (24 * 1024) / (2 * (size_t)c.sc * c.n3 * sizeof(float2)));
    c.zc = std::max(1, std::min(c.zc, c.nchunk));
    c.sb = batch_reg_slab(c) ? (c.n1 <= 48 ? 8 : 4) : 1;
    b->lds = std::max(b->lds, batch_cell_lds(c));
  }
  if (b->lds > 160 * 1024) return fail(PW_ERR_UNSUPPORTED, "batch: a job needs %zu bytes of LDS", b->lds);
  struct J {
    int4 j;
    double cost;
  };
  std::vector<J> st[4];
  for (int i = 0; i < b->ncell; i++) {
    const BatchCell& c = b->h_cells[i];
    const int ngrp = (c.nchunk + c.zc - 1) / c.zc;
    const double zcost = (double)c.zc * c.sc * c.n3 * std::log2((double)c.n3 + 1);
    for (int bb = 0; bb < c.nb; bb++)
      for (int g = 0; g < ngrp; g++) {
        st[0].push_back({int4{kJobInvZ, i, bb, g}, zcost});
        st[2].push_back({int4{kJobFwdZ, i, bb, g}, zcost});
      }
    for (int p0 = 0; p0 < c.nbeta; p0 += 16)
      for (int b0 = 0; b0 < c.nb; b0 += 16) st[0].push_back({int4{kJobProj, i, p0, b0}, 2.0 * c.npw * 16});
    const double scost = (double)c.sb * c.n1 * c.n2 * std::log2((double)c.n1 * c.n2);
    for (int z = 0; z < c.n3; z++)
      for (int bb = 0; bb < c.nb; bb += c.sb) st[1].push_back({int4{kJobSlab, i, z, bb}, scost});
    const int m = 2 * c.npw;
    for (int r0 = 0; r0 < m; r0 += 64)
      for (int b0 = 0; b0 < c.nb; b0 += 16) st[3].push_back({int4{kJobApply, i, r0, b0}, 64.0 * (c.nbeta + 8)});
  }
  b->h_jobs.clear();
  for (int s = 0; s < 4; s++) {
    std::stable_sort(st[s].begin(), st[s].end(), [](const J& x, const J& y) { return x.cost > y.cost; });
    b->joff[s] = (int)b->h_jobs.size();
    for (const J& j : st[s]) b->h_jobs.push_back(j.j);
  }
  b->joff[4] = (int)b->h_jobs.size();
  // the grid: every workgroup resident at once (the barriers need it)
  int ncu = 0, per_cu = 0;
  PW_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, b->device));
  PW_TRY(ensure_lds((const void*)k_batch, b->lds));
  PW_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, (const void*)k_batch, kBatchThreads, b->lds));
  if (per_cu < 1) return fail(PW_ERR_UNSUPPORTED, "batch: the persistent kernel does not fit a CU");
  const int full = ncu * per_cu;
  b->nwg = b->nwg_force > 0 ? std::min(b->nwg_force, full) : full;
  return PW_OK;
}

static void batch_free_device(pw_batch* b) {
  void* p[] = {b->d_cells, b->d_jobs, b->d_bar, b->d_coef, b->d_YA, b->d_YB, b->d_HY, b->d_T, b->d_W};
  for (void* q : p)
    if (q) (void)hipFree(q);
  b->d_cells = nullptr;
  b->d_jobs = nullptr;
  b->d_bar = nullptr;
  b->d_coef = nullptr;
  b->d_YA = b->d_YB = b->d_HY = b->d_T = nullptr;
  b->d_W = nullptr;
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
  b->gexec = nullptr;
  b->gkey = b->gwarm = -1;
}

static int batch_upload(pw_batch* b) {
  batch_free_device(b);
  // per-cell stick and projection buffers
  long tsum = 0, wsum = 0;
  for (const BatchCell& c : b->h_cells) {
    tsum += (long)c.nb * c.n3 * c.nstick;
    wsum += (long)c.nbeta * c.nb;
  }
  if (hipMalloc((void**)&b->d_T, sizeof(float2) * std::max(tsum, 1L)) != hipSuccess ||
      hipMalloc((void**)&b->d_W, sizeof(float) * std::max(wsum, 1L)) != hipSuccess ||
      hipMalloc((void**)&b->d_YA, sizeof(float2) * b->xtotal) != hipSuccess ||
      hipMalloc((void**)&b->d_YB, sizeof(float2) * b->xtotal) != hipSuccess ||
      hipMalloc((void**)&b->d_HY, sizeof(float2) * b->xtotal) != hipSuccess)
    return fail(PW_ERR_NOMEM, "batch: cannot allocate the batch buffers");
  tsum = wsum = 0;
  for (BatchCell& c : b->h_cells) {
    c.T = b->d_T + tsum;
    c.W = b->d_W + wsum;
    tsum += (long)c.nb * c.n3 * c.nstick;
    wsum += (long)c.nbeta * c.nb;
  }
  PW_TRY(batch_plan(b));
  PW_HIP(hipMalloc((void**)&b->d_cells, sizeof(BatchCell) * b->ncell));
  PW_HIP(hipMemcpy(b->d_cells, b->h_cells.data(), sizeof(BatchCell) * b->ncell, hipMemcpyHostToDevice));
  PW_HIP(hipMalloc((void**)&b->d_jobs, sizeof(int4) * std::max<size_t>(b->h_jobs.size(), 1)));
  PW_HIP(hipMemcpy(b->d_jobs, b->h_jobs.data(), sizeof(int4) * b->h_jobs.size(), hipMemcpyHostToDevice));
  PW_HIP(hipMalloc((void**)&b->d_bar, sizeof(unsigned) * 4));
  PW_HIP(hipMemset(b->d_bar, 0, sizeof(unsigned) * 4));
  PW_HIP(hipMalloc((void**)&b->d_coef, sizeof(float) * 3 * kMaxDegree * b->ncell));
  PW_HIP(hipMemset(b->d_coef, 0, sizeof(float) * 3 * kMaxDegree * b->ncell));
  return PW_OK;
}

static int batch_launch(pw_batch* b, const void* X, void* HX, int degree) {
  hipStream_t s = b->stream;
  BatchArgs g;
  g.cells = b->d_cells;
  g.jobs = b->d_jobs;
  for (int i = 0; i < 5; i++) g.joff[i] = b->joff[i];
  g.bar = b->d_bar;
  g.X = (float2*)X;
  g.YA = b->d_YA;
  g.YB = b->d_YB;
  g.HY = b->d_HY;
  g.HX = (float2*)HX;
  g.coefs = b->d_coef;
  g.degree = degree;
  g.xtotal = b->xtotal;
  g.stamps = b->d_stamps;
  if (b->d_stamps) b->stamp_steps = degree > 0 ? degree : 1;
  hipLaunchKernelGGL(k_batch_reset, dim3(1), dim3(64), 0, s, b->d_bar);
  hipLaunchKernelGGL(k_batch, dim3(b->nwg), dim3(kBatchThreads), b->lds, s, g);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

// After the launch: the barrier error word (a barrier that did not fill).
static int batch_check(pw_batch* b) {
  unsigned h[2] = {0, 0};
  PW_HIP(hipMemcpyAsync(h, b->d_bar, sizeof(h), hipMemcpyDeviceToHost, b->stream));
  PW_HIP(hipStreamSynchronize(b->stream));
  if (h[1] != 0u) return fail(PW_ERR_HIP, "batch: a device-wide barrier timed out (workgroups not co-resident)");
  return PW_OK;
}

// The reset and the persistent launch, captured into a graph on the second
// call of a shape (degree, X, HX) and replayed from the third on.
static int batch_run(pw_batch* b, const void* X, void* HX, int degree) {
  const long key = degree;
  if (!b->use_graphs) {
    PW_TRY(batch_launch(b, X, HX, degree));
    return batch_check(b);
  }
  if (b->gexec && b->gkey == key && b->gX == X && b->gHX == HX) {
    PW_HIP(hipEventRecord(b->gev[0], b->stream));
    PW_HIP(hipStreamWaitEvent(b->gstream, b->gev[0], 0));
    PW_HIP(hipGraphLaunch(b->gexec, b->gstream));
    PW_HIP(hipEventRecord(b->gev[1], b->gstream));
    PW_HIP(hipStreamWaitEvent(b->stream, b->gev[1], 0));
    return batch_check(b);
  }
  if (b->gwarm != key || b->gX != X || b->gHX != HX) {
    b->gwarm = key;
    b->gX = X;
    b->gHX = HX;
    PW_TRY(batch_launch(b, X, HX, degree));
    return batch_check(b);
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
  return batch_check(b);
}

}  // namespace pw

extern "C" int pw_batch_destroy(pw_batch* b) {
  if (!b) return PW_OK;
  batch_free_device(b);
  if (b->d_stamps) (void)hipFree(b->d_stamps);
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
    if (nb[i] < 1 || npw[i] < 1 || nbeta[i] < 0) {
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
  const int rc = batch_upload(b);
  if (rc != PW_OK) {
    pw_batch_destroy(b);
    return rc;
  }
  *out = b;
  return PW_OK;
}

extern "C" int pw_batch_info(const pw_batch* b, long* xtotal, int* njobs, int* nwg, int* lds) {
  if (!b) return fail(PW_ERR_ARG, "null batch");
  if (xtotal) *xtotal = b->xtotal;
  if (njobs) *njobs = b->joff[4];
  if (nwg) *nwg = b->nwg;
  if (lds) *lds = (int)b->lds;
  return PW_OK;
}

extern "C" int pw_batch_tune(pw_batch* b, int rp, int zc, int nwg) {
  if (!b || rp < 0 || zc < 0 || nwg < 0) return fail(PW_ERR_ARG, "pw_batch_tune: bad argument");
  b->rp_force = rp;
  b->zc_force = zc;
  b->nwg_force = nwg;
  return batch_upload(b);
}

// Stage timing: the launch stamps the device wall clock (100 MHz) at its
// start and after every stage barrier; pw_batch_timing sums the last
// launch's stage durations by stage.
extern "C" int pw_batch_set_timing(pw_batch* b, int enable) {
  if (!b) return fail(PW_ERR_ARG, "null batch");
  if (b->gexec) (void)hipGraphExecDestroy(b->gexec);  // the stamps pointer is baked into a graph
  b->gexec = nullptr;
  b->gkey = b->gwarm = -1;
  if (enable && !b->d_stamps) PW_HIP(hipMalloc((void**)&b->d_stamps, sizeof(unsigned long long) * (4 * kMaxDegree + 1)));
  if (!enable && b->d_stamps) {
    (void)hipFree(b->d_stamps);
    b->d_stamps = nullptr;
  }
  b->stamp_steps = 0;
  return PW_OK;
}

extern "C" int pw_batch_timing(pw_batch* b, double* stage_ms) {
  if (!b || !stage_ms) return fail(PW_ERR_ARG, "pw_batch_timing: bad argument");
  for (int i = 0; i < 4; i++) stage_ms[i] = 0.0;
  if (!b->d_stamps || b->stamp_steps == 0) return PW_OK;
  std::vector<unsigned long long> h(4 * (size_t)b->stamp_steps + 1);
  PW_HIP(hipStreamSynchronize(b->stream));
  PW_HIP(hipMemcpy(h.data(), b->d_stamps, sizeof(unsigned long long) * h.size(), hipMemcpyDeviceToHost));
  for (int k = 1; k < (int)h.size(); k++) stage_ms[(k - 1) % 4] += (double)(h[k] - h[k - 1]) * 1e-5;
  return PW_OK;
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

// Autotuner of the batched launch (see pw_tune.cpp for the single-cell
// plans): x row pairs per slab job, stick chunks per z job and the size of
// the grid, timed on the batch's own H X over every cell; the fastest is
// kept and recorded in cache_path under a key of the device and the batch's
// shapes.
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
  if (cache_path) {
    FILE* f = fopen(cache_path, "r");
    if (f) {
      char line[512];
      int r = 0, z = 0, n = 0, found = 0;
      while (fgets(line, sizeof(line), f)) {
        char k[400];
        int r2, z2, n2;
        if (sscanf(line, "%399s %d %d %d", k, &r2, &z2, &n2) == 4 && key == k) {
          r = r2; z = z2; n = n2; found = 1;
        }
      }
      fclose(f);
      if (found) {
        b->rp_force = r;
        b->zc_force = z;
        b->nwg_force = n;
        PW_TRY(batch_upload(b));
        char s[160];
        snprintf(s, sizeof(s), "batch rp %d zc %d nwg %d (cached)", r, z, n);
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
  hipEvent_t e0, e1;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
    return done(fail(PW_ERR_HIP, "pw_batch_autotune: event creation failed"));
  int ncu = 0;
  (void)hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, b->device);
  double tbest = 1e30;
  int br = 0, bz = 0, bn = 0, tried = 0;
  int rc = PW_OK;
  for (int r : {0, 8, 4})
    for (int z : {0, 1, 2})
      for (int nw : {0, ncu}) {
        b->rp_force = r;
        b->zc_force = z;
        b->nwg_force = nw;
        if ((rc = batch_upload(b)) != PW_OK) break;
        if ((rc = batch_launch(b, X, HX, 0)) != PW_OK || (rc = batch_check(b)) != PW_OK) break;
        float best = 1e30f;
        for (int rep = 0; rep < 2 && rc == PW_OK; rep++) {
          (void)hipEventRecord(e0, b->stream);
          rc = batch_launch(b, X, HX, 0);
          (void)hipEventRecord(e1, b->stream);
          if (rc == PW_OK) rc = batch_check(b);
          float t = 0.f;
          (void)hipEventElapsedTime(&t, e0, e1);
          best = std::min(best, t);
        }
        if (rc != PW_OK) break;
        tried++;
        if (best < tbest) {
          tbest = best;
          br = r;
          bz = z;
          bn = nw;
        }
      }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc != PW_OK) return done(rc);
  b->rp_force = br;
  b->zc_force = bz;
  b->nwg_force = bn;
  PW_TRY(batch_upload(b));
  if (cache_path) {
    FILE* f = fopen(cache_path, "a");
    if (f) {
      fprintf(f, "%s %d %d %d\n", key.c_str(), br, bz, bn);
      fclose(f);
    }
  }
  char s[200];
  snprintf(s, sizeof(s), "batch rp %d zc %d nwg %d (%.3f ms H X over %d layouts)", br, bz, bn, tbest, tried);
  if (summary && len > 0) snprintf(summary, (size_t)len, "%s", s);
  return done(PW_OK);
}
