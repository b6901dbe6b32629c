// The xy slab transform with every 1D FFT in one thread's registers, as
// workgroup-level device code: slab_xy_reg_body (one z plane of SB bands;
// the body of pw_hpsi.hip's k_slab_xy_reg, also run by the batched
// persistent kernel) and slab_xy_reg_persistent (a workgroup loops over
// (plane, band group) items and prefetches the next item's sticks and V
// plane into registers while it transforms the current one), both built
// from the phases of SlabReg.
//
// For compile-time N1 = N2: one z plane of SB bands per call (bands b0 ..).
// V's plane is staged once in LDS as (row 2q, row 2q + 1) pairs, one 64-bit
// word per x, pair rows padded to N1 + 1 words so that neighbouring threads
// hit distinct banks; each band's y-space lives in LDS, and the five phases
// (scatter, y IFFT per column, x c2r / V / r2c per row pair, y FFT per
// column, gather) are separated by one barrier each.  In the row-pair phase a
// thread holds N1 complex values; its LDS reads are issued in chunks fenced
// by scheduling barriers so that the compiler does not hoist all of them next
// to the spectrum (that spills past the 256 VGPRs of two waves per SIMD).
// NT threads per workgroup (blockDim.x == NT).
// LY lanes share each y column's transform (col_fft_lanes): LY = 1 keeps
// k_slab_xy's arithmetic and order; LY = 2 or 4 splits every y transform
// into LY shorter ones combined across lanes, so that SB nk columns occupy
// LY SB nk threads.  t1 may equal t2 (each
// plane slice is read whole before it is written).  LDS (smem_raw, 16-byte
// aligned): SB nk N2 complex + N2/2 (N1 + 1) complex + 2 nstick ints.
#pragma once
#include "pw_fft.h"

namespace pw {

template <typename T, int N1, int N2, int SB, int LY = 1, int NT = 256>
struct SlabReg {
  static_assert(N1 % 2 == 0 && N2 % 2 == 0, "even N1 and N2");
  typedef typename cx<T>::t C;
  static constexpr int H1 = N1 / 2, NPAIR = N2 / 2, LV = N1 + 1, KC = 8;
  static constexpr int KS = 16;                        // prefetched sticks per thread
  static constexpr int VB = (N1 * N2 + NT - 1) / NT;   // prefetched V values per thread
  C* Y;        // SB x nk x N2
  C* vp;       // NPAIR x LV pairs (v[2q][x], v[2q+1][x])
  T* vs;
  int* spos;   // nstick
  int* smpos;  // nstick
  int n3, nstick, kxmax, nk, nb;
  struct Pre {  // one item's global data, in registers
    C sk[KS];
    T vb[VB];
  };

  __device__ __forceinline__ SlabReg(unsigned char* smem_raw, int n3_, int nstick_, int kxmax_, int nb_)
      : n3(n3_), nstick(nstick_), kxmax(kxmax_), nk(kxmax_ + 1), nb(nb_) {
    Y = reinterpret_cast<C*>(smem_raw);
    vp = Y + (size_t)SB * nk * N2;
    vs = reinterpret_cast<T*>(vp);
    spos = reinterpret_cast<int*>(vp + (size_t)NPAIR * LV);
    smpos = spos + nstick;
  }

  // Global loads of item (z, b0): the first KS sticks per thread of the
  // plane's ns x nstick sticks and the V plane, all issued before any use.
  // Element e is (s, st) = (e / nstick, e % nstick), advanced incrementally
  // (no integer division per element).
  __device__ __forceinline__ void load(int z, int b0, const C* t1, const T* __restrict__ vtot, Pre& r) const {
    const int t = threadIdx.x, tot = min(SB, nb - b0) * nstick;
    int s_ = t / nstick, st_ = t - s_ * nstick;
#pragma unroll
    for (int u = 0; u < KS; u++) {
      const int e = t + NT * u;
      r.sk[u] = e < tot ? t1[((size_t)(b0 + s_) * n3 + z) * nstick + st_] : C{0, 0};
      st_ += NT;
      while (st_ >= nstick) { st_ -= nstick; s_++; }
    }
    const T* vz = vtot + (size_t)z * N1 * N2;
#pragma unroll
    for (int u = 0; u < VB; u++) {
      const int e = t + NT * u;
      r.vb[u] = e < N1 * N2 ? vz[e] : (T)0;
    }
  }

  // Stick positions (and mirror positions) of the plane into LDS; read after
  // the next barrier.
  __device__ __forceinline__ void stage_pos(const int* __restrict__ pos, const int* __restrict__ mpos) const {
    const int t = threadIdx.x;
    for (int e0 = 0; e0 < nstick; e0 += 4 * NT) {
      int pb[4], mb[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int e = e0 + t + NT * u;
        pb[u] = e < nstick ? pos[e] : 0;
        mb[u] = e < nstick ? mpos[e] : -1;
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int e = e0 + t + NT * u;
        if (e < nstick) { spos[e] = pb[u]; smpos[e] = mb[u]; }
      }
    }
  }

  // V plane into LDS as row-pair words, the y-space zeroed, barrier, the
  // sticks scattered to their (kx, y) slots (and conjugate mirrors of the
  // kx = 0 plane), barrier.  Sticks beyond KS per thread load here.
  __device__ __forceinline__ void stage(int z, int b0, const C* t1, const Pre& r) const {
    const int t = threadIdx.x, tot = min(SB, nb - b0) * nstick;
#pragma unroll
    for (int u = 0; u < VB; u++) {
      const int e = t + NT * u;
      if (e < N1 * N2) {
        const int y = e / N1, x = e - y * N1;
        vs[2 * ((y >> 1) * LV + x) + (y & 1)] = r.vb[u];
      }
    }
    for (int e = t; e < SB * nk * N2; e += NT) Y[e] = C{0, 0};
    __syncthreads();
    {
      int s_ = t / nstick, st_ = t - s_ * nstick;
#pragma unroll
      for (int u = 0; u < KS; u++) {
        const int e = t + NT * u;
        if (e < tot) {
          C* Ys = Y + (size_t)s_ * nk * N2;
          Ys[spos[st_]] = r.sk[u];
          const int mp = smpos[st_];
          if (mp >= 0) Ys[mp] = C{r.sk[u].x, -r.sk[u].y};
        }
        st_ += NT;
        while (st_ >= nstick) { st_ -= nstick; s_++; }
      }
    }
    for (int e0 = KS * NT; e0 < tot; e0 += 8 * NT) {  // more than KS sticks per thread
      C vb8[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int e = e0 + t + NT * u;
        const int s_ = e / nstick, st_ = e - s_ * nstick;
        vb8[u] = e < tot ? t1[((size_t)(b0 + s_) * n3 + z) * nstick + st_] : C{0, 0};
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int e = e0 + t + NT * u;
        if (e < tot) {
          const int s_ = e / nstick, st_ = e - s_ * nstick;
          C* Ys = Y + (size_t)s_ * nk * N2;
          Ys[spos[st_]] = vb8[u];
          const int mp = smpos[st_];
          if (mp >= 0) Ys[mp] = C{vb8[u].x, -vb8[u].y};
        }
      }
    }
    __syncthreads();
  }

  // y IFFT, x c2r / V / r2c per row pair, y FFT; ns bands
  __device__ __forceinline__ void compute(int ns) const {
    const int t = threadIdx.x;
  // inverse y on the kx <= kxmax columns
  for (int c = t / LY; c < ns * nk; c += NT / LY) col_fft_lanes<T, N2, 1, LY>(Y + (size_t)c * N2, t % LY);
  __syncthreads();
  // per row pair: Hermitian-extended spectra packed as X1 + i X2, c2r, x V/N,
  // r2c, unpack kx <= kxmax; a thread reads and writes only its two rows,
  // which sit side by side in LDS (one 128-bit access per kx)
  for (int c = t; c < ns * NPAIR; c += NT) {
    const int s = c / NPAIR, q = c - s * NPAIR;
    const int y1 = 2 * q;
    C* Ys = Y + (size_t)s * nk * N2;
    C a[N1];
#pragma unroll
    for (int k0 = 0; k0 <= H1; k0 += KC) {
#pragma unroll
      for (int kk = k0; kk < k0 + KC; kk++) {
        if (kk > H1) continue;
        struct { T x, y, z, w; } p = {0, 0, 0, 0};
        if (kk <= kxmax) {
          if constexpr (sizeof(T) == 4) {
            const float4 q = *reinterpret_cast<const float4*>(Ys + kk * N2 + y1);
            p = {q.x, q.y, q.z, q.w};
          } else {
            const C q1 = Ys[kk * N2 + y1], q2 = Ys[kk * N2 + y1 + 1];
            p = {q1.x, q1.y, q2.x, q2.y};
          }
        }
        if (kk == 0 || kk == H1) { p.y = 0; p.w = 0; }
        // row y1 spectrum (p.x, p.y), row y2 (p.z, p.w); the kx > H1 half is
        // their conjugates
        a[kk] = C{p.x - p.w, p.y + p.z};
        if (kk != 0 && kk != H1) a[N1 - kk] = C{p.x + p.w, p.z - p.y};
      }
    }
    pin_regs(a);
    reg_fft<T, N1, 1>(a);
    pin_regs(a);
    const C* vq = vp + q * LV;
#pragma unroll
    for (int x = 0; x < N1; x++) {
      const C v = vq[x];
      a[x] = C{a[x].x * v.x, a[x].y * v.y};
    }
    pin_regs(a);
    reg_fft<T, N1, -1>(a);
    pin_regs(a);
#pragma unroll
    for (int k = 0; k <= H1; k++) {
      if (k <= kxmax) {
        const C w = a[k];
        const C wc = a[k == 0 ? 0 : N1 - k];
        const C o1{(T)0.5 * (w.x + wc.x), (T)0.5 * (w.y - wc.y)}, o2{(T)0.5 * (w.y + wc.y), (T)0.5 * (wc.x - w.x)};
        if constexpr (sizeof(T) == 4) {
          *reinterpret_cast<float4*>(Ys + k * N2 + y1) = make_float4(o1.x, o1.y, o2.x, o2.y);
        } else {
          Ys[k * N2 + y1] = o1;
          Ys[k * N2 + y1 + 1] = o2;
        }
      }
    }
  }
  __syncthreads();
  // forward y
  for (int c = t / LY; c < ns * nk; c += NT / LY) col_fft_lanes<T, N2, -1, LY>(Y + (size_t)c * N2, t % LY);
  __syncthreads();
  }

  // the sticks back to global memory, 8 LDS reads per thread in flight, then
  // the 8 stores; barrier (the LDS is free again)
  __device__ __forceinline__ void gather(int z, int b0, C* t2) const {
    const int t = threadIdx.x, tot = min(SB, nb - b0) * nstick;
    int s_ = t / nstick, st_ = t - s_ * nstick;
    for (int e0 = t; e0 < tot; e0 += 8 * NT) {
      C g[8];
      long idx[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int e = e0 + NT * u;
        idx[u] = e < tot ? ((long)(b0 + s_) * n3 + z) * nstick + st_ : -1;
        g[u] = e < tot ? Y[(size_t)s_ * nk * N2 + spos[st_]] : C{0, 0};
        st_ += NT;
        while (st_ >= nstick) { st_ -= nstick; s_++; }
      }
#pragma unroll
      for (int u = 0; u < 8; u++)
        if (idx[u] >= 0) t2[idx[u]] = g[u];
    }
    __syncthreads();
  }
};

// One z plane (z) of SB bands from b0 (the caller may reuse the LDS at once).
template <typename T, int N1, int N2, int SB, int LY = 1, int NT = 256>
__device__ __forceinline__ void slab_xy_reg_body(int z, int b0, const typename cx<T>::t* t1, typename cx<T>::t* t2,
                                                 int n3, int nstick, int kxmax, int nb, const int* __restrict__ pos,
                                                 const int* __restrict__ mpos, const T* __restrict__ vtot,
                                                 unsigned char* smem_raw) {
  typedef SlabReg<T, N1, N2, SB, LY, NT> S;
  const S sl(smem_raw, n3, nstick, kxmax, nb);
  typename S::Pre r;
  sl.load(z, b0, t1, vtot, r);
  sl.stage_pos(pos, mpos);
  sl.stage(z, b0, t1, r);
  sl.compute(min(SB, nb - b0));
  sl.gather(z, b0, t2);
}

// Items i = blockIdx.x, blockIdx.x + gridDim.x, ... < n3 ceil(nb / SB),
// item i = (plane i % n3, bands SB (i / n3) ..): while the workgroup
// transforms item i, the sticks and V plane of its next item are already in
// flight into registers, and the stores of item i drain during the next
// item's phases.
template <typename T, int N1, int N2, int SB, int LY = 1, int NT = 256>
__device__ __forceinline__ void slab_xy_reg_persistent(const typename cx<T>::t* t1, typename cx<T>::t* t2, int n3,
                                                       int nstick, int kxmax, int nb, const int* __restrict__ pos,
                                                       const int* __restrict__ mpos, const T* __restrict__ vtot,
                                                       unsigned char* smem_raw) {
  typedef SlabReg<T, N1, N2, SB, LY, NT> S;
  const S sl(smem_raw, n3, nstick, kxmax, nb);
  const int items = n3 * ((nb + SB - 1) / SB);
  int i = blockIdx.x;
  if (i >= items) return;  // uniform over the workgroup
  typename S::Pre r;
  sl.load(i % n3, (i / n3) * SB, t1, vtot, r);
  sl.stage_pos(pos, mpos);
  for (; i < items; i += gridDim.x) {
    const int z = i % n3, b0 = (i / n3) * SB;
    sl.stage(z, b0, t1, r);
    const int in = i + gridDim.x;
    if (in < items) sl.load(in % n3, (in / n3) * SB, t1, vtot, r);
    sl.compute(min(SB, nb - b0));
    sl.gather(z, b0, t2);
  }
}

}  // namespace pw
