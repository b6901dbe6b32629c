// Dense symmetric eigensolver of the Rayleigh-Ritz and SolveGenEig problems
// (n up to 1024) on gfx950.  Replaces the lapack::Syevd('V','U') calls of the
// reference (eigensolver.cpp:224-297, lobpcg.cpp:536-548): eigenvalues
// ascending, eigenvectors in place of A.
//
// At these sizes every library path is latency-bound (rocSOLVER's dsyevd
// launches several kernels per Householder step).  Here the sequential part
// runs inside one workgroup and everything else is parallel:
//
//   k_sym_from_upper  mirror the upper triangle (Syevd 'U' reads only it)
//   k_sytrd           Householder tridiagonalisation A = Q T Q^T, blocked
//                     as LAPACK dsytrd/dlatrd 'L': per column a reflector
//                     and a read-only symmetric product with the L2-resident
//                     trailing matrix, per panel one rank-2 nbk update; one
//                     1024-thread workgroup
//   k_stebz           all eigenvalues of T: 512-way multisection on Sturm
//                     counts, one 256-thread workgroup per eigenvalue, two
//                     interleaved counts per lane (LAPACK dstebz tolerances)
//   k_stein_lds       eigenvectors of T: inverse iteration on LU with
//   (k_stein)         partial pivoting of T - lambda I, one workgroup (lane
//                     0, scratch in LDS) per eigenvalue up to n = 768, one
//                     thread per eigenvalue beyond; modified Gram-Schmidt within tight clusters
//                     (gaps <= 1e-5 ||T||; LAPACK dstein uses 1e-3, which
//                     chains the dense Rayleigh-Ritz spectra into long serial
//                     runs); scratch interleaved by thread so that lanes
//                     stepping through the same row coalesce
//   k_larft           the triangular factor T of every compact-WY chunk of
//                     reflectors (dlarft), one workgroup per chunk
//   k_ormtr           Z <- Q Z_T chunk by chunk (dlarfb): every workgroup
//                     applies all chunks to its own block of kOrmCols (2)
//                     columns held in LDS
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>

#include "pw_internal.h"
#include "pw_mfma.h"

namespace pw {

constexpr int kTrdThreads = 1024;  // k_sytrd workgroup
constexpr int kTrdWaves = kTrdThreads / 64;
constexpr int kOrmCols = 2;         // columns per k_ormtr workgroup (n / 2 workgroups)
constexpr int kEigMaxN = 1024;      // LDS bound of k_sytrd
constexpr int kSteinLdsMaxN = 768;  // 10 n doubles of k_stein_lds within 64 KiB

// Block-wide sum, identical order on every call (deterministic), returned to
// all threads.  sred holds one double per wavefront.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* sred) {
  v = wave_sum(v);
  lds_barrier();
  if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = v;
  lds_barrier();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < NT / 64; w++) s += sred[w];
  return s;
}

// Per-matrix strides of a batch of independent problems (all 0 for one
// problem): A the matrix, v the d, e and tau vectors, w the eigenvalues, T
// the compact-WY factors, Z the eigenvector blocks.  The batch index is
// blockIdx.y, or blockIdx.x in the single-workgroup k_sytrd_reg.
struct EigBatch {
  long A = 0, v = 0, w = 0, T = 0, Z = 0;
};

__global__ void k_sym_from_upper(double* __restrict__ A, int n, int lda, long sA) {
  A += blockIdx.y * sA;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)n * n; e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e % n), j = (int)(e / n);
    if (i > j) A[i + (size_t)lda * j] = A[j + (size_t)lda * i];
  }
}

// A (full symmetric storage) -> T = Q^T A Q: d (n), e (n-1), tau (n-1); the
// reflector v_k (v_k[0] = 1) overwrites A[k+1:n, k].  H_k = I - tau_k v v^T.
//
// Blocked as LAPACK dsytrd/dlatrd ('L'), panels of nbk columns: inside a
// panel the trailing matrix is only read.  Column k is first brought up to
// date with the panel's earlier reflectors, then p = A22 v - V (W^T v) -
// W (V^T v) from the stale A22 and the panel's V, W (kept in LDS), and the
// rank-2 nbk update A22 -= V W^T + W V^T runs once per panel.  The symmetric
// product streams A22 from L2 in batches of 16 independent loads per lane;
// its waves split rows in 64-row chunks and columns in G groups with
// G * chunks <= 16, so the partial sums fit 1024 doubles.
__global__ void __launch_bounds__(kTrdThreads)
k_sytrd(double* __restrict__ A, int n, int lda, int nbk, double* __restrict__ d, double* __restrict__ e,
        double* __restrict__ tau) {
  extern __shared__ double sm[];
  double* Vp = sm;                       // nbk columns of n (global row index)
  double* Wp = Vp + (size_t)nbk * n;
  double* sv = Wp + (size_t)nbk * n;    // v, n (local row index i = r - k - 1)
  double* red = sv + n;                  // 1024 partial sums
  double* al = red + 1024;               // W^T v, nbk
  double* be = al + nbk;                 // V^T v, nbk
  __shared__ double sred[kTrdWaves];
  __shared__ double sbc[2];              // alpha, diagonal of column k
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  for (int p0 = 0; p0 < n - 2; p0 += nbk) {
    const int nbp = min(nbk, n - 2 - p0);
    // column p0 (rows p0.. n-1) for the first step; later columns are
    // prefetched one step ahead (stale within the panel, so the early copy is
    // the value the step needs).  n <= 1024 rows: one per thread.
    double xpre = (p0 + t < n) ? A[(size_t)lda * p0 + p0 + t] : 0.0;
    for (int jj = 0; jj < nbp; jj++) {
      const int k = p0 + jj, m = n - k - 1;
      double* colk = A + (size_t)lda * k;
      // column k (rows k..n-1) brought up to date, and ||x[1:]||^2 of the
      // reflector input x = A[k+1:, k].  Thread t owns row k + t in every
      // phase of the step, so its global writes are read back only by itself.
      double s = 0.0;
      {
        const int r = k + t;
        double x = xpre;
        if (r < n) {
          if (jj > 0) {
            double u = 0.0;
#pragma unroll 8
            for (int q = 0; q < jj; q++) u += Vp[q * n + r] * Wp[q * n + k] + Wp[q * n + r] * Vp[q * n + k];
            x -= u;
          }
          if (r >= k + 2) s += x * x;
          if (r == k) sbc[1] = x;
          if (r == k + 1) sbc[0] = x;
        }
        xpre = x;  // the updated entry, reused below by the same thread
      }
      s = block_sum<kTrdThreads>(s, sred);
      const double alpha = sbc[0];
      double tk = 0.0, beta = alpha, scal = 0.0;
      if (s > 0.0) {
        beta = -copysign(sqrt(alpha * alpha + s), alpha);
        tk = (beta - alpha) / beta;
        scal = 1.0 / (alpha - beta);
      }
      if (t == 0) {
        d[k] = sbc[1];
        e[k] = beta;
        tau[k] = tk;
      }
      {
        const int r = k + t;
        if (r > k && r < n) {
          const double vi = (r == k + 1) ? 1.0 : xpre * scal;
          sv[r - k - 1] = vi;
          colk[r] = vi;  // reflector storage, read only by the back-transform
          Vp[jj * n + r] = vi;
        }
      }
      lds_barrier();
      if (tk == 0.0) {  // block-uniform: H_k = I, w = 0
        for (int r = k + 1 + t; r < n; r += kTrdThreads) Wp[jj * n + r] = 0.0;
        if (jj + 1 < nbp) xpre = (k + 1 + t < n) ? A[(size_t)lda * (k + 1) + k + 1 + t] : 0.0;
        lds_barrier();
        continue;
      }
      // partial sums of A22 v (stale trailing block) and the panel dots
      const int nrc = (m + 63) >> 6;
      const int G = max(1, kTrdWaves / nrc);
      const double* A22 = A + (size_t)lda * (k + 1) + (k + 1);
      if (wv < nrc * G) {
        const int rc = wv / G, g = wv - rc * G;
        const int i = rc * 64 + lane;
        double acc = 0.0;
        if (i < m) {
          for (int j = g; j < m; j += 16 * G) {
            double a[16];
#pragma unroll
            for (int u = 0; u < 16; u++) {
              const int jc = j + u * G;
              a[u] = jc < m ? A22[(size_t)lda * jc + i] : 0.0;
            }
#pragma unroll
            for (int u = 0; u < 16; u++) {
              const int jc = j + u * G;
              if (jc < m) acc += a[u] * sv[jc];
            }
          }
        }
        red[g * (nrc * 64) + i] = acc;
      }
      // next column, issued after this step's matrix loads have been consumed
      if (jj + 1 < nbp) xpre = (k + 1 + t < n) ? A[(size_t)lda * (k + 1) + k + 1 + t] : 0.0;
      for (int q = wv; q < jj; q += kTrdWaves) {
        double sa = 0.0, sb = 0.0;
        for (int i = lane; i < m; i += 64) {
          const double vi = sv[i];
          sa += Wp[q * n + k + 1 + i] * vi;
          sb += Vp[q * n + k + 1 + i] * vi;
        }
        sa = wave_sum(sa);
        sb = wave_sum(sb);
        if (lane == 0) {
          al[q] = sa;
          be[q] = sb;
        }
      }
      lds_barrier();
      // p = tau (A22 v - V al - W be), w = p - (tau/2)(p.v) v
      double pv = 0.0;
      double pl[(1024 + kTrdThreads - 1) / kTrdThreads];
      for (int i = t, c = 0; i < m; i += kTrdThreads, c++) {
        double p = 0.0;
        for (int g2 = 0; g2 < G; g2++) p += red[g2 * (nrc * 64) + i];
        const int r = k + 1 + i;
#pragma unroll 8
        for (int q = 0; q < jj; q++) p -= Vp[q * n + r] * al[q] + Wp[q * n + r] * be[q];
        p *= tk;
        pl[c] = p;
        pv += p * sv[i];
      }
      pv = block_sum<kTrdThreads>(pv, sred);
      const double h = 0.5 * tk * pv;
      for (int i = t, c = 0; i < m; i += kTrdThreads, c++) Wp[jj * n + k + 1 + i] = pl[c] - h * sv[i];
      lds_barrier();
    }
    // the trailing update below reads V, W from LDS and rewrites A in global
    // memory; the full barrier after it publishes A to every wavefront
    // trailing update of rows and columns >= p1 in 4 x 4 register tiles
    const int p1 = p0 + nbp, mt = n - p1, nt = (mt + 3) >> 2;
    for (int tile = t; tile < nt * nt; tile += kTrdThreads) {
      const int r0 = p1 + 4 * (tile % nt), c0 = p1 + 4 * (tile / nt);
      double acc[4][4];
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int b = 0; b < 4; b++) acc[a][b] = 0.0;
      for (int q = 0; q < nbp; q++) {
        double vr[4], wr[4], vc[4], wc[4];
#pragma unroll
        for (int a = 0; a < 4; a++) {
          const bool okr = r0 + a < n, okc = c0 + a < n;
          vr[a] = okr ? Vp[q * n + r0 + a] : 0.0;
          wr[a] = okr ? Wp[q * n + r0 + a] : 0.0;
          vc[a] = okc ? Vp[q * n + c0 + a] : 0.0;
          wc[a] = okc ? Wp[q * n + c0 + a] : 0.0;
        }
#pragma unroll
        for (int a = 0; a < 4; a++)
#pragma unroll
          for (int b = 0; b < 4; b++) acc[a][b] += vr[a] * wc[b] + wr[a] * vc[b];
      }
#pragma unroll
      for (int b = 0; b < 4; b++)
#pragma unroll
        for (int a = 0; a < 4; a++)
          if (r0 + a < n && c0 + b < n) A[(size_t)lda * (c0 + b) + r0 + a] -= acc[a][b];
    }
    __syncthreads();
  }
  if (t == 0) {
    d[n - 2] = A[(size_t)lda * (n - 2) + (n - 2)];
    d[n - 1] = A[(size_t)lda * (n - 1) + (n - 1)];
    e[n - 2] = A[(size_t)lda * (n - 2) + (n - 1)];
    tau[n - 2] = 0.0;
  }
}

// ------------------------------------------------------------------------
// n <= 256: the same reduction unblocked (LAPACK dsytd2 'L' order) with the
// matrix on chip.  The strictly lower 32 x 32 blocks (at most 28) sit in the
// MFMA accumulator layout of eight wavefronts (block b, numbered row by row,
// belongs to wave b % 8); diagonal block I lives in LDS (full, kept
// symmetric) and is handled by wave I.  Per column k, three barriers:
//   (a) column k goes to LDS from its owners (two buffers by k's parity,
//       so that the next column's write cannot meet a slow wave's read),
//   (b) every wave forms the reflector's scalars (norm, beta, tau) itself,
//       with the same lanes and order, so all hold identical values, and
//       its own LDS copy of the reflector v (zero outside k+1 .. n-1),
//   (c) every wave sums A v over its blocks that meet the trailing matrix:
//       row sums across the 16 lanes of a DPP row, column sums (the
//       transposed half of an off-diagonal block) across lanes l, l+16,
//       l+32, l+48 by v_permlane16_swap / v_permlane32_swap; each block's
//       row and column sums go to their own 32-entry rows of a table (no
//       read-modify-write), with this wave's share of v^T A v,
//   (d) wave I forms w = p - (tau/2)(p.v) v, p = tau A v, on rows of
//       block I, adding the table's terms of that row block in a fixed
//       order,
//   (e) A -= v w^T + w v^T: two VALU FMAs per register element,
//       read-modify-write on the LDS diagonal blocks.
// No global memory traffic after the load except the reflector and d, e,
// tau.  Same arithmetic as k_sytrd up to the summation order.
constexpr int kTr2MaxNB = 8;
constexpr int kTr2Waves = 8;  // = kTr2MaxNB: wave I owns diagonal block I
constexpr int kTr2Slots = (kTr2MaxNB * (kTr2MaxNB - 1) / 2 + kTr2Waves - 1) / kTr2Waves;  // 4
constexpr int kTr2Ld = 33;    // row stride of an LDS diagonal block
constexpr int kTr2NBlk = kTr2MaxNB * (kTr2MaxNB - 1) / 2;  // strictly lower blocks (28)
// partial-sum table rows of 32: diagonal block I, then per strictly lower
// block b its row sums (rows of I) and its column sums (rows of J)
constexpr int kTr2TabRows = kTr2MaxNB + 2 * kTr2NBlk;

// strictly lower block b (I > J), numbered row by row: b = I (I - 1) / 2 + J
__device__ __forceinline__ void tr_block(int b, int& I, int& J) {
  int i = 1;
  while ((i + 1) * i / 2 <= b) i++;
  I = i;
  J = b - i * (i - 1) / 2;
}

// Sums over the 16 lanes of each DPP row of eight values r[0..7] per lane,
// by a butterfly that halves the values at every stage: lanes l, l ^ 8
// (row_ror 8), then l, 7 - l within each half row (row_half_mirror), then
// l ^ 2 and l ^ 1 (quad_perm).  Lanes 2 j and 2 j + 1 of the row end with
// the total of r[j], j = (l % 16) / 2.
__device__ __forceinline__ double row16_sum8(const double (&r)[8], int lr) {
  const bool b3 = lr & 8, b2 = lr & 4, b1 = lr & 2;
  double s1[4];
#pragma unroll
  for (int j = 0; j < 4; j++) {
    const double keep = b3 ? r[4 + j] : r[j], send = b3 ? r[j] : r[4 + j];
    s1[j] = keep + dpp_f64<0x128, 0xf>(send);  // row_ror:8
  }
  double s2[2];
#pragma unroll
  for (int j = 0; j < 2; j++) {
    const double keep = b2 ? s1[2 + j] : s1[j], send = b2 ? s1[j] : s1[2 + j];
    s2[j] = keep + dpp_f64<0x141, 0xf>(send);  // row_half_mirror
  }
  const double keep = b1 ? s2[1] : s2[0], send = b1 ? s2[0] : s2[1];
  double s3 = keep + dpp_f64<0x4E, 0xf>(send);  // quad_perm [2,3,0,1]
  return s3 + dpp_f64<0xB1, 0xf>(s3);            // quad_perm [1,0,3,2]
}

// Sum over the four 16-lane rows of a wavefront, in every lane:
// v_permlane16_swap pairs rows (0, 1) and (2, 3), v_permlane32_swap the two
// halves (VALU lane swaps; lane layouts checked on gfx950).
__device__ __forceinline__ double half_sum(double v) {
  const unsigned lo = (unsigned)__double2loint(v), hi = (unsigned)__double2hiint(v);
  auto l32 = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
  auto h32 = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
  return __hiloint2double((int)h32[0], (int)l32[0]) + __hiloint2double((int)h32[1], (int)l32[1]);
}
__device__ __forceinline__ double row4_sum(double v) {
  const unsigned lo = (unsigned)__double2loint(v), hi = (unsigned)__double2hiint(v);
  auto l16 = __builtin_amdgcn_permlane16_swap(lo, lo, false, false);
  auto h16 = __builtin_amdgcn_permlane16_swap(hi, hi, false, false);
  return half_sum(__hiloint2double((int)h16[0], (int)l16[0]) + __hiloint2double((int)h16[1], (int)l16[1]));
}

// Column c, rows r0 .. n-1, of the on-chip matrix into col[] (slot s of
// this wave holds block (bI[s], bJ[s]), bJ[s] < 0 for none).
__device__ __forceinline__ void tr_column(const d4v (&a)[kTr2Slots][4], const int (&bI)[kTr2Slots],
                                          const int (&bJ)[kTr2Slots], const double* Dg, int w, int lane, int n,
                                          int c, int r0, double* col) {
  const int lr = lane & 15, lk = lane >> 4;
  const int K = c >> 5, kt = (c >> 4) & 1, kl = c & 15;
  if (w == K && lane < 32) {
    const int row = 32 * K + lane;
    if (row >= r0 && row < n) col[row] = Dg[(size_t)K * 32 * kTr2Ld + lane * kTr2Ld + (c & 31)];
  }
#pragma unroll
  for (int s = 0; s < kTr2Slots; s++) {
    const int I = bI[s];
    if (bJ[s] != K || 32 * I + 31 < r0 || lr != kl) continue;
#pragma unroll
    for (int ti = 0; ti < 2; ti++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int row = 32 * I + 16 * ti + 4 * i + lk;
        const double val = kt ? a[s][ti + 2][i] : a[s][ti][i];
        if (row >= r0 && row < n) col[row] = val;
      }
  }
}

__global__ void __launch_bounds__(512)
k_sytrd_reg(double* __restrict__ A, int n, int lda, double* __restrict__ d, double* __restrict__ e,
            double* __restrict__ tau, EigBatch bs) {
  extern __shared__ double sm[];
  A += blockIdx.x * bs.A;
  d += blockIdx.x * bs.v;
  e += blockIdx.x * bs.v;
  tau += blockIdx.x * bs.v;
  double* xb = sm;                             // 2 x 256: column k (buffer k % 2)
  double* sw = xb + 512;                       // 256: w
  double* tab = sw + 256;                      // kTr2TabRows x 32: partial sums of A v
  double* dots = tab + kTr2TabRows * 32;       // kTr2Waves: partial sums of v^T A v
  double* Dg = dots + kTr2Waves;               // kTr2MaxNB x 32 x kTr2Ld: diagonal blocks
  double* stg = Dg + kTr2MaxNB * 32 * kTr2Ld;  // load phase: kTr2Waves x 32 x 33 staging
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int NB = (n + 31) / 32, nblk = NB * (NB - 1) / 2;
  double* Dw = Dg + (size_t)w * 32 * kTr2Ld;   // this wave's diagonal block
  // diagonal block w (zero beyond n): lane = row, two columns per load
  if (w < NB) {
    double v16[16];
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int row = 32 * w + (lane & 31), col = 32 * w + 2 * u + (lane >> 5);
      v16[u] = (row < n && col < n) ? A[row + (size_t)lda * col] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 16; u++) Dw[(lane & 31) * kTr2Ld + 2 * u + (lane >> 5)] = v16[u];
  }
  // off-diagonal blocks: the same loads, then through this wave's staging
  // area into the accumulator layout; block coordinates per slot (wave
  // uniform, scalar registers)
  int bI[kTr2Slots], bJ[kTr2Slots], bB[kTr2Slots];
#pragma unroll
  for (int s = 0; s < kTr2Slots; s++) {
    const int b = __builtin_amdgcn_readfirstlane(w) + kTr2Waves * s;
    int I = 1, J = -1;
    if (b < nblk) tr_block(b, I, J);
    bI[s] = I;
    bJ[s] = J;
    bB[s] = b;
  }
  d4v a[kTr2Slots][4];
  double* st = stg + w * 32 * 33;
#pragma unroll
  for (int s = 0; s < kTr2Slots; s++) {
    const int I = bI[s], J = bJ[s];
    double v16[16];
#pragma unroll
    for (int u = 0; u < 16; u++) {
      const int row = 32 * I + (lane & 31), col = 32 * J + 2 * u + (lane >> 5);
      v16[u] = (J >= 0 && row < n && col < n) ? A[row + (size_t)lda * col] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 16; u++) st[(lane & 31) * 33 + 2 * u + (lane >> 5)] = v16[u];
    wave_lds_sync();
    blk_from_lds(a[s], st, 33, lane);
    wave_lds_sync();
  }
  __syncthreads();
  double* sv = stg + w * 32 * 33;  // the loop's per-wave copy of v
  for (int k = 0; k < n - 2; k++) {
    double* x = xb + 256 * (k & 1);  // the other buffer may still be read by (e) of column k - 1
    // (a) column k, rows >= k
    tr_column(a, bI, bJ, Dg, w, lane, n, k, k, x);
    __syncthreads();
    // (b) the reflector's scalars in every wave (same lanes, same summation
    // order: identical values); v(r) = 1 at r = k + 1, x[r] / (alpha - beta)
    // below, 0 above
    // one batch of LDS reads: lane l holds rows l + 64 u
    double xr[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int r = lane + 64 * u;
      xr[u] = r < n ? x[r] : 0.0;
    }
    const double alpha = x[k + 1], xkk = x[k];
    double ss = 0.0;
#pragma unroll
    for (int u = 0; u < 4; u++) ss += lane + 64 * u > k + 1 ? xr[u] * xr[u] : 0.0;
    ss = wave_sum(ss);
    double tk = 0.0, beta = alpha, scal = 0.0;
    if (ss > 0.0) {
      beta = -copysign(sqrt(alpha * alpha + ss), alpha);
      tk = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    // this wave's copy of v (in its load staging area)
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const int r = lane + 64 * u;
      const double v = r == k + 1 ? 1.0 : ((r > k + 1 && r < n) ? xr[u] * scal : 0.0);
      sv[r] = v;
      if (w == 0 && r > k && r < n) A[r + (size_t)lda * k] = v;  // reflector storage for the back-transform
    }
    if (w == 0 && lane == 0) {
      d[k] = xkk;
      e[k] = beta;
      tau[k] = tk;
    }
    if (tk == 0.0) continue;  // uniform: x is already (alpha, 0, ..., 0) below row k
    wave_lds_sync();
    auto vat = [&](int r) -> double { return sv[r]; };
    // (c) A v over the blocks meeting the trailing matrix, each block's row
    // and column sums stored once in the table (no read-modify-write), and
    // this wave's share of v^T A v
    double dp = 0.0;
    if (w < NB && 32 * w + 31 > k) {
      // diagonal block: lane (row, half) sums 16 columns
      const int r = lane & 31, h = lane >> 5;
      double acc = 0.0;
#pragma unroll
      for (int j = 0; j < 16; j++) acc += Dw[r * kTr2Ld + 16 * h + j] * vat(32 * w + 16 * h + j);
      acc = half_sum(acc);
      if (h == 0) {
        tab[w * 32 + r] = acc;
        dp += acc * vat(32 * w + r);
      }
    }
#pragma unroll
    for (int s = 0; s < kTr2Slots; s++) {
      const int I = bI[s], J = bJ[s], b = bB[s];
      if (J < 0 || 32 * J + 31 <= k) continue;  // no block, or v is zero on all its columns
      const double vc0 = vat(32 * J + lr), vc1 = vat(32 * J + 16 + lr);
      double rp[8];
#pragma unroll
      for (int ti = 0; ti < 2; ti++)
#pragma unroll
        for (int i = 0; i < 4; i++) rp[4 * ti + i] = a[s][ti][i] * vc0 + a[s][ti + 2][i] * vc1;
      const double rs = row16_sum8(rp, lr);
      if ((lr & 1) == 0) {
        const int j = lr >> 1;  // = 4 ti + i
        const int rl = 16 * (j >> 2) + 4 * (j & 3) + lk;
        tab[(kTr2MaxNB + b) * 32 + rl] = rs;
        dp += rs * vat(32 * I + rl);
      }
      double cp0 = 0.0, cp1 = 0.0;
#pragma unroll
      for (int ti = 0; ti < 2; ti++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const double vr = vat(32 * I + 16 * ti + 4 * i + lk);
          cp0 += a[s][ti][i] * vr;
          cp1 += a[s][ti + 2][i] * vr;
        }
      cp0 = row4_sum(cp0);
      cp1 = row4_sum(cp1);
      if (lk == 0) {
        tab[(kTr2MaxNB + kTr2NBlk + b) * 32 + lr] = cp0;
        tab[(kTr2MaxNB + kTr2NBlk + b) * 32 + 16 + lr] = cp1;
        dp += cp0 * vc0 + cp1 * vc1;
      }
      // one block at a time: keeps the scheduler from overlapping blocks
#pragma unroll
      for (int q = 0; q < 4; q++) asm volatile("" : "+v"(a[s][q]));
    }
    dp = wave_sum(dp);
    if (lane == 0) dots[w] = dp;
    __syncthreads();
    // (d) rows 32 w .. 32 w + 31: p = tau A v (the table's terms of the row
    // block in a fixed order), w = p - (tau/2)(p.v) v with p.v = tau v^T A v
    if (w < NB && 32 * w + 31 > k && lane < 32) {
      double vav = 0.0;
#pragma unroll
      for (int ww = 0; ww < kTr2Waves; ww++) vav += dots[ww];
      const double h = 0.5 * tk * (tk * vav);
      // terms in a fixed order: row sums of blocks (w, J), J = jmin .. w-1,
      // then column sums of blocks (I, w), I = w+1 .. NB-1
      double p = tab[w * 32 + lane];
      for (int J = (k + 1) >> 5; J < w; J++) p += tab[(kTr2MaxNB + w * (w - 1) / 2 + J) * 32 + lane];
      for (int I = w + 1; I < NB; I++) p += tab[(kTr2MaxNB + kTr2NBlk + I * (I - 1) / 2 + w) * 32 + lane];
      const int r = 32 * w + lane;
      sw[r] = (r > k && r < n) ? tk * p - h * vat(r) : 0.0;
    }
    __syncthreads();
    // (e) A -= v w^T + w v^T
    if (w < NB && 32 * w + 31 > k) {
      // reads of each 4 columns issued together, then the 4 writes
      const int r = lane & 31, h = lane >> 5;
      const double vr = vat(32 * w + r), wr = sw[32 * w + r];
#pragma unroll
      for (int j0 = 0; j0 < 16; j0 += 4) {
        double dv[4], wc[4], vc[4];
#pragma unroll
        for (int j = 0; j < 4; j++) {
          const int c = 16 * h + j0 + j;
          dv[j] = Dw[r * kTr2Ld + c];
          wc[j] = sw[32 * w + c];
          vc[j] = vat(32 * w + c);
        }
#pragma unroll
        for (int j = 0; j < 4; j++) Dw[r * kTr2Ld + 16 * h + j0 + j] = dv[j] - (vr * wc[j] + wr * vc[j]);
      }
    }
    // register tiles by VALU FMAs, two per element (an FP64 MFMA issues no
    // faster than the VALU on gfx950, and this product would use K = 2 of
    // its 4): lane holds rows 32 I + 16 ti + 4 i + lk, columns 32 J + 16 tj
    // + lr in a[s][ti + 2 tj][i]
#pragma unroll
    for (int s = 0; s < kTr2Slots; s++) {
      const int I = bI[s], J = bJ[s];
      if (J < 0 || 32 * J + 31 <= k) continue;
      double vc[2], wc[2];
#pragma unroll
      for (int tj = 0; tj < 2; tj++) {
        const int c = 32 * J + 16 * tj + lr;
        vc[tj] = vat(c);
        wc[tj] = sw[c];
      }
#pragma unroll
      for (int ti = 0; ti < 2; ti++) {  // one row tile at a time (register pressure)
        double vr[4], wr[4];
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int r = 32 * I + 16 * ti + 4 * i + lk;
          vr[i] = vat(r);
          wr[i] = sw[r];
        }
#pragma unroll
        for (int tj = 0; tj < 2; tj++)
#pragma unroll
          for (int i = 0; i < 4; i++) {
            const double x = fma(-vr[i], wc[tj], a[s][ti + 2 * tj][i]);
            a[s][ti + 2 * tj][i] = fma(-wr[i], vc[tj], x);
          }
#pragma unroll
        for (int tj = 0; tj < 2; tj++) asm volatile("" : "+v"(a[s][ti + 2 * tj]));
      }
    }
    wave_lds_sync();  // this wave's diagonal-block writes before the next column read
  }
  // the last 2 x 2: columns n - 2 (rows n-2, n-1) and n - 1, into buffers
  // that the last update does not read (x of column n - 2 and the table)
  double* xl = xb + 256 * ((n - 2) & 1);
  tr_column(a, bI, bJ, Dg, w, lane, n, n - 2, n - 2, xl);
  tr_column(a, bI, bJ, Dg, w, lane, n, n - 1, n - 1, tab);
  __syncthreads();
  if (t == 0) {
    d[n - 2] = xl[n - 2];
    d[n - 1] = tab[n - 1];
    e[n - 2] = xl[n - 1];
    tau[n - 2] = 0.0;
  }
}

// 1 / p from v_rcp_f64 and two Newton steps (within an ulp or so of the
// rounded quotient for the normal p that reach it; callers keep |p| away
// from zero).  Five dependent operations against eleven for an IEEE divide.
__device__ __forceinline__ double fast_rcp(double p) {
  double r = __builtin_amdgcn_rcp(p);
  r = fma(r, fma(-p, r, 1.0), r);
  return fma(r, fma(-p, r, 1.0), r);
}

// Numbers of eigenvalues of T below x0 and below x1 (Sturm sequences of the
// LDL^T pivots, dstebz's recurrence with pivmin), the two chains
// interleaved so that each hides the other's latency.
// d and e are read at the same index in every lane (scalar loads through
// the constant cache; e^2 formed here as the LDS copy of k_stebz forms it).
__device__ __forceinline__ void sturm_count2(const double* __restrict__ d, const double* __restrict__ e, int n,
                                             double x0, double x1, double pivmin, int& c0, int& c1) {
  double q0 = d[0] - x0, q1 = d[0] - x1;
  if (fabs(q0) < pivmin) q0 = -pivmin;
  if (fabs(q1) < pivmin) q1 = -pivmin;
  int n0 = q0 < 0.0, n1 = q1 < 0.0;
#pragma unroll 8
  for (int i = 1; i < n; i++) {
    const double di = d[i], ei = e[i - 1] * e[i - 1];
    q0 = (di - x0) - ei * fast_rcp(q0);
    q1 = (di - x1) - ei * fast_rcp(q1);
    if (fabs(q0) < pivmin) q0 = -pivmin;
    if (fabs(q1) < pivmin) q1 = -pivmin;
    n0 += q0 < 0.0;
    n1 += q1 < 0.0;
  }
  c0 = n0;
  c1 = n1;
}

// Gershgorin interval and pivmin of T, shared by k_stebz and k_stein.
__device__ void tri_bounds(const double* d, const double* e, int n, double* sred, double* gl, double* gu,
                           double* onenrm, double* e2max) {
  double lo = DBL_MAX, hi = -DBL_MAX, nrm = 0.0, em = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double r = (i > 0 ? fabs(e[i - 1]) : 0.0) + (i < n - 1 ? fabs(e[i]) : 0.0);
    lo = fmin(lo, d[i] - r);
    hi = fmax(hi, d[i] + r);
    nrm = fmax(nrm, fabs(d[i]) + r);
    if (i < n - 1) em = fmax(em, e[i] * e[i]);
  }
  // block min/max through shuffles and LDS (blockDim = 256)
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    lo = fmin(lo, __shfl_xor(lo, o, 64));
    hi = fmax(hi, __shfl_xor(hi, o, 64));
    nrm = fmax(nrm, __shfl_xor(nrm, o, 64));
    em = fmax(em, __shfl_xor(em, o, 64));
  }
  const int nw = blockDim.x >> 6, w = threadIdx.x >> 6;
  if ((threadIdx.x & 63) == 0) {
    sred[4 * w] = lo;
    sred[4 * w + 1] = hi;
    sred[4 * w + 2] = nrm;
    sred[4 * w + 3] = em;
  }
  __syncthreads();
  lo = sred[0]; hi = sred[1]; nrm = sred[2]; em = sred[3];
  for (int q = 1; q < nw; q++) {
    lo = fmin(lo, sred[4 * q]);
    hi = fmax(hi, sred[4 * q + 1]);
    nrm = fmax(nrm, sred[4 * q + 2]);
    em = fmax(em, sred[4 * q + 3]);
  }
  *gl = lo; *gu = hi; *onenrm = nrm; *e2max = em;
}

// Eigenvalue j = blockIdx.x of T by multisection: the 256 threads count at
// 512 points per round (two per lane, adjacent), which narrows the bracket
// 513-fold, about 6 rounds to dstebz's tolerance.  Every thread forms the
// points from the same bracket with the same expression, so the new bracket
// needs only the index of the first point above lambda_j (a wavefront
// ballot, then the minimum over the four wavefronts through LDS).
constexpr int kBzPts = 512;
__global__ void __launch_bounds__(256)
k_stebz(const double* __restrict__ dg, const double* __restrict__ eg, int n, double* __restrict__ w, EigBatch bs) {
  dg += blockIdx.y * bs.v;
  eg += blockIdx.y * bs.v;
  w += blockIdx.y * bs.w;
  __shared__ double sred[16];
  __shared__ int sfirst[2][4];
  double gl, gu, nrm, em;
  tri_bounds(dg, eg, n, sred, &gl, &gu, &nrm, &em);
  __syncthreads();
  const double eps = DBL_EPSILON;
  const double pivmin = DBL_MIN * fmax(1.0, em);
  const double tnorm = fmax(fabs(gl), fabs(gu));
  const double wid = gu - gl;
  gl -= 2.0 * n * eps * wid + 2.0 * pivmin;  // dstebz's widening
  gu += 2.0 * n * eps * wid + 2.0 * pivmin;
  const double atol = eps * tnorm, rtol = 2.0 * eps;
  const int j = blockIdx.x;
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  double lo = gl, hi = gu;
  auto point = [&](int p) { return lo + (hi - lo) * (double)(p + 1) / (double)(kBzPts + 1); };
  for (int it = 0; it < 16; it++) {
    if (hi - lo <= fmax(fmax(atol, pivmin), rtol * fmax(fabs(lo), fabs(hi)))) break;  // uniform
    int c0, c1;
    sturm_count2(dg, eg, n, point(2 * t), point(2 * t + 1), pivmin, c0, c1);
    // first point p with count >= j + 1 (lambda_j < x_p); kBzPts if none
    const unsigned long long m1 = __ballot(c1 >= j + 1);
    int first = kBzPts;
    if (m1) {
      const int l = __ffsll((long long)m1) - 1;
      const int a0 = __shfl(c0, l, 64);
      first = 2 * (64 * wv + l) + (a0 >= j + 1 ? 0 : 1);
    }
    if (lane == 0) sfirst[it & 1][wv] = first;
    __syncthreads();  // the other parity's slots are free: every thread read them before this barrier
    const int* sf = sfirst[it & 1];
    const int f = min(min(sf[0], sf[1]), min(sf[2], sf[3]));
    const double nhi = f < kBzPts ? point(f) : hi;
    const double nlo = f > 0 ? point(f - 1) : lo;
    lo = nlo;
    hi = nhi;
  }
  if (t == 0) w[j] = 0.5 * (lo + hi);
}

__device__ __forceinline__ double start_value(int j, int i) {
  // deterministic pseudo-random start vector in [-1, 1)
  uint32_t h = (uint32_t)j * 0x9E3779B1u ^ ((uint32_t)i + 0x7F4A7C15u) * 0x85EBCA77u;
  h ^= h >> 15;
  h *= 0x2C1B3C6Du;
  h ^= h >> 12;
  h *= 0x297A2D39u;
  h ^= h >> 15;
  return (double)h * (2.0 / 4294967296.0) - 1.0;
}

// Per-thread rows of the inverse-iteration scratch: array a of thread q at
// S + (a * n) * nthr, row i at [i * nthr + q] (lanes of a wavefront stepping
// through the same row touch consecutive addresses).
struct SteinRows {
  double* __restrict__ r0;   // 1 / pivot
  double* __restrict__ u1;   // first superdiagonal of U
  double* __restrict__ u2;   // second superdiagonal of U (row interchanges)
  double* __restrict__ ml;   // multiplier
  double* __restrict__ pv;   // 1 where rows i, i+1 were interchanged
  double* __restrict__ x;    // iterate
  double* __restrict__ y;    // forward-eliminated right-hand side
};

constexpr int kSteinB = 16;  // rows per batch of loads

// LU with partial pivoting of T - lam I (dlagtf's elimination order), fused
// with the forward elimination of the start vector (start_value(jj, .)):
// the first inverse iteration's first half.  d and e are read kSteinB rows
// ahead into registers, so a batch costs one LDS round trip (the row
// stores of the previous batch may alias them and would otherwise order
// every load behind them); one reciprocal per row (the multiplier is the
// numerator times 1 / pivot).  Branch-free selects.
__device__ void stein_factor(const double* d, const double* e, int n, double lam, double pert, int jj,
                             SteinRows R, int nthr, int q) {
  double c = d[0] - lam, s1 = n > 1 ? e[0] : 0.0, sub = s1;
  double y0 = start_value(jj, 0);
  for (int i0 = 0; i0 < n - 1; i0 += kSteinB) {
    double db[kSteinB], eb[kSteinB];
    // unconditional loads (a guarded load becomes a branch and a wait on
    // every row's stores); whole batches at affine rows (one base register
    // and immediate offsets), the last one at clamped rows (not used past
    // the end)
    if (i0 + kSteinB <= n - 2) {
#pragma unroll
      for (int u = 0; u < kSteinB; u++) {
        db[u] = d[i0 + u + 1];
        eb[u] = e[i0 + u + 1];
      }
    } else {
#pragma unroll
      for (int u = 0; u < kSteinB; u++) {
        const int i = min(i0 + u, n - 2);
        db[u] = d[i + 1];
        eb[u] = e[min(i + 1, n - 2)];
      }
    }
#pragma unroll
    for (int u = 0; u < kSteinB; u++) {
      const int i = i0 + u;
      if (i < n - 1) {
        const double dn = db[u] - lam, sup = i + 2 < n ? eb[u] : 0.0;
        const size_t o = (size_t)i * nthr + q;
        const bool sw = fabs(sub) > fabs(c);
        // without interchange the pivot c is kept away from zero: +-pert
        // when |c| < pert (+pert for c = +-0), else c
        const double cp = sw ? c : (c >= 0.0 ? 1.0 : -1.0) * fmax(fabs(c), pert);
        const double piv = sw ? sub : cp;
        const double rp = fast_rcp(piv);
        const double mlt = (sw ? cp : sub) * rp;
        R.r0[o] = rp;
        R.u1[o] = sw ? dn : s1;
        R.u2[o] = sw ? sup : 0.0;
        R.ml[o] = mlt;
        R.pv[o] = sw ? 1.0 : 0.0;
        const double cn = sw ? s1 - mlt * dn : dn - mlt * s1;
        s1 = sw ? -mlt * sup : sup;
        c = cn;
        sub = sup;
        const double y1 = start_value(jj, i + 1);
        R.y[o] = sw ? y1 : y0;
        y0 = sw ? y0 - mlt * y1 : y1 - mlt * y0;
      }
    }
  }
  if (fabs(c) < pert) c = c >= 0.0 ? pert : -pert;
  const size_t o = (size_t)(n - 1) * nthr + q;
  R.r0[o] = 1.0 / c; R.u1[o] = 0.0; R.u2[o] = 0.0;
  R.y[o] = y0;
}

// Forward elimination x * scale -> y with the factor's multipliers and
// interchanges.  Rows go in batches of kSteinB: all loads of a batch are
// issued before any of its stores, so a batch costs one memory round trip
// (the compiler cannot reorder across the previous batch's stores by
// itself).
__device__ void stein_forward(int n, double scale, SteinRows R, int nthr, int q) {
  double y0 = R.x[q] * scale;
  for (int i0 = 0; i0 < n - 1; i0 += kSteinB) {
    double xb[kSteinB], mb[kSteinB], pb[kSteinB];
    if (i0 + kSteinB - 1 <= n - 2) {  // whole batch: affine rows
#pragma unroll
      for (int u = 0; u < kSteinB; u++) {
        const size_t o = (size_t)(i0 + u) * nthr + q;
        xb[u] = R.x[o + nthr];
        mb[u] = R.ml[o];
        pb[u] = R.pv[o];
      }
    } else {
#pragma unroll
      for (int u = 0; u < kSteinB; u++) {
        const size_t o = (size_t)min(i0 + u, n - 2) * nthr + q;  // clamped: unconditional loads
        xb[u] = R.x[o + nthr];
        mb[u] = R.ml[o];
        pb[u] = R.pv[o];
      }
    }
#pragma unroll
    for (int u = 0; u < kSteinB; u++) {
      const int i = i0 + u;
      if (i < n - 1) {
        const double y1 = xb[u] * scale, mlt = mb[u];
        const bool sw = pb[u] != 0.0;
        R.y[(size_t)i * nthr + q] = sw ? y1 : y0;
        y0 = sw ? y0 - mlt * y1 : y1 - mlt * y0;
      }
    }
  }
  R.y[(size_t)(n - 1) * nthr + q] = y0;
}

// Back substitution U x = y; returns ||x||^2.
__device__ double stein_back(int n, SteinRows R, int nthr, int q) {
  double xn1 = 0.0, xn2 = 0.0, ss = 0.0;
  for (int i0 = n - 1; i0 >= 0; i0 -= kSteinB) {
    double yb[kSteinB], ab[kSteinB], bb[kSteinB], rb[kSteinB];
    if (i0 - (kSteinB - 1) >= 0) {  // whole batch: affine rows
#pragma unroll
      for (int u = 0; u < kSteinB; u++) {
        const size_t o = (size_t)(i0 - u) * nthr + q;
        yb[u] = R.y[o];
        ab[u] = R.u1[o];
        bb[u] = R.u2[o];
        rb[u] = R.r0[o];
      }
    } else {
#pragma unroll
      for (int u = 0; u < kSteinB; u++) {
        const size_t o = (size_t)max(i0 - u, 0) * nthr + q;  // clamped: unconditional loads
        yb[u] = R.y[o];
        ab[u] = R.u1[o];
        bb[u] = R.u2[o];
        rb[u] = R.r0[o];
      }
    }
#pragma unroll
    for (int u = 0; u < kSteinB; u++) {
      const int i = i0 - u;
      if (i >= 0) {
        const double xi = (yb[u] - ab[u] * xn1 - bb[u] * xn2) * rb[u];
        R.x[(size_t)i * nthr + q] = xi;
        ss += xi * xi;
        xn2 = xn1;
        xn1 = xi;
      }
    }
  }
  return ss;
}

// Eigenvector(s) of T for eigenvalue j into the columns of Zt (ld n): two
// inverse iterations from a pseudo-random start.  Tight clusters
// (consecutive gaps <= 1e-5 ||T||, which includes exact degeneracies) are
// taken by the caller of their first member, sequentially with modified
// Gram-Schmidt against the earlier members on every iteration (LAPACK
// dstein's scheme).  Outside them the vectors are orthogonal to ~eps ||T|| /
// gap <= 1e-10.  R holds the caller's scratch rows (stride nthr, column q).
__device__ void stein_cluster(const double* d, const double* e, const double* __restrict__ w, int n, int j,
                              double nrm, double* __restrict__ Zt, SteinRows R, int nthr, int q) {
  const double eps = DBL_EPSILON;
  const double ctol = 1e-5 * nrm;
  const double pert = eps * fmax(nrm, DBL_MIN);
  if (j > 0 && w[j] - w[j - 1] <= ctol) return;  // member of an earlier caller's cluster
  int j1 = j;
  while (j1 + 1 < n && w[j1 + 1] - w[j1] <= ctol) j1++;
  double lamprev = 0.0;
  for (int jj = j; jj <= j1; jj++) {
    double lam = w[jj];
    if (jj > j && lam - lamprev < 10.0 * eps * fabs(lam)) lam = lamprev + 10.0 * eps * fabs(lam);
    lamprev = lam;
    // the first iteration's forward elimination rides on the factorisation
    stein_factor(d, e, n, lam, pert, jj, R, nthr, q);
    const int iters = (j1 > j) ? 4 : 2;
    double scale = 1.0;
    for (int it = 0; it < iters; it++) {
      if (it > 0) stein_forward(n, scale, R, nthr, q);
      double ss = stein_back(n, R, nthr, q);
      if (jj > j) {
        // orthogonalise against the cluster's earlier vectors (unit norm)
        for (int kk = j; kk < jj; kk++) {
          const double* zk = Zt + (size_t)n * kk;
          double dot = 0.0;
          for (int i = 0; i < n; i++) dot += zk[i] * R.x[(size_t)i * nthr + q];
          for (int i = 0; i < n; i++) R.x[(size_t)i * nthr + q] -= dot * zk[i];
        }
        ss = 0.0;
        for (int i = 0; i < n; i++) {
          const double v = R.x[(size_t)i * nthr + q];
          ss += v * v;
        }
      }
      scale = ss > 0.0 ? 1.0 / sqrt(ss) : 1.0;
    }
    double* zj = Zt + (size_t)n * jj;
    for (int i0 = 0; i0 < n; i0 += kSteinB) {
      double xb[kSteinB];
#pragma unroll
      for (int u = 0; u < kSteinB; u++) xb[u] = (i0 + u < n) ? R.x[(size_t)(i0 + u) * nthr + q] : 0.0;
#pragma unroll
      for (int u = 0; u < kSteinB; u++)
        if (i0 + u < n) zj[i0 + u] = xb[u] * scale;
    }
  }
}

// One thread per eigenvalue, scratch rows in global memory S (7 n^2
// doubles, interleaved by thread so that lanes stepping through the same
// row coalesce).  blockDim = 64.  For n beyond the LDS form below.
__global__ void __launch_bounds__(64)
k_stein(const double* __restrict__ dg, const double* __restrict__ eg, const double* __restrict__ w, int n,
        double* __restrict__ Zt, double* __restrict__ S) {
  extern __shared__ double sde[];
  double* d = sde;
  double* e = d + n;
  __shared__ double sred[16];
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    d[i] = dg[i];
    if (i < n - 1) e[i] = eg[i];
  }
  double gl, gu, nrm, em;
  tri_bounds(dg, eg, n, sred, &gl, &gu, &nrm, &em);  // ends with a barrier
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const size_t an = (size_t)n * n;
  SteinRows R{S, S + an, S + 2 * an, S + 3 * an, S + 4 * an, S + 5 * an, S + 6 * an};
  stein_cluster(d, e, w, n, j, nrm, Zt, R, n, j);
}

// One eigenvalue per workgroup (one wavefront; lane 0 iterates), its
// scratch (seven fields per row, rows interleaved) in LDS next to d and e: every step of the serial
// factorisation and solves is an LDS round trip instead of an L2 one, and
// the n workgroups spread over the CUs.  10 n doubles of LDS (d, e and
// eight doubles per row).
__global__ void __launch_bounds__(64)
k_stein_lds(const double* __restrict__ dg, const double* __restrict__ eg, const double* __restrict__ w, int n,
            double* __restrict__ Zt, EigBatch bs) {
  extern __shared__ double sde[];
  dg += blockIdx.y * bs.v;
  eg += blockIdx.y * bs.v;
  w += blockIdx.y * bs.w;
  Zt += blockIdx.y * bs.Z;
  double* d = sde;
  double* e = d + n;
  double* S = e + n;
  __shared__ double sred[16];
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    d[i] = dg[i];
    if (i < n - 1) e[i] = eg[i];
  }
  double gl, gu, nrm, em;
  tri_bounds(dg, eg, n, sred, &gl, &gu, &nrm, &em);  // ends with a barrier
  if (threadIdx.x != 0) return;
  // the seven fields of row i at S + 8 i + f (the row stride as the
  // "thread count" of the shared indexing: every access of a batch is one
  // base address plus an immediate offset)
  SteinRows R{S, S + 1, S + 2, S + 3, S + 4, S + 5, S + 6};
  stein_cluster(d, e, w, n, blockIdx.x, nrm, Zt, R, 8, 0);
}

// V of the chunk of reflectors kb .. kb + b - 1 into LDS, row-major with the
// padded stride nbw + 1 (row r, reflector j; 1 on row kb + 1 + j, zeros
// above).  Consecutive threads walk down one reflector column (coalesced
// loads), 8 loads in flight per thread; the padded stride keeps the
// transposed LDS stores conflict-free.  blockDim = 256.
__device__ __forceinline__ void stage_v_chunk(const double* __restrict__ A, int lda, int n, int nbw, int kb, int b,
                                              double* Vs) {
  const int lvs = nbw + 1, t = threadIdx.x;
  const int nr = n - kb, tot = nr * nbw;
  for (int e0 = 0; e0 < tot; e0 += 8 * 256) {
    double b8[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int e = e0 + u * 256 + t;
      const int j = e / nr, r = kb + e % nr, k = kb + j;
      b8[u] = (e < tot && j < b) ? ((r == k + 1) ? 1.0 : (r > k + 1 ? A[(size_t)lda * k + r] : 0.0)) : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int e = e0 + u * 256 + t;
      if (e < tot) Vs[(size_t)(kb + e % nr) * lvs + e / nr] = b8[u];
    }
  }
}

// Triangular factor T of each compact-WY chunk (LAPACK dlarft 'F', 'C'):
// H_kb ... H_{kb+b-1} = I - V T V^T with T_jj = tau_j and
// T[0:j, j] = -tau_j T[0:j, 0:j] (V^T V)[0:j, j].  One workgroup per chunk,
// all chunks at once; T (nbw x nbw, row-major) to Tg + chunk * nbw^2.
// blockDim = 256.
__global__ void __launch_bounds__(256)
k_larft(const double* __restrict__ A, int lda, const double* __restrict__ tau, int n, int nbw,
        double* __restrict__ Tg, EigBatch bs) {
  extern __shared__ double sm[];
  A += blockIdx.y * bs.A;
  tau += blockIdx.y * bs.v;
  Tg += blockIdx.y * bs.T;
  const int lvs = nbw + 1;
  double* Vs = sm;                                 // n x nbw (rows kb.. used)
  double* Gs = Vs + (size_t)n * lvs;               // nbw x nbw: V^T V
  double* Ts = Gs + (size_t)nbw * nbw;
  const int nref = n - 2;
  const int kb = blockIdx.x * nbw, b = min(nref, kb + nbw) - kb;
  const int t = threadIdx.x;
  stage_v_chunk(A, lda, n, nbw, kb, b, Vs);
  __syncthreads();
  // G = V^T V, strictly upper part (rows of V below kb + 1 + j are zero in
  // column j): 4 independent partial sums per entry
  for (int e = t; e < nbw * nbw; e += blockDim.x) {
    const int i = e / nbw, j = e % nbw;
    double g0 = 0.0, g1 = 0.0, g2 = 0.0, g3 = 0.0;
    if (i < j && j < b) {
      int r = kb + 1 + j;
      for (; r + 3 < n; r += 4) {
        g0 += Vs[(size_t)r * lvs + i] * Vs[(size_t)r * lvs + j];
        g1 += Vs[(size_t)(r + 1) * lvs + i] * Vs[(size_t)(r + 1) * lvs + j];
        g2 += Vs[(size_t)(r + 2) * lvs + i] * Vs[(size_t)(r + 2) * lvs + j];
        g3 += Vs[(size_t)(r + 3) * lvs + i] * Vs[(size_t)(r + 3) * lvs + j];
      }
      for (; r < n; r++) g0 += Vs[(size_t)r * lvs + i] * Vs[(size_t)r * lvs + j];
    }
    Gs[e] = (g0 + g1) + (g2 + g3);
  }
  __syncthreads();
  // Thread i owns row i of T and reads only that row, so the recurrence
  // needs no synchronisation.
  if (t < nbw) {
    for (int j = 0; j < nbw; j++) {
      double x = 0.0;
      if (t < j && j < b)
        for (int q = t; q < j; q++) x += Ts[t * nbw + q] * Gs[q * nbw + j];
      Ts[t * nbw + j] = (j >= b || t > j) ? 0.0 : (t == j ? tau[kb + j] : -tau[kb + j] * x);
    }
  }
  __syncthreads();
  double* Tc = Tg + (size_t)blockIdx.x * nbw * nbw;
  for (int e = t; e < nbw * nbw; e += blockDim.x) Tc[e] = Ts[e];
}

// Z <- H_0 H_1 ... H_{n-3} Z for the kOrmCols columns of this workgroup, in
// compact-WY chunks of nbw reflectors (LAPACK dlarfb with the T of k_larft):
// for the chunk H_kb ... H_{ke-1} = I - V T V^T,  Z <- Z - V (T (V^T Z)),
// chunks from the last to the first.  V, T and the kOrmCols-column block of Z live
// in LDS; no per-reflector reductions.  blockDim = 256.
__global__ void __launch_bounds__(256)
k_ormtr(const double* __restrict__ A, int lda, const double* __restrict__ Tg, int n, int nbw,
        double* __restrict__ Z, int ldz, EigBatch bs) {
  extern __shared__ double sm[];
  A += blockIdx.y * bs.A;
  Tg += blockIdx.y * bs.T;
  Z += blockIdx.y * bs.Z;
  double* zs = sm;                                 // n x kOrmCols, row-major
  const int lvs = nbw + 1;                         // padded row stride of Vs
  double* Vs = zs + (size_t)n * kOrmCols;          // n x nbw, row-major (row r, reflector j)
  double* Ts = Vs + (size_t)n * lvs;               // nbw x nbw
  double* W1 = Ts + (size_t)nbw * nbw;             // nbw x kOrmCols: V^T Z
  double* W2 = W1 + (size_t)nbw * kOrmCols;        // nbw x kOrmCols: T W1
  const int c0 = blockIdx.x * kOrmCols;
  const int nc = min(kOrmCols, n - c0);
  const int t = threadIdx.x;
  // staged 8 loads per thread at a time (one memory round trip per batch)
  for (int e0 = 0; e0 < n * kOrmCols; e0 += 8 * 256) {
    double b8[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int e = e0 + u * 256 + t, r = e / kOrmCols, c = e % kOrmCols;
      b8[u] = (e < n * kOrmCols && c < nc) ? Z[r + (size_t)ldz * (c0 + c)] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; u++)
      if (e0 + u * 256 + t < n * kOrmCols) zs[e0 + u * 256 + t] = b8[u];
  }
  const int nref = n - 2;  // reflectors 0..n-3 (tau[n-2] = 0)
  const int nchunk = (nref + nbw - 1) / nbw;
  for (int ch = nchunk - 1; ch >= 0; ch--) {
    const int kb = ch * nbw, ke = min(nref, kb + nbw), b = ke - kb;
    __syncthreads();
    {
      const double* Tc = Tg + (size_t)ch * nbw * nbw;
      double tb[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int e = t + 256 * u;
        tb[u] = e < nbw * nbw ? Tc[e] : 0.0;
      }
      stage_v_chunk(A, lda, n, nbw, kb, b, Vs);
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const int e = t + 256 * u;
        if (e < nbw * nbw) Ts[e] = tb[u];
      }
    }
    __syncthreads();
    // W1 = V^T Z (rows of V above kb + 1 + j are zero in column j), each
    // of the nbw x kOrmCols dot products split over S consecutive lanes
    // (rows strided by S), then summed across them
    {
      const int D = nbw * kOrmCols;
      int S = 1;
      while (S < 16 && 2 * S * D <= (int)blockDim.x) S *= 2;
      for (int g0 = 0; g0 < D * S; g0 += blockDim.x) {
        const int g = g0 + t, e = g / S, p = g % S;
        double w0 = 0.0, w1 = 0.0;
        const int j = e / kOrmCols, c = e % kOrmCols;
        if (e < D && j < b) {
          int r = kb + 1 + j + p;
          for (; r + S < n; r += 2 * S) {
            w0 += Vs[(size_t)r * lvs + j] * zs[r * kOrmCols + c];
            w1 += Vs[(size_t)(r + S) * lvs + j] * zs[(r + S) * kOrmCols + c];
          }
          if (r < n) w0 += Vs[(size_t)r * lvs + j] * zs[r * kOrmCols + c];
        }
        w0 += w1;
        for (int o = 1; o < S; o <<= 1) w0 += __shfl_xor(w0, o, 64);  // S divides 64: groups stay in a wavefront
        if (e < D && p == 0) W1[e] = w0;
      }
    }
    __syncthreads();
    // W2 = T W1
    for (int e = t; e < nbw * kOrmCols; e += blockDim.x) {
      const int j = e / kOrmCols, c = e % kOrmCols;
      double w = 0.0;
      if (j < b)
        for (int i = j; i < b; i++) w += Ts[j * nbw + i] * W1[i * kOrmCols + c];
      W2[e] = w;
    }
    __syncthreads();
    // Z -= V W2 on rows kb+1 .. n-1
    for (int e = t; e < (n - kb - 1) * kOrmCols; e += blockDim.x) {
      const int r = kb + 1 + e / kOrmCols, c = e % kOrmCols;
      double u = 0.0;
      for (int j = 0; j < b; j++) u += Vs[(size_t)r * lvs + j] * W2[j * kOrmCols + c];
      zs[r * kOrmCols + c] -= u;
    }
  }
  __syncthreads();
  for (int e = t; e < n * kOrmCols; e += blockDim.x) {
    const int r = e / kOrmCols, c = e % kOrmCols;
    if (c < nc) Z[r + (size_t)ldz * (c0 + c)] = zs[e];
  }
}

// Eigen-decomposition of the symmetric n x n A (upper triangle used, lda) in
// place: A <- eigenvectors, w (device) <- eigenvalues ascending.
int eig_sym(pw_ctx* c, int n, double* A, int lda, double* w) {
  hipStream_t s = c->stream;
  if (n < 1 || n > kEigMaxN) return fail(PW_ERR_UNSUPPORTED, "eig_sym: n = %d outside [1, %d]", n, kEigMaxN);
  if (n == 1) {
    const double one = 1.0;
    PW_HIP(hipMemcpyAsync(w, A, sizeof(double), hipMemcpyDeviceToDevice, s));
    PW_HIP(hipMemcpyAsync(A, &one, sizeof(double), hipMemcpyHostToDevice, s));
    PW_HIP(hipStreamSynchronize(s));
    return PW_OK;
  }
  const size_t nn = (size_t)n * n;
  // compact-WY chunks of the back-transform: nbw reflectors each
  const int nbw = std::max(1, std::min(32, 8192 / n));
  const int nchunk = (n - 2 + nbw - 1) / nbw;
  const size_t need = sizeof(double) * (3 * (size_t)n + nn + 7 * nn + (size_t)nchunk * nbw * nbw);
  if (need > c->eig_ws_bytes) {
    if (c->d_eig_ws) (void)hipFree(c->d_eig_ws);
    c->d_eig_ws = nullptr;
    c->eig_ws_bytes = 0;
    PW_HIP(hipMalloc(&c->d_eig_ws, need));
    c->eig_ws_bytes = need;
  }
  double* d = (double*)c->d_eig_ws;
  double* e = d + n;
  double* tau = e + n;
  double* Zt = tau + n;
  double* S = Zt + nn;
  double* Tw = S + 7 * nn;
  const int g1 = (int)std::min<size_t>((nn + 255) / 256, 4096);
  hipLaunchKernelGGL(k_sym_from_upper, dim3(g1), dim3(256), 0, s, A, n, lda, 0L);
  // panel width: the V and W panels (2 nbk n doubles) fit LDS next to v and
  // the partial sums (n = 256: 32; 512: 16; 1024: 8)
  if (n <= 32 * kTr2MaxNB) {
    const size_t lds_reg = sizeof(double) * (3 * 256 + (size_t)kTr2TabRows * 32 + kTr2Waves +
                                             (size_t)kTr2MaxNB * 32 * kTr2Ld + (size_t)kTr2Waves * 32 * 33);
    PW_TRY(ensure_lds((const void*)k_sytrd_reg, lds_reg));
    hipLaunchKernelGGL(k_sytrd_reg, dim3(1), dim3(64 * kTr2Waves), lds_reg, s, A, n, lda, d, e, tau, EigBatch{});
  } else {
    const int nbk = std::max(1, std::min(32, 8192 / n));
    const size_t lds_trd = sizeof(double) * ((size_t)2 * nbk * n + n + 1024 + 2 * (size_t)nbk);
    PW_TRY(ensure_lds((const void*)k_sytrd, lds_trd + 1024));  // + its static LDS
    hipLaunchKernelGGL(k_sytrd, dim3(1), dim3(kTrdThreads), lds_trd, s, A, n, lda, nbk, d, e, tau);
  }
  // the compact-WY factors (reflectors only) on the second stream while
  // bisection and inverse iteration (d, e only) run on this one
  const bool side = nchunk > 0;
  if (side) {
    PW_TRY(side_stream(c));
    PW_HIP(hipEventRecord(c->ev_fork, s));
    PW_HIP(hipStreamWaitEvent(c->side, c->ev_fork, 0));
    const size_t lds_t = sizeof(double) * ((size_t)n * (nbw + 1) + 2 * (size_t)nbw * nbw);
    PW_TRY(ensure_lds((const void*)k_larft, lds_t));
    hipLaunchKernelGGL(k_larft, dim3(nchunk), dim3(256), lds_t, c->side, (const double*)A, lda, (const double*)tau,
                       n, nbw, Tw, EigBatch{});
    PW_HIP(hipEventRecord(c->ev_join, c->side));
  }
  hipLaunchKernelGGL(k_stebz, dim3(n), dim3(256), 0, s, (const double*)d,
                     (const double*)e, n, w, EigBatch{});
  if (n <= kSteinLdsMaxN)
    hipLaunchKernelGGL(k_stein_lds, dim3(n), dim3(64), sizeof(double) * 10 * n, s, (const double*)d,
                       (const double*)e, (const double*)w, n, Zt, EigBatch{});
  else
    hipLaunchKernelGGL(k_stein, dim3((n + 63) / 64), dim3(64), sizeof(double) * 2 * n, s, (const double*)d,
                       (const double*)e, (const double*)w, n, Zt, S);
  if (side) PW_HIP(hipStreamWaitEvent(s, c->ev_join, 0));
  // compact-WY chunk: V (n x nbw) and T next to the Z block in LDS
  const size_t lds_orm = sizeof(double) * ((size_t)n * kOrmCols + (size_t)n * (nbw + 1) + (size_t)nbw * nbw +
                                           2 * (size_t)nbw * kOrmCols);
  PW_TRY(ensure_lds((const void*)k_ormtr, lds_orm));
  hipLaunchKernelGGL(k_ormtr, dim3((n + kOrmCols - 1) / kOrmCols), dim3(256), lds_orm, s, (const double*)A, lda,
                     (const double*)Tw, n, nbw, Zt, n, EigBatch{});
  PW_HIP(hipGetLastError());
  PW_HIP(hipMemcpy2DAsync(A, sizeof(double) * lda, Zt, sizeof(double) * n, sizeof(double) * n, n,
                          hipMemcpyDeviceToDevice, s));
  return PW_OK;
}

// ------------------------------------------------------------------------
// Batched generalized problems A x = lambda B x (Sygvd itype 1, 'V', 'U';
// the PPCG sub-block problems of ppcg.cpp:438-460), order n <= 256:
// B = U^T U (k_potrf), C = U^{-T} A U^{-1} by two right triangular solves
// around a transpose, the eigenpairs of C by the kernels above with a batch
// index, x = U^{-1} y (k_backsub_upper).  Only the upper triangles of A and
// B are read.

// T_b = A_b^T, n x n (T with leading dimension n)
__global__ void k_transpose_batched(const double* __restrict__ A, int lda, long sA, double* __restrict__ T, int n) {
  A += blockIdx.y * sA;
  T += (size_t)blockIdx.y * n * n;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)n * n; e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e % n), j = (int)(e / n);
    T[i + (size_t)n * j] = A[j + (size_t)lda * i];
  }
}

// Z = U^{-1} V by block back substitution, Z_I = Dinv_I (V_I - sum_{K>I}
// U_IK Z_K) with Dinv_I = U_II^{-1} (row-major 32 x 32 blocks from
// k_potrf); one workgroup per (problem, 32-column tile), the tile of Z in
// LDS (n x 33).  Z may alias V.
__global__ void __launch_bounds__(256)
k_backsub_upper(const double* __restrict__ U, int ldu, long sU, const double* __restrict__ Dinv, long sD,
                const double* V, int ldv, long sV, double* Z, int ldz, long sZ, int n) {
  extern __shared__ double zs[];  // n x 33 (rows of the tile), then the 32 x 33 block R
  double* rs = zs + (size_t)n * 33;
  const int t = threadIdx.x, j0 = blockIdx.x * 32, nc = min(32, n - j0);
  U += blockIdx.y * sU;
  Dinv += blockIdx.y * sD;
  V += blockIdx.y * sV;
  Z += blockIdx.y * sZ;
  const int NB = (n + 31) / 32;
  for (int I = NB - 1; I >= 0; I--) {
    const int r0 = 32 * I, nr = min(32, n - r0);
    for (int e = t; e < 32 * 32; e += 256) {
      const int i = e >> 5, jj = e & 31;
      double acc = 0.0;
      if (i < nr && jj < nc) {
        acc = V[(r0 + i) + (size_t)ldv * (j0 + jj)];
        for (int k = r0 + 32; k < n; k++) acc -= U[(r0 + i) + (size_t)ldu * k] * zs[k * 33 + jj];
      }
      rs[i * 33 + jj] = acc;
    }
    __syncthreads();
    const double* Di = Dinv + (size_t)I * 32 * 32;
    for (int e = t; e < 32 * 32; e += 256) {
      const int i = e >> 5, jj = e & 31;
      double acc = 0.0;
      for (int q = 0; q < 32; q++) acc += Di[i * 32 + q] * rs[q * 33 + jj];
      if (i < nr) zs[(r0 + i) * 33 + jj] = acc;
    }
    __syncthreads();
  }
  for (int e = t; e < n * 32; e += 256) {
    const int i = e % n, jj = e / n;
    if (jj < nc) Z[i + (size_t)ldz * (j0 + jj)] = zs[i * 33 + jj];
  }
}

int sygvd_batched(pw_ctx* c, int n, double* A, int lda, long sA, double* B, int ldb, long sB, double* w, long sw,
                  int* info, int nbatch) {
  // n >= 2: the tridiagonalisation ends on a 2 x 2 block
  if (n < 2 || n > 32 * kTr2MaxNB) return fail(PW_ERR_UNSUPPORTED, "sygvd_batched: n = %d (2 .. 256)", n);
  if (nbatch < 1) return PW_OK;
  hipStream_t s = c->stream;
  const size_t nn = (size_t)n * n, ndi = chol_dinv_doubles(n);
  const int nbw = std::max(1, std::min(32, 8192 / n));
  const int nchunk = (n - 2 + nbw - 1) / nbw;
  const size_t ntw = (size_t)std::max(1, nchunk) * nbw * nbw;
  const size_t per = ndi + nn + nn + 3 * (size_t)n + ntw;  // Dinv, C, Zt, d e tau, Tw
  const size_t need = sizeof(double) * per * nbatch;
  if (need > c->geig_ws_bytes) {
    if (c->d_geig_ws) (void)hipFree(c->d_geig_ws);
    c->d_geig_ws = nullptr;
    c->geig_ws_bytes = 0;
    PW_HIP(hipMalloc(&c->d_geig_ws, need));
    c->geig_ws_bytes = need;
  }
  double* Dinv = (double*)c->d_geig_ws;
  double* C = Dinv + ndi * nbatch;
  double* Zt = C + nn * nbatch;
  double* d = Zt + nn * nbatch;
  double* e = d + (size_t)n * nbatch;
  double* tau = e + (size_t)n * nbatch;
  double* Tw = tau + (size_t)n * nbatch;
  const int g1 = (int)std::min<size_t>((nn + 255) / 256, 4096);
  PW_TRY(chol_upper_batched(c, n, B, ldb, sB, Dinv, (long)ndi, info, nbatch));
  hipLaunchKernelGGL(k_sym_from_upper, dim3(g1, nbatch), dim3(256), 0, s, A, n, lda, sA);
  PW_TRY(trsm_right_upper_batched(c, n, n, A, lda, sA, B, ldb, sB, Dinv, (long)ndi, nbatch));  // A U^{-1}
  hipLaunchKernelGGL(k_transpose_batched, dim3(g1, nbatch), dim3(256), 0, s, (const double*)A, lda, sA, C, n);
  PW_TRY(trsm_right_upper_batched(c, n, n, C, n, (long)nn, B, ldb, sB, Dinv, (long)ndi, nbatch));  // U^{-T} A U^{-1}
  hipLaunchKernelGGL(k_sym_from_upper, dim3(g1, nbatch), dim3(256), 0, s, C, n, n, (long)nn);
  EigBatch bs;
  bs.A = (long)nn;
  bs.v = n;
  bs.w = sw;
  bs.T = (long)ntw;
  bs.Z = (long)nn;
  const size_t lds_reg = sizeof(double) * (3 * 256 + (size_t)kTr2TabRows * 32 + kTr2Waves +
                                           (size_t)kTr2MaxNB * 32 * kTr2Ld + (size_t)kTr2Waves * 32 * 33);
  PW_TRY(ensure_lds((const void*)k_sytrd_reg, lds_reg));
  hipLaunchKernelGGL(k_sytrd_reg, dim3(nbatch), dim3(64 * kTr2Waves), lds_reg, s, C, n, n, d, e, tau, bs);
  if (nchunk > 0) {
    const size_t lds_t = sizeof(double) * ((size_t)n * (nbw + 1) + 2 * (size_t)nbw * nbw);
    PW_TRY(ensure_lds((const void*)k_larft, lds_t));
    hipLaunchKernelGGL(k_larft, dim3(nchunk, nbatch), dim3(256), lds_t, s, (const double*)C, n, (const double*)tau,
                       n, nbw, Tw, bs);
  }
  hipLaunchKernelGGL(k_stebz, dim3(n, nbatch), dim3(256), 0, s, (const double*)d,
                     (const double*)e, n, w, bs);
  hipLaunchKernelGGL(k_stein_lds, dim3(n, nbatch), dim3(64), sizeof(double) * 10 * n, s, (const double*)d,
                     (const double*)e, (const double*)w, n, Zt, bs);
  const size_t lds_orm = sizeof(double) * ((size_t)n * kOrmCols + (size_t)n * (nbw + 1) + (size_t)nbw * nbw +
                                           2 * (size_t)nbw * kOrmCols);
  PW_TRY(ensure_lds((const void*)k_ormtr, lds_orm));
  hipLaunchKernelGGL(k_ormtr, dim3((n + kOrmCols - 1) / kOrmCols, nbatch), dim3(256), lds_orm, s, (const double*)C,
                     n, (const double*)Tw, n, nbw, Zt, n, bs);
  // x = U^{-1} y into A
  const size_t lds_b = sizeof(double) * ((size_t)n * 33 + 32 * 33);
  PW_TRY(ensure_lds((const void*)k_backsub_upper, lds_b));
  hipLaunchKernelGGL(k_backsub_upper, dim3((n + 31) / 32, nbatch), dim3(256), lds_b, s, (const double*)B, ldb, sB,
                     (const double*)Dinv, (long)ndi, (const double*)Zt, n, (long)nn, A, lda, sA, n);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

}  // namespace pw
