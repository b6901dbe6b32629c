// Tall-skinny GEMMs of the hot path on MFMA (gfx950), FP64 and FP32.
//
//   TN:  C(n1 x n2) = A^T B, A: K x n1, B: K x n2, K = 2 npw long
//        (Gram blocks Re(X^H Y), projections Re(beta^H X); blas.cpp:997-1029,
//        eigensolver.cpp:471-515, lobpcg.cpp:227-419).  Split-K: each
//        workgroup reduces one K chunk of one 64x64 tile into a partial
//        buffer; a second kernel sums the partials in a fixed order, so the
//        result is deterministic.
//   NN:  Y(m x n) = alpha A(m x k) B(k x n) + beta Y  (subspace rotations,
//        beta * W; blas.cpp:1031-1067, lobpcg.cpp:445-526), 64 x 64 tiles;
//        the projector apply inside H Y takes an epilogue that adds H's
//        other terms or forms a Chebyshev step.
//
// Both use 64x64 workgroup tiles, 4 waves of 32x32, and the 16x16x4 MFMA
// (v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32).  Lane layout of the
// 16x16x4 MFMA: A operand lane l = (row l%16, k l/16), B operand lane l =
// (k l/16, col l%16).  Accumulator lane l register i holds column l%16 and
// row 4(l/16)+i for the FP32 MFMA but row 4i + l/16 for the FP64 one
// (checked on gfx950 with a known 16x16x4 product).
#include <algorithm>
#include <cstdlib>
#include <string>

#include "pw_internal.h"

namespace pw {

typedef double d4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

template <typename T> struct acc4;
template <> struct acc4<double> { typedef d4v t; };
template <> struct acc4<float> { typedef f4v t; };

__device__ __forceinline__ d4v mfma16(double a, double b, d4v c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f4v mfma16(float a, float b, f4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

constexpr int BM = 64, BN = 64, BK = 32;

// Block -> (row tile, column tile, K split) for the split-K products, which
// launch ti x tj tiles x nsplit splits as one row of blocks.  Blocks are
// dispatched round-robin over the 8 XCDs (block b to XCD b % 8), each with
// its own L2; giving every XCD one contiguous range of (split, tile) pairs
// puts the ti x tj tiles of a K chunk, which read the same A and B rows, in
// one L2 instead of spreading them over all eight.
struct TnBlock {
  int i, j, split;
};
__device__ __forceinline__ TnBlock tn_block(int n1, int n2, int K, int kc, int tb) {
  const int ti = (n1 + tb - 1) / tb, tj = (n2 + tb - 1) / tb, tiles = ti * tj;
  const int n = tiles * ((K + kc - 1) / kc), b = blockIdx.x;
  const int q = n >> 3, r = n & 7, x = b & 7;
  const int L = x * q + min(x, r) + (b >> 3);
  const int tile = L % tiles;
  return TnBlock{tile % ti, tile / ti, L / tiles};
}

template <typename T>
__device__ __forceinline__ int acc_row(int lane, int r) {
  return sizeof(T) == 8 ? 4 * r + (lane >> 4) : 4 * (lane >> 4) + r;
}

// NN products accumulate FP64 tiles transposed (mfma(b, a) = (A B)^T): the
// 16x16x4f64 accumulator then holds, in register r, 16 consecutive rows of
// column 4 r + lane / 16 in lanes 16 j .. 16 j + 15, so the epilogue reads
// and writes 128-byte runs; FP32 tiles keep A B (four consecutive rows of
// column lane % 16 per lane).  (row, column) of accumulator register r:
__device__ __forceinline__ d4v mfma_nn(double a, double b, d4v c) { return mfma16(b, a, c); }
__device__ __forceinline__ f4v mfma_nn(float a, float b, f4v c) { return mfma16(a, b, c); }
template <typename T>
__device__ __forceinline__ int nn_row(int lane, int r) {
  return sizeof(T) == 8 ? (lane & 15) : 4 * (lane >> 4) + r;
}
template <typename T>
__device__ __forceinline__ int nn_col(int lane, int r) {
  return sizeof(T) == 8 ? 4 * r + (lane >> 4) : (lane & 15);
}

// Partial C tile of A^T B over K rows [kb, ke): P[split](n1 x n2, ld n1).
// TM = 1: 64 x 64 tile, each wave 32 x 32; TM = 2: 128 x 128 tile, each wave
// 64 x 64 (twice the flops per loaded byte, for large outputs).
// Software-pipelined: the next K chunk is loaded into registers while the
// MFMAs run on the current one in LDS (dynamic LDS: 2 x 64 TM x (BK + 1)).
// Chunk loads are column runs: load q of wave w reads column 16 TM w + 2 q +
// lane/32, rows lane%32 (one 32-row segment per half wave), for A and for B.
template <typename T, int TM>
__global__ void __launch_bounds__(256)
k_gemm_tn_partial(const T* __restrict__ A, int lda, const T* __restrict__ B, int ldb, int K, int n1, int n2,
                  int kc, T* __restrict__ P, const T* const* __restrict__ Ab, const T* const* __restrict__ Bb,
                  long sP) {
  typedef typename acc4<T>::t V;
  if (Ab) {  // product blockIdx.y of a batch: operands from the pointer tables
    A = Ab[blockIdx.y];
    B = Bb[blockIdx.y];
    P += blockIdx.y * sP;
  }
  constexpr int TB = BM * TM, NQ = 8 * TM, NA = 2 * TM;
  extern __shared__ __align__(16) unsigned char gemm_smem[];
  T (*As)[BK + 1] = reinterpret_cast<T (*)[BK + 1]>(gemm_smem);
  T (*Bs)[BK + 1] = As + TB;
  const TnBlock tb = tn_block(n1, n2, K, kc, TB);
  const int i0 = tb.i * TB, j0 = tb.j * TB, split = tb.split;
  const int kb = split * kc, ke = min(K, kb + kc);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wi = (w & 1) * 32 * TM, wj = (w >> 1) * 32 * TM;
  V acc[NA][NA];
#pragma unroll
  for (int a = 0; a < NA; a++)
#pragma unroll
    for (int b = 0; b < NA; b++) acc[a][b] = V{0, 0, 0, 0};
  const int lr = lane & 31, lcol = 16 * TM * w + (lane >> 5);  // column of load q: lcol + 2 q
  T ra[NQ], rb[NQ];
  // unconditional loads at clamped indices, the zero fill outside the
  // operands at the LDS store (a select on a loaded value makes the
  // compiler wait for the load where it is issued)
  auto load = [&](int k0) {
    const int r = min(k0 + lr, ke - 1);
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      const int ci = min(i0 + lcol + 2 * q, n1 - 1), cj = min(j0 + lcol + 2 * q, n2 - 1);
      ra[q] = A[(size_t)ci * lda + r];
      rb[q] = B[(size_t)cj * ldb + r];
    }
  };
  load(kb);
  for (int k0 = kb; k0 < ke; k0 += BK) {
    const bool okr = k0 + lr < ke;
#pragma unroll
    for (int q = 0; q < NQ; q++) {
      As[lcol + 2 * q][lr] = (okr && i0 + lcol + 2 * q < n1) ? ra[q] : (T)0;
      Bs[lcol + 2 * q][lr] = (okr && j0 + lcol + 2 * q < n2) ? rb[q] : (T)0;
    }
    lds_barrier();
    if (k0 + BK < ke) load(k0 + BK);  // in flight during the MFMAs below and the barrier after them
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T av[NA], bv[NA];
#pragma unroll
      for (int a = 0; a < NA; a++) av[a] = As[wi + a * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int b = 0; b < NA; b++) bv[b] = Bs[wj + b * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int a = 0; a < NA; a++)
#pragma unroll
        for (int b = 0; b < NA; b++) acc[a][b] = mfma16(av[a], bv[b], acc[a][b]);
    }
    lds_barrier();
  }
  T* Ps = P + (size_t)split * n1 * n2;
#pragma unroll
  for (int a = 0; a < NA; a++)
#pragma unroll
    for (int b = 0; b < NA; b++) {
      const int col = j0 + wj + b * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = i0 + wi + a * 16 + acc_row<T>(lane, r);
        if (row < n1 && col < n2) Ps[row + (size_t)n1 * col] = acc[a][b][r];
      }
    }
}

// ---- split BF16 products (FP32 data, FP32 accumulation) ----
// Each FP32 operand x is split into NS BF16 terms rounded to nearest even:
// NS = 2: x = hi + lo (16 significant bits), NS = 3: x = hi + mid + lo (24
// bits, FP32's).  A product a b is the sum of the term products down to
// 2^-8 (NS - 1) relative: NS = 2: a_lo b_hi + a_hi b_lo + a_hi b_hi; NS = 3:
// a_lo b_hi + a_hi b_lo + a_mid b_mid + a_mid b_hi + a_hi b_mid + a_hi b_hi
// (smallest first), on v_mfma_f32_16x16x32_bf16, whose operand lane l holds
// row (or column) l % 16 and k = 8 (l / 16) .. 8 (l / 16) + 7 and whose
// accumulator is laid out as v_mfma_f32_16x16x4f32's.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

// Round to nearest even: one v_cvt_pk_bf16_f32 per pair of values on
// gfx950; x - hi is exact in FP32, so the terms sum to x up to the last
// term's rounding.
template <int NS>
__device__ __forceinline__ void bf16_split(float x, unsigned short (&t)[NS]) {
  float r = x;
#pragma unroll
  for (int i = 0; i < NS; i++) {
    const __bf16 h = (__bf16)r;
    t[i] = __builtin_bit_cast(unsigned short, h);
    r -= (float)h;
  }
}
// Four values (two pairs) into NS packed BF16 term pairs: term i of x0, x1
// in t[i].x (x0 in the low half), of x2, x3 in t[i].y; the same terms as
// bf16_split (one v_cvt_pk_bf16_f32 per pair and term).
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
template <int NS>
__device__ __forceinline__ void bf16_split4(float2 p, float2 q, uint2 (&t)[NS]) {
  float r0 = p.x, r1 = p.y, r2 = q.x, r3 = q.y;
#pragma unroll
  for (int i = 0; i < NS; i++) {
    const bf16x2 h01 = {(__bf16)r0, (__bf16)r1}, h23 = {(__bf16)r2, (__bf16)r3};
    const unsigned u = __builtin_bit_cast(unsigned, h01), v = __builtin_bit_cast(unsigned, h23);
    t[i].x = u;
    t[i].y = v;
    if (i + 1 < NS) {  // residuals from the packed words (a BF16 is the top half of its FP32)
      r0 -= __uint_as_float(u << 16);
      r1 -= __uint_as_float(u & 0xffff0000u);
      r2 -= __uint_as_float(v << 16);
      r3 -= __uint_as_float(v & 0xffff0000u);
    }
  }
}
__device__ __forceinline__ f4v mfma_b16(const unsigned short* a, const unsigned short* b, f4v acc) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(*reinterpret_cast<const bf16x8*>(a),
                                                 *reinterpret_cast<const bf16x8*>(b), acc, 0, 0, 0);
}
// a[i], b[i]: the i-th BF16 term tiles (16 x 32 operands at the lane's
// offset); NS term tiles each, at a common offset.
template <int NS>
__device__ __forceinline__ f4v mfma_split(const unsigned short* const (&a)[NS], const unsigned short* const (&b)[NS],
                                          f4v acc) {
  if constexpr (NS == 3) {
    acc = mfma_b16(a[2], b[0], acc);
    acc = mfma_b16(a[0], b[2], acc);
    acc = mfma_b16(a[1], b[1], acc);
  }
  acc = mfma_b16(a[1], b[0], acc);
  acc = mfma_b16(a[0], b[1], acc);
  return mfma_b16(a[0], b[0], acc);
}
constexpr int BKP = 40;  // LDS row stride of the bf16 tiles (80 bytes: 16-byte aligned rows)
// The term pairs of mfma_split in its order (smallest first): NS = 3 uses
// all six, NS = 2 the last three.  Loops over (pair, tile) in this order
// give every tile the same sequence of products while consecutive MFMAs
// write different accumulators.
__device__ constexpr int kSplitPairA[6] = {2, 0, 1, 1, 0, 0};
__device__ constexpr int kSplitPairB[6] = {0, 2, 1, 0, 1, 0};

// k_gemm_tn_partial with FP32 inputs split into NS BF16 terms: the same
// tiles, loads and split-K partials; each 32-row K chunk is one 16x16x32
// product per kept pairing of terms instead of eight 16x16x4 FP32 products.
template <int TM, int NS>
__global__ void __launch_bounds__(256)
k_gemm_tn_partial_bf(const float* __restrict__ A, int lda, const float* __restrict__ B, int ldb, int K, int n1,
                     int n2, int kc, float* __restrict__ P) {
  constexpr int TB = BM * TM, NQ = 8 * TM, NA = 2 * TM;
  extern __shared__ __align__(16) unsigned char gemm_smem[];
  unsigned short* As = reinterpret_cast<unsigned short*>(gemm_smem);  // NS term tiles of A, then of B
  unsigned short* Bs = As + NS * TB * BKP;
  const TnBlock tb = tn_block(n1, n2, K, kc, TB);
  const int i0 = tb.i * TB, j0 = tb.j * TB, split = tb.split;
  const int kb = split * kc, ke = min(K, kb + kc);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wi = (w & 1) * 32 * TM, wj = (w >> 1) * 32 * TM;
  f4v acc[NA][NA];
#pragma unroll
  for (int a = 0; a < NA; a++)
#pragma unroll
    for (int b = 0; b < NA; b++) acc[a][b] = f4v{0, 0, 0, 0};
  // thread t loads rows 4 (t % 8) .. + 3 of the chunk (two float pairs) of
  // columns t / 8 + 32 q: unconditional loads at clamped indices (K, kb and
  // kc are even, so a pair is wholly inside or outside [kb, ke)), zeroed
  // outside the operands
  constexpr int NC = 2 * TM;  // columns per thread and operand
  const int k4 = 4 * (t & 7), lc = t >> 3;
  float2 ra[NC][2], rb[NC][2];  // the next chunk in flight during the current one's products
  auto load = [&](int k0, float2 (&ra)[NC][2], float2 (&rb)[NC][2]) {
#pragma unroll
    for (int q = 0; q < NC; q++) {
      const int ci = i0 + lc + 32 * q, cj = j0 + lc + 32 * q;
      const float* pa = A + (size_t)min(ci, n1 - 1) * lda;
      const float* pb = B + (size_t)min(cj, n2 - 1) * ldb;
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int rc = min(k0 + k4 + 2 * h, ke - 2);
        ra[q][h] = *reinterpret_cast<const float2*>(pa + rc);
        rb[q][h] = *reinterpret_cast<const float2*>(pb + rc);
      }
    }
  };
  // the zero fill outside the operands at use, not at load (a select on a
  // loaded value makes the compiler wait for the load where it is issued)
  auto mask = [&](int k0, float2 (&ra)[NC][2], float2 (&rb)[NC][2]) {
#pragma unroll
    for (int q = 0; q < NC; q++) {
      const int ci = i0 + lc + 32 * q, cj = j0 + lc + 32 * q;
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int r = k0 + k4 + 2 * h;
        if (!(ci < n1 && r < ke)) ra[q][h] = float2{0.f, 0.f};
        if (!(cj < n2 && r < ke)) rb[q][h] = float2{0.f, 0.f};
      }
    }
  };
  load(kb, ra, rb);
  const int fr = lane & 15, fk = 8 * (lane >> 4);
  // one chunk: its operands split into LDS, the next chunk's loads into the
  // same registers, the products
  auto chunk = [&](int k0, float2 (&ra)[NC][2], float2 (&rb)[NC][2]) {
    mask(k0, ra, rb);
#pragma unroll
    for (int q = 0; q < NC; q++) {
      uint2 ta[NS], tb[NS];
      bf16_split4<NS>(ra[q][0], ra[q][1], ta);
      bf16_split4<NS>(rb[q][0], rb[q][1], tb);
#pragma unroll
      for (int i = 0; i < NS; i++) {
        *reinterpret_cast<uint2*>(As + i * TB * BKP + (lc + 32 * q) * BKP + k4) = ta[i];
        *reinterpret_cast<uint2*>(Bs + i * TB * BKP + (lc + 32 * q) * BKP + k4) = tb[i];
      }
    }
    lds_barrier();
    if (k0 + BK < ke) load(k0 + BK, ra, rb);
    bf16x8 fa[NA][NS], fb[NA][NS];
#pragma unroll
    for (int a = 0; a < NA; a++)
#pragma unroll
      for (int i = 0; i < NS; i++) {
        fa[a][i] = *reinterpret_cast<const bf16x8*>(As + i * TB * BKP + (wi + a * 16 + fr) * BKP + fk);
        fb[a][i] = *reinterpret_cast<const bf16x8*>(Bs + i * TB * BKP + (wj + a * 16 + fr) * BKP + fk);
      }
    // term pair outer, tiles inner: mfma_split's order per tile
#pragma unroll
    for (int p = 6 - (NS == 3 ? 6 : 3); p < 6; p++)
#pragma unroll
      for (int a = 0; a < NA; a++)
#pragma unroll
        for (int b = 0; b < NA; b++)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(fa[a][kSplitPairA[p]], fb[b][kSplitPairB[p]],
                                                               acc[a][b], 0, 0, 0);
    lds_barrier();
  };
  for (int k0 = kb; k0 < ke; k0 += BK) chunk(k0, ra, rb);
  float* Ps = P + (size_t)split * n1 * n2;
#pragma unroll
  for (int a = 0; a < NA; a++)
#pragma unroll
    for (int b = 0; b < NA; b++) {
      const int col = j0 + wj + b * 16 + (lane & 15);
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = i0 + wi + a * 16 + acc_row<float>(lane, r);
        if (row < n1 && col < n2) Ps[row + (size_t)n1 * col] = acc[a][b][r];
      }
    }
}

// C[i + ldc j] = rowscale[i] * sum_s P[s][i + n1 j] (+ beta C), rowscale may
// be null.  The partials are summed in split order (deterministic), eight
// loads in flight at a time.
template <typename T>
__global__ void k_gemm_tn_reduce(const T* __restrict__ P, int nsplit, int n1, int n2, T* __restrict__ C, int ldc,
                                 const T* __restrict__ rowscale, T beta, T* const* __restrict__ Cb, long sP) {
  if (Cb) {  // product blockIdx.y of a batch
    C = Cb[blockIdx.y];
    P += blockIdx.y * sP;
  }
  const long total = (long)n1 * n2;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e % n1), j = (int)(e / n1);
    T s = 0;
    for (int q0 = 0; q0 < nsplit; q0 += 8) {
      T v[8];
#pragma unroll
      for (int u = 0; u < 8; u++) v[u] = (q0 + u < nsplit) ? P[(size_t)(q0 + u) * total + e] : (T)0;
#pragma unroll
      for (int u = 0; u < 8; u++) s += v[u];
    }
    if (rowscale) s *= rowscale[i];
    T* c = C + i + (size_t)ldc * j;
    *c = (beta == (T)0) ? s : s + beta * *c;
  }
}

// Tile origin of block blockIdx.x in the one-row grid of ti x tj BM x BN
// output tiles of the NN products: the tj column tiles of a row tile, which
// read the same rows of A, are consecutive items, and xcd_item puts runs of
// consecutive items on one XCD (one L2).
__device__ __forceinline__ void nn_tile_origin(int m, int n, int& i0, int& j0) {
  const int ti = (m + BM - 1) / BM, tj = (n + BN - 1) / BN;
  const int item = xcd_item(blockIdx.x, ti * tj), i = item / tj;
  i0 = i * BM;
  j0 = (item - i * tj) * BN;
}

// Y = alpha A B + beta Y; A m x k (lda), B k x n (ldb), Y m x n (ldy).
// Pipelined like k_gemm_tn_partial.  A chunk (64 rows x 32 k): load q of
// wave w reads k = 8 w + q, rows lane (one 64-row run per wave); B chunk
// (32 k x 64 columns): column 16 w + 2 q + lane/32, k lane%32.
template <typename T>
__global__ void __launch_bounds__(256)
k_gemm_nn(const T* __restrict__ A, int lda, const T* __restrict__ B, int ldb, T* __restrict__ Y, int ldy, int m,
          int n, int k, T alpha, T beta) {
  typedef typename acc4<T>::t V;
  __shared__ T As[BK][BM + 1];
  __shared__ T Bs[BN][BK + 1];
  int i0, j0;
  nn_tile_origin(m, n, i0, j0);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wi = (w & 1) * 32, wj = (w >> 1) * 32;
  V acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) acc[a][b] = V{0, 0, 0, 0};
  const int bk = lane & 31, bcol = 16 * w + (lane >> 5);
  const int ar = i0 + lane;
  T ra[8], rb[8];
  // unconditional loads at clamped indices, the zero fill at the LDS store
  auto load = [&](int k0) {
    const int arc = min(ar, m - 1);
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int kk = min(k0 + 8 * w + q, k - 1);
      ra[q] = A[(size_t)kk * lda + arc];
      const int cj = min(j0 + bcol + 2 * q, n - 1), kq = min(k0 + bk, k - 1);
      rb[q] = B[(size_t)cj * ldb + kq];
    }
  };
  load(0);
  for (int k0 = 0; k0 < k; k0 += BK) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      As[8 * w + q][lane] = (ar < m && k0 + 8 * w + q < k) ? ra[q] : (T)0;
      Bs[bcol + 2 * q][bk] = (j0 + bcol + 2 * q < n && k0 + bk < k) ? rb[q] : (T)0;
    }
    lds_barrier();
    if (k0 + BK < k) load(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T av[2], bv[2];
#pragma unroll
      for (int a = 0; a < 2; a++) av[a] = As[kk + (lane >> 4)][wi + a * 16 + (lane & 15)];
#pragma unroll
      for (int b = 0; b < 2; b++) bv[b] = Bs[wj + b * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = mfma_nn(av[a], bv[b], acc[a][b]);
    }
    lds_barrier();
  }
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = i0 + wi + a * 16 + nn_row<T>(lane, r), col = j0 + wj + b * 16 + nn_col<T>(lane, r);
        if (row < m && col < n) {
          T* y = Y + row + (size_t)ldy * col;
          const T v = alpha * acc[a][b][r];
          *y = (beta == (T)0) ? v : v + beta * *y;
        }
      }
}

// Epilogue of the projector apply inside H Y (acc = (beta w) W for one
// output element e = row + ld col):
//   kCheb == false: out[e] = H[e] + acc                    (H Y += beta w W)
//   kCheb == true:  out[e] = s1 (H[e] + acc - c Yc[e]) - s2 Yp[e]
// with (c, s1, s2) = coef[3 step ..] and Yp == null meaning s2 = 0: one step
// of the scaled Chebyshev three-term recurrence fused with H Y's last term.
// out may alias H or Yp, not Yc.
// The plain product uses the same kernels with out[e] = alpha acc + beta H[e]
// (H read only when beta != 0; H may alias out).
template <typename T>
struct NnEpilogue {
  const T* H;
  const T* Yc;
  const T* Yp;
  T* out;
  int ld;
  const T* coef;
  int step;
  T alpha, beta;
};

template <typename T, bool kCheb>
__device__ __forceinline__ void nn_epilogue(const NnEpilogue<T>& ep, size_t e, T acc) {
  if constexpr (kCheb) {
    const T h = ep.H[e] + acc;
    const T c = ep.coef[3 * ep.step], s1 = ep.coef[3 * ep.step + 1], s2 = ep.coef[3 * ep.step + 2];
    T v = (h - c * ep.Yc[e]) * s1;
    if (ep.Yp) v -= s2 * ep.Yp[e];
    ep.out[e] = v;
  } else {
    const T v = ep.alpha * acc;
    ep.out[e] = ep.beta == (T)0 ? v : v + ep.beta * ep.H[e];
  }
}

// Y-tile product A (m x k) B (k x n) with FP32 inputs split into NS BF16
// terms, FP32 accumulation, and the epilogue above.  64 x 64 tiles, four
// waves of 32 x 32, 32-deep K chunks: A chunk loads are 64-row runs (lane =
// row, 8 k per thread), B chunk loads 8 contiguous k of one column per
// thread; both are split and stored as k-contiguous BF16 rows so that one
// 16-byte LDS read is one MFMA operand.  The finished tile goes through LDS
// so that the epilogue reads and writes 16 contiguous rows per thread.
template <bool kCheb, int NS>
__global__ void __launch_bounds__(256)
k_gemm_nn_bf(const float* __restrict__ A, int lda, const float* __restrict__ B, int ldb, int m, int n, int k,
             NnEpilogue<float> ep) {
  __shared__ __align__(16) unsigned short Sh[2 * NS * 64 * BKP > 64 * 65 * 2 ? 2 * NS * 64 * BKP : 64 * 65 * 2];
  unsigned short* As = Sh;                 // NS term tiles of A (64 rows x 32 k)
  unsigned short* Bs = Sh + NS * 64 * BKP;  // NS term tiles of B (64 columns x 32 k)
  int i0, j0;
  nn_tile_origin(m, n, i0, j0);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wi = (w & 1) * 32, wj = (w >> 1) * 32;
  f4v acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) acc[a][b] = f4v{0, 0, 0, 0};
  const int arow = i0 + (t & 63), ak = 8 * (t >> 6);
  const int bcol = j0 + (t >> 2), bk = 8 * (t & 3);
  const int fr = lane & 15, fk = 8 * (lane >> 4);
  auto pack = [](const unsigned short* v) {
    return uint4{v[0] | ((unsigned)v[1] << 16), v[2] | ((unsigned)v[3] << 16), v[4] | ((unsigned)v[5] << 16),
                 v[6] | ((unsigned)v[7] << 16)};
  };
  for (int k0 = 0; k0 < k; k0 += BK) {
    unsigned short ta[NS][8], tb[NS][8];
#pragma unroll
    for (int j = 0; j < 8; j++) {
      const int ka = k0 + ak + j, kb2 = k0 + bk + j;
      const float x = (arow < m && ka < k) ? A[(size_t)ka * lda + arow] : 0.f;
      const float y = (bcol < n && kb2 < k) ? B[(size_t)bcol * ldb + kb2] : 0.f;
      unsigned short sx[NS], sy[NS];
      bf16_split<NS>(x, sx);
      bf16_split<NS>(y, sy);
#pragma unroll
      for (int i = 0; i < NS; i++) {
        ta[i][j] = sx[i];
        tb[i][j] = sy[i];
      }
    }
#pragma unroll
    for (int i = 0; i < NS; i++) {
      *reinterpret_cast<uint4*>(As + i * 64 * BKP + (t & 63) * BKP + ak) = pack(ta[i]);
      *reinterpret_cast<uint4*>(Bs + i * 64 * BKP + (t >> 2) * BKP + bk) = pack(tb[i]);
    }
    __syncthreads();
#pragma unroll
    for (int a = 0; a < 2; a++)
#pragma unroll
      for (int b = 0; b < 2; b++) {
        const int oa = (wi + a * 16 + fr) * BKP + fk, ob = (wj + b * 16 + fr) * BKP + fk;
        const unsigned short* pa[NS];
        const unsigned short* pb[NS];
#pragma unroll
        for (int i = 0; i < NS; i++) {
          pa[i] = As + i * 64 * BKP + oa;
          pb[i] = Bs + i * 64 * BKP + ob;
        }
        acc[a][b] = mfma_split<NS>(pa, pb, acc[a][b]);
      }
    __syncthreads();
  }
  // tile (64 x 64 floats, column stride 65) through LDS
  float* Ct = reinterpret_cast<float*>(Sh);
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int i = 0; i < 4; i++)
        Ct[(wj + b * 16 + (lane & 15)) * 65 + wi + a * 16 + acc_row<float>(lane, i)] = acc[a][b][i];
  __syncthreads();
  const int col = j0 + (t >> 2), r0 = 16 * (t & 3);
  if (col < n) {
    for (int rr = 0; rr < 16; rr++) {
      const int row = i0 + r0 + rr;
      if (row < m) nn_epilogue<float, kCheb>(ep, (size_t)col * ep.ld + row, Ct[(t >> 2) * 65 + r0 + rr]);
    }
  }
}

// ---- products with operands split ahead of time ----
// Operand tile layout (both operands): for a 16-row (or 16-column) tile r
// and a 32-deep K chunk q, the NS_MAX = 3 term blocks of 64 lanes x 8 BF16
// (one 16-byte v_mfma_f32_16x16x32_bf16 operand per lane: row / column
// l % 16, k = 32 q + 8 (l / 16) .. + 7) sit at ((r nkc + q) 3 + i) 512
// BF16 elements, so that a wave loads an operand with one 1 KiB read.  Rows,
// columns and k beyond the matrix are zero.  The 2-term split is the first
// two terms of the 3-term one.
constexpr int kSplitTerms = 3;
typedef unsigned short u16x8 __attribute__((ext_vector_type(8)));

// One operand group (r, q) per lane-block: A column-major (ld lda): rows
// of A are the tile dimension (transpose = false), or columns of B
// (transpose = true, B column-major with ld ldb: B(k, col) at col ldb + k).
template <bool kCols>
__global__ void k_split_operand(const float* __restrict__ X, int ldx, int nr, int K, int nkc,
                                unsigned short* __restrict__ out, long ngroups) {
  const long g = (long)blockIdx.x * (blockDim.x / 64) + (threadIdx.x >> 6);
  if (g >= ngroups) return;
  const int lane = threadIdx.x & 63;
  const int r = (int)(g / nkc), q = (int)(g - (long)r * nkc);
  const int idx = 16 * r + (lane & 15), k0 = 32 * q + 8 * (lane >> 4);
  unsigned short t[kSplitTerms][8];
#pragma unroll
  for (int j = 0; j < 8; j++) {
    const int k = k0 + j;
    float x = 0.f;
    if (idx < nr && k < K) x = kCols ? X[(size_t)idx * ldx + k] : X[(size_t)k * ldx + idx];
    unsigned short v[kSplitTerms];
    bf16_split<kSplitTerms>(x, v);
#pragma unroll
    for (int i = 0; i < kSplitTerms; i++) t[i][j] = v[i];
  }
#pragma unroll
  for (int i = 0; i < kSplitTerms; i++) {
    u16x8 v;
#pragma unroll
    for (int j = 0; j < 8; j++) v[j] = t[i][j];
    *reinterpret_cast<u16x8*>(out + ((size_t)g * kSplitTerms + i) * 512 + 8 * lane) = v;
  }
}


// out = epilogue(A B) with both operands split ahead of time (layout above;
// A: nrt = ceil(m / 16) row tiles, B: ceil(n / 16) column tiles, nkc
// chunks).  A workgroup takes 64 rows x 256 columns: wave w the 64 x 64
// block of columns 64 w .., as 4 x 4 MFMA tiles whose operands come
// straight from global memory (A from HBM, shared by the four waves through
// the caches; B from L2), the next chunk's operands in flight during the
// current chunk's products.  (Staging the A slab in LDS measured slower:
// 42 against 38 us at 13134 x 256 x 256.)  The epilogue works on the
// accumulator layout: a lane's four rows of one column are contiguous, read
// and written as two float pairs (m and ld even).
template <bool kCheb, int NS, int RT = 4>
__global__ void __launch_bounds__(256, 1)
k_gemm_nn_pre(const unsigned short* __restrict__ Ab, const unsigned short* __restrict__ Bb, int m, int n, int nkc,
              NnEpilogue<float> ep) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int nrt = (m + 15) >> 4, nct = (n + 15) >> 4;
  const int rt0 = RT * blockIdx.x, ct0 = 16 * blockIdx.y + 4 * w;
  const int na = min(RT, nrt - rt0), nbt = min(4, nct - ct0);
  if (nbt <= 0) return;
  const bf16x8* A8 = reinterpret_cast<const bf16x8*>(Ab);
  const bf16x8* B8 = reinterpret_cast<const bf16x8*>(Bb);
  f4v acc[RT][4];
#pragma unroll
  for (int a = 0; a < RT; a++)
#pragma unroll
    for (int b = 0; b < 4; b++) acc[a][b] = f4v{0, 0, 0, 0};
  bf16x8 av[RT][NS], bv[4][NS];
  auto load = [&](int q, bf16x8 (&xa)[RT][NS], bf16x8 (&xb)[4][NS]) {
#pragma unroll
    for (int a = 0; a < RT; a++)
#pragma unroll
      for (int i = 0; i < NS; i++)
        xa[a][i] = a < na ? A8[((size_t)((rt0 + a) * nkc + q) * kSplitTerms + i) * 64 + lane] : bf16x8{};
#pragma unroll
    for (int b = 0; b < 4; b++)
#pragma unroll
      for (int i = 0; i < NS; i++)
        xb[b][i] = b < nbt ? B8[((size_t)((ct0 + b) * nkc + q) * kSplitTerms + i) * 64 + lane] : bf16x8{};
  };
  load(0, av, bv);
  for (int q = 0; q < nkc; q++) {
    bf16x8 an[RT][NS], bn[4][NS];
    if (q + 1 < nkc) load(q + 1, an, bn);
    // term pair outer, tiles inner: the order of mfma_split per tile, while
    // consecutive MFMAs write different accumulators
#pragma unroll
    for (int p = 6 - (NS == 3 ? 6 : 3); p < 6; p++)
#pragma unroll
      for (int a = 0; a < RT; a++)
#pragma unroll
        for (int b = 0; b < 4; b++)
          acc[a][b] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[a][kSplitPairA[p]], bv[b][kSplitPairB[p]],
                                                               acc[a][b], 0, 0, 0);
    if (q + 1 < nkc) {
#pragma unroll
      for (int a = 0; a < 4; a++)
#pragma unroll
        for (int i = 0; i < NS; i++) {
          if (a < RT) av[a][i] = an[a][i];
          bv[a][i] = bn[a][i];
        }
    }
  }
  float cc = 0.f, s1 = 0.f, s2 = 0.f;
  if constexpr (kCheb) {
    cc = ep.coef[3 * ep.step];
    s1 = ep.coef[3 * ep.step + 1];
    s2 = ep.coef[3 * ep.step + 2];
  }
#pragma unroll
  for (int b = 0; b < 4; b++) {
    const int col = 16 * (ct0 + b) + (lane & 15);
    if (b >= nbt || col >= n) continue;
#pragma unroll
    for (int a = 0; a < RT; a++) {
      const int r0 = 16 * (rt0 + a) + 4 * (lane >> 4);
      if (a >= na) continue;
#pragma unroll
      for (int h = 0; h < 2; h++) {
        const int row = r0 + 2 * h;
        if (row >= m) continue;
        const size_t e = (size_t)col * ep.ld + row;
        const float x0 = acc[a][b][2 * h], x1 = acc[a][b][2 * h + 1];
        float2 o;
        if constexpr (kCheb) {
          const float2 hv = *reinterpret_cast<const float2*>(ep.H + e);
          const float2 yc = *reinterpret_cast<const float2*>(ep.Yc + e);
          o.x = (hv.x + x0 - cc * yc.x) * s1;
          o.y = (hv.y + x1 - cc * yc.y) * s1;
          if (ep.Yp) {
            const float2 yp = *reinterpret_cast<const float2*>(ep.Yp + e);
            o.x -= s2 * yp.x;
            o.y -= s2 * yp.y;
          }
        } else {
          o.x = ep.alpha * x0;
          o.y = ep.alpha * x1;
          if (ep.beta != 0.f) {
            const float2 hv = *reinterpret_cast<const float2*>(ep.H + e);
            o.x += ep.beta * hv.x;
            o.y += ep.beta * hv.y;
          }
        }
        *reinterpret_cast<float2*>(ep.out + e) = o;
      }
    }
  }
}

// The same product on v_mfma_f64_16x16x4f64 / v_mfma_f32_16x16x4f32
// (pipelined like k_gemm_nn), with the same epilogue.
template <typename T, bool kCheb>
__global__ void __launch_bounds__(256)
k_gemm_nn_ep(const T* __restrict__ A, int lda, const T* __restrict__ B, int ldb, int m, int n, int k,
             NnEpilogue<T> ep) {
  typedef typename acc4<T>::t V;
  __shared__ T As[BK][BM + 1];
  __shared__ T Bs[BN][BK + 1];
  int i0, j0;
  nn_tile_origin(m, n, i0, j0);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int wi = (w & 1) * 32, wj = (w >> 1) * 32;
  V acc[2][2];
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) acc[a][b] = V{0, 0, 0, 0};
  const int bk = lane & 31, bcol = 16 * w + (lane >> 5);
  const int ar = i0 + lane;
  T ra[8], rb[8];
  // unconditional loads at clamped indices, the zero fill at the LDS store
  auto load = [&](int k0) {
    const int arc = min(ar, m - 1);
#pragma unroll
    for (int q = 0; q < 8; q++) {
      const int kk = min(k0 + 8 * w + q, k - 1);
      ra[q] = A[(size_t)kk * lda + arc];
      const int cj = min(j0 + bcol + 2 * q, n - 1), kq = min(k0 + bk, k - 1);
      rb[q] = B[(size_t)cj * ldb + kq];
    }
  };
  load(0);
  for (int k0 = 0; k0 < k; k0 += BK) {
#pragma unroll
    for (int q = 0; q < 8; q++) {
      As[8 * w + q][lane] = (ar < m && k0 + 8 * w + q < k) ? ra[q] : (T)0;
      Bs[bcol + 2 * q][bk] = (j0 + bcol + 2 * q < n && k0 + bk < k) ? rb[q] : (T)0;
    }
    lds_barrier();
    if (k0 + BK < k) load(k0 + BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T av[2], bv[2];
#pragma unroll
      for (int a = 0; a < 2; a++) av[a] = As[kk + (lane >> 4)][wi + a * 16 + (lane & 15)];
#pragma unroll
      for (int b = 0; b < 2; b++) bv[b] = Bs[wj + b * 16 + (lane & 15)][kk + (lane >> 4)];
#pragma unroll
      for (int a = 0; a < 2; a++)
#pragma unroll
        for (int b = 0; b < 2; b++) acc[a][b] = mfma_nn(av[a], bv[b], acc[a][b]);
    }
    lds_barrier();
  }
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = i0 + wi + a * 16 + nn_row<T>(lane, r), col = j0 + wj + b * 16 + nn_col<T>(lane, r);
        if (row < m && col < n) nn_epilogue<T, kCheb>(ep, (size_t)col * ep.ld + row, acc[a][b][r]);
      }
}

// Elementwise form of the epilogue (no projectors: acc = 0).
template <typename T, bool kCheb>
__global__ void k_nn_epilogue_only(long m, int n, NnEpilogue<T> ep) {
  const long total = m * n;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const long col = e / m, row = e - col * m;
    nn_epilogue<T, kCheb>(ep, (size_t)col * ep.ld + row, (T)0);
  }
}

// Single-column products (the one-band H x of Lanczos): GEMV kernels.
// C(n1) = A^T b over K: workgroup (i, split) sums column i of A against b
// over one K range, 256 threads with 8 loads in flight each; partials go
// through k_gemm_tn_reduce like the split-K GEMM.
template <typename T>
__global__ void __launch_bounds__(256)
k_gemv_t_partial(const T* __restrict__ A, int lda, const T* __restrict__ b, int K, int kc, int n1,
                 T* __restrict__ P) {
  const int i = blockIdx.x, split = blockIdx.y;
  const int kb = split * kc, ke = min(K, kb + kc);
  const T* a = A + (size_t)i * lda;
  T s = 0;
  for (int r0 = kb + threadIdx.x; r0 < ke; r0 += 8 * 256) {
    T x[8], y[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int r = r0 + 256 * u;
      x[u] = r < ke ? a[r] : (T)0;
      y[u] = r < ke ? b[r] : (T)0;
    }
#pragma unroll
    for (int u = 0; u < 8; u++) s += x[u] * y[u];
  }
  __shared__ T red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) P[(size_t)split * n1 + i] = red[0];
}

// y = alpha A x + beta y, A m x k: workgroup of NW waves per 64 rows; wave
// w sums columns j = w (mod NW) for its 64 rows (16 loads in flight per
// lane), the NW partial rows meet in LDS in a fixed order.  NW = 16 for
// k >= 128: one round of loads per wave, four times the waves in flight
// (the single-column products of the Lanczos steps are latency-bound).
template <typename T, int NW>
__global__ void __launch_bounds__(64 * NW)
k_gemv_n(const T* __restrict__ A, int lda, const T* __restrict__ x, T* __restrict__ y, int m, int k, T alpha,
         T beta) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int row = blockIdx.x * 64 + lane;
  T s = 0;
  if (row < m) {
    for (int j0 = w; j0 < k; j0 += 16 * NW) {
      T a[16];
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const int j = j0 + NW * u;
        a[u] = j < k ? A[(size_t)j * lda + row] : (T)0;
      }
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const int j = j0 + NW * u;
        if (j < k) s += a[u] * x[j];
      }
    }
  }
  __shared__ T red[NW][64];
  red[w][lane] = s;
  __syncthreads();
  if (w == 0 && row < m) {
    T v = 0;
#pragma unroll
    for (int q = 0; q < NW; q++) v += red[q][lane];
    v *= alpha;
    y[row] = (beta == (T)0) ? v : v + beta * y[row];
  }
}

template <typename T>
static void launch_gemv_n(pw_ctx* c, const T* A, int lda, const T* x, T* y, int m, int k, T alpha, T beta) {
  if (k >= 128)
    hipLaunchKernelGGL((k_gemv_n<T, 16>), dim3((m + 63) / 64), dim3(1024), 0, c->stream, A, lda, x, y, m, k, alpha,
                       beta);
  else
    hipLaunchKernelGGL((k_gemv_n<T, 4>), dim3((m + 63) / 64), dim3(256), 0, c->stream, A, lda, x, y, m, k, alpha,
                       beta);
}

// The context's split-K partial buffer, grown to need bytes (allocates).
static int grow_gemm_ws(pw_ctx* c, size_t need) {
  if (need <= c->gemm_ws_bytes) return PW_OK;
  invalidate_graph(c);
  if (c->d_gemm_ws) (void)hipFree(c->d_gemm_ws);
  c->d_gemm_ws = nullptr;
  c->gemm_ws_bytes = 0;
  PW_HIP(hipMalloc(&c->d_gemm_ws, need));
  c->gemm_ws_bytes = need;
  return PW_OK;
}

// ws == null: the partials go to the context's buffer (grown on demand);
// else to ws (ws_bytes long), with fewer K splits when the usual count does
// not fit and PW_ERR_UNSUPPORTED when not even one split does.
template <typename T>
int gemm_tn_ws(pw_ctx* c, int K, int n1, int n2, const T* A, int lda, const T* B, int ldb, T* C, int ldc,
               const T* rowscale, T beta, int split, void* ws, size_t ws_bytes) {
  if (n1 <= 0 || n2 <= 0) return PW_OK;
  if (n2 == 1) {
    int nsplit = std::max(1, std::min(64, (K + 2047) / 2048));
    if (ws) nsplit = (int)std::min<size_t>(nsplit, ws_bytes / (sizeof(T) * (size_t)n1));
    if (nsplit < 1) return fail(PW_ERR_UNSUPPORTED, "projection of %d columns exceeds the %zu-byte scratch", n1, ws_bytes);
    const int kc = (K + nsplit - 1) / nsplit;
    const size_t need = sizeof(T) * (size_t)nsplit * n1;
    if (!ws) PW_TRY(grow_gemm_ws(c, need));
    T* P = (T*)(ws ? ws : c->d_gemm_ws);
    hipLaunchKernelGGL(k_gemv_t_partial<T>, dim3(n1, nsplit), dim3(256), 0, c->stream, A, lda, B, K, kc, n1, P);
    const int g = std::max(1, std::min((n1 + 255) / 256, 4096));
    hipLaunchKernelGGL(k_gemm_tn_reduce<T>, dim3(g), dim3(256), 0, c->stream, (const T*)P, nsplit, n1, 1, C, ldc,
                       rowscale, beta, (T* const*)nullptr, 0L);
    PW_HIP(hipGetLastError());
    return PW_OK;
  }
  // outputs of 256 or more 64 x 64 tiles (1024 x 1024 and up) take
  // 128 x 128 tiles
  const pw_ctx::Plan& pl = c->plan[sizeof(T) == 8 ? PW_F64 : PW_F32];
  const int tm = pl.tn_tm > 0 ? pl.tn_tm : (((n1 + BM - 1) / BM) * ((n2 + BN - 1) / BN) >= 256 ? 2 : 1);
  const int ti = (n1 + BM * tm - 1) / (BM * tm), tj = (n2 + BN * tm - 1) / (BN * tm);
  const int tiles = ti * tj;
  // about 1024 partial workgroups, but at most 256 K-splits for one or two
  // output tiles (the sb x sb Gram blocks of PPCG over a long K): the
  // reduction then reads 256 partials per entry, not ~1000
  const int cap = pl.tn_cap > 0 ? pl.tn_cap : (tiles <= 2 ? 256 : 1024);
  int nsplit = std::max(1, std::min((cap + tiles - 1) / tiles, (K + 255) / 256));
  const size_t per_split = sizeof(T) * (size_t)n1 * n2;
  if (ws) {
    if (per_split > ws_bytes)
      return fail(PW_ERR_UNSUPPORTED, "a %d x %d projection exceeds the %zu-byte scratch", n1, n2, ws_bytes);
    nsplit = (int)std::min<size_t>(nsplit, ws_bytes / per_split);
  }
  int kc = (K + nsplit - 1) / nsplit;
  kc = (kc + BK - 1) / BK * BK;
  nsplit = (K + kc - 1) / kc;
  const size_t need = per_split * nsplit;
  if (!ws) PW_TRY(grow_gemm_ws(c, need));
  T* P = (T*)(ws ? ws : c->d_gemm_ws);
  if constexpr (sizeof(T) == 4) {
    // the split kernels load float pairs: K, lda and ldb even
    if ((split == 2 || split == 3) && !(K & 1) && !(lda & 1) && !(ldb & 1)) {
      const size_t ldsb = sizeof(unsigned short) * 2 * split * (size_t)(BM * tm) * BKP;
      const dim3 gp(ti * tj * nsplit);
      if (split == 3 && tm == 2) {
        PW_TRY(ensure_lds((const void*)k_gemm_tn_partial_bf<2, 3>, ldsb));
        hipLaunchKernelGGL((k_gemm_tn_partial_bf<2, 3>), gp, dim3(256), ldsb, c->stream, A, lda, B, ldb, K, n1, n2,
                           kc, P);
      } else if (split == 3) {
        hipLaunchKernelGGL((k_gemm_tn_partial_bf<1, 3>), gp, dim3(256), ldsb, c->stream, A, lda, B, ldb, K, n1, n2,
                           kc, P);
      } else if (tm == 2) {
        PW_TRY(ensure_lds((const void*)k_gemm_tn_partial_bf<2, 2>, ldsb));
        hipLaunchKernelGGL((k_gemm_tn_partial_bf<2, 2>), gp, dim3(256), ldsb, c->stream, A, lda, B, ldb, K, n1, n2,
                           kc, P);
      } else {
        hipLaunchKernelGGL((k_gemm_tn_partial_bf<1, 2>), gp, dim3(256), ldsb, c->stream, A, lda, B, ldb, K, n1, n2,
                           kc, P);
      }
      const long total = (long)n1 * n2;
      const int g = (int)std::min<long>((total + 255) / 256, 4096);
      hipLaunchKernelGGL(k_gemm_tn_reduce<T>, dim3(g), dim3(256), 0, c->stream, (const T*)P, nsplit, n1, n2, C, ldc,
                         rowscale, beta, (T* const*)nullptr, 0L);
      PW_HIP(hipGetLastError());
      return PW_OK;
    }
  }
  const size_t lds = sizeof(T) * 2 * (size_t)(BM * tm) * (BK + 1);
  if (tm == 2) {
    PW_TRY(ensure_lds((const void*)k_gemm_tn_partial<T, 2>, lds));
    hipLaunchKernelGGL((k_gemm_tn_partial<T, 2>), dim3(ti * tj * nsplit), dim3(256), lds, c->stream, A, lda, B, ldb,
                       K, n1, n2, kc, P, (const T* const*)nullptr, (const T* const*)nullptr, 0L);
  } else {
    hipLaunchKernelGGL((k_gemm_tn_partial<T, 1>), dim3(ti * tj * nsplit), dim3(256), lds, c->stream, A, lda, B, ldb,
                       K, n1, n2, kc, P, (const T* const*)nullptr, (const T* const*)nullptr, 0L);
  }
  const long total = (long)n1 * n2;
  const int g = (int)std::min<long>((total + 255) / 256, 4096);
  hipLaunchKernelGGL(k_gemm_tn_reduce<T>, dim3(g), dim3(256), 0, c->stream, (const T*)P, nsplit, n1, n2, C, ldc,
                     rowscale, beta, (T* const*)nullptr, 0L);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

template <typename T>
int gemm_tn(pw_ctx* c, int K, int n1, int n2, const T* A, int lda, const T* B, int ldb, T* C, int ldc,
            const T* rowscale, T beta, int split) {
  return gemm_tn_ws<T>(c, K, n1, n2, A, lda, B, ldb, C, ldc, rowscale, beta, split, nullptr, 0);
}

template <typename T>
int gemm_nn(pw_ctx* c, int m, int n, int k, T alpha, const T* A, int lda, const T* B, int ldb, T beta, T* Y,
            int ldy) {
  if (m <= 0 || n <= 0) return PW_OK;
  if (n == 1) {
    launch_gemv_n<T>(c, A, lda, B, Y, m, k, alpha, beta);
    PW_HIP(hipGetLastError());
    return PW_OK;
  }
  hipLaunchKernelGGL(k_gemm_nn<T>, dim3(((m + BM - 1) / BM) * ((n + BN - 1) / BN)), dim3(256), 0, c->stream, A, lda, B,
                     ldb, Y, ldy, m, n, k, alpha, beta);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

// out (m x n) = epilogue(H, A B) for the projector apply of H Y: FP32 on
// split-BF16 MFMA (FP32 MFMA when the context's split products are off), FP64
// on FP64 MFMA; A = beta w (m x k), B = W (k x n).
template <typename T>
int gemm_nn_ep(pw_ctx* c, int m, int n, int k, const T* A, int lda, const T* B, int ldb, const T* H, const T* Yc,
               const T* Yp, T* out, int ld, const T* coef, int step, bool cheb, int split) {
  if (m <= 0 || n <= 0) return PW_OK;
  NnEpilogue<T> ep{H, Yc, Yp, out, ld, coef, step, (T)1, (T)1};
  const dim3 grid(((m + BM - 1) / BM) * ((n + BN - 1) / BN));  // nn_tile_origin's order
  bool done = false;
  if (!cheb && n == 1 && k > 0 && H == out) {
    // one column (the Lanczos vector): out += A b
    launch_gemv_n<T>(c, A, lda, B, out, m, k, (T)1, (T)1);
    done = true;
  }
  if (k <= 0) {
    const int g = (int)std::min<long>(((long)m * n + 255) / 256, 8192);
    if (cheb)
      hipLaunchKernelGGL((k_nn_epilogue_only<T, true>), dim3(g), dim3(256), 0, c->stream, (long)m, n, ep);
    else
      hipLaunchKernelGGL((k_nn_epilogue_only<T, false>), dim3(g), dim3(256), 0, c->stream, (long)m, n, ep);
    done = true;
  }
  if constexpr (sizeof(T) == 4) {
    if (!done && (split == 2 || split == 3)) {
      if (cheb && split == 3)
        hipLaunchKernelGGL((k_gemm_nn_bf<true, 3>), grid, dim3(256), 0, c->stream, A, lda, B, ldb, m, n, k, ep);
      else if (split == 3)
        hipLaunchKernelGGL((k_gemm_nn_bf<false, 3>), grid, dim3(256), 0, c->stream, A, lda, B, ldb, m, n, k, ep);
      else if (cheb)
        hipLaunchKernelGGL((k_gemm_nn_bf<true, 2>), grid, dim3(256), 0, c->stream, A, lda, B, ldb, m, n, k, ep);
      else
        hipLaunchKernelGGL((k_gemm_nn_bf<false, 2>), grid, dim3(256), 0, c->stream, A, lda, B, ldb, m, n, k, ep);
      done = true;
    }
  }
  if (!done) {
    if (cheb)
      hipLaunchKernelGGL((k_gemm_nn_ep<T, true>), grid, dim3(256), 0, c->stream, A, lda, B, ldb, m, n, k, ep);
    else
      hipLaunchKernelGGL((k_gemm_nn_ep<T, false>), grid, dim3(256), 0, c->stream, A, lda, B, ldb, m, n, k, ep);
  }
  PW_HIP(hipGetLastError());
  return PW_OK;
}

// A (m x k, column-major, ld lda) split ahead of time into *out (the tile
// layout of k_gemm_nn_pre; allocated here, freed by the caller).
int split_a_operand(pw_ctx* c, const float* A, int lda, int m, int k, unsigned short** out) {
  *out = nullptr;
  if (m <= 0 || k <= 0) return PW_OK;
  const int nrt = (m + 15) / 16, nkc = (k + 31) / 32;
  const long ng = (long)nrt * nkc;
  PW_HIP(hipMalloc((void**)out, sizeof(unsigned short) * (size_t)ng * kSplitTerms * 512));
  hipLaunchKernelGGL(k_split_operand<false>, dim3((unsigned)((ng + 3) / 4)), dim3(256), 0, c->stream, A, lda, m, k,
                     nkc, *out, ng);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

size_t split_b_bytes(int n, int k) {
  return sizeof(unsigned short) * (size_t)((n + 15) / 16) * ((k + 31) / 32) * kSplitTerms * 512;
}

// out (m x n) = epilogue(A B) on k_gemm_nn_pre: Ab = A split by
// split_a_operand, B (k x n, ld ldb) split here into Bws (split_b_bytes(n,
// k) bytes); NS = split terms (2 or 3).  m and ld even, H / Yc / Yp / out
// 8-byte aligned.
int gemm_nn_pre(pw_ctx* c, int m, int n, int k, const unsigned short* Ab, const float* B, int ldb,
                unsigned short* Bws, const float* H, const float* Yc, const float* Yp, float* out, int ld,
                const float* coef, int step, bool cheb, int split) {
  if (m <= 0 || n <= 0) return PW_OK;
  if ((m & 1) || (ld & 1)) return fail(PW_ERR_ARG, "gemm_nn_pre: odd row count or leading dimension");
  const int nkc = (k + 31) / 32, nct = (n + 15) / 16, nrt = (m + 15) / 16;
  const long ng = (long)nct * nkc;
  hipLaunchKernelGGL(k_split_operand<true>, dim3((unsigned)((ng + 3) / 4)), dim3(256), 0, c->stream, B, ldb, n, k,
                     nkc, Bws, ng);
  NnEpilogue<float> ep{H, Yc, Yp, out, ld, coef, step, 1.f, 1.f};
  // 16 RT rows per workgroup: RT = 2 doubles the grid of RT = 4 (206 -> 411
  // workgroups at 13134 rows x 256 columns) at fewer registers per wave, for
  // grids that would give fewer than two workgroups per CU; larger ones keep
  // RT = 4 (twice the reuse of each B fragment)
  const char* rte = std::getenv("PW_NNPRE_RT");
  int ncu = 256;
  PW_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, c->device));
  const long g4 = (long)((nrt + 3) / 4) * ((nct + 15) / 16);
  const int rt = (rte && std::atoi(rte) == 4) ? 4 : (rte && std::atoi(rte) == 2) ? 2 : (g4 < 2L * ncu ? 2 : 4);
  const dim3 grid((nrt + rt - 1) / rt, (nct + 15) / 16);
  auto kern = rt == 4 ? (cheb ? (split == 3 ? k_gemm_nn_pre<true, 3, 4> : k_gemm_nn_pre<true, 2, 4>)
                              : (split == 3 ? k_gemm_nn_pre<false, 3, 4> : k_gemm_nn_pre<false, 2, 4>))
                      : (cheb ? (split == 3 ? k_gemm_nn_pre<true, 3, 2> : k_gemm_nn_pre<true, 2, 2>)
                              : (split == 3 ? k_gemm_nn_pre<false, 3, 2> : k_gemm_nn_pre<false, 2, 2>));
  hipLaunchKernelGGL(kern, grid, dim3(256), 0, c->stream, Ab, Bws, m, n, nkc, ep);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

template int gemm_nn_ep<double>(pw_ctx*, int, int, int, const double*, int, const double*, int, const double*,
                                const double*, const double*, double*, int, const double*, int, bool, int);
template int gemm_nn_ep<float>(pw_ctx*, int, int, int, const float*, int, const float*, int, const float*,
                               const float*, const float*, float*, int, const float*, int, bool, int);
// nbatch products C_b = A_b^T B_b (n1 x n2, K rows; A_b, B_b, C_b from
// device pointer tables dA, dB, dC, one leading dimension each): one
// launch of split-K partial tiles for all products and one fixed-order
// reduction, each about 2048 workgroups in all.
template <typename T>
int gemm_tn_batched(pw_ctx* c, int K, int n1, int n2, const T* const* dA, int lda, const T* const* dB, int ldb,
                    T* const* dC, int ldc, int nbatch) {
  if (n1 <= 0 || n2 <= 0 || nbatch <= 0) return PW_OK;
  const int ti = (n1 + BM - 1) / BM, tj = (n2 + BN - 1) / BN, tiles = ti * tj;
  int nsplit = std::max(1, std::min((2048 + tiles * nbatch - 1) / (tiles * nbatch), (K + 255) / 256));
  int kc = (K + nsplit - 1) / nsplit;
  kc = (kc + BK - 1) / BK * BK;
  nsplit = (K + kc - 1) / kc;
  const long per = (long)n1 * n2 * nsplit;  // partials of one product
  PW_TRY(grow_gemm_ws(c, sizeof(T) * (size_t)per * nbatch));
  T* P = (T*)c->d_gemm_ws;
  const size_t lds = sizeof(T) * 2 * (size_t)BM * (BK + 1);
  hipLaunchKernelGGL((k_gemm_tn_partial<T, 1>), dim3(tiles * nsplit, nbatch), dim3(256), lds, c->stream,
                     (const T*)nullptr, lda, (const T*)nullptr, ldb, K, n1, n2, kc, P, dA, dB, per);
  const long total = (long)n1 * n2;
  const int g = (int)std::max<long>(1, std::min<long>((total + 255) / 256, 4096));
  hipLaunchKernelGGL(k_gemm_tn_reduce<T>, dim3(g, nbatch), dim3(256), 0, c->stream, (const T*)P, nsplit, n1, n2,
                     (T*)nullptr, ldc, (const T*)nullptr, (T)0, dC, per);
  PW_HIP(hipGetLastError());
  return PW_OK;
}
template int gemm_tn_batched<double>(pw_ctx*, int, int, int, const double* const*, int, const double* const*, int,
                                     double* const*, int, int);

template int gemm_tn_ws<double>(pw_ctx*, int, int, int, const double*, int, const double*, int, double*, int,
                                const double*, double, int, void*, size_t);
template int gemm_tn_ws<float>(pw_ctx*, int, int, int, const float*, int, const float*, int, float*, int,
                               const float*, float, int, void*, size_t);
template int gemm_tn<double>(pw_ctx*, int, int, int, const double*, int, const double*, int, double*, int,
                             const double*, double, int);
template int gemm_tn<float>(pw_ctx*, int, int, int, const float*, int, const float*, int, float*, int,
                            const float*, float, int);
template int gemm_nn<double>(pw_ctx*, int, int, int, double, const double*, int, const double*, int, double,
                             double*, int);
template int gemm_nn<float>(pw_ctx*, int, int, int, float, const float*, int, const float*, int, float, float*,
                            int);

}  // namespace pw
