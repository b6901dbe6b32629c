// Cholesky factor S = U^T U of the overlap matrix and its inverse W = U^{-1}
// for the orthogonalisation X <- X U^{-1} (eigensolver.cpp:386-469: Potrf
// 'U' then Trsm 'R','U','N','N'; here X U^{-1} = X W goes through the MFMA
// GEMM).  n <= 256, one 512-thread workgroup, no library launches.
//
// The strictly upper blocks of S (32 x 32, at most 28) stay in the MFMA
// accumulator registers of wavefronts 1..7 for the whole factorisation
// (block b, numbered column by column, belongs to wave 1 + b % 7); the
// diagonal blocks live in LDS; wave 0 holds no blocks and spends its
// registers on the diagonal factorisations.  Per block column P
// (right-looking, LAPACK dpotrf's blocked order):
//   wave 0 factors the diagonal block in LDS (unblocked Cholesky, then the
//   inverse of its 32 x 32 factor by back-substitution),
//   the block row becomes U_PJ = U_PP^{-T} A_PJ (MFMA, A_PJ staged in LDS),
//   the trailing blocks take A_IJ -= U_PI^T U_PJ (MFMA from the LDS panel;
//   diagonal block I by wave 1 + I % 7, through LDS).
// The two branches meet at the same three barriers per block column.
// W by block rows from the last, all eight waves: W_II = U_II^{-1},
//   W_IJ = -U_II^{-1} sum_{K=I+1..J} U_IK W_KJ.
// A non-positive pivot stops with info = its 1-based column (LAPACK's info).
#include "pw_internal.h"

namespace pw {

typedef double d4v __attribute__((ext_vector_type(4)));

constexpr int kChB = 32;                                   // block size
constexpr int kChMaxNB = 8;                                // n <= 256
constexpr int kChWaves = 8;                                // 512 threads
constexpr int kChOwners = kChWaves - 1;                    // waves 1..7 hold the blocks
constexpr int kChSlots = (kChMaxNB * (kChMaxNB - 1) / 2 + kChOwners - 1) / kChOwners;  // 4 blocks per wave
constexpr int kChLdp = kChB * kChMaxNB + 1;                // row stride of the LDS panel
constexpr int kChLdd = kChB + 1;                           // row stride of a 32 x 32 LDS block

__device__ __forceinline__ d4v ch_mfma(double a, double b, d4v c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// block b of the strictly upper block triangle, numbered column by column
// (b = J (J - 1) / 2 + I, I < J)
__device__ __forceinline__ void ch_block(int b, int& I, int& J) {
  int j = 1;
  while ((j + 1) * j / 2 <= b) j++;
  J = j;
  I = b - j * (j - 1) / 2;
}

// Accumulator layout of v_mfma_f64_16x16x4f64: tile q = ti + 2 tj of a
// 32 x 32 block, lane l, register i holds element (16 ti + 4 i + l / 16,
// 16 tj + l % 16).  Row-major LDS copies of a block:
__device__ __forceinline__ void ch_to_lds(const d4v (&a)[4], double* M, int ldm, int lane) {
#pragma unroll
  for (int q = 0; q < 4; q++)
#pragma unroll
    for (int i = 0; i < 4; i++) M[(16 * (q & 1) + 4 * i + (lane >> 4)) * ldm + 16 * (q >> 1) + (lane & 15)] = a[q][i];
}
__device__ __forceinline__ void ch_from_lds(d4v (&a)[4], const double* M, int ldm, int lane) {
#pragma unroll
  for (int q = 0; q < 4; q++)
#pragma unroll
    for (int i = 0; i < 4; i++) a[q][i] = M[(16 * (q & 1) + 4 * i + (lane >> 4)) * ldm + 16 * (q >> 1) + (lane & 15)];
}

// S (n x n, upper triangle read, ld) -> U in its upper triangle; W (n x n,
// ldw) <- U^{-1} (zeros below the diagonal); *info = 0 or the failing column.
__global__ void __launch_bounds__(512)
k_potrf_inv(double* __restrict__ S, int n, int ld, double* __restrict__ W, int ldw, int* __restrict__ info) {
  extern __shared__ double sm[];
  double* Pn = sm;                                // 32 x kChLdp: block row of U (global columns)
  double* Dm = Pn + kChB * kChLdp;                // 32 x kChLdd: W_II in the W phase
  double* Di = Dm + kChB * kChLdd;                // 32 x kChLdd: inverse of the diagonal U block
  double* Dg = Di + kChB * kChLdd;                // kChMaxNB x 32 x kChLdd: diagonal blocks
  double* Sc = Dg;                                // W phase: per-wave scratch (the diagonal blocks are done)
  double* rk = Dg + (size_t)kChMaxNB * kChB * kChLdd;  // 32: current row of the diagonal factorisation
  __shared__ int sflag;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int NB = (n + kChB - 1) / kChB, nblk = NB * (NB - 1) / 2;
  if (t == 0) sflag = 0;
  // diagonal blocks into LDS (upper triangle mirrored; identity outside n)
  for (int e = t; e < NB * kChB * kChB; e += blockDim.x) {
    const int I = e / (kChB * kChB), r = e & 31, cc = (e >> 5) & 31;
    const int row = kChB * I + r, col = kChB * I + cc;
    double v;
    if (row < n && col < n)
      v = row <= col ? S[row + (size_t)ld * col] : S[col + (size_t)ld * row];
    else
      v = row == col ? 1.0 : 0.0;
    Dg[(size_t)I * kChB * kChLdd + r * kChLdd + cc] = v;
  }
  if (w == 0) {
    const int c = lane & 31, h = lane >> 5;
    for (int P = 0; P < NB; P++) {
      double* Dm = Dg + (size_t)P * kChB * kChLdd;  // diagonal block P, becomes U_PP
      __syncthreads();  // diagonal block P up to date
      // unblocked Cholesky: lane (column c, half h) holds rows 16 h ..
      // 16 h + 15 of column c; row k goes through rk to the other lanes.
      double d[16];
#pragma unroll
      for (int ii = 0; ii < 16; ii++) d[ii] = Dm[(16 * h + ii) * kChLdd + c];
      int bad = 0;
#pragma unroll
      for (int k = 0; k < kChB; k++) {
        wave_lds_sync();  // the previous step's reads of rk are done
        if (h == (k >> 4)) rk[c] = d[k & 15];
        wave_lds_sync();
        const double dkk = rk[k];
        const bool okp = dkk > 0.0;  // false for NaN too
        if (!okp && bad == 0) bad = kChB * P + k + 1;
        const double ukk = okp ? sqrt(dkk) : 1.0;
        const double inv = 1.0 / ukk;
        const double uc = rk[c] * inv;
        if (h == (k >> 4)) d[k & 15] = c > k ? uc : (c == k ? ukk : 0.0);
#pragma unroll
        for (int ii = 0; ii < 16; ii++) {
          const int i = 16 * h + ii;
          if (i > k && i <= c) d[ii] -= rk[i] * inv * uc;
        }
      }
#pragma unroll
      for (int ii = 0; ii < 16; ii++) Dm[(16 * h + ii) * kChLdd + c] = (16 * h + ii <= c) ? d[ii] : 0.0;
      wave_lds_sync();
      // inverse of the factor, column c by back-substitution (lanes 0..31)
      if (h == 0) {
#pragma unroll
        for (int i = kChB - 1; i >= 0; i--) {
          double x = (i == c) ? 1.0 : 0.0;
#pragma unroll
          for (int k = i + 1; k < kChB; k++) x -= Dm[i * kChLdd + k] * Di[k * kChLdd + c];
          Di[i * kChLdd + c] = x / Dm[i * kChLdd + i];
        }
      }
      if (lane == 0 && bad) sflag = bad;
      __syncthreads();  // U_PP, its inverse and the flag published
      if (sflag) {
        if (t == 0) *info = sflag;
        return;
      }
      // W's diagonal block P
      for (int e = lane; e < kChB * kChB; e += 64) {
        const int r = e & 31, cc = e >> 5;
        const int row = kChB * P + r, col = kChB * P + cc;
        if (row < n && col < n) W[row + (size_t)ldw * col] = Di[r * kChLdd + cc];
      }
      __syncthreads();  // block row P of U in the panel
    }
  } else {
    const int wo = w - 1;
    d4v acc[kChSlots][4];
    // owned blocks; outside n the matrix is the identity (decoupled, so the
    // factor and the inverse restricted to n are exact)
#pragma unroll
    for (int s = 0; s < kChSlots; s++) {
      const int b = wo + kChOwners * s;
      int I = 0, J = 0;
      if (b < nblk) ch_block(b, I, J);
#pragma unroll
      for (int q = 0; q < 4; q++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int row = kChB * I + 16 * (q & 1) + 4 * i + lk, col = kChB * J + 16 * (q >> 1) + lr;
          double v = 0.0;
          if (b < nblk) {
            if (row < n && col < n)
              v = row <= col ? S[row + (size_t)ld * col] : S[col + (size_t)ld * row];
            else
              v = row == col ? 1.0 : 0.0;
          }
          acc[s][q][i] = v;
        }
    }
    for (int P = 0; P < NB; P++) {
      __syncthreads();  // diagonal block P up to date
      __syncthreads();  // U_PP, its inverse and the flag published
      if (sflag) return;
      // block row P: U_PJ = Di^T A_PJ (A_PJ staged in the panel)
#pragma unroll
      for (int s = 0; s < kChSlots; s++) {
        const int b = wo + kChOwners * s;
        if (b >= nblk) continue;
        int I, J;
        ch_block(b, I, J);
        if (I != P) continue;
        double* PJ = Pn + kChB * J;
        ch_to_lds(acc[s], PJ, kChLdp, lane);
        wave_lds_sync();
#pragma unroll
        for (int q = 0; q < 4; q++) acc[s][q] = d4v{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int k0 = 0; k0 < kChB; k0 += 4) {
          double a[2], bb[2];
#pragma unroll
          for (int u = 0; u < 2; u++) {
            a[u] = Di[(k0 + lk) * kChLdd + 16 * u + lr];
            bb[u] = PJ[(k0 + lk) * kChLdp + 16 * u + lr];
          }
#pragma unroll
          for (int q = 0; q < 4; q++) acc[s][q] = ch_mfma(a[q & 1], bb[q >> 1], acc[s][q]);
        }
        wave_lds_sync();  // every lane's operand reads of A_PJ are done
        ch_to_lds(acc[s], PJ, kChLdp, lane);
      }
      __syncthreads();  // block row P of U in the panel
      // trailing blocks: A_IJ -= U_PI^T U_PJ
#pragma unroll
      for (int s = 0; s < kChSlots; s++) {
        const int b = wo + kChOwners * s;
        if (b >= nblk) continue;
        int I, J;
        ch_block(b, I, J);
        if (I <= P) continue;
#pragma unroll
        for (int k0 = 0; k0 < kChB; k0 += 4) {
          double a[2], bb[2];
#pragma unroll
          for (int u = 0; u < 2; u++) {
            a[u] = -Pn[(k0 + lk) * kChLdp + kChB * I + 16 * u + lr];
            bb[u] = Pn[(k0 + lk) * kChLdp + kChB * J + 16 * u + lr];
          }
#pragma unroll
          for (int q = 0; q < 4; q++) acc[s][q] = ch_mfma(a[q & 1], bb[q >> 1], acc[s][q]);
        }
      }
      // diagonal blocks I > P owned by this wave, through LDS
      for (int I = P + 1; I < NB; I++) {
        if (I % kChOwners != wo) continue;
        double* DI = Dg + (size_t)I * kChB * kChLdd;
        d4v dacc[4];
        ch_from_lds(dacc, DI, kChLdd, lane);
#pragma unroll
        for (int k0 = 0; k0 < kChB; k0 += 4) {
          double a[2], bb[2];
#pragma unroll
          for (int u = 0; u < 2; u++) {
            a[u] = -Pn[(k0 + lk) * kChLdp + kChB * I + 16 * u + lr];
            bb[u] = Pn[(k0 + lk) * kChLdp + kChB * I + 16 * u + lr];
          }
#pragma unroll
          for (int q = 0; q < 4; q++) dacc[q] = ch_mfma(a[q & 1], bb[q >> 1], dacc[q]);
        }
        ch_to_lds(dacc, DI, kChLdd, lane);
      }
    }
    // U into the upper triangle of S (diagonal blocks from LDS)
    for (int e = t - 64; e < NB * kChB * kChB; e += blockDim.x - 64) {
      const int I = e / (kChB * kChB), r = e & 31, cc = (e >> 5) & 31;
      const int row = kChB * I + r, col = kChB * I + cc;
      if (row < n && col < n && r <= cc) S[row + (size_t)ld * col] = Dg[(size_t)I * kChB * kChLdd + r * kChLdd + cc];
    }
#pragma unroll
    for (int s = 0; s < kChSlots; s++) {
      const int b = wo + kChOwners * s;
      if (b >= nblk) continue;
      int I, J;
      ch_block(b, I, J);
#pragma unroll
      for (int q = 0; q < 4; q++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int row = kChB * I + 16 * (q & 1) + 4 * i + lk, col = kChB * J + 16 * (q >> 1) + lr;
          if (row < n && col < n) S[row + (size_t)ld * col] = acc[s][q][i];
        }
    }
  }
  // W below the diagonal blocks is zero
  for (int e = t; e < n * n; e += blockDim.x) {
    const int row = e % n, col = e / n;
    if (row / kChB > col / kChB) W[row + (size_t)ldw * col] = 0.0;
  }
  __syncthreads();
  // W by block rows from the last
  for (int I = NB - 2; I >= 0; I--) {
    // block row I of U (columns beyond block I) and W_II into LDS
    const int ncol = kChB * (NB - 1 - I);
    for (int e = t; e < kChB * ncol; e += blockDim.x) {
      const int r = e & 31, col = kChB * (I + 1) + (e >> 5), row = kChB * I + r;
      Pn[r * kChLdp + col] = (row < n && col < n) ? S[row + (size_t)ld * col] : 0.0;
    }
    for (int e = t; e < kChB * kChB; e += blockDim.x) {
      const int r = e & 31, cc = e >> 5, row = kChB * I + r, col = kChB * I + cc;
      Dm[r * kChLdd + cc] = (row < n && col < n) ? W[row + (size_t)ldw * col] : (r == cc ? 1.0 : 0.0);
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kChSlots; s++) {
      const int b = w + kChWaves * s;
      if (b >= nblk) continue;
      int Ib, J;
      ch_block(b, Ib, J);
      if (Ib != I || J == I) continue;
      // T = sum_{K=I+1..J} U_IK W_KJ
      d4v tacc[4];
#pragma unroll
      for (int q = 0; q < 4; q++) tacc[q] = d4v{0.0, 0.0, 0.0, 0.0};
      for (int K = I + 1; K <= J; K++) {
        double bw[8][2];
#pragma unroll
        for (int kq = 0; kq < 8; kq++)
#pragma unroll
          for (int u = 0; u < 2; u++) {
            const int row = kChB * K + 4 * kq + lk, col = kChB * J + 16 * u + lr;
            bw[kq][u] = (row < n && col < n) ? W[row + (size_t)ldw * col] : (row == col ? 1.0 : 0.0);
          }
#pragma unroll
        for (int kq = 0; kq < 8; kq++) {
          double a[2];
#pragma unroll
          for (int u = 0; u < 2; u++) a[u] = Pn[(16 * u + lr) * kChLdp + kChB * K + 4 * kq + lk];
#pragma unroll
          for (int q = 0; q < 4; q++) tacc[q] = ch_mfma(a[q & 1], bw[kq][q >> 1], tacc[q]);
        }
      }
      // W_IJ = -W_II T
      double* Sw = Sc + (size_t)w * kChB * kChLdd;
      wave_lds_sync();  // a previous block's reads of Sw are done
      ch_to_lds(tacc, Sw, kChLdd, lane);
      wave_lds_sync();
      d4v o[4];
#pragma unroll
      for (int q = 0; q < 4; q++) o[q] = d4v{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int k0 = 0; k0 < kChB; k0 += 4) {
        double a[2], bb[2];
#pragma unroll
        for (int u = 0; u < 2; u++) {
          a[u] = -Dm[(16 * u + lr) * kChLdd + k0 + lk];
          bb[u] = Sw[(k0 + lk) * kChLdd + 16 * u + lr];
        }
#pragma unroll
        for (int q = 0; q < 4; q++) o[q] = ch_mfma(a[q & 1], bb[q >> 1], o[q]);
      }
#pragma unroll
      for (int q = 0; q < 4; q++)
#pragma unroll
        for (int i = 0; i < 4; i++) {
          const int row = kChB * I + 16 * (q & 1) + 4 * i + lk, col = kChB * J + 16 * (q >> 1) + lr;
          if (row < n && col < n) W[row + (size_t)ldw * col] = o[q][i];
        }
    }
    __syncthreads();
  }
  if (t == 0) *info = 0;
}

// Host side: U (in S's upper triangle) and W = U^{-1}; the device info word
// is 0 or the failing column.  n <= kChB * kChMaxNB.
int chol_upper_inv(pw_ctx* c, int n, double* S, int ld, double* W, int ldw, int* info) {
  if (n < 1 || n > kChB * kChMaxNB) return fail(PW_ERR_UNSUPPORTED, "chol_upper_inv: n = %d", n);
  const size_t lds = sizeof(double) * ((size_t)kChB * kChLdp + 2 * (size_t)kChB * kChLdd +
                                       (size_t)kChMaxNB * kChB * kChLdd + kChB);
  PW_TRY(ensure_lds((const void*)k_potrf_inv, lds + 256));
  hipLaunchKernelGGL(k_potrf_inv, dim3(1), dim3(64 * kChWaves), lds, c->stream, S, n, ld, W, ldw, info);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

int chol_inv_max_n() { return kChB * kChMaxNB; }

}  // namespace pw
