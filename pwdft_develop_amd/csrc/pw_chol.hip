// Cholesky QR pieces of the orthogonalisation X <- X U^{-1}, S = X^T X =
// U^T U (eigensolver.cpp:386-469: Potrf 'U' then Trsm 'R','U','N','N'), for
// up to 256 bands, without library launches:
//
//   k_potrf            S -> U (upper triangle of S) and the inverses of the
//                      32 x 32 diagonal blocks of U; one 512-thread workgroup
//   k_trsm_right_upper X <- X U^{-1} by block substitution, rows of X
//                      independent: each wavefront owns 16 rows
//
// k_potrf keeps the strictly upper 32 x 32 blocks of S (at most 28) in the
// MFMA accumulator registers of wavefronts 1..7 for the whole factorisation
// (block b, numbered column by column, belongs to wave 1 + b % 7); the
// diagonal blocks live in LDS; wave 0 holds no blocks and spends its
// registers on the sequential part.  Per block column P (right-looking,
// LAPACK dpotrf's blocked order):
//   wave 0 factors the diagonal block (unblocked Cholesky, pivots broadcast
//   from registers, rows through LDS) and inverts its 32 x 32 factor,
//   the block row becomes U_PJ = U_PP^{-T} A_PJ (MFMA, A_PJ staged in LDS),
//   the trailing blocks take A_IJ -= U_PI^T U_PJ (MFMA from the LDS panel;
//   diagonal block I by wave 1 + I % 7, through LDS).
// The two branches meet at the same three barriers per block column.  A
// non-positive pivot stops with info = its 1-based column (LAPACK's info).
#include "pw_internal.h"
#include "pw_mfma.h"

namespace pw {

constexpr int kChB = 32;                                   // block size
constexpr int kChMaxNB = 8;                                // n <= 256
constexpr int kChWaves = 8;                                // 512 threads
constexpr int kChOwners = kChWaves - 1;                    // waves 1..7 hold the blocks
constexpr int kChSlots = (kChMaxNB * (kChMaxNB - 1) / 2 + kChOwners - 1) / kChOwners;  // 4 blocks per wave
constexpr int kChLdp = kChB * kChMaxNB + 1;                // row stride of the LDS panel
constexpr int kChLdd = kChB + 1;                           // row stride of a 32 x 32 LDS block

// block b of the strictly upper block triangle, numbered column by column
// (b = J (J - 1) / 2 + I, I < J)
__device__ __forceinline__ void ch_block(int b, int& I, int& J) {
  int j = 1;
  while ((j + 1) * j / 2 <= b) j++;
  J = j;
  I = b - j * (j - 1) / 2;
}

// S (n x n, upper triangle read, ld) -> U in its upper triangle; Dinv
// (NB blocks of 32 x 32, row-major) <- inverses of U's diagonal blocks
// (identity beyond n); *info = 0 or the failing column.
__global__ void __launch_bounds__(512)
k_potrf(double* __restrict__ S, int n, int ld, double* __restrict__ Dinv, int* __restrict__ info, long sS, long sD) {
  extern __shared__ double sm[];
  // one problem per workgroup of a batch (strides sS, sD; info one word each)
  S += blockIdx.x * sS;
  Dinv += blockIdx.x * sD;
  info += blockIdx.x;
  double* Pn = sm;                                // 32 x kChLdp: block row of U (global columns)
  double* Di = Pn + kChB * kChLdp;                // 32 x kChLdd: inverse of the diagonal U block
  double* Dg = Di + kChB * kChLdd;                // kChMaxNB x 32 x kChLdd: diagonal blocks
  double* rk = Dg + (size_t)kChMaxNB * kChB * kChLdd;  // 32: current row of the diagonal factorisation
  __shared__ int sflag;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int NB = (n + kChB - 1) / kChB, nblk = NB * (NB - 1) / 2;
  if (t == 0) sflag = 0;
  // diagonal blocks into LDS (upper triangle mirrored; identity outside n),
  // 8 loads in flight per thread
  for (int e0 = 0; e0 < NB * kChB * kChB; e0 += 8 * 512) {
    double v8[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int e = e0 + 512 * u + t, I = e / (kChB * kChB), r = e & 31, cc = (e >> 5) & 31;
      const int row = kChB * I + r, col = kChB * I + cc;
      double v = row == col ? 1.0 : 0.0;
      if (e < NB * kChB * kChB && row < n && col < n)
        v = row <= col ? S[row + (size_t)ld * col] : S[col + (size_t)ld * row];
      v8[u] = v;
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const int e = e0 + 512 * u + t;
      if (e < NB * kChB * kChB)
        Dg[(size_t)(e / (kChB * kChB)) * kChB * kChLdd + (e & 31) * kChLdd + ((e >> 5) & 31)] = v8[u];
    }
  }
  if (w == 0) {
    const int c = lane & 31, h = lane >> 5;
    for (int P = 0; P < NB; P++) {
      double* Dm = Dg + (size_t)P * kChB * kChLdd;  // diagonal block P, becomes U_PP
      __syncthreads();                               // diagonal block P up to date
      // unblocked Cholesky: lane (column c, half h) holds rows 16 h ..
      // 16 h + 15 of column c.  The pivot of step k is read from its lane's
      // register; row k goes to the other lanes through rk.
      double d[16];
#pragma unroll
      for (int ii = 0; ii < 16; ii++) d[ii] = Dm[(16 * h + ii) * kChLdd + c];
      int bad = 0;
#pragma unroll
      for (int k = 0; k < kChB; k++) {
        const int hk = k >> 4, rr = k & 15;
        const double dkk = lane_bcast(d[rr], 32 * hk + k);
        if (h == hk) rk[c] = d[rr];
        const bool okp = dkk > 0.0;  // false for NaN too
        if (!okp && bad == 0) bad = kChB * P + k + 1;
        const double ukk = okp ? sqrt(dkk) : 1.0;
        const double inv = 1.0 / ukk;
        wave_lds_sync();
        const double uc = rk[c] * inv;
        if (h == hk) d[rr] = c > k ? uc : (c == k ? ukk : 0.0);
        // branch-free, so that the 16 loads issue together (the next step's
        // store to rk may alias them, so the compiler keeps it after them;
        // the hardware runs one wavefront's LDS operations in order)
#pragma unroll
        for (int ii = 0; ii < 16; ii++) {
          const int i = 16 * h + ii;
          const double nd = d[ii] - rk[i] * inv * uc;
          d[ii] = (i > k && i <= c) ? nd : d[ii];
        }
        // step boundary: without it the scheduler interleaves the unrolled
        // steps and runs out of registers
#pragma unroll
        for (int ii = 0; ii < 16; ii++) asm volatile("" : "+v"(d[ii]));
        asm volatile("" ::: "memory");
      }
#pragma unroll
      for (int ii = 0; ii < 16; ii++) Dm[(16 * h + ii) * kChLdd + c] = (16 * h + ii <= c) ? d[ii] : 0.0;
      wave_lds_sync();
      // reciprocals of the diagonal of U_PP
      if (h == 0) rk[c] = 1.0 / Dm[c * kChLdd + c];
      wave_lds_sync();
      // inverse of the factor, column c by back-substitution, the column in
      // registers (lanes 0..31)
      if (h == 0) {
        double x[kChB];
#pragma unroll
        for (int i = kChB - 1; i >= 0; i--) {
          asm volatile("" ::: "memory");  // one row's loads at a time
          double s0 = (i == c) ? 1.0 : 0.0, s1 = 0.0;
#pragma unroll
          for (int k = i + 1; k < kChB; k++) {
            const double u = Dm[i * kChLdd + k];
            if ((k - i) & 1) s0 -= u * x[k]; else s1 -= u * x[k];
          }
          x[i] = (s0 + s1) * rk[i];
        }
#pragma unroll
        for (int i = 0; i < kChB; i++) Di[i * kChLdd + c] = x[i];
      }
      if (lane == 0 && bad) sflag = bad;
      __syncthreads();  // U_PP, its inverse and the flag published
      if (sflag) {
        if (t == 0) *info = sflag;
        return;
      }
      double* DP = Dinv + (size_t)P * kChB * kChB;
      for (int e = lane; e < kChB * kChB; e += 64) DP[e] = Di[(e >> 5) * kChLdd + (e & 31)];
      __syncthreads();  // block row P of U in the panel
    }
    // U_PP into the upper triangle of S
    for (int e = lane; e < NB * kChB * kChB; e += 64) {
      const int I = e / (kChB * kChB), r = e & 31, cc = (e >> 5) & 31;
      const int row = kChB * I + r, col = kChB * I + cc;
      if (row < n && col < n && r <= cc) S[row + (size_t)ld * col] = Dg[(size_t)I * kChB * kChLdd + r * kChLdd + cc];
    }
  } else {
    const int wo = w - 1;
    d4v acc[kChSlots][4];
    // owned blocks: coalesced loads (lane = row, two columns per load) into
    // the registers, then through this wave's LDS staging area into the
    // accumulator layout; outside n the matrix is the identity (decoupled,
    // so the factor restricted to n is exact; off-diagonal blocks are 0)
#pragma unroll
    for (int s = 0; s < kChSlots; s++) {
      const int b = wo + kChOwners * s;
      int I = 0, J = 1;
      if (b < nblk) ch_block(b, I, J);
#pragma unroll
      for (int u = 0; u < 16; u++) {
        const int row = kChB * I + (lane & 31), col = kChB * J + 2 * u + (lane >> 5);
        acc[s][u >> 2][u & 3] = (b < nblk && row < n && col < n) ? S[row + (size_t)ld * col] : 0.0;
      }
    }
    double* stg = Pn + (size_t)wo * kChB * kChLdd;  // 7 x 32 x 33 <= 32 x kChLdp
#pragma unroll
    for (int s = 0; s < kChSlots; s++) {
#pragma unroll
      for (int u = 0; u < 16; u++) stg[(lane & 31) * kChLdd + 2 * u + (lane >> 5)] = acc[s][u >> 2][u & 3];
      wave_lds_sync();
      blk_from_lds(acc[s], stg, kChLdd, lane);
      wave_lds_sync();
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
        blk_to_lds(acc[s], PJ, kChLdp, lane);
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
          for (int q = 0; q < 4; q++) acc[s][q] = mfma_f64(a[q & 1], bb[q >> 1], acc[s][q]);
        }
        wave_lds_sync();  // every lane's operand reads of A_PJ are done
        blk_to_lds(acc[s], PJ, kChLdp, lane);
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
          for (int q = 0; q < 4; q++) acc[s][q] = mfma_f64(a[q & 1], bb[q >> 1], acc[s][q]);
        }
      }
      // diagonal blocks I > P owned by this wave, through LDS
      for (int I = P + 1; I < NB; I++) {
        if (I % kChOwners != wo) continue;
        double* DI = Dg + (size_t)I * kChB * kChLdd;
        d4v dacc[4];
        blk_from_lds(dacc, DI, kChLdd, lane);
#pragma unroll
        for (int k0 = 0; k0 < kChB; k0 += 4) {
          double a[2], bb[2];
#pragma unroll
          for (int u = 0; u < 2; u++) {
            a[u] = -Pn[(k0 + lk) * kChLdp + kChB * I + 16 * u + lr];
            bb[u] = Pn[(k0 + lk) * kChLdp + kChB * I + 16 * u + lr];
          }
#pragma unroll
          for (int q = 0; q < 4; q++) dacc[q] = mfma_f64(a[q & 1], bb[q >> 1], dacc[q]);
        }
        blk_to_lds(dacc, DI, kChLdd, lane);
      }
    }
    // off-diagonal blocks of U into S
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
  if (t == 0) *info = 0;
}

// X (m x n, ldx) <- X U^{-1}, transposed: Y^T = U^{-T} X^T by block rows,
//   Y_J^T = Dinv_J^T (X_J^T - sum_{K<J} U_KJ^T Y_K^T).
// Each wavefront owns 16 rows of X, held as the 32 x 16 blocks Y_K^T in
// MFMA accumulator registers: in the v_mfma_f64_16x16x4f64 layout, register
// (ti, i) of a block is exactly the B operand of its k-step 16 ti + 4 i, so
// the substitution reads only U and Dinv, which the workgroup stages in LDS
// one block column at a time.  256 threads = 64 rows of X per workgroup.
__global__ void __launch_bounds__(256, 2)  // <= 256 registers: two waves per SIMD
k_trsm_right_upper(double* __restrict__ X, int m, int ldx, int n, const double* __restrict__ U, int ldu,
                   const double* __restrict__ Dinv, long sX, long sU, long sD) {
  extern __shared__ double sm[];
  // problem blockIdx.y of a batch (strides sX, sU, sD)
  X += blockIdx.y * sX;
  U += blockIdx.y * sU;
  Dinv += blockIdx.y * sD;
  double* Us = sm;                                            // (32 J) x kChLdd: U_{<J,J}
  double* Ds = Us + (size_t)kChB * (kChMaxNB - 1) * kChLdd;  // 32 x kChLdd: Dinv_J
  const int NB = (n + kChB - 1) / kChB;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const int row = blockIdx.x * 64 + 16 * w + lr;  // this lane's row of X
  d4v yt[kChMaxNB][2];
#pragma unroll
  for (int K = 0; K < kChMaxNB; K++)
#pragma unroll
    for (int ti = 0; ti < 2; ti++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int col = kChB * K + 16 * ti + 4 * i + lk;
        yt[K][ti][i] = (K < NB && row < m && col < n) ? X[row + (size_t)ldx * col] : 0.0;
      }
#pragma unroll
  for (int J = 0; J < kChMaxNB; J++) {
    if (J >= NB) break;
    __syncthreads();  // the previous block column's staging is consumed
    // U_{<J,J} and Dinv_J, 8 loads in flight per thread
    const int nu = kChB * J * kChB;
    for (int e0 = 0; e0 < nu + kChB * kChB; e0 += 8 * 256) {
      double v8[8];
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int e = e0 + 256 * u + t;
        double v = 0.0;
        if (e < nu) {
          const int k = e % (kChB * J), col = kChB * J + e / (kChB * J);
          if (col < n) v = U[k + (size_t)ldu * col];
        } else if (e < nu + kChB * kChB) {
          v = Dinv[(size_t)J * kChB * kChB + (e - nu)];
        }
        v8[u] = v;
      }
#pragma unroll
      for (int u = 0; u < 8; u++) {
        const int e = e0 + 256 * u + t;
        if (e < nu)
          Us[(e % (kChB * J)) * kChLdd + e / (kChB * J)] = v8[u];
        else if (e < nu + kChB * kChB)
          Ds[((e - nu) >> 5) * kChLdd + ((e - nu) & 31)] = v8[u];
      }
    }
    __syncthreads();
    // C = X_J^T - sum_{K<J} U_KJ^T Y_K^T
    d4v c[2] = {yt[J][0], yt[J][1]};
#pragma unroll
    for (int K = 0; K < J; K++)
#pragma unroll
      for (int kq = 0; kq < 8; kq++) {
#pragma unroll
        for (int ti = 0; ti < 2; ti++)
          c[ti] = mfma_f64(-Us[(kChB * K + 4 * kq + lk) * kChLdd + 16 * ti + lr], yt[K][kq >> 2][kq & 3], c[ti]);
      }
    // Y_J^T = Dinv_J^T C
    d4v o[2] = {d4v{0.0, 0.0, 0.0, 0.0}, d4v{0.0, 0.0, 0.0, 0.0}};
#pragma unroll
    for (int kq = 0; kq < 8; kq++)
#pragma unroll
      for (int ti = 0; ti < 2; ti++)
        o[ti] = mfma_f64(Ds[(4 * kq + lk) * kChLdd + 16 * ti + lr], c[kq >> 2][kq & 3], o[ti]);
    yt[J][0] = o[0];
    yt[J][1] = o[1];
  }
#pragma unroll
  for (int K = 0; K < kChMaxNB; K++)
#pragma unroll
    for (int ti = 0; ti < 2; ti++)
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int col = kChB * K + 16 * ti + 4 * i + lk;
        if (K < NB && row < m && col < n) X[row + (size_t)ldx * col] = yt[K][ti][i];
      }
}

// Host side.  S = U^T U (U into S's upper triangle) and the inverses of U's
// diagonal blocks into Dinv (chol_dinv_doubles(n) doubles); the device info
// word is 0 or the failing column.  n <= chol_inv_max_n().
int chol_upper(pw_ctx* c, int n, double* S, int ld, double* Dinv, int* info) {
  if (n < 1 || n > kChB * kChMaxNB) return fail(PW_ERR_UNSUPPORTED, "chol_upper: n = %d", n);
  const size_t lds = sizeof(double) * ((size_t)kChB * kChLdp + (size_t)kChB * kChLdd +
                                       (size_t)kChMaxNB * kChB * kChLdd + kChB);
  PW_TRY(ensure_lds((const void*)k_potrf, lds + 256));
  hipLaunchKernelGGL(k_potrf, dim3(1), dim3(64 * kChWaves), lds, c->stream, S, n, ld, Dinv, info, 0L, 0L);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

// X (m x n) <- X U^{-1} with U and Dinv from chol_upper.
int trsm_right_upper(pw_ctx* c, int m, int n, double* X, int ldx, const double* U, int ldu, const double* Dinv) {
  if (n < 1 || n > kChB * kChMaxNB) return fail(PW_ERR_UNSUPPORTED, "trsm_right_upper: n = %d", n);
  const size_t lds = sizeof(double) * ((size_t)kChB * (kChMaxNB - 1) * kChLdd + (size_t)kChB * kChLdd);
  PW_TRY(ensure_lds((const void*)k_trsm_right_upper, lds));
  hipLaunchKernelGGL(k_trsm_right_upper, dim3((m + 63) / 64), dim3(256), lds, c->stream, X, m, ldx, n, U, ldu,
                     Dinv, 0L, 0L, 0L);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

// nbatch problems at strides sS (matrices) and sD (Dinv); info[b] per problem.
int chol_upper_batched(pw_ctx* c, int n, double* S, int ld, long sS, double* Dinv, long sD, int* info, int nbatch) {
  if (n < 1 || n > kChB * kChMaxNB) return fail(PW_ERR_UNSUPPORTED, "chol_upper: n = %d", n);
  const size_t lds = sizeof(double) * ((size_t)kChB * kChLdp + (size_t)kChB * kChLdd +
                                       (size_t)kChMaxNB * kChB * kChLdd + kChB);
  PW_TRY(ensure_lds((const void*)k_potrf, lds + 256));
  hipLaunchKernelGGL(k_potrf, dim3(nbatch), dim3(64 * kChWaves), lds, c->stream, S, n, ld, Dinv, info, sS, sD);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

int trsm_right_upper_batched(pw_ctx* c, int m, int n, double* X, int ldx, long sX, const double* U, int ldu,
                             long sU, const double* Dinv, long sD, int nbatch) {
  if (n < 1 || n > kChB * kChMaxNB) return fail(PW_ERR_UNSUPPORTED, "trsm_right_upper: n = %d", n);
  const size_t lds = sizeof(double) * ((size_t)kChB * (kChMaxNB - 1) * kChLdd + (size_t)kChB * kChLdd);
  PW_TRY(ensure_lds((const void*)k_trsm_right_upper, lds));
  hipLaunchKernelGGL(k_trsm_right_upper, dim3((m + 63) / 64, nbatch), dim3(256), lds, c->stream, X, m, ldx, n, U,
                     ldu, Dinv, sX, sU, sD);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

int chol_inv_max_n() { return kChB * kChMaxNB; }
size_t chol_dinv_doubles(int n) { return (size_t)((n + kChB - 1) / kChB) * kChB * kChB; }

}  // namespace pw
