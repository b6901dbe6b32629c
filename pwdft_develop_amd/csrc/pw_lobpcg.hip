// LOBPCG on gfx950: EigenSolver::LOBPCGSolveReal, sphere-cut branch
// (src/EigenSolver/lobpcg.cpp:119-576) on one GPU, with
// EigenSolver::SolveGenEig's sequential branch (src/EigenSolver/
// eigensolver.cpp:224-297) for the (3w)^2 generalised problem.
//
// Blocks S = (X | W | P) and AS = (AX | AW | AP) live in HBM as real
// (2 npw) x w panels; every Gram block is a split-K MFMA GEMM written straight
// into AMat/BMat (lda = 3w); the small dense steps use rocSOLVER (dsyevd) and
// the eigenvector sign convention of the oracle (largest |component| > 0).
// One rank: the reference's Alltoall transposes and MPI reductions are
// identities here.
#include <algorithm>
#include <cmath>
#include <vector>

#include <rocsolver/rocsolver.h>

#include "pw_internal.h"

namespace pw {

__global__ void k_sym_upper_to_lower(double* __restrict__ M, int n, int ld) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)n * n; e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e % n), j = (int)(e / n);
    if (i > j) M[i + (size_t)ld * j] = M[j + (size_t)ld * i];
  }
}

// M(i,j) *= s[i] * (both ? s[j] : 1) for i, j < n
__global__ void k_scale_sym(double* __restrict__ M, int n, int ld, const double* __restrict__ s, int both) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)n * n; e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e % n), j = (int)(e / n);
    M[i + (size_t)ld * j] *= both ? s[i] * s[j] : s[i];
  }
}

// X[:, j] *= 1 / sqrt(nrm2[j])
__global__ void k_inv_norm_cols(double* __restrict__ X, long m, int n, const double* __restrict__ nrm2) {
  const long total = m * n;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x)
    X[e] *= 1.0 / sqrt(nrm2[e / m]);
}

__global__ void k_copy2d(double* __restrict__ dst, int ldd, const double* __restrict__ src, int lds, int rows,
                         int cols) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)rows * cols; e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e % rows), j = (int)(e / rows);
    dst[i + (size_t)ldd * j] = src[i + (size_t)lds * j];
  }
}

// column sum of squares (k_col_sumsq with theta == null) -- local copy
__global__ void k_col_sumsq2(const double* __restrict__ X, long m, double* __restrict__ out) {
  const int j = blockIdx.x;
  const double* x = X + (size_t)j * m;
  double s = 0.0;
  for (long i = threadIdx.x; i < m; i += blockDim.x) s += x[i] * x[i];
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = red[0];
}

// sign convention shared with pw_solver.hip (largest |q_i| positive)
__global__ void k_sign_fix2(double* __restrict__ Q, int n, int ld) {
  const int j = blockIdx.x;
  double* q = Q + (size_t)j * ld;
  double best = -1.0;
  int bi = 0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    const double a = fabs(q[i]);
    if (a > best) { best = a; bi = i; }
  }
  __shared__ double sb[256];
  __shared__ int si[256];
  sb[threadIdx.x] = best;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      const double o = sb[threadIdx.x + w];
      const int oi = si[threadIdx.x + w];
      if (o > sb[threadIdx.x] || (o == sb[threadIdx.x] && oi < si[threadIdx.x])) {
        sb[threadIdx.x] = o;
        si[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  const double sg = q[si[0]] < 0.0 ? -1.0 : 1.0;
  __syncthreads();
  if (sg < 0.0)
    for (int i = threadIdx.x; i < n; i += blockDim.x) q[i] = -q[i];
}

static inline int g1d(long n) { return (int)std::min<long>((n + 255) / 256, 4096); }

struct LobpcgWs {
  double *S = nullptr, *AS = nullptr, *Xtemp = nullptr;
  double *AMat = nullptr, *BMat = nullptr, *T1 = nullptr, *XTX = nullptr;
  double *w = nullptr, *e = nullptr, *red = nullptr, *inv = nullptr;
  int* info = nullptr;
  void free_all() {
    void* p[] = {S, AS, Xtemp, AMat, BMat, T1, XTX, w, e, red, inv, info};
    for (void* q : p)
      if (q) (void)hipFree(q);
  }
};

static int syevd_fixed(pw_ctx* c, int n, double* A, int ld, double* w, double* e, int* info) {
  return sym_eig(c, n, A, ld, w, e, info);
}

// SolveGenEig sequential branch on the device; eigenvectors overwrite AMat
// (numCol x numKeep), eigenvalues to eig (device, >= width entries).
static int solve_gen_eig(pw_ctx* c, LobpcgWs& W, int lda, int numCol, int width, double* eigd) {
  hipStream_t s = c->stream;
  hipLaunchKernelGGL(k_sym_upper_to_lower, dim3(g1d((long)numCol * numCol)), dim3(256), 0, s, W.AMat, numCol, lda);
  hipLaunchKernelGGL(k_sym_upper_to_lower, dim3(g1d((long)numCol * numCol)), dim3(256), 0, s, W.BMat, numCol, lda);
  PW_TRY(syevd_fixed(c, numCol, W.BMat, lda, W.w, W.e, W.info));
  std::vector<double> sig(numCol);
  PW_HIP(hipMemcpyAsync(sig.data(), W.w, sizeof(double) * numCol, hipMemcpyDeviceToHost, s));
  PW_HIP(hipStreamSynchronize(s));
  int numKeep = 0;
  for (int i = numCol - 1; i >= 0; i--) {
    if (sig[i] / sig[numCol - 1] > 1e-8) numKeep++;
    else break;
  }
  if (numKeep < width)
    return fail(PW_ERR_LINALG, "SolveGenEig: width = %d numKeep = %d: there are not enough number of columns.",
                width, numKeep);
  std::vector<double> inv(numKeep);
  for (int i = 0; i < numKeep; i++) inv[i] = 1.0 / std::sqrt(sig[i + numCol - numKeep]);
  PW_HIP(hipMemcpyAsync(W.inv, inv.data(), sizeof(double) * numKeep, hipMemcpyHostToDevice, s));
  const double* Uk = W.BMat + (size_t)lda * (numCol - numKeep);
  // T1 = A * U_keep ; A = U_keep^T T1 ; A(i,j) *= inv_i inv_j
  PW_TRY(gemm_nn<double>(c, numCol, numKeep, numCol, 1.0, W.AMat, lda, Uk, lda, 0.0, W.T1, lda));
  PW_TRY(gemm_tn<double>(c, numCol, numKeep, numKeep, Uk, lda, W.T1, lda, W.AMat, lda, nullptr, 0.0));
  hipLaunchKernelGGL(k_scale_sym, dim3(g1d((long)numKeep * numKeep)), dim3(256), 0, s, W.AMat, numKeep, lda,
                     (const double*)W.inv, 1);
  PW_TRY(syevd_fixed(c, numKeep, W.AMat, lda, eigd, W.e, W.info));
  hipLaunchKernelGGL(k_scale_sym, dim3(g1d((long)numKeep * numKeep)), dim3(256), 0, s, W.AMat, numKeep, lda,
                     (const double*)W.inv, 0);
  PW_TRY(gemm_nn<double>(c, numCol, numKeep, numKeep, 1.0, Uk, lda, W.AMat, lda, 0.0, W.T1, lda));
  hipLaunchKernelGGL(k_copy2d, dim3(g1d((long)numCol * numKeep)), dim3(256), 0, s, W.AMat, lda,
                     (const double*)W.T1, lda, numCol, numKeep);
  return PW_OK;
}

}  // namespace pw

using namespace pw;

extern "C" int pw_lobpcg(pw_ctx* c, int nb, int numEig, int scfIter, int eigMaxIter, double eigMinTolerance,
                         double eigTolerance, void* Xio, double* eigval, double* resval, int* iters) {
  (void)eigMinTolerance;
  if (!c || !Xio || !eigval || !resval || nb < 1 || nb > c->max_bands)
    return fail(PW_ERR_ARG, "pw_lobpcg: bad argument");
  if (numEig > nb) return fail(PW_ERR_ARG, "Number of eigenvalues requested = %d which is larger than the number "
                                           "of columns in psi = %d", numEig, nb);
  const int width = nb, lda = 3 * width, m = 2 * c->npw;
  const size_t blk = (size_t)m * width;
  hipStream_t s = c->stream;
  LobpcgWs W;
  auto cleanup = [&](int rc) { W.free_all(); return rc; };
#define ALLOC(p, n)                                                                     \
  if (hipMalloc((void**)&(p), sizeof(*(p)) * (size_t)(n)) != hipSuccess)               \
    return cleanup(fail(PW_ERR_NOMEM, "pw_lobpcg: cannot allocate %s", #p));
  ALLOC(W.S, 3 * blk);
  ALLOC(W.AS, 3 * blk);
  ALLOC(W.Xtemp, blk);
  ALLOC(W.AMat, (size_t)lda * lda);
  ALLOC(W.BMat, (size_t)lda * lda);
  ALLOC(W.T1, (size_t)lda * lda);
  ALLOC(W.XTX, (size_t)width * width);
  ALLOC(W.w, lda);
  ALLOC(W.e, lda);
  ALLOC(W.red, lda);
  ALLOC(W.inv, lda);
  ALLOC(W.info, 4);
#undef ALLOC
#define TRY(call)                      \
  do {                                 \
    int rc_ = (call);                  \
    if (rc_ != PW_OK) return cleanup(rc_); \
  } while (0)
#define HIPTRY(call)                                                                    \
  do {                                                                                  \
    hipError_t e_ = (call);                                                             \
    if (e_ != hipSuccess) return cleanup(fail(PW_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_))); \
  } while (0)
  double *X = W.S, *Wb = W.S + blk, *P = W.S + 2 * blk;
  double *AX = W.AS, *AW = W.AS + blk, *AP = W.AS + 2 * blk;
  HIPTRY(hipMemsetAsync(W.AMat, 0, sizeof(double) * lda * lda, s));
  HIPTRY(hipMemsetAsync(W.BMat, 0, sizeof(double) * lda * lda, s));
  HIPTRY(hipMemcpyAsync(X, Xio, sizeof(double) * blk, hipMemcpyDeviceToDevice, s));
  if (scfIter == 1) TRY(pw_orthogonalize(c, width, X));
  TRY(pw_apply_h(c, PW_F64, width, X, AX));

  std::vector<double> res(width), nrm(width);
  double resMax = 0.0;
  int iter = 0;
  do {
    iter++;
    const int numSet = (iter == 1) ? 2 : 3;
    // X'AX -> XTX and AMat(0,0)
    TRY(gemm_tn<double>(c, m, width, width, X, m, AX, m, W.XTX, width, nullptr, 0.0));
    hipLaunchKernelGGL(k_copy2d, dim3(g1d((long)width * width)), dim3(256), 0, s, W.AMat, lda,
                       (const double*)W.XTX, width, width, width);
    // R = AX - X (X'AX)
    HIPTRY(hipMemcpyAsync(W.Xtemp, AX, sizeof(double) * blk, hipMemcpyDeviceToDevice, s));
    TRY(gemm_nn<double>(c, m, width, width, -1.0, X, m, W.AMat, lda, 1.0, W.Xtemp, m));
    hipLaunchKernelGGL(k_col_sumsq2, dim3(width), dim3(256), 0, s, (const double*)W.Xtemp, (long)m, W.red);
    std::vector<double> diag(width);
    HIPTRY(hipMemcpyAsync(res.data(), W.red, sizeof(double) * width, hipMemcpyDeviceToHost, s));
    HIPTRY(hipMemcpy2DAsync(diag.data(), sizeof(double), W.XTX, sizeof(double) * (width + 1), sizeof(double), width,
                            hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    for (int k = 0; k < width; k++) res[k] = std::sqrt(res[k]) / std::max(1.0, std::fabs(diag[k]));
    resMax = *std::max_element(res.begin(), res.begin() + std::max(1, numEig));
    // W = T R, normalised
    HIPTRY(hipMemsetAsync(Wb, 0, sizeof(double) * blk, s));
    TRY(pw_teter_precond(c, PW_F64, width, W.Xtemp, Wb));
    hipLaunchKernelGGL(k_col_sumsq2, dim3(width), dim3(256), 0, s, (const double*)Wb, (long)m, W.red);
    hipLaunchKernelGGL(k_inv_norm_cols, dim3(g1d((long)blk)), dim3(256), 0, s, Wb, (long)m, width,
                       (const double*)W.red);
    if (numSet == 3) {
      hipLaunchKernelGGL(k_col_sumsq2, dim3(width), dim3(256), 0, s, (const double*)P, (long)m, W.red);
      hipLaunchKernelGGL(k_inv_norm_cols, dim3(g1d((long)blk)), dim3(256), 0, s, P, (long)m, width,
                         (const double*)W.red);
      hipLaunchKernelGGL(k_inv_norm_cols, dim3(g1d((long)blk)), dim3(256), 0, s, AP, (long)m, width,
                         (const double*)W.red);
    }
    TRY(pw_apply_h(c, PW_F64, width, Wb, AW));
    auto gram = [&](double* M, int r0, int c0, const double* A, const double* B) {
      return gemm_tn<double>(c, m, width, width, A, m, B, m, M + r0 + (size_t)lda * c0, lda, nullptr, 0.0);
    };
    TRY(gram(W.AMat, 0, width, X, AW));
    TRY(gram(W.AMat, width, width, Wb, AW));
    if (numSet == 3) {
      TRY(gram(W.AMat, 0, 2 * width, X, AP));
      TRY(gram(W.AMat, width, 2 * width, Wb, AP));
      TRY(gram(W.AMat, 2 * width, 2 * width, P, AP));
    }
    TRY(gram(W.BMat, 0, 0, X, X));
    TRY(gram(W.BMat, 0, width, X, Wb));
    TRY(gram(W.BMat, width, width, Wb, Wb));
    if (numSet == 3) {
      TRY(gram(W.BMat, 0, 2 * width, X, P));
      TRY(gram(W.BMat, width, 2 * width, Wb, P));
      TRY(gram(W.BMat, 2 * width, 2 * width, P, P));
    }
    const int numCol = (numSet == 3) ? 3 * width : 2 * width;
    TRY(solve_gen_eig(c, W, lda, numCol, width, W.red));
    const double* CX = W.AMat;
    const double* CW = W.AMat + width;
    const double* CP = W.AMat + 2 * width;
    if (numSet == 2) {
      TRY(gemm_nn<double>(c, m, width, width, 1.0, X, m, CX, lda, 0.0, W.Xtemp, m));
      TRY(gemm_nn<double>(c, m, width, width, 1.0, Wb, m, CW, lda, 1.0, W.Xtemp, m));
      HIPTRY(hipMemcpyAsync(X, W.Xtemp, sizeof(double) * blk, hipMemcpyDeviceToDevice, s));
      HIPTRY(hipMemcpyAsync(P, Wb, sizeof(double) * blk, hipMemcpyDeviceToDevice, s));
      TRY(gemm_nn<double>(c, m, width, width, 1.0, AX, m, CX, lda, 0.0, W.Xtemp, m));
      TRY(gemm_nn<double>(c, m, width, width, 1.0, AW, m, CW, lda, 1.0, W.Xtemp, m));
      HIPTRY(hipMemcpyAsync(AX, W.Xtemp, sizeof(double) * blk, hipMemcpyDeviceToDevice, s));
      HIPTRY(hipMemcpyAsync(AP, AW, sizeof(double) * blk, hipMemcpyDeviceToDevice, s));
    } else {
      TRY(gemm_nn<double>(c, m, width, width, 1.0, Wb, m, CW, lda, 0.0, W.Xtemp, m));
      TRY(gemm_nn<double>(c, m, width, width, 1.0, P, m, CP, lda, 1.0, W.Xtemp, m));
      HIPTRY(hipMemcpyAsync(P, W.Xtemp, sizeof(double) * blk, hipMemcpyDeviceToDevice, s));
      TRY(gemm_nn<double>(c, m, width, width, 1.0, X, m, CX, lda, 1.0, W.Xtemp, m));
      HIPTRY(hipMemcpyAsync(X, W.Xtemp, sizeof(double) * blk, hipMemcpyDeviceToDevice, s));
      TRY(gemm_nn<double>(c, m, width, width, 1.0, AW, m, CW, lda, 0.0, W.Xtemp, m));
      TRY(gemm_nn<double>(c, m, width, width, 1.0, AP, m, CP, lda, 1.0, W.Xtemp, m));
      HIPTRY(hipMemcpyAsync(AP, W.Xtemp, sizeof(double) * blk, hipMemcpyDeviceToDevice, s));
      TRY(gemm_nn<double>(c, m, width, width, 1.0, AX, m, CX, lda, 1.0, W.Xtemp, m));
      HIPTRY(hipMemcpyAsync(AX, W.Xtemp, sizeof(double) * blk, hipMemcpyDeviceToDevice, s));
    }
  } while (iter < eigMaxIter && resMax > eigTolerance);

  // post-processing (lobpcg.cpp:536-548): eigen-decomposition of the X'AX of
  // the last iteration's start, X <- X C
  TRY(syevd_fixed(c, width, W.XTX, width, W.w, W.e, W.info));
  TRY(gemm_nn<double>(c, m, width, width, 1.0, X, m, W.XTX, width, 0.0, (double*)Xio, m));
  HIPTRY(hipMemcpyAsync(eigval, W.w, sizeof(double) * width, hipMemcpyDeviceToHost, s));
  HIPTRY(hipStreamSynchronize(s));
  for (int k = 0; k < width; k++) resval[k] = res[k];
  if (iters) *iters = iter;
  return cleanup(PW_OK);
#undef TRY
#undef HIPTRY
}
