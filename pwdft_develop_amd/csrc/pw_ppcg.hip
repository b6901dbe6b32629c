// PPCG on gfx950: EigenSolver::PPCGSolveReal (src/EigenSolver/ppcg.cpp), the
// sphere-cut branch (:119-576) on the (2 npw)-row real view of the bands and
// the real-space branch (:683-1246, the same steps) on the ntot-row
// coarse-grid bands, one GPU.
//
// S = (X | W | P) and AS = (AX | AW | AP) live in HBM as real (2 npw) x w
// panels.  Per iteration: residual and Teter preconditioner, one H*W, the
// projections of W and P against X (split-K MFMA GEMMs), then for every
// sub-block of sbSize columns the (2 or 3 sbSize)^2 Gram blocks are written
// by GEMMs straight into a stack of small matrices, all sub-blocks' Sygvd
// problems go to rocSOLVER's strided-batched dsygvd in one call (plus one for
// a ragged last block), and the block updates p <- w cw (+ p cp),
// x <- x cx + p are small GEMMs.  Cholesky QR closes the iteration.  One
// rank: the reference's Alltoall transposes and MPI reductions are
// identities here.  Eigenvector signs follow the oracle (largest |component|
// of each sub-block eigenvector positive).
#include <algorithm>
#include <cmath>
#include <functional>
#include <vector>

#include <rocsolver/rocsolver.h>

#include "pw_internal.h"

namespace pw {

// nrm2[j] = sum_i (X[i,j] - (theta ? theta[j] Y[i,j] : 0))^2, one block per
// column, 8 loads in flight per thread
__global__ void __launch_bounds__(256)
k_pp_colnorm2(const double* __restrict__ X, const double* __restrict__ Y, const double* __restrict__ theta, long m,
              double* __restrict__ nrm2) {
  const int j = blockIdx.x;
  const double* x = X + (size_t)j * m;
  const double* y = Y ? Y + (size_t)j * m : nullptr;
  const double th = theta ? theta[j] : 0.0;
  double s = 0.0;
  for (long i0 = threadIdx.x; i0 < m; i0 += 8 * 256) {
    double a[8], b[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const long i = i0 + 256L * u;
      a[u] = i < m ? x[i] : 0.0;
      b[u] = (y && i < m) ? y[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const double v = a[u] - th * b[u];
      s += v * v;
    }
  }
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) nrm2[j] = red[0];
}

// X[:, j] /= sqrt(nrm2[j]) and Y[:, j] likewise
__global__ void k_pp_scale2(double* __restrict__ X, double* __restrict__ Y, long m, int n,
                            const double* __restrict__ nrm2) {
  const long total = m * n;
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const double s = 1.0 / sqrt(nrm2[e / m]);
    X[e] *= s;
    Y[e] *= s;
  }
}

// res[k] = sqrt(nrm2[k]) / max(1, |d[k]|), d[k] = D[k * ldd] (diagonal or vector)
__global__ void k_pp_resnorm(const double* __restrict__ nrm2, const double* __restrict__ D, int ldd, int n,
                             double* __restrict__ res) {
  for (int k = blockIdx.x * blockDim.x + threadIdx.x; k < n; k += gridDim.x * blockDim.x)
    res[k] = sqrt(nrm2[k]) / fmax(1.0, fabs(D[(size_t)k * ldd]));
}

// Sign rule of the oracle: column j's largest-magnitude entry (first index
// on ties) made positive; one block per column.
__global__ void k_pp_sign_fix(double* __restrict__ Q, int n, int ld) {
  double* q = Q + (size_t)blockIdx.x * ld;
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
  for (int w = 128; w > 0; w >>= 1) {
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

static inline int grid1(long n) { return (int)std::max<long>(1, std::min<long>((n + 255) / 256, 8192)); }

struct PpcgWs {
  double *S = nullptr, *AS = nullptr, *Xtemp = nullptr, *tmp = nullptr;
  double *XTX = nullptr, *AM = nullptr, *BM = nullptr, *eig = nullptr, *E = nullptr;
  double *nrm = nullptr, *res = nullptr, *w = nullptr, *dinv = nullptr;
  int* info = nullptr;
  void free_all() {
    void* p[] = {S, AS, Xtemp, tmp, XTX, AM, BM, eig, E, nrm, res, w, dinv, info};
    for (void* q : p)
      if (q) (void)hipFree(q);
  }
};

}  // namespace pw

using namespace pw;

namespace {

// The operator a PPCG run works with: y = H x, and w += T r (the Teter
// preconditioner), on real m x nb device blocks.
struct PpcgOps {
  std::function<int(int, const double*, double*)> apply_h, precond;
};

// The PPCG iteration on real m-row blocks with the linear algebra of c
// (stream, rocBLAS handle, GEMM and eigensolver workspaces).
int ppcg_core(pw_ctx* c, int m, const PpcgOps& ops, int nb, int numEig, int scfIter, int eigMaxIter,
              double eigTolerance, int sbSize, void* Xio, double* eigval, double* resval, int* iters) {
  const int width = nb;
  const int sbs = std::min(sbSize, width);
  const int nsb = (width + sbs - 1) / sbs;
  const int sb3 = 3 * sbs;
  const size_t blk = (size_t)m * width, mat = (size_t)sb3 * sb3;
  hipStream_t s = c->stream;
  PpcgWs W;
  auto cleanup = [&](int rc) { W.free_all(); return rc; };
#define ALLOC(p, n)                                                                     \
  if (hipMalloc((void**)&(p), sizeof(*(p)) * (size_t)(n)) != hipSuccess)               \
    return cleanup(fail(PW_ERR_NOMEM, "pw_ppcg: cannot allocate %s", #p));
  ALLOC(W.S, 3 * blk);
  ALLOC(W.AS, 3 * blk);
  ALLOC(W.Xtemp, blk);
  ALLOC(W.tmp, (size_t)m * sbs);
  ALLOC(W.XTX, (size_t)width * width);
  ALLOC(W.AM, mat * nsb);
  ALLOC(W.BM, mat * nsb);
  ALLOC(W.eig, (size_t)sb3 * nsb);
  ALLOC(W.E, (size_t)sb3 * nsb);
  ALLOC(W.nrm, width);
  ALLOC(W.res, width);
  ALLOC(W.w, width);
  ALLOC(W.info, nsb + 1);
  const bool own_chol = width <= chol_inv_max_n();
  if (own_chol) ALLOC(W.dinv, chol_dinv_doubles(width));
#undef ALLOC
#define TRY(call)                          \
  do {                                     \
    int rc_ = (call);                      \
    if (rc_ != PW_OK) return cleanup(rc_); \
  } while (0)
#define HIPTRY(call)                                                                                     \
  do {                                                                                                   \
    hipError_t e_ = (call);                                                                              \
    if (e_ != hipSuccess) return cleanup(fail(PW_ERR_HIP, "%s: %s", #call, hipGetErrorString(e_)));     \
  } while (0)
#define BLASTRY(call)                                                                        \
  do {                                                                                       \
    if ((call) != rocblas_status_success) return cleanup(fail(PW_ERR_HIP, "%s failed", #call)); \
  } while (0)
  double *X = W.S, *Wb = W.S + blk, *P = W.S + 2 * blk;
  double *AX = W.AS, *AW = W.AS + blk, *AP = W.AS + 2 * blk;
  const double one = 1.0;
  HIPTRY(hipMemsetAsync(W.S, 0, sizeof(double) * 3 * blk, s));
  HIPTRY(hipMemsetAsync(W.AS, 0, sizeof(double) * 3 * blk, s));
  HIPTRY(hipMemcpyAsync(X, Xio, sizeof(double) * blk, hipMemcpyDeviceToDevice, s));
  // Cholesky QR of X (and of AX by the same factor when it exists): U in
  // XTX.  Up to 256 bands the factor and the solves are k_potrf and
  // k_trsm_right_upper, and the pivot check (info in W.info[nsb]) waits for
  // the next host read (the residual norms of the next iteration, or the
  // end); beyond, rocSOLVER potrf with an immediate check and rocBLAS trsm.
  auto chol_check = [&]() -> int {
    int h = 0;
    PW_HIP(hipMemcpyAsync(&h, W.info + nsb, sizeof(int), hipMemcpyDeviceToHost, s));
    PW_HIP(hipStreamSynchronize(s));
    if (h != 0) return fail(PW_ERR_LINALG, "Potrf: overlap matrix is not positive definite (info = %d)", h);
    return PW_OK;
  };
  auto chol_qr = [&](bool with_ax) -> int {
    PW_TRY(gemm_tn<double>(c, m, width, width, X, m, X, m, W.XTX, width, nullptr, 0.0));
    if (own_chol) {
      PW_TRY(chol_upper(c, width, W.XTX, width, W.dinv, W.info + nsb));
      PW_TRY(trsm_right_upper(c, m, width, X, m, W.XTX, width, W.dinv));
      if (with_ax) PW_TRY(trsm_right_upper(c, m, width, AX, m, W.XTX, width, W.dinv));
      return PW_OK;
    }
    PW_BLAS(rocsolver_dpotrf(c->blas, rocblas_fill_upper, width, W.XTX, width, W.info + nsb));
    PW_TRY(chol_check());
    PW_BLAS(rocblas_dtrsm(c->blas, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                          rocblas_diagonal_non_unit, m, width, &one, W.XTX, width, X, m));
    if (with_ax)
      PW_BLAS(rocblas_dtrsm(c->blas, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                            rocblas_diagonal_non_unit, m, width, &one, W.XTX, width, AX, m));
    return PW_OK;
  };
  if (scfIter == 1) {
    TRY(chol_qr(false));
    if (own_chol) TRY(chol_check());
  }
  TRY(ops.apply_h(width, X, AX));

  int iter = 0;
  double resMax = 0.0;
  std::vector<double> hres(width);
  do {
    iter++;
    const int numSet = (iter == 1) ? 2 : 3;
    // XTX = X'AX, R = AX - X XTX in Xtemp, residual norms (ppcg.cpp:228-266)
    TRY(gemm_tn<double>(c, m, width, width, X, m, AX, m, W.XTX, width, nullptr, 0.0));
    HIPTRY(hipMemcpyAsync(W.Xtemp, AX, sizeof(double) * blk, hipMemcpyDeviceToDevice, s));
    TRY(gemm_nn<double>(c, m, width, width, -1.0, X, m, W.XTX, width, 1.0, W.Xtemp, m));
    hipLaunchKernelGGL(k_pp_colnorm2, dim3(width), dim3(256), 0, s, (const double*)W.Xtemp, (const double*)nullptr,
                       (const double*)nullptr, (long)m, W.nrm);
    hipLaunchKernelGGL(k_pp_resnorm, dim3(grid1(width)), dim3(256), 0, s, (const double*)W.nrm,
                       (const double*)W.XTX, width + 1, width, W.res);
    HIPTRY(hipMemcpyAsync(hres.data(), W.res, sizeof(double) * width, hipMemcpyDeviceToHost, s));
    HIPTRY(hipStreamSynchronize(s));
    if (own_chol && iter > 1) TRY(chol_check());  // the previous iteration's Cholesky QR
    resMax = *std::max_element(hres.begin(), hres.begin() + std::max(numEig, 1));
    // W = T R, AW = A W (ppcg.cpp:272-293)
    HIPTRY(hipMemsetAsync(Wb, 0, sizeof(double) * blk, s));
    TRY(ops.precond(width, W.Xtemp, Wb));
    TRY(ops.apply_h(width, Wb, AW));
    // W -= X (X'W), AW -= AX (X'W), normalise (ppcg.cpp:295-322)
    TRY(gemm_tn<double>(c, m, width, width, X, m, Wb, m, W.XTX, width, nullptr, 0.0));
    TRY(gemm_nn<double>(c, m, width, width, -1.0, X, m, W.XTX, width, 1.0, Wb, m));
    TRY(gemm_nn<double>(c, m, width, width, -1.0, AX, m, W.XTX, width, 1.0, AW, m));
    hipLaunchKernelGGL(k_pp_colnorm2, dim3(width), dim3(256), 0, s, (const double*)Wb, (const double*)nullptr,
                       (const double*)nullptr, (long)m, W.nrm);
    hipLaunchKernelGGL(k_pp_scale2, dim3(grid1((long)blk)), dim3(256), 0, s, Wb, AW, (long)m, width,
                       (const double*)W.nrm);
    if (numSet == 3) {  // ppcg.cpp:325-350
      TRY(gemm_tn<double>(c, m, width, width, X, m, P, m, W.XTX, width, nullptr, 0.0));
      TRY(gemm_nn<double>(c, m, width, width, -1.0, X, m, W.XTX, width, 1.0, P, m));
      TRY(gemm_nn<double>(c, m, width, width, -1.0, AX, m, W.XTX, width, 1.0, AP, m));
      hipLaunchKernelGGL(k_pp_colnorm2, dim3(width), dim3(256), 0, s, (const double*)P, (const double*)nullptr,
                         (const double*)nullptr, (long)m, W.nrm);
      hipLaunchKernelGGL(k_pp_scale2, dim3(grid1((long)blk)), dim3(256), 0, s, P, AP, (long)m, width,
                         (const double*)W.nrm);
    }
    // sub-block Gram matrices (ppcg.cpp:353-436): block k at AM + k mat, ld 3 sb_k
    HIPTRY(hipMemsetAsync(W.AM, 0, sizeof(double) * mat * nsb, s));
    HIPTRY(hipMemsetAsync(W.BM, 0, sizeof(double) * mat * nsb, s));
    for (int k = 0; k < nsb; k++) {
      const int sb = (k == nsb - 1 && width % sbs) ? width % sbs : sbs;
      const int ld = 3 * sb;
      const size_t off = (size_t)m * sbs * k;
      double* Ak = W.AM + mat * k;
      double* Bk = W.BM + mat * k;
      const double *x = X + off, *w = Wb + off, *p = P + off, *ax = AX + off, *aw = AW + off, *ap = AP + off;
      auto g = [&](const double* a, const double* b, double* M, int r0, int c0) {
        return gemm_tn<double>(c, m, sb, sb, a, m, b, m, M + r0 + (size_t)ld * c0, ld, nullptr, 0.0);
      };
      TRY(g(x, ax, Ak, 0, 0));
      TRY(g(w, aw, Ak, sb, sb));
      TRY(g(x, aw, Ak, 0, sb));
      TRY(g(x, x, Bk, 0, 0));
      TRY(g(w, w, Bk, sb, sb));
      TRY(g(x, w, Bk, 0, sb));
      if (numSet == 3) {
        TRY(g(p, ap, Ak, 2 * sb, 2 * sb));
        TRY(g(x, ap, Ak, 0, 2 * sb));
        TRY(g(w, ap, Ak, sb, 2 * sb));
        TRY(g(p, p, Bk, 2 * sb, 2 * sb));
        TRY(g(x, p, Bk, 0, 2 * sb));
        TRY(g(w, p, Bk, sb, 2 * sb));
      }
    }
    // all full sub-blocks in one batched Sygvd, the ragged last one alone
    {
      const int nfull = (width % sbs) ? nsb - 1 : nsb;
      const int dimf = (numSet == 3) ? 3 * sbs : 2 * sbs;
      // orders up to 256: the batched solver of pw_eig.hip (one launch per
      // stage for all sub-blocks); beyond, rocSOLVER
      if (nfull > 0) {
        if (dimf <= chol_inv_max_n())
          TRY(sygvd_batched(c, dimf, W.AM, sb3, (long)mat, W.BM, sb3, (long)mat, W.eig, sb3, W.info, nfull));
        else
          BLASTRY(rocsolver_dsygvd_strided_batched(c->blas, rocblas_eform_ax, rocblas_evect_original,
                                                   rocblas_fill_upper, dimf, W.AM, sb3, (rocblas_stride)mat, W.BM,
                                                   sb3, (rocblas_stride)mat, W.eig, sb3, W.E, sb3, W.info, nfull));
      }
      if (nfull < nsb) {
        const int sb = width % sbs, dim = (numSet == 3) ? 3 * sb : 2 * sb;
        if (dim <= chol_inv_max_n())
          TRY(sygvd_batched(c, dim, W.AM + mat * nfull, 3 * sb, 0, W.BM + mat * nfull, 3 * sb, 0,
                            W.eig + (size_t)sb3 * nfull, 0, W.info + nfull, 1));
        else
          BLASTRY(rocsolver_dsygvd(c->blas, rocblas_eform_ax, rocblas_evect_original, rocblas_fill_upper, dim,
                                   W.AM + mat * nfull, 3 * sb, W.BM + mat * nfull, 3 * sb,
                                   W.eig + (size_t)sb3 * nfull, W.E + (size_t)sb3 * nfull, W.info + nfull));
      }
      std::vector<int> hinfo(nsb);
      HIPTRY(hipMemcpyAsync(hinfo.data(), W.info, sizeof(int) * nsb, hipMemcpyDeviceToHost, s));
      HIPTRY(hipStreamSynchronize(s));
      for (int k = 0; k < nsb; k++)
        if (hinfo[k] != 0) return cleanup(fail(PW_ERR_LINALG, "sygvd returned with info = %d", hinfo[k]));
    }
    // block updates (ppcg.cpp:462-527)
    for (int k = 0; k < nsb; k++) {
      const int sb = (k == nsb - 1 && width % sbs) ? width % sbs : sbs;
      const int ld = 3 * sb, dim = (numSet == 3) ? 3 * sb : 2 * sb;
      const size_t off = (size_t)m * sbs * k;
      double* Zk = W.AM + mat * k;
      hipLaunchKernelGGL(k_pp_sign_fix, dim3(sb), dim3(256), 0, s, Zk, dim, ld);
      const double* cx = Zk;
      const double* cw = Zk + sb;
      const double* cp = Zk + 2 * sb;
      double *x = X + off, *w = Wb + off, *p = P + off, *ax = AX + off, *aw = AW + off, *ap = AP + off;
      if (numSet == 3) {
        TRY(gemm_nn<double>(c, m, sb, sb, 1.0, p, m, cp, ld, 0.0, W.tmp, m));
        TRY(gemm_nn<double>(c, m, sb, sb, 1.0, w, m, cw, ld, 1.0, W.tmp, m));
        HIPTRY(hipMemcpyAsync(p, W.tmp, sizeof(double) * m * sb, hipMemcpyDeviceToDevice, s));
        TRY(gemm_nn<double>(c, m, sb, sb, 1.0, ap, m, cp, ld, 0.0, W.tmp, m));
        TRY(gemm_nn<double>(c, m, sb, sb, 1.0, aw, m, cw, ld, 1.0, W.tmp, m));
        HIPTRY(hipMemcpyAsync(ap, W.tmp, sizeof(double) * m * sb, hipMemcpyDeviceToDevice, s));
      } else {
        TRY(gemm_nn<double>(c, m, sb, sb, 1.0, w, m, cw, ld, 0.0, p, m));
        TRY(gemm_nn<double>(c, m, sb, sb, 1.0, aw, m, cw, ld, 0.0, ap, m));
      }
      HIPTRY(hipMemcpyAsync(W.tmp, p, sizeof(double) * m * sb, hipMemcpyDeviceToDevice, s));
      TRY(gemm_nn<double>(c, m, sb, sb, 1.0, x, m, cx, ld, 1.0, W.tmp, m));
      HIPTRY(hipMemcpyAsync(x, W.tmp, sizeof(double) * m * sb, hipMemcpyDeviceToDevice, s));
      HIPTRY(hipMemcpyAsync(W.tmp, ap, sizeof(double) * m * sb, hipMemcpyDeviceToDevice, s));
      TRY(gemm_nn<double>(c, m, sb, sb, 1.0, ax, m, cx, ld, 1.0, W.tmp, m));
      HIPTRY(hipMemcpyAsync(ax, W.tmp, sizeof(double) * m * sb, hipMemcpyDeviceToDevice, s));
    }
    TRY(chol_qr(true));  // ppcg.cpp:530-544
  } while (iter < eigMaxIter && resMax > eigTolerance);

  // post-processing (ppcg.cpp:551-633): Syevd of X'AX, rotation, residuals
  TRY(gemm_tn<double>(c, m, width, width, X, m, AX, m, W.XTX, width, nullptr, 0.0));
  TRY(sym_eig(c, width, W.XTX, width, W.w, W.E, W.info));
  TRY(gemm_nn<double>(c, m, width, width, 1.0, X, m, W.XTX, width, 0.0, W.Xtemp, m));
  HIPTRY(hipMemcpyAsync(X, W.Xtemp, sizeof(double) * blk, hipMemcpyDeviceToDevice, s));
  TRY(gemm_nn<double>(c, m, width, width, 1.0, AX, m, W.XTX, width, 0.0, W.Xtemp, m));
  HIPTRY(hipMemcpyAsync(AX, W.Xtemp, sizeof(double) * blk, hipMemcpyDeviceToDevice, s));
  hipLaunchKernelGGL(k_pp_colnorm2, dim3(width), dim3(256), 0, s, (const double*)AX, (const double*)X,
                     (const double*)W.w, (long)m, W.nrm);
  hipLaunchKernelGGL(k_pp_resnorm, dim3(grid1(width)), dim3(256), 0, s, (const double*)W.nrm, (const double*)W.w, 1,
                     width, W.res);
  HIPTRY(hipMemcpyAsync(Xio, X, sizeof(double) * blk, hipMemcpyDeviceToDevice, s));
  HIPTRY(hipMemcpyAsync(eigval, W.w, sizeof(double) * width, hipMemcpyDeviceToHost, s));
  HIPTRY(hipMemcpyAsync(resval, W.res, sizeof(double) * width, hipMemcpyDeviceToHost, s));
  HIPTRY(hipStreamSynchronize(s));
  if (own_chol) TRY(chol_check());  // the last iteration's Cholesky QR
  if (iters) *iters = iter;
#undef TRY
#undef HIPTRY
#undef BLASTRY
  return cleanup(PW_OK);
}

}  // namespace

extern "C" int pw_ppcg(pw_ctx* c, int nb, int numEig, int scfIter, int eigMaxIter, double eigMinTolerance,
                       double eigTolerance, int sbSize, void* Xio, double* eigval, double* resval, int* iters) {
  (void)eigMinTolerance;
  if (!c || !Xio || !eigval || !resval || nb < 1 || nb > c->max_bands || sbSize < 1)
    return fail(PW_ERR_ARG, "pw_ppcg: bad argument");
  if (numEig > nb) return fail(PW_ERR_ARG, "Number of eigenvalues requested  = %d which is larger than the number "
                                           "of columns in psi = %d", numEig, nb);
  PpcgOps ops;
  ops.apply_h = [c](int w, const double* x, double* y) { return pw_apply_h(c, PW_F64, w, x, y); };
  ops.precond = [c](int w, const double* r, double* y) { return pw_teter_precond(c, PW_F64, w, r, y); };
  return ppcg_core(c, 2 * c->npw, ops, nb, numEig, scfIter, eigMaxIter, eigTolerance, sbSize, Xio, eigval, resval,
                   iters);
}

extern "C" int pw_rs_ppcg(pw_rs* rs, int nb, int numEig, int scfIter, int eigMaxIter, double eigMinTolerance,
                          double eigTolerance, int sbSize, double* X, double* eigval, double* resval, int* iters) {
  (void)eigMinTolerance;
  if (!rs || !X || !eigval || !resval || nb < 1 || sbSize < 1) return fail(PW_ERR_ARG, "pw_rs_ppcg: bad argument");
  if (numEig > nb) return fail(PW_ERR_ARG, "Number of eigenvalues requested  = %d which is larger than the number "
                                           "of columns in psi = %d", numEig, nb);
  long ntot = 0;
  PW_TRY(pw_rs_info(rs, &ntot, nullptr, nullptr));
  pw_ctx* la = nullptr;
  PW_TRY(rs_linalg_ctx(rs, nb, &la));
  PpcgOps ops;
  ops.apply_h = [rs](int w, const double* x, double* y) { return pw_rs_mult_spinor(rs, w, x, y); };
  ops.precond = [rs](int w, const double* r, double* y) { return pw_rs_teter_precond(rs, w, r, y); };
  return ppcg_core(la, (int)ntot, ops, nb, numEig, scfIter, eigMaxIter, eigTolerance, sbSize, X, eigval, resval,
                   iters);
}
