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
__global__ void k_pp_sign_fix(double* __restrict__ Q, int n, int ld, long sQ) {
  double* q = Q + blockIdx.y * sQ + (size_t)blockIdx.x * ld;  // column blockIdx.x of matrix blockIdx.y
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

// PPCG block update (ppcg.cpp:462-527) of sub-block blockIdx.y, rows
// kPpRows blockIdx.x ..: with c_x, c_w, c_p the first sb columns of rows
// 0, sb and 2 sb of the sub-block's eigenvector matrix Z (ld ldz),
// p <- p c_p + w c_w (w c_w when numSet = 2), ap likewise, x <- p + x c_x,
// ax <- ap + ax c_x.  The sub-block's columns start at column off0 + sbs k
// of each m-row block; the tile's rows and the coefficients are staged in
// LDS first, so the block is updated in place.  sb <= kPpMaxSb.
constexpr int kPpRows = 32, kPpMaxSb = 32;
__global__ void __launch_bounds__(256)
k_pp_update(double* __restrict__ X, double* __restrict__ Wb, double* __restrict__ P, double* __restrict__ AX,
            double* __restrict__ AW, double* __restrict__ AP, int m, int sb, int sbs, int col0,
            const double* __restrict__ Z, long sZ, int ldz, int numSet) {
  constexpr int LD = kPpMaxSb + 1;
  __shared__ double cf[3][kPpMaxSb][LD];    // c_x, c_w, c_p (row i, column j)
  __shared__ double rw[6][kPpRows][LD];     // x, w, p, ax, aw, ap rows of the tile
  const int t = threadIdx.x, k = blockIdx.y, r0 = blockIdx.x * kPpRows;
  const int nr = min(kPpRows, m - r0);
  const double* Zk = Z + k * sZ;
  for (int e = t; e < 3 * sb * sb; e += 256) {
    const int q = e / (sb * sb), ij = e - q * sb * sb, j = ij / sb, i = ij - j * sb;
    cf[q][i][j] = Zk[(size_t)q * sb + i + (size_t)ldz * j];
  }
  const size_t off = (size_t)m * (col0 + (size_t)sbs * k);
  double* arr[6] = {X + off, Wb + off, P + off, AX + off, AW + off, AP + off};
  for (int e = t; e < 6 * sb * kPpRows; e += 256) {
    const int a = e / (sb * kPpRows), rem = e - a * sb * kPpRows, j = rem / kPpRows, r = rem - j * kPpRows;
    rw[a][r][j] = (r < nr && (numSet == 3 || (a != 2 && a != 5))) ? arr[a][r0 + r + (size_t)m * j] : 0.0;
  }
  __syncthreads();
  for (int e = t; e < sb * kPpRows; e += 256) {
    const int j = e / kPpRows, r = e - j * kPpRows;
    if (r >= nr) continue;
    double pn = 0.0, apn = 0.0, xs = 0.0, axs = 0.0;
    if (numSet == 3)
      for (int i = 0; i < sb; i++) {
        pn += rw[2][r][i] * cf[2][i][j];
        apn += rw[5][r][i] * cf[2][i][j];
      }
    for (int i = 0; i < sb; i++) {
      pn += rw[1][r][i] * cf[1][i][j];
      apn += rw[4][r][i] * cf[1][i][j];
    }
    for (int i = 0; i < sb; i++) {
      xs += rw[0][r][i] * cf[0][i][j];
      axs += rw[3][r][i] * cf[0][i][j];
    }
    const size_t g = r0 + r + (size_t)m * j;
    arr[2][g] = pn;
    arr[5][g] = apn;
    arr[0][g] = pn + xs;
    arr[3][g] = apn + axs;
  }
}

struct PpcgWs {
  double *S = nullptr, *AS = nullptr, *Xtemp = nullptr, *tmp = nullptr;
  double *XTX = nullptr, *AM = nullptr, *BM = nullptr, *eig = nullptr, *E = nullptr;
  double *nrm = nullptr, *res = nullptr, *w = nullptr, *dinv = nullptr;
  int* info = nullptr;
  void** gtab = nullptr;  // device pointer tables of the sub-block Gram products
  void free_all() {
    void* p[] = {S, AS, Xtemp, tmp, XTX, AM, BM, eig, E, nrm, res, w, dinv, info, gtab};
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
  // sub-block Gram products (ppcg.cpp:353-436) as pointer tables, built once:
  // per (numSet 2 or 3) x (full sub-blocks, the ragged last one) a list of
  // (a, b, C) with C = a^T b written into block (r0, c0) of AM or BM
  const int nfull = (width % sbs) ? nsb - 1 : nsb, sbr = width % sbs;
  struct GramList {
    std::vector<const double*> a, b;
    std::vector<double*> cc;
    size_t at = 0;  // offset in the device table
  };
  GramList gl[2][2];  // [numSet - 2][ragged]
  for (int ns = 2; ns <= 3; ns++)
    for (int k = 0; k < nsb; k++) {
      const bool rag = k >= nfull;
      const int sb = rag ? sbr : sbs, ld = 3 * sb;
      const size_t off = (size_t)m * sbs * k;
      const double *x = X + off, *w = Wb + off, *p = P + off, *ax = AX + off, *aw = AW + off, *ap = AP + off;
      double* Ak = W.AM + mat * k;
      double* Bk = W.BM + mat * k;
      GramList& L = gl[ns - 2][rag ? 1 : 0];
      auto add = [&](const double* a, const double* b, double* M, int r0, int c0) {
        L.a.push_back(a);
        L.b.push_back(b);
        L.cc.push_back(M + r0 + (size_t)ld * c0);
      };
      add(x, ax, Ak, 0, 0);
      add(w, aw, Ak, sb, sb);
      add(x, aw, Ak, 0, sb);
      add(x, x, Bk, 0, 0);
      add(w, w, Bk, sb, sb);
      add(x, w, Bk, 0, sb);
      if (ns == 3) {
        add(p, ap, Ak, 2 * sb, 2 * sb);
        add(x, ap, Ak, 0, 2 * sb);
        add(w, ap, Ak, sb, 2 * sb);
        add(p, p, Bk, 2 * sb, 2 * sb);
        add(x, p, Bk, 0, 2 * sb);
        add(w, p, Bk, sb, 2 * sb);
      }
    }
  {
    std::vector<void*> h;
    for (auto& row : gl)
      for (GramList& L : row) {
        L.at = h.size();
        for (size_t i = 0; i < L.a.size(); i++) h.push_back((void*)L.a[i]);
        for (size_t i = 0; i < L.a.size(); i++) h.push_back((void*)L.b[i]);
        for (size_t i = 0; i < L.a.size(); i++) h.push_back((void*)L.cc[i]);
      }
    if (hipMalloc((void**)&W.gtab, sizeof(void*) * std::max<size_t>(1, h.size())) != hipSuccess)
      return cleanup(fail(PW_ERR_NOMEM, "pw_ppcg: cannot allocate the Gram tables"));
    HIPTRY(hipMemcpy(W.gtab, h.data(), sizeof(void*) * h.size(), hipMemcpyHostToDevice));
  }
  auto gram = [&](int ns, int rag) -> int {
    const GramList& L = gl[ns - 2][rag];
    const int cnt = (int)L.a.size();
    if (cnt == 0) return PW_OK;
    const int sb = rag ? sbr : sbs;
    void** t = W.gtab + L.at;
    return gemm_tn_batched<double>(c, m, sb, sb, (const double* const*)t, m, (const double* const*)(t + cnt), m,
                                   (double* const*)(t + 2 * cnt), 3 * sb, cnt);
  };
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
    TRY(gram(numSet, 0));
    TRY(gram(numSet, 1));
    // all full sub-blocks in one batched Sygvd, the ragged last one alone
    {
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
    // block updates (ppcg.cpp:462-527): the sign rule on every sub-block's
    // eigenvectors, then p, ap, x, ax in one fused pass per sub-block shape
    if (nfull > 0)
      hipLaunchKernelGGL(k_pp_sign_fix, dim3(sbs, nfull), dim3(256), 0, s, W.AM, (numSet == 3) ? 3 * sbs : 2 * sbs,
                         3 * sbs, (long)mat);
    if (nfull < nsb)
      hipLaunchKernelGGL(k_pp_sign_fix, dim3(sbr, 1), dim3(256), 0, s, W.AM + mat * nfull,
                         (numSet == 3) ? 3 * sbr : 2 * sbr, 3 * sbr, 0L);
    if (sbs <= kPpMaxSb) {
      const int rt = (m + kPpRows - 1) / kPpRows;
      if (nfull > 0)
        hipLaunchKernelGGL(k_pp_update, dim3(rt, nfull), dim3(256), 0, s, X, Wb, P, AX, AW, AP, m, sbs, sbs, 0,
                           (const double*)W.AM, (long)mat, 3 * sbs, numSet);
      if (nfull < nsb)
        hipLaunchKernelGGL(k_pp_update, dim3(rt, 1), dim3(256), 0, s, X, Wb, P, AX, AW, AP, m, sbr, sbs,
                           sbs * nfull, (const double*)(W.AM + mat * nfull), 0L, 3 * sbr, numSet);
      HIPTRY(hipGetLastError());
    } else {
      for (int k = 0; k < nsb; k++) {
        const int sb = (k == nsb - 1 && width % sbs) ? width % sbs : sbs;
        const int ld = 3 * sb;
        const size_t off = (size_t)m * sbs * k;
        double* Zk = W.AM + mat * k;
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
