// Dense symmetric eigensolver for the Rayleigh-Ritz problems (n up to a few
// thousand) on gfx950: block one-sided Jacobi (Hestenes) on B = A + sigma I.
//
// The lapack::Syevd calls of the reference (eigensolver.cpp:224-297,
// lobpcg.cpp:536-548) are latency-bound at these sizes on a GPU: a
// tridiagonal reduction needs n sequential matrix-vector steps.  Block
// one-sided Jacobi needs only (2k - 1) independent launches per sweep for 2k
// column blocks: in each step every workgroup takes one pair of 16-column
// blocks, forms their 32x32 Gram in LDS, diagonalises it with parallel cyclic
// Jacobi rotations, and rotates those 32 columns of B and of V.  Columns of
// B V become orthogonal, so V holds the eigenvectors; eigenvalues are the
// Rayleigh quotients v^T A v.  Warm Rayleigh-Ritz matrices (nearly diagonal)
// converge in a few sweeps.  sigma = -(Gershgorin lower bound) keeps B SPD.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <numeric>
#include <vector>

#include "pw_internal.h"

namespace pw {

constexpr int JB = 16;        // columns per block
constexpr int JM = 2 * JB;    // columns per pair

__device__ __forceinline__ void atomic_max_pos(double* addr, double v) {
  unsigned long long* a = reinterpret_cast<unsigned long long*>(addr);
  unsigned long long nv = __double_as_longlong(v);
  unsigned long long old = *a;
  while (__longlong_as_double(old) < v) {
    unsigned long long prev = atomicCAS(a, old, nv);
    if (prev == old) break;
    old = prev;
  }
}

// One step: workgroup w rotates block pair (pairs[2w], pairs[2w+1]).
__global__ void __launch_bounds__(256)
k_bosj_step(double* __restrict__ B, double* __restrict__ V, int n, int ld, const int* __restrict__ pairs,
            double tol, double* __restrict__ offmax) {
  __shared__ double G[JM][JM + 1];
  __shared__ double Q[JM][JM + 1];
  __shared__ int cols[JM];
  __shared__ double red[256];
  __shared__ double cs[JB], sn[JB];
  __shared__ int rp[JB], rq[JB];
  __shared__ int busy;
  const int I = pairs[2 * blockIdx.x], J = pairs[2 * blockIdx.x + 1];
  const int t = threadIdx.x;
  if (t < JM) {
    const int blk = (t < JB) ? I : J;
    const int c = blk * JB + (t % JB);
    cols[t] = (c < n) ? c : -1;
  }
  __syncthreads();
  // compact the valid columns
  __shared__ int m;
  if (t == 0) {
    int k = 0;
    for (int q = 0; q < JM; q++)
      if (cols[q] >= 0) cols[k++] = cols[q];
    m = k;
  }
  __syncthreads();
  const int mm = m;
  if (mm < 2) return;
  // Gram of the pair's columns (upper triangle, mirrored)
  const int npq = mm * (mm + 1) / 2;
  for (int e = t; e < npq; e += blockDim.x) {
    int p = 0, r = e;
    while (r >= mm - p) { r -= mm - p; p++; }
    const int q = p + r;
    const double* bp = B + (size_t)ld * cols[p];
    const double* bq = B + (size_t)ld * cols[q];
    double s = 0.0;
    for (int i = 0; i < n; i++) s += bp[i] * bq[i];
    G[p][q] = s;
    G[q][p] = s;
  }
  __syncthreads();
  double off = 0.0;
  for (int e = t; e < mm * mm; e += blockDim.x) {
    const int p = e / mm, q = e % mm;
    if (p != q) off = fmax(off, fabs(G[p][q]) / sqrt(G[p][p] * G[q][q]));
  }
  red[t] = off;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (t < w) red[t] = fmax(red[t], red[t + w]);
    __syncthreads();
  }
  const double offw = red[0];
  if (t == 0) atomic_max_pos(offmax, offw);
  if (!(offw > tol)) return;
  // Q = I; parallel cyclic Jacobi on G (round-robin over an even count)
  for (int e = t; e < JM * JM; e += blockDim.x) Q[e / JM][e % JM] = (e / JM == e % JM) ? 1.0 : 0.0;
  const int me = (mm + 1) & ~1;
  const int half = me / 2;
  for (int sweep = 0; sweep < 30; sweep++) {
    if (t == 0) busy = 0;
    __syncthreads();
    for (int round = 0; round < me - 1; round++) {
      if (t < half) {
        auto pos = [&](int j) { return j == 0 ? 0 : 1 + ((j - 1 + round) % (me - 1)); };
        int p = pos(t), q = pos(me - 1 - t);
        if (p > q) { const int x = p; p = q; q = x; }
        double c = 1.0, s = 0.0;
        if (q < mm) {
          const double gpq = G[p][q];
          if (fabs(gpq) > 1e-300 && fabs(gpq) > 1e-15 * sqrt(fabs(G[p][p] * G[q][q]))) {
            const double tau = (G[q][q] - G[p][p]) / (2.0 * gpq);
            const double tt = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + tt * tt);
            s = tt * c;
            busy = 1;
          }
        }
        rp[t] = p; rq[t] = q; cs[t] = c; sn[t] = s;
      }
      __syncthreads();
      // G <- G J and Q <- Q J (columns p, q of every row)
      for (int e = t; e < half * JM; e += blockDim.x) {
        const int i = e / JM, r = e % JM;
        const int p = rp[i], q = rq[i];
        if (q >= mm) continue;
        const double c = cs[i], s = sn[i];
        if (r < mm) {
          const double a = G[r][p], b = G[r][q];
          G[r][p] = c * a - s * b;
          G[r][q] = s * a + c * b;
        }
        const double qa = Q[r][p], qb = Q[r][q];
        Q[r][p] = c * qa - s * qb;
        Q[r][q] = s * qa + c * qb;
      }
      __syncthreads();
      // G <- J^T G (rows p, q of every column)
      for (int e = t; e < half * mm; e += blockDim.x) {
        const int i = e / mm, col = e % mm;
        const int p = rp[i], q = rq[i];
        if (q >= mm) continue;
        const double c = cs[i], s = sn[i];
        const double a = G[p][col], b = G[q][col];
        G[p][col] = c * a - s * b;
        G[q][col] = s * a + c * b;
      }
      __syncthreads();
    }
    if (!busy) break;
    __syncthreads();
  }
  // rotate the pair's columns of B and V row by row
  for (int i = t; i < n; i += blockDim.x) {
    double bv[JM], vv[JM];
#pragma unroll
    for (int q = 0; q < JM; q++) {
      bv[q] = (q < mm) ? B[i + (size_t)ld * cols[q]] : 0.0;
      vv[q] = (q < mm) ? V[i + (size_t)ld * cols[q]] : 0.0;
    }
#pragma unroll
    for (int p = 0; p < JM; p++) {
      if (p >= mm) break;
      double sb = 0.0, sv = 0.0;
#pragma unroll
      for (int q = 0; q < JM; q++) {
        sb += bv[q] * Q[q][p];
        sv += vv[q] * Q[q][p];
      }
      B[i + (size_t)ld * cols[p]] = sb;
      V[i + (size_t)ld * cols[p]] = sv;
    }
  }
}

// B = A + sigma I (upper triangle of A mirrored), V = I
__global__ void k_bosj_init(const double* __restrict__ A, int lda, int n, double sigma, double* __restrict__ B,
                            double* __restrict__ V) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)n * n; e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e % n), j = (int)(e / n);
    const double a = (i <= j) ? A[i + (size_t)lda * j] : A[j + (size_t)lda * i];
    B[e] = a + (i == j ? sigma : 0.0);
    V[e] = (i == j) ? 1.0 : 0.0;
  }
}

// row r: a_rr - sum_{j != r} |a_rj| (upper triangle mirrored)
__global__ void k_gersh(const double* __restrict__ A, int lda, int n, double* __restrict__ out) {
  for (int r = blockIdx.x * blockDim.x + threadIdx.x; r < n; r += gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int j = 0; j < n; j++) {
      if (j == r) continue;
      s += fabs((r <= j) ? A[r + (size_t)lda * j] : A[j + (size_t)lda * r]);
    }
    out[r] = A[r + (size_t)lda * r] - s;
  }
}

// theta_j = v_j . (A v)_j
__global__ void k_col_dot(const double* __restrict__ V, const double* __restrict__ AV, int n, double* __restrict__ th) {
  const int j = blockIdx.x;
  double s = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) s += V[i + (size_t)n * j] * AV[i + (size_t)n * j];
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) th[j] = red[0];
}

// Out[:, j] = V[:, perm[j]]
__global__ void k_gather_cols(const double* __restrict__ V, const int* __restrict__ perm, int n,
                              double* __restrict__ out, int ldo) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)n * n; e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e % n), j = (int)(e / n);
    out[i + (size_t)ldo * j] = V[i + (size_t)n * perm[j]];
  }
}

// Eigen-decomposition of the symmetric n x n A (upper triangle, lda):
// eigenvalues ascending to w (host), eigenvectors to Z (device, ldz).
int eig_sym(pw_ctx* c, int n, const double* A, int lda, double* w, double* Z, int ldz) {
  hipStream_t s = c->stream;
  if (n == 1) {
    PW_HIP(hipMemcpyAsync(w, A, sizeof(double), hipMemcpyDeviceToHost, s));
    const double one = 1.0;
    PW_HIP(hipMemcpyAsync(Z, &one, sizeof(double), hipMemcpyHostToDevice, s));
    PW_HIP(hipStreamSynchronize(s));
    return PW_OK;
  }
  const size_t nn = (size_t)n * n;
  const int nblk = (n + JB - 1) / JB;
  const int nb2 = nblk + (nblk & 1);
  const int steps = nb2 - 1, npair = nb2 / 2;
  // round-robin schedule of block pairs, all steps
  std::vector<int> sched((size_t)steps * npair * 2);
  {
    std::vector<int> idx(nb2);
    std::iota(idx.begin(), idx.end(), 0);
    for (int st = 0; st < steps; st++) {
      for (int i = 0; i < npair; i++) {
        sched[((size_t)st * npair + i) * 2] = idx[i];
        sched[((size_t)st * npair + i) * 2 + 1] = idx[nb2 - 1 - i];
      }
      const int last = idx[nb2 - 1];
      for (int k = nb2 - 1; k > 1; k--) idx[k] = idx[k - 1];
      idx[1] = last;
    }
  }
  double *B = nullptr, *V = nullptr, *T = nullptr, *th = nullptr, *off = nullptr;
  int *dsched = nullptr, *dperm = nullptr;
  auto done = [&](int rc) {
    void* p[] = {B, V, T, th, off, dsched, dperm};
    for (void* q : p)
      if (q) (void)hipFree(q);
    return rc;
  };
  if (hipMalloc((void**)&B, sizeof(double) * nn) != hipSuccess || hipMalloc((void**)&V, sizeof(double) * nn) != hipSuccess ||
      hipMalloc((void**)&T, sizeof(double) * nn) != hipSuccess || hipMalloc((void**)&th, sizeof(double) * n) != hipSuccess ||
      hipMalloc((void**)&off, sizeof(double)) != hipSuccess ||
      hipMalloc((void**)&dsched, sizeof(int) * sched.size()) != hipSuccess ||
      hipMalloc((void**)&dperm, sizeof(int) * n) != hipSuccess)
    return done(fail(PW_ERR_NOMEM, "eig_sym: scratch allocation failed"));
  if (hipMemcpyAsync(dsched, sched.data(), sizeof(int) * sched.size(), hipMemcpyHostToDevice, s) != hipSuccess)
    return done(fail(PW_ERR_HIP, "eig_sym: upload failed"));
  // shift: -(Gershgorin lower bound) + a margin keeps B SPD
  std::vector<double> g(n);
  hipLaunchKernelGGL(k_gersh, dim3((n + 255) / 256), dim3(256), 0, s, A, lda, n, th);
  if (hipMemcpyAsync(g.data(), th, sizeof(double) * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return done(fail(PW_ERR_HIP, "eig_sym: Gershgorin bound failed"));
  const double gmin = *std::min_element(g.begin(), g.end());
  double scale = 0.0;
  for (double x : g) scale = std::max(scale, std::fabs(x));
  const double sigma = -gmin + 1e-3 * std::max(scale, 1.0);
  hipLaunchKernelGGL(k_bosj_init, dim3((int)std::min<size_t>((nn + 255) / 256, 4096)), dim3(256), 0, s, A, lda, n,
                     sigma, B, V);
  const double tol = 1e-13;  // relative orthogonality of the columns of B V
  int sweep = 0;
  for (; sweep < 40; sweep++) {
    const double zero = 0.0;
    if (hipMemcpyAsync(off, &zero, sizeof(double), hipMemcpyHostToDevice, s) != hipSuccess)
      return done(fail(PW_ERR_HIP, "eig_sym: reset failed"));
    for (int st = 0; st < steps; st++)
      hipLaunchKernelGGL(k_bosj_step, dim3(npair), dim3(256), 0, s, B, V, n, n,
                         (const int*)(dsched + (size_t)st * npair * 2), tol, off);
    double h = 0.0;
    if (hipMemcpyAsync(&h, off, sizeof(double), hipMemcpyDeviceToHost, s) != hipSuccess ||
        hipStreamSynchronize(s) != hipSuccess)
      return done(fail(PW_ERR_HIP, "eig_sym: sweep failed"));
    if (!(h > tol)) break;
  }
  if (sweep == 40) return done(fail(PW_ERR_LINALG, "eig_sym: Jacobi did not converge in 40 sweeps"));
  // Rayleigh quotients of the converged vectors, sorted ascending: the
  // rotated B is (A + sigma I) V, so theta_j = v_j . B_j - sigma
  hipLaunchKernelGGL(k_col_dot, dim3(n), dim3(256), 0, s, (const double*)V, (const double*)B, n, th);
  std::vector<double> h(n);
  if (hipMemcpyAsync(h.data(), th, sizeof(double) * n, hipMemcpyDeviceToHost, s) != hipSuccess ||
      hipStreamSynchronize(s) != hipSuccess)
    return done(fail(PW_ERR_HIP, "eig_sym: eigenvalue readback failed"));
  for (double& x : h) x -= sigma;
  std::vector<int> perm(n);
  std::iota(perm.begin(), perm.end(), 0);
  std::stable_sort(perm.begin(), perm.end(), [&](int a, int b) { return h[a] < h[b]; });
  for (int j = 0; j < n; j++) w[j] = h[perm[j]];
  if (hipMemcpyAsync(dperm, perm.data(), sizeof(int) * n, hipMemcpyHostToDevice, s) != hipSuccess)
    return done(fail(PW_ERR_HIP, "eig_sym: permutation upload failed"));
  hipLaunchKernelGGL(k_gather_cols, dim3((int)std::min<size_t>((nn + 255) / 256, 4096)), dim3(256), 0, s,
                     (const double*)V, (const int*)dperm, n, Z, ldz);
  if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return done(fail(PW_ERR_HIP, "eig_sym: kernel failed"));
  return done(PW_OK);
}

}  // namespace pw
