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
//   k_sytrd           Householder tridiagonalisation A = Q T Q^T in the
//                     order of LAPACK dsytd2 'L': n-2 steps of reflector,
//                     symmetric matrix-vector product and rank-2 update,
//                     one 1024-thread workgroup on the L2-resident matrix
//   k_stebz           all eigenvalues of T: 64-way multisection on Sturm
//                     counts, one wavefront per eigenvalue (LAPACK dstebz
//                     tolerances)
//   k_stein           eigenvectors of T: inverse iteration on LU with
//                     partial pivoting of T - lambda I, modified
//                     Gram-Schmidt within clusters of eigenvalues closer than
//                     1e-3 ||T|| (LAPACK dstein's rule), one thread per
//                     cluster, scratch interleaved by thread so that lanes
//                     stepping through the same row coalesce
//   k_ormtr           Z <- Q Z_T: every workgroup applies all reflectors to
//                     its own block of 8 columns held in LDS
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <cstdint>

#include "pw_internal.h"

namespace pw {

constexpr int kTrdThreads = 1024;  // k_sytrd workgroup
constexpr int kTrdWaves = kTrdThreads / 64;
constexpr int kOrmCols = 8;         // columns per k_ormtr workgroup
constexpr int kEigMaxN = 1024;      // LDS bound of k_sytrd

// Block-wide sum, identical order on every call (deterministic), returned to
// all threads.  sred holds one double per wavefront.
template <int NT>
__device__ __forceinline__ double block_sum(double v, double* sred) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) sred[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < NT / 64; w++) s += sred[w];
  return s;
}

__global__ void k_sym_from_upper(double* __restrict__ A, int n, int lda) {
  for (long e = blockIdx.x * (long)blockDim.x + threadIdx.x; e < (long)n * n; e += (long)gridDim.x * blockDim.x) {
    const int i = (int)(e % n), j = (int)(e / n);
    if (i > j) A[i + (size_t)lda * j] = A[j + (size_t)lda * i];
  }
}

// A (full symmetric storage) -> T = Q^T A Q: d (n), e (n-1), tau (n-1); the
// reflector v_k (v_k[0] = 1) overwrites A[k+1:n, k].  H_k = I - tau_k v v^T.
__global__ void __launch_bounds__(kTrdThreads)
k_sytrd(double* __restrict__ A, int n, int lda, double* __restrict__ d, double* __restrict__ e,
        double* __restrict__ tau) {
  extern __shared__ double sm[];
  double* sv = sm;           // v, n
  double* sw = sv + n;       // p then w, n
  double* red = sw + n;      // per-wave partial products, kTrdWaves x n
  __shared__ double sred[kTrdWaves];
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  for (int k = 0; k < n - 2; k++) {
    const int m = n - k - 1;
    double* col = A + (size_t)lda * k + (k + 1);
    double* A22 = A + (size_t)lda * (k + 1) + (k + 1);
    double s = 0.0;
    for (int i = 1 + t; i < m; i += kTrdThreads) {
      const double x = col[i];
      s += x * x;
    }
    s = block_sum<kTrdThreads>(s, sred);
    const double alpha = col[0];
    double tk = 0.0, beta = alpha, scal = 0.0;
    if (s > 0.0) {
      beta = -copysign(sqrt(alpha * alpha + s), alpha);
      tk = (beta - alpha) / beta;
      scal = 1.0 / (alpha - beta);
    }
    __syncthreads();  // every thread has read col[0] before it is overwritten
    for (int i = t; i < m; i += kTrdThreads) {
      const double vi = (i == 0) ? 1.0 : col[i] * scal;
      sv[i] = vi;
      col[i] = vi;
    }
    if (t == 0) {
      d[k] = A[(size_t)lda * k + k];
      e[k] = beta;
      tau[k] = tk;
    }
    __syncthreads();
    if (tk == 0.0) continue;  // block-uniform
    // p = tau A22 v: wave wv sums columns j = wv (mod waves), lanes run down rows
    for (int c0 = 0; c0 < m; c0 += 64) {
      const int i = c0 + lane;
      double acc = 0.0;
      if (i < m) {
#pragma unroll 4
        for (int j = wv; j < m; j += kTrdWaves) acc += A22[(size_t)lda * j + i] * sv[j];
        red[wv * n + i] = acc;
      }
    }
    __syncthreads();
    double pv = 0.0;
    for (int i = t; i < m; i += kTrdThreads) {
      double p = 0.0;
#pragma unroll
      for (int w2 = 0; w2 < kTrdWaves; w2++) p += red[w2 * n + i];
      p *= tk;
      sw[i] = p;
      pv += p * sv[i];
    }
    pv = block_sum<kTrdThreads>(pv, sred);
    const double h = 0.5 * tk * pv;
    for (int i = t; i < m; i += kTrdThreads) sw[i] -= h * sv[i];
    __syncthreads();
    // A22 -= v w^T + w v^T
    for (int c0 = 0; c0 < m; c0 += 64) {
      const int i = c0 + lane;
      if (i < m) {
        const double vi = sv[i], wi = sw[i];
#pragma unroll 4
        for (int j = wv; j < m; j += kTrdWaves) {
          double* a = A22 + (size_t)lda * j + i;
          *a -= vi * sw[j] + wi * sv[j];
        }
      }
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

// Number of eigenvalues of T below x (Sturm sequence of the LDL^T pivots).
__device__ __forceinline__ int sturm_count(const double* d, const double* e2, int n, double x, double pivmin) {
  double q = d[0] - x;
  if (fabs(q) < pivmin) q = -pivmin;
  int c = q < 0.0;
  for (int i = 1; i < n; i++) {
    q = (d[i] - x) - e2[i - 1] / q;
    if (fabs(q) < pivmin) q = -pivmin;
    c += q < 0.0;
  }
  return c;
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

// Eigenvalue j = (block * waves + wave) of T by multisection: 64 Sturm counts
// per round narrow the bracket 65-fold.  blockDim = 256.
__global__ void __launch_bounds__(256)
k_stebz(const double* __restrict__ dg, const double* __restrict__ eg, int n, double* __restrict__ w) {
  extern __shared__ double smd[];
  double* d = smd;
  double* e2 = d + n;
  __shared__ double sred[16];
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    d[i] = dg[i];
    if (i < n - 1) e2[i] = eg[i] * eg[i];
  }
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
  const int j = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (j >= n) return;  // whole wavefront
  double lo = gl, hi = gu;
  for (int it = 0; it < 48; it++) {
    if (hi - lo <= fmax(fmax(atol, pivmin), rtol * fmax(fabs(lo), fabs(hi)))) break;
    const double x = lo + (hi - lo) * (double)(lane + 1) / 65.0;
    const int cnt = sturm_count(d, e2, n, x, pivmin);
    const unsigned long long mask = __ballot(cnt >= j + 1);  // lambda_j < x
    const int first = mask ? __ffsll((long long)mask) - 1 : 64;
    const double xh = __shfl(x, first < 64 ? first : 63, 64);
    const double xl = __shfl(x, first > 0 ? first - 1 : 0, 64);
    const double nhi = first < 64 ? xh : hi;
    const double nlo = first > 0 ? xl : lo;
    lo = nlo;
    hi = nhi;
  }
  if (lane == 0) w[j] = 0.5 * (lo + hi);
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

// Eigenvectors of T (column jj of Zt, ld n) for the cluster that starts at
// thread j.  Scratch: 6 arrays of n rows, entry (a, i) of thread q at
// S[(a * n + i) * nthr + q].
__global__ void __launch_bounds__(64)
k_stein(const double* __restrict__ d, const double* __restrict__ e, const double* __restrict__ w, int n,
        double* __restrict__ Zt, double* __restrict__ S) {
  __shared__ double sred[16];
  double gl, gu, nrm, em;
  tri_bounds(d, e, n, sred, &gl, &gu, &nrm, &em);
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const double eps = DBL_EPSILON;
  const double ortol = 1e-3 * nrm;
  const double pert = eps * fmax(nrm, DBL_MIN);
  if (j > 0 && w[j] - w[j - 1] <= ortol) return;  // not a cluster start
  int j1 = j;
  while (j1 + 1 < n && w[j1 + 1] - w[j1] <= ortol) j1++;
  const int nthr = n;
  auto at = [&](int a, int i) -> double& { return S[((size_t)a * n + i) * nthr + j]; };
  // arrays: 0 u0 (pivot), 1 u1, 2 u2, 3 multiplier, 4 pivoting flag, 5 x
  double lamprev = 0.0;
  for (int jj = j; jj <= j1; jj++) {
    double lam = w[jj];
    if (jj > j && lam - lamprev < 10.0 * eps * fabs(lam)) lam = lamprev + 10.0 * eps * fabs(lam);
    lamprev = lam;
    // LU with partial pivoting of T - lam I
    double c = d[0] - lam, s1 = n > 1 ? e[0] : 0.0;
    for (int i = 0; i < n - 1; i++) {
      const double sub = e[i], dn = d[i + 1] - lam, sup = (i + 1 < n - 1) ? e[i + 1] : 0.0;
      if (fabs(sub) > fabs(c)) {
        const double mlt = c / sub;
        at(0, i) = sub; at(1, i) = dn; at(2, i) = sup; at(3, i) = mlt; at(4, i) = 1.0;
        c = s1 - mlt * dn;
        s1 = -mlt * sup;
      } else {
        if (fabs(c) < pert) c = c >= 0.0 ? pert : -pert;
        const double mlt = sub / c;
        at(0, i) = c; at(1, i) = s1; at(2, i) = 0.0; at(3, i) = mlt; at(4, i) = 0.0;
        c = dn - mlt * s1;
        s1 = sup;
      }
    }
    if (fabs(c) < pert) c = c >= 0.0 ? pert : -pert;
    at(0, n - 1) = c; at(1, n - 1) = 0.0; at(2, n - 1) = 0.0;
    for (int i = 0; i < n; i++) at(5, i) = start_value(jj, i);
    const int iters = (j1 > j) ? 5 : 3;
    for (int it = 0; it < iters; it++) {
      // scale to unit max norm, forward elimination, back substitution
      double xm = 0.0;
      for (int i = 0; i < n; i++) xm = fmax(xm, fabs(at(5, i)));
      const double inv = xm > 0.0 ? 1.0 / xm : 1.0;
      double y0 = at(5, 0) * inv;
      for (int i = 0; i < n - 1; i++) {
        const double y1 = at(5, i + 1) * inv, mlt = at(3, i);
        if (at(4, i) != 0.0) {
          at(5, i) = y1;
          y0 = y0 - mlt * y1;
        } else {
          at(5, i) = y0;
          y0 = y1 - mlt * y0;
        }
      }
      at(5, n - 1) = y0;
      double xn1 = 0.0, xn2 = 0.0;
      for (int i = n - 1; i >= 0; i--) {
        const double xi = (at(5, i) - at(1, i) * xn1 - at(2, i) * xn2) / at(0, i);
        at(5, i) = xi;
        xn2 = xn1;
        xn1 = xi;
      }
      // orthogonalise against the cluster's earlier vectors
      for (int kk = j; kk < jj; kk++) {
        const double* zk = Zt + (size_t)n * kk;
        double dot = 0.0;
        for (int i = 0; i < n; i++) dot += zk[i] * at(5, i);
        for (int i = 0; i < n; i++) at(5, i) -= dot * zk[i];
      }
      double nn2 = 0.0;
      for (int i = 0; i < n; i++) nn2 += at(5, i) * at(5, i);
      const double sc = nn2 > 0.0 ? 1.0 / sqrt(nn2) : 0.0;
      for (int i = 0; i < n; i++) at(5, i) *= sc;
    }
    double* zj = Zt + (size_t)n * jj;
    for (int i = 0; i < n; i++) zj[i] = at(5, i);
  }
}

// Z <- H_0 H_1 ... H_{n-3} Z for the kOrmCols columns of this workgroup.
__global__ void __launch_bounds__(256)
k_ormtr(const double* __restrict__ V, int ldv, const double* __restrict__ tau, int n, double* __restrict__ Z,
        int ldz) {
  extern __shared__ double zs[];  // n x kOrmCols, row-major
  __shared__ double sred[4][kOrmCols];
  __shared__ double dots[kOrmCols];
  const int c0 = blockIdx.x * kOrmCols;
  const int nc = min(kOrmCols, n - c0);
  const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
  for (int e = t; e < n * kOrmCols; e += blockDim.x) {
    const int r = e / kOrmCols, c = e % kOrmCols;
    zs[e] = (c < nc) ? Z[r + (size_t)ldz * (c0 + c)] : 0.0;
  }
  __syncthreads();
  for (int k = n - 3; k >= 0; k--) {
    const double tk = tau[k];
    if (tk == 0.0) continue;  // block-uniform
    const double* v = V + (size_t)ldv * k;  // v[r] for r in k+1..n-1, v[k+1] = 1
    double part[kOrmCols];
#pragma unroll
    for (int c = 0; c < kOrmCols; c++) part[c] = 0.0;
    for (int r = k + 1 + t; r < n; r += blockDim.x) {
      const double vr = v[r];
#pragma unroll
      for (int c = 0; c < kOrmCols; c++) part[c] += vr * zs[r * kOrmCols + c];
    }
#pragma unroll
    for (int c = 0; c < kOrmCols; c++) {
      double x = part[c];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) x += __shfl_xor(x, o, 64);
      if (lane == 0) sred[wv][c] = x;
    }
    __syncthreads();
    if (t < kOrmCols) dots[t] = tk * (sred[0][t] + sred[1][t] + sred[2][t] + sred[3][t]);
    __syncthreads();
    for (int r = k + 1 + t; r < n; r += blockDim.x) {
      const double vr = v[r];
#pragma unroll
      for (int c = 0; c < kOrmCols; c++) zs[r * kOrmCols + c] -= vr * dots[c];
    }
    __syncthreads();
  }
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
  const size_t need = sizeof(double) * (3 * (size_t)n + nn + 6 * nn);
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
  const int g1 = (int)std::min<size_t>((nn + 255) / 256, 4096);
  hipLaunchKernelGGL(k_sym_from_upper, dim3(g1), dim3(256), 0, s, A, n, lda);
  const size_t lds_trd = sizeof(double) * (size_t)(2 + kTrdWaves) * n;
  PW_TRY(ensure_lds((const void*)k_sytrd, lds_trd + 1024));  // + its static LDS
  hipLaunchKernelGGL(k_sytrd, dim3(1), dim3(kTrdThreads), lds_trd, s, A, n, lda, d, e, tau);
  hipLaunchKernelGGL(k_stebz, dim3((n + 3) / 4), dim3(256), sizeof(double) * 2 * n, s, (const double*)d,
                     (const double*)e, n, w);
  hipLaunchKernelGGL(k_stein, dim3((n + 63) / 64), dim3(64), 0, s, (const double*)d, (const double*)e,
                     (const double*)w, n, Zt, S);
  const size_t lds_orm = sizeof(double) * (size_t)n * kOrmCols;
  PW_TRY(ensure_lds((const void*)k_ormtr, lds_orm + 1024));  // + its static LDS
  hipLaunchKernelGGL(k_ormtr, dim3((n + kOrmCols - 1) / kOrmCols), dim3(256), lds_orm, s, (const double*)A, lda,
                     (const double*)tau, n, Zt, n);
  PW_HIP(hipGetLastError());
  PW_HIP(hipMemcpy2DAsync(A, sizeof(double) * lda, Zt, sizeof(double) * n, sizeof(double) * n, n,
                          hipMemcpyDeviceToDevice, s));
  return PW_OK;
}

}  // namespace pw
