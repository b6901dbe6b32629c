// Eigensolver layer on gfx950: Lanczos spectral bounds, the scaled Chebyshev
// filter around H*psi, Cholesky-QR orthonormalisation and Rayleigh-Ritz.
//
// The reference defines no ChebFSI (FirstChebyStep / GeneralChebyStep are
// compiled out at src/SCF/scf.cpp:1306-1316, 1321-1332); the method follows
// Zhou, Saad, Tiago & Chelikowsky, J. Comput. Phys. 219 (2006) 172, with
// the k-step Lanczos upper bound of Zhou & Li, Linear Algebra Appl. 435
// (2011) 480.  Orthonormalisation is EigenSolver::Orthogonalize's Cholesky QR
// (src/EigenSolver/eigensolver.cpp:386-469).  Band blocks are handled as
// real (2 npw) x nb matrices: every overlap is Re(a^H b).
#include <algorithm>
#include <cmath>
#include <vector>

#include <rocsolver/rocsolver.h>

#include "pw_internal.h"

namespace pw {

template <typename A, typename B>
__global__ void k_convert(B* __restrict__ out, const A* __restrict__ in, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    out[i] = (B)in[i];
}

// out[j] = sum_i (X[i,j] - theta_j * Y[i,j])^2 (theta == nullptr: sum X^2); one block per column
__global__ void k_col_sumsq(const double* __restrict__ X, const double* __restrict__ Y,
                            const double* __restrict__ theta, long m, double* __restrict__ out) {
  const int j = blockIdx.x;
  const double* x = X + (size_t)j * m;
  const double* y = Y ? Y + (size_t)j * m : nullptr;
  const double th = theta ? theta[j] : 0.0;
  double s = 0.0;
  for (long i = threadIdx.x; i < m; i += blockDim.x) {
    double v = x[i];
    if (y) v -= th * y[i];
    s += v * v;
  }
  __shared__ double red[256];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[j] = red[0];
}

// Sign convention of eigenvectors: the largest-magnitude component of each
// column is made positive (first index on ties).  One block per column.
__global__ void k_sign_fix(double* __restrict__ Q, int n, int ld) {
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

// a = dot(x, y) over n reals into out[0]: one 1024-thread block, each
// thread streaming its elements in batches of 8 independent loads, then a
// fixed-order tree (deterministic).
__global__ void __launch_bounds__(1024) k_dot(const double* __restrict__ x, const double* __restrict__ y, long n,
                                              double* out) {
  double s = 0.0;
  for (long i0 = threadIdx.x; i0 < n; i0 += 8L * blockDim.x) {
    double a[8], b[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const long i = i0 + (long)u * blockDim.x;
      a[u] = i < n ? x[i] : 0.0;
      b[u] = i < n ? y[i] : 0.0;
    }
#pragma unroll
    for (int u = 0; u < 8; u++) s += a[u] * b[u];
  }
  __shared__ double red[1024];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

// f = f - a v - b vp ; then out = ||f||^2 is computed separately
__global__ void k_lanczos_update(double* __restrict__ f, const double* __restrict__ v,
                                 const double* __restrict__ vp, const double* __restrict__ ab, long n) {
  const double a = ab[0], b = ab[1];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    f[i] -= a * v[i] + b * vp[i];
}

// hist = (alpha, ||f||^2); ab[1] = ||f|| becomes the next step's beta
__global__ void k_lanczos_record(double* __restrict__ ab, double* __restrict__ hist) {
  hist[0] = ab[0];
  hist[1] = ab[2];
  ab[1] = sqrt(ab[2]);
}

__global__ void k_scale_to(double* __restrict__ out, const double* __restrict__ in, const double* __restrict__ nrm2,
                           long n) {
  const double s = 1.0 / sqrt(nrm2[0]);
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    out[i] = in[i] * s;
}

// FP32 vectors of the mixed-precision Lanczos bound: the same three
// operations, dot products accumulated in FP64.
__global__ void __launch_bounds__(1024) k_dot_f(const float* __restrict__ x, const float* __restrict__ y, long n,
                                                double* out) {
  double s = 0.0;
  for (long i0 = threadIdx.x; i0 < n; i0 += 8L * blockDim.x) {
    float a[8], b[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
      const long i = i0 + (long)u * blockDim.x;
      a[u] = i < n ? x[i] : 0.f;
      b[u] = i < n ? y[i] : 0.f;
    }
#pragma unroll
    for (int u = 0; u < 8; u++) s += (double)a[u] * (double)b[u];
  }
  __shared__ double red[1024];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) red[threadIdx.x] += red[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0];
}

__global__ void k_lanczos_update_f(float* __restrict__ f, const float* __restrict__ v, const float* __restrict__ vp,
                                   const double* __restrict__ ab, long n) {
  const float a = (float)ab[0], b = (float)ab[1];
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    f[i] -= a * v[i] + b * vp[i];
}

__global__ void k_scale_to_f(float* __restrict__ out, const float* __restrict__ in, const double* __restrict__ nrm2,
                             long n) {
  const float s = (float)(1.0 / sqrt(nrm2[0]));
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    out[i] = in[i] * s;
}

// One FP32 Lanczos vector step in a single workgroup, the vectors held in
// registers (n <= kLzMaxPer * 1024): alpha = v.f, f <- f - alpha v - beta
// v_prev (beta = ab[1]), hist = (alpha, ||f||^2), ab[1] = ||f||, and unless
// this is the last step v_prev <- f / ||f|| (the next v).  Fixed-order
// block sums (deterministic).
constexpr int kLzMaxPer = 16;
__device__ __forceinline__ double lz_block_sum(double v, double* red) {
  v = wave_sum(v);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < 16; w++) s += red[w];
  return s;
}
template <typename T>
__global__ void __launch_bounds__(1024)
k_lanczos_step(const T* __restrict__ v, const T* __restrict__ f, T* __restrict__ vp, double* __restrict__ ab,
               double* __restrict__ hist, long n, int last) {
  __shared__ double red[16];
  const int t = threadIdx.x;
  T xv[kLzMaxPer], xf[kLzMaxPer];
#pragma unroll
  for (int u = 0; u < kLzMaxPer; u++) {
    const long i = t + 1024L * u;
    xv[u] = i < n ? v[i] : (T)0;
    xf[u] = i < n ? f[i] : (T)0;
  }
  double sa = 0.0;
#pragma unroll
  for (int u = 0; u < kLzMaxPer; u++) sa += (double)xv[u] * (double)xf[u];
  const double alpha = lz_block_sum(sa, red);
  const T a = (T)alpha, b = (T)ab[1];
  double sf = 0.0;
  // v_prev in batches of four loads (registers for three full vectors would
  // not fit 1024 threads in FP64)
#pragma unroll
  for (int u0 = 0; u0 < kLzMaxPer; u0 += 4) {
    asm volatile("" ::: "memory");  // one batch of loads at a time
    T xp[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
      const long i = t + 1024L * (u0 + u);
      xp[u] = i < n ? vp[i] : (T)0;
    }
#pragma unroll
    for (int u = 0; u < 4; u++) {
      xf[u0 + u] -= a * xv[u0 + u] + b * xp[u];
      sf += (double)xf[u0 + u] * (double)xf[u0 + u];
    }
  }
  const double nf = lz_block_sum(sf, red);
  if (t == 0) {
    hist[0] = alpha;
    hist[1] = nf;
    ab[1] = sqrt(nf);
  }
  if (!last) {
    const T sc = (T)(1.0 / sqrt(nf));
#pragma unroll
    for (int u = 0; u < kLzMaxPer; u++) {
      const long i = t + 1024L * u;
      if (i < n) vp[i] = xf[u] * sc;
    }
  }
}

static inline void launch_dot(bool f32, const void* x, const void* y, long n, double* out, hipStream_t s) {
  if (f32)
    hipLaunchKernelGGL(k_dot_f, dim3(1), dim3(1024), 0, s, (const float*)x, (const float*)y, n, out);
  else
    hipLaunchKernelGGL(k_dot, dim3(1), dim3(1024), 0, s, (const double*)x, (const double*)y, n, out);
}

static inline int grid_for(long n) { return (int)std::min<long>((n + 255) / 256, 8192); }

// Symmetric eigen-decomposition in place: A (n x n, upper triangle, lda)
// <- eigenvectors (largest |component| positive), w (device) <- eigenvalues
// ascending.  The tridiagonal solver of pw_eig.hip up to n = 512, rocSOLVER
// dsyevd beyond; PW_EIGEN=rocsolver sends every n to rocSOLVER, PW_EIGEN=own
// keeps the own solver up to n = 1024.
int sym_eig(pw_ctx* c, int n, double* A, int lda, double* w, double* work, int* info) {
  hipStream_t s = c->stream;
  // own single-workgroup solver up to n = 512; rocSOLVER's multi-workgroup
  // dsyevd beyond (n = 768 / 1024: 17 / 23 ms against 37 / 79 ms).
  // PW_EIGEN=rocsolver / own moves the switch to 0 / 1024.
  const char* e = getenv("PW_EIGEN");
  const int own_max = (e && std::string(e) == "rocsolver") ? 0 : (e && std::string(e) == "own") ? 1024 : 512;
  if (n > own_max) {
    PW_BLAS(rocsolver_dsyevd(c->blas, rocblas_evect_original, rocblas_fill_upper, n, A, lda, w, work, info));
    int h = 0;
    PW_HIP(hipMemcpyAsync(&h, info, sizeof(int), hipMemcpyDeviceToHost, s));
    PW_HIP(hipStreamSynchronize(s));
    if (h != 0) return fail(PW_ERR_LINALG, "Syevd failed (info = %d)", h);
  } else {
    PW_TRY(eig_sym(c, n, A, lda, w));
  }
  hipLaunchKernelGGL(k_sign_fix, dim3(n), dim3(256), 0, s, A, n, lda);
  return PW_OK;
}

static int ensure_ws(pw_ctx* c) {
  if (c->ws_nb) return PW_OK;
  const size_t blk = sizeof(double) * 2 * (size_t)c->npw * c->max_bands;
  for (int i = 0; i < 6; i++)
    if (hipMalloc(&c->d_ws[i], blk) != hipSuccess)
      return fail(PW_ERR_NOMEM, "eigensolver workspace: cannot allocate %zu bytes", blk);
  const size_t nn = (size_t)c->max_bands * c->max_bands;
  PW_HIP(hipMalloc((void**)&c->d_S, sizeof(double) * nn));
  PW_HIP(hipMalloc((void**)&c->d_S2, sizeof(double) * nn));
  PW_HIP(hipMalloc((void**)&c->d_D, sizeof(double) * c->max_bands));
  PW_HIP(hipMalloc((void**)&c->d_E, sizeof(double) * c->max_bands));
  PW_HIP(hipMalloc((void**)&c->d_red, sizeof(double) * (std::max(c->max_bands, 128) + 8)));
  PW_HIP(hipMalloc((void**)&c->d_info, sizeof(int) * 4));
  PW_HIP(hipMalloc((void**)&c->d_chol, sizeof(double) * chol_dinv_doubles(chol_inv_max_n())));
  PW_HIP(hipMalloc((void**)&c->d_lz, sizeof(double) * 4 * 2 * (size_t)c->npw));
  PW_HIP(hipMalloc((void**)&c->d_coef, sizeof(double) * 192));
  PW_HIP(hipMalloc((void**)&c->d_coef_f, sizeof(float) * 192));
  c->ws_nb = c->max_bands;
  return PW_OK;
}

}  // namespace pw

using namespace pw;

extern "C" int pw_free_workspace(pw_ctx* c) {
  if (!c) return fail(PW_ERR_ARG, "null context");
  for (int i = 0; i < 6; i++)
    if (c->d_ws[i]) { (void)hipFree(c->d_ws[i]); c->d_ws[i] = nullptr; }
  void* p[] = {c->d_S, c->d_S2, c->d_D, c->d_E, c->d_red, c->d_info, c->d_lz, c->d_coef, c->d_coef_f, c->d_chol};
  for (void* q : p)
    if (q) (void)hipFree(q);
  c->d_S = c->d_S2 = c->d_D = c->d_E = c->d_red = c->d_lz = c->d_chol = nullptr;
  c->d_coef = nullptr;
  c->d_coef_f = nullptr;
  c->d_info = nullptr;
  c->ws_nb = 0;
  if (c->d_eig_ws) (void)hipFree(c->d_eig_ws);
  c->d_eig_ws = nullptr;
  c->eig_ws_bytes = 0;
  if (c->d_geig_ws) (void)hipFree(c->d_geig_ws);
  c->d_geig_ws = nullptr;
  c->geig_ws_bytes = 0;
  // the captured filter graph refers to the workspace: drop it too
  if (c->fg_exec) (void)hipGraphExecDestroy(c->fg_exec);
  c->fg_exec = nullptr;
  if (c->lz_exec) (void)hipGraphExecDestroy(c->lz_exec);
  c->lz_exec = nullptr;
  c->lz_key = c->lz_warm = -1;
  c->fg_key = c->fg_warm = -1;
  if (c->gstream) (void)hipStreamDestroy(c->gstream);
  c->gstream = nullptr;
  for (auto& e : c->gev)
    if (e) { (void)hipEventDestroy(e); e = nullptr; }
  return PW_OK;
}

// Capture body() (stream-ordered work on c->stream) on the context's
// private stream into *exec, replacing what it held.
template <typename Body>
static int graph_capture(pw_ctx* c, hipGraphExec_t* exec, Body body) {
  if (*exec) {
    (void)hipGraphExecDestroy(*exec);
    *exec = nullptr;
  }
  if (!c->gstream) PW_HIP(hipStreamCreateWithFlags(&c->gstream, hipStreamNonBlocking));
  if (!c->gev[0]) {
    PW_HIP(hipEventCreateWithFlags(&c->gev[0], hipEventDisableTiming));
    PW_HIP(hipEventCreateWithFlags(&c->gev[1], hipEventDisableTiming));
  }
  hipStream_t user = c->stream;
  const bool timing = c->timing;
  c->timing = false;  // no event nodes in the graph
  c->stream = c->gstream;
  hipGraph_t g = nullptr;
  int rc = PW_OK;
  if (hipStreamBeginCapture(c->gstream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
    rc = fail(PW_ERR_HIP, "hipStreamBeginCapture failed");
  } else {
    rc = body();
    const hipError_t e = hipStreamEndCapture(c->gstream, &g);
    if (rc == PW_OK && e != hipSuccess) rc = fail(PW_ERR_HIP, "hipStreamEndCapture: %s", hipGetErrorString(e));
    if (rc == PW_OK && hipGraphInstantiate(exec, g, nullptr, nullptr, 0) != hipSuccess)
      rc = fail(PW_ERR_HIP, "hipGraphInstantiate failed");
  }
  if (g) (void)hipGraphDestroy(g);
  c->stream = user;
  c->timing = timing;
  return rc;
}

// Replay: ordered after the work already queued on the context's stream,
// and the stream waits for the replay.
static int graph_launch(pw_ctx* c, hipGraphExec_t exec) {
  PW_HIP(hipEventRecord(c->gev[0], c->stream));
  PW_HIP(hipStreamWaitEvent(c->gstream, c->gev[0], 0));
  PW_HIP(hipGraphLaunch(exec, c->gstream));
  PW_HIP(hipEventRecord(c->gev[1], c->gstream));
  PW_HIP(hipStreamWaitEvent(c->stream, c->gev[1], 0));
  return PW_OK;
}

__global__ void k_zero_d(double* __restrict__ p, long n) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) p[i] = 0.0;
}

// k Lanczos steps (vectors in dtype) from the normalised copy of v0 held at
// d_lz + 3 m; alpha_j and ||f_j||^2 to hist = d_red + 4.  Three rotating
// vectors (v, v_prev, f), no copies, zeroing by kernel: only stream-ordered
// kernel work, safe to capture.
static int lanczos_body(pw_ctx* c, int dtype, int k) {
  const bool f32 = dtype == PW_F32;
  const long m = 2L * c->npw;
  const size_t es = f32 ? sizeof(float) : sizeof(double);
  char* base = (char*)c->d_lz;
  void* B[3] = {base, base + es * m, base + 2 * es * m};
  const void* v0 = base + 3 * es * m;
  double* ab = c->d_red;  // ab[0] = alpha, ab[1] = beta, ab[2] = ||f||^2
  double* hist = ab + 4;  // 2k <= 128 entries; d_red holds max(max_bands, 128) + 8
  hipStream_t s = c->stream;
  const dim3 g(grid_for(m)), b(256);
  // v_prev = 0 (m = 2 npw is even: m / 2 doubles hold m floats)
  hipLaunchKernelGGL(k_zero_d, dim3(grid_for(m)), dim3(256), 0, s, (double*)B[1], f32 ? m / 2 : m);
  hipLaunchKernelGGL(k_zero_d, dim3(1), dim3(64), 0, s, ab, 4L);
  launch_dot(f32, v0, v0, m, ab + 2, s);
  if (f32)
    hipLaunchKernelGGL(k_scale_to_f, g, b, 0, s, (float*)B[0], (const float*)v0, (const double*)(ab + 2), m);
  else
    hipLaunchKernelGGL(k_scale_to, g, b, 0, s, (double*)B[0], (const double*)v0, (const double*)(ab + 2), m);
  int iv = 0, ivp = 1;
  void* f = B[2];
  for (int j = 0; j < k; j++) {
    void* v = B[iv];
    PW_TRY(pw_apply_h(c, dtype, 1, v, f));
    if (f32 && m <= 1024L * kLzMaxPer) {
      // FP32: the vector step fused into one launch (three vectors fit one
      // workgroup's registers); the next v goes over v_prev
      hipLaunchKernelGGL(k_lanczos_step<float>, dim3(1), dim3(1024), 0, s, (const float*)v, (const float*)f,
                         (float*)B[ivp], ab, hist + 2 * j, m, j + 1 == k ? 1 : 0);
      std::swap(iv, ivp);
      continue;
    }
    launch_dot(f32, v, f, m, ab, s);
    if (f32)
      hipLaunchKernelGGL(k_lanczos_update_f, g, b, 0, s, (float*)f, (const float*)v, (const float*)B[ivp],
                         (const double*)ab, m);
    else
      hipLaunchKernelGGL(k_lanczos_update, g, b, 0, s, (double*)f, (const double*)v, (const double*)B[ivp],
                         (const double*)ab, m);
    launch_dot(f32, f, f, m, ab + 2, s);
    hipLaunchKernelGGL(k_lanczos_record, dim3(1), dim3(1), 0, s, ab, hist + 2 * j);
    if (j + 1 < k) {
      // v_prev <- v ; v <- f / beta, written over the old v_prev
      if (f32)
        hipLaunchKernelGGL(k_scale_to_f, g, b, 0, s, (float*)B[ivp], (const float*)f, (const double*)(ab + 2), m);
      else
        hipLaunchKernelGGL(k_scale_to, g, b, 0, s, (double*)B[ivp], (const double*)f, (const double*)(ab + 2), m);
      std::swap(iv, ivp);
    }
  }
  PW_HIP(hipGetLastError());
  return PW_OK;
}

// Lanczos bound in dtype (PW_F64 for the C-ABI entry point; the mixed sweep
// runs it in PW_F32): theta (k Ritz values of T_k, ascending) and the upper
// bound theta_k + |beta_k|.
static int lanczos_run(pw_ctx* c, int dtype, int k, const void* v0, double* theta, double* b_up) {
  if (!c || k < 1 || !v0 || !theta || !b_up) return fail(PW_ERR_ARG, "pw_lanczos: bad argument");
  if (k > 64) return fail(PW_ERR_ARG, "pw_lanczos: k > 64");
  PW_TRY(ensure_ws(c));
  const long m = 2L * c->npw;
  hipStream_t s = c->stream;
  if (dtype == PW_F32)
    hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(m)), dim3(256), 0, s, (float*)c->d_lz + 3 * m,
                       (const double*)v0, m);
  else
    PW_HIP(hipMemcpyAsync(c->d_lz + 3 * m, v0, sizeof(double) * m, hipMemcpyDeviceToDevice, s));
  // the steps are replayed from a graph after one eager call of this shape
  const long key = 100L * dtype + k;
  if (!c->use_graphs) {
    PW_TRY(lanczos_body(c, dtype, k));
  } else if (c->lz_exec && c->lz_key == key) {
    PW_TRY(graph_launch(c, c->lz_exec));
  } else if (c->lz_warm == key) {
    PW_TRY(graph_capture(c, &c->lz_exec, [&] { return lanczos_body(c, dtype, k); }));
    c->lz_key = key;
    PW_TRY(graph_launch(c, c->lz_exec));
  } else {
    PW_TRY(lanczos_body(c, dtype, k));
    c->lz_warm = key;
  }
  std::vector<double> al(k), be(k);
  std::vector<double> h(2 * k);
  PW_HIP(hipMemcpyAsync(h.data(), c->d_red + 4, sizeof(double) * 2 * k, hipMemcpyDeviceToHost, s));
  PW_HIP(hipStreamSynchronize(s));
  for (int j = 0; j < k; j++) {
    al[j] = h[2 * j];
    be[j] = std::sqrt(h[2 * j + 1]);
  }
  // Ritz values of T_k (symmetric tridiagonal, k <= 64): Jacobi rotations on host
  std::vector<double> T((size_t)k * k, 0.0);
  for (int j = 0; j < k; j++) {
    T[j + (size_t)k * j] = al[j];
    if (j + 1 < k) T[j + (size_t)k * (j + 1)] = T[(j + 1) + (size_t)k * j] = be[j];
  }
  for (int sweep = 0; sweep < 100; sweep++) {
    double off = 0.0;
    for (int q = 0; q < k; q++)
      for (int p = 0; p < q; p++) off += T[p + (size_t)k * q] * T[p + (size_t)k * q];
    if (off < 1e-30) break;
    for (int q = 1; q < k; q++)
      for (int p = 0; p < q; p++) {
        const double apq = T[p + (size_t)k * q];
        if (std::fabs(apq) < 1e-300) continue;
        const double app = T[p + (size_t)k * p], aqq = T[q + (size_t)k * q];
        const double tau = (aqq - app) / (2.0 * apq);
        const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
        const double cs = 1.0 / std::sqrt(1.0 + t * t), sn = t * cs;
        for (int r = 0; r < k; r++) {  // columns p, q
          const double a = T[r + (size_t)k * p], b = T[r + (size_t)k * q];
          T[r + (size_t)k * p] = cs * a - sn * b;
          T[r + (size_t)k * q] = sn * a + cs * b;
        }
        for (int r = 0; r < k; r++) {  // rows p, q
          const double a = T[p + (size_t)k * r], b = T[q + (size_t)k * r];
          T[p + (size_t)k * r] = cs * a - sn * b;
          T[q + (size_t)k * r] = sn * a + cs * b;
        }
      }
  }
  for (int j = 0; j < k; j++) theta[j] = T[j + (size_t)k * j];
  std::sort(theta, theta + k);
  *b_up = theta[k - 1] + std::fabs(be[k - 1]);
  return PW_OK;
}

extern "C" int pw_lanczos(pw_ctx* c, int k, const void* v0, double* theta, double* b_up) {
  return lanczos_run(c, PW_F64, k, v0, theta, b_up);
}

void pw::cheb_coef(double a, double b, double aL, int m, double* coef) {
  const double e = (b - a) / 2.0, cc = (b + a) / 2.0;
  double sigma = e / (aL - cc);
  const double tau = 2.0 / sigma;
  coef[0] = cc;
  coef[1] = sigma / e;
  coef[2] = 0.0;
  for (int it = 2; it <= m; it++) {
    const double sn = 1.0 / (tau - sigma);
    coef[3 * (it - 1)] = cc;
    coef[3 * (it - 1) + 1] = 2.0 * sn / e;
    coef[3 * (it - 1) + 2] = sigma * sn;
    sigma = sn;
  }
}

// Filter body on the fixed buffers: input d_ws[0], three rotating buffers
// d_ws[1..3], H Y in d_ws[4]; *result = buffer holding p_m(H) X.  Only
// stream-ordered work: safe to capture.
static int cheb_body(pw_ctx* c, int dtype, int nb, int m, void** result) {
  const long n = 2L * c->npw * nb;
  void* bufs[3] = {c->d_ws[1], c->d_ws[2], c->d_ws[3]};
  void* X = c->d_ws[0];
  void* HY = c->d_ws[4];
  hipStream_t s = c->stream;
  void* Xp = X;       // X_{k-1}
  void* Y = bufs[0];  // X_k
  (void)n;
  (void)s;
  // each step: the local terms into HY, then the projector product whose
  // epilogue forms the recurrence (no separate pass over the block)
  PW_TRY(cheb_step_apply(c, dtype, nb, X, HY, nullptr, Y, 0));
  int nxt = 1;
  for (int it = 2; it <= m; it++) {
    void* Yn = bufs[nxt];
    PW_TRY(cheb_step_apply(c, dtype, nb, Y, HY, Xp, Yn, it - 1));
    // rotate: Xp <- Y, Y <- Yn over the three scratch buffers
    Xp = Y;
    Y = Yn;
    nxt = (nxt + 1) % 3;
    if (bufs[nxt] == Xp) nxt = (nxt + 1) % 3;
    if (bufs[nxt] == Y) nxt = (nxt + 1) % 3;
  }
  PW_HIP(hipGetLastError());
  *result = Y;
  return PW_OK;
}

// Chebyshev filter of degree m on X (dtype blocks, in place).
extern "C" int pw_cheb_filter(pw_ctx* c, int dtype, int nb, int m, double a, double b, double aL, void* X) {
  if (!c || !X || nb < 1 || nb > c->max_bands || m < 1) return fail(PW_ERR_ARG, "pw_cheb_filter: bad argument");
  if (dtype != PW_F64 && dtype != PW_F32) return fail(PW_ERR_ARG, "unknown dtype %d", dtype);
  if (!(b > a) || !(aL < a)) return fail(PW_ERR_ARG, "pw_cheb_filter: need aL < a < b");
  if (m > 64) return fail(PW_ERR_ARG, "pw_cheb_filter: degree > 64");
  PW_TRY(ensure_ws(c));
  const long n = 2L * c->npw * nb;
  const size_t bytes = (dtype == PW_F64 ? 8 : 4) * (size_t)n;
  // scaled-recurrence coefficients (c, s1, s2) per step, uploaded before the
  // (possibly graph-replayed) filter body reads them
  cheb_coef(a, b, aL, m, c->h_coef);
  for (int i = 0; i < 3 * m; i++) c->h_coef_f[i] = (float)c->h_coef[i];
  hipStream_t s = c->stream;
  PW_HIP(hipMemcpyAsync(c->d_coef, c->h_coef, sizeof(double) * 3 * m, hipMemcpyHostToDevice, s));
  PW_HIP(hipMemcpyAsync(c->d_coef_f, c->h_coef_f, sizeof(float) * 3 * m, hipMemcpyHostToDevice, s));
  void* X0 = c->d_ws[0];
  if (X != X0) PW_HIP(hipMemcpyAsync(X0, X, bytes, hipMemcpyDeviceToDevice, s));
  void* result = nullptr;
  const long key = ((long)dtype << 40) | ((long)m << 32) | (long)nb;
  if (!c->use_graphs) {
    PW_TRY(cheb_body(c, dtype, nb, m, &result));
  } else if (c->fg_exec && c->fg_key == key) {
    PW_TRY(graph_launch(c, c->fg_exec));
    result = c->fg_result;
  } else if (c->fg_warm == key) {
    PW_TRY(graph_capture(c, &c->fg_exec, [&] { return cheb_body(c, dtype, nb, m, &c->fg_result); }));
    c->fg_key = key;
    PW_TRY(graph_launch(c, c->fg_exec));
    result = c->fg_result;
  } else {
    // first call of this shape runs eagerly (allocating every lazily sized
    // scratch); the next one is captured and replayed from then on
    PW_TRY(cheb_body(c, dtype, nb, m, &result));
    c->fg_warm = key;
  }
  if (X != result) PW_HIP(hipMemcpyAsync(X, result, bytes, hipMemcpyDeviceToDevice, s));
  PW_HIP(hipGetLastError());
  return PW_OK;
}

extern "C" int pw_syevd(pw_ctx* c, int n, double* A, int lda, double* w) {
  if (!c || !A || !w || n < 1 || lda < n) return fail(PW_ERR_ARG, "pw_syevd: bad argument");
  double* work = nullptr;
  int* info = nullptr;
  if (hipMalloc((void**)&work, sizeof(double) * n) != hipSuccess || hipMalloc((void**)&info, sizeof(int)) != hipSuccess) {
    if (work) (void)hipFree(work);
    return fail(PW_ERR_NOMEM, "pw_syevd: scratch allocation failed");
  }
  int rc = sym_eig(c, n, A, lda, w, work, info);
  if (hipStreamSynchronize(c->stream) != hipSuccess && rc == PW_OK) rc = fail(PW_ERR_HIP, "pw_syevd: kernel failed");
  (void)hipFree(work);
  (void)hipFree(info);
  return rc;
}

extern "C" int pw_sygvd_batched(pw_ctx* c, int n, double* A, int lda, long strideA, double* B, int ldb,
                                long strideB, double* w, long strideW, int* info, int nbatch) {
  if (!c || !A || !B || !w || !info || n < 1 || lda < n || ldb < n || nbatch < 1)
    return fail(PW_ERR_ARG, "pw_sygvd_batched: bad argument");
  int rc = sygvd_batched(c, n, A, lda, strideA, B, ldb, strideB, w, strideW, info, nbatch);
  if (hipStreamSynchronize(c->stream) != hipSuccess && rc == PW_OK)
    rc = fail(PW_ERR_HIP, "pw_sygvd_batched: kernel failed");
  return rc;
}

extern "C" int pw_set_graphs(pw_ctx* c, int enable) {
  if (!c) return fail(PW_ERR_ARG, "null context");
  c->use_graphs = enable != 0;
  return PW_OK;
}

// Cholesky QR of X (FP64, in place): S = X^T X, S = U^T U, X <- X U^{-1}.
// Up to 256 bands the factor comes from one single-workgroup kernel and
// X U^{-1} from a row-parallel block substitution; beyond, rocSOLVER potrf +
// rocBLAS trsm.
// When the overlap matrix is not positive definite the status is
// PW_ERR_LINALG and X is unspecified.
extern "C" int pw_orthogonalize(pw_ctx* c, int nb, void* X) {
  if (!c || !X || nb < 1 || nb > c->max_bands) return fail(PW_ERR_ARG, "pw_orthogonalize: bad argument");
  PW_TRY(ensure_ws(c));
  const int m = 2 * c->npw;
  const double one = 1.0, zero = 0.0;
  PW_TRY(gemm_tn<double>(c, m, nb, nb, (const double*)X, m, (const double*)X, m, c->d_S, nb, nullptr, zero));
  if (nb <= chol_inv_max_n()) {
    PW_TRY(chol_upper(c, nb, c->d_S, nb, c->d_chol, c->d_info));
    PW_TRY(trsm_right_upper(c, m, nb, (double*)X, m, c->d_S, nb, c->d_chol));
    int info = 0;
    PW_HIP(hipMemcpyAsync(&info, c->d_info, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    PW_HIP(hipStreamSynchronize(c->stream));
    if (info != 0) return fail(PW_ERR_LINALG, "Potrf: overlap matrix is not positive definite (info = %d)", info);
    return PW_OK;
  }
  PW_BLAS(rocsolver_dpotrf(c->blas, rocblas_fill_upper, nb, c->d_S, nb, c->d_info));
  int info = 0;
  PW_HIP(hipMemcpyAsync(&info, c->d_info, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  PW_HIP(hipStreamSynchronize(c->stream));
  if (info != 0) return fail(PW_ERR_LINALG, "Potrf: overlap matrix is not positive definite (info = %d)", info);
  PW_BLAS(rocblas_dtrsm(c->blas, rocblas_side_right, rocblas_fill_upper, rocblas_operation_none,
                        rocblas_diagonal_non_unit, m, nb, &one, c->d_S, nb, (double*)X, m));
  return PW_OK;
}

// HX = H X on c->stream, in FP64 or (hdtype PW_F32) through an FP32 copy.
static int rr_apply_h(pw_ctx* c, int hdtype, int nb, const void* X, double* HX) {
  const long n = 2L * c->npw * nb;
  if (hdtype == PW_F64) return pw_apply_h(c, PW_F64, nb, X, HX);
  float* Xf = (float*)c->d_ws[1];
  float* HXf = (float*)c->d_ws[2];
  hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(n)), dim3(256), 0, c->stream, Xf, (const double*)X, n);
  PW_TRY(pw_apply_h(c, PW_F32, nb, Xf, HXf));
  hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for(n)), dim3(256), 0, c->stream, HX, (const float*)HXf, n);
  PW_HIP(hipGetLastError());
  return PW_OK;
}

// Rayleigh-Ritz on an FP64 block: Cholesky QR, H X, X^T H X, eigen-
// decomposition, rotation; eig (ascending) and residual norms to host.
// hdtype selects the precision of the H X product (PW_F64 or PW_F32).
// Up to 256 bands H is applied to the block Y before its orthonormalisation
// and the factor carried over, H (Y U^{-1}) = (H Y) U^{-1}: the product runs
// on a second stream while S = Y^T Y is formed and the single-workgroup
// Cholesky factors it, then X = Y U^{-1} and HX = (H Y) U^{-1} (the
// factor's pivot check rides on the final host read).
extern "C" int pw_rayleigh_ritz(pw_ctx* c, int hdtype, int nb, void* X, double* eig, double* res) {
  if (!c || !X || !eig || !res || nb < 1 || nb > c->max_bands) return fail(PW_ERR_ARG, "pw_rayleigh_ritz: bad argument");
  const int m = 2 * c->npw;
  const long n = (long)m * nb;
  const double one = 1.0, zero = 0.0;
  hipStream_t s = c->stream;
  PW_TRY(ensure_ws(c));
  double* HX = (double*)c->d_ws[4];
  double* T = (double*)c->d_ws[5];
  const bool fork = nb <= chol_inv_max_n();
  if (fork) {
    PW_TRY(side_stream(c));
    PW_HIP(hipEventRecord(c->ev_fork, s));
    PW_HIP(hipStreamWaitEvent(c->side, c->ev_fork, 0));
    c->stream = c->side;
    const int rc = rr_apply_h(c, hdtype, nb, X, HX);
    c->stream = s;
    PW_TRY(rc);
    PW_HIP(hipEventRecord(c->ev_join, c->side));
    // S = Y^T Y concurrently, its split-K partials in T (free until the
    // rotation) since H Y's projector product uses the context's buffer
    PW_TRY(gemm_tn_ws<double>(c, m, nb, nb, (const double*)X, m, (const double*)X, m, c->d_S, nb, nullptr, zero, 0,
                              T, sizeof(double) * (size_t)n));
    PW_TRY(chol_upper(c, nb, c->d_S, nb, c->d_chol, c->d_info + 1));  // word 0 is the eigensolver's
    PW_HIP(hipStreamWaitEvent(s, c->ev_join, 0));
    // X U^-1 and (H Y) U^-1 in one launch (batch index 0: X, 1: HX at the
    // element offset between the two blocks; the same U and Dinv): twice the
    // workgroups of one 13134-row solve, which alone leaves CUs idle
    PW_TRY(trsm_right_upper_batched(c, m, nb, (double*)X, m, (long)(HX - (double*)X), c->d_S, nb, 0L, c->d_chol,
                                    0L, 2));
  } else {
    PW_TRY(pw_orthogonalize(c, nb, X));
    PW_TRY(rr_apply_h(c, hdtype, nb, X, HX));
  }
  PW_TRY(gemm_tn<double>(c, m, nb, nb, (const double*)X, m, HX, m, c->d_S, nb, nullptr, zero));
  PW_TRY(sym_eig(c, nb, c->d_S, nb, c->d_D, c->d_E, c->d_info));
  PW_TRY(gemm_nn<double>(c, m, nb, nb, one, (const double*)X, m, c->d_S, nb, zero, T, m));
  PW_HIP(hipMemcpyAsync(X, T, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
  PW_TRY(gemm_nn<double>(c, m, nb, nb, one, HX, m, c->d_S, nb, zero, T, m));
  hipLaunchKernelGGL(k_col_sumsq, dim3(nb), dim3(256), 0, s, (const double*)T, (const double*)X,
                     (const double*)c->d_D, (long)m, c->d_red);
  std::vector<double> r2(nb);
  int info = 0;
  PW_HIP(hipMemcpyAsync(eig, c->d_D, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
  PW_HIP(hipMemcpyAsync(r2.data(), c->d_red, sizeof(double) * nb, hipMemcpyDeviceToHost, s));
  if (fork) PW_HIP(hipMemcpyAsync(&info, c->d_info + 1, sizeof(int), hipMemcpyDeviceToHost, s));
  PW_HIP(hipStreamSynchronize(s));
  if (info != 0) return fail(PW_ERR_LINALG, "Potrf: overlap matrix is not positive definite (info = %d)", info);
  for (int j = 0; j < nb; j++) res[j] = std::sqrt(r2[j]);
  return PW_OK;
}

// One ChebFSI sweep on an FP64 block X: Lanczos(k) bounds from v0 and the
// degree-m filter, both in the precision of mode, then Cholesky QR +
// Rayleigh-Ritz (FP64 H X).
extern "C" int pw_chebfsi_sweep(pw_ctx* c, int mode, int nb, int m, int k, const void* v0,
                                const double* prev_eig, void* X, double* eig, double* res, double* bounds) {
  if (!c || !X || !v0 || !eig || !res) return fail(PW_ERR_ARG, "pw_chebfsi_sweep: null argument");
  if (mode != PW_F64 && mode != PW_F32) return fail(PW_ERR_ARG, "pw_chebfsi_sweep: unknown mode %d", mode);
  std::vector<double> theta(k > 0 ? k : 1);
  double bup = 0.0;
  PW_TRY(lanczos_run(c, mode, k, v0, theta.data(), &bup));
  double a, aL;
  if (prev_eig) {
    a = prev_eig[nb - 1];
    aL = prev_eig[0];
  } else {
    aL = theta[0];
    a = theta[0] + 0.1 * (bup - theta[0]);
  }
  if (bounds) { bounds[0] = a; bounds[1] = bup; bounds[2] = aL; }
  if (mode == PW_F64) {
    PW_TRY(pw_cheb_filter(c, PW_F64, nb, m, a, bup, aL, X));
  } else {
    const long n = 2L * c->npw * nb;
    float* Xf = (float*)c->d_ws[0];
    hipLaunchKernelGGL((k_convert<double, float>), dim3(grid_for(n)), dim3(256), 0, c->stream, Xf,
                       (const double*)X, n);
    PW_TRY(pw_cheb_filter(c, PW_F32, nb, m, a, bup, aL, Xf));
    hipLaunchKernelGGL((k_convert<float, double>), dim3(grid_for(n)), dim3(256), 0, c->stream, (double*)X,
                       (const float*)Xf, n);
  }
  return pw_rayleigh_ritz(c, PW_F64, nb, X, eig, res);
}
