// Batched Stockham FFT engine in LDS for gfx950 (radix 4, 2, 3, 5), shared
// by the H*psi and exact-exchange kernels.  Pencils of length n with stride
// ld live in LDS; one pass reads R values per butterfly, applies directly
// computed twiddles tw[m] = exp(+2 pi i m / n) and writes the autosorted
// result to the other buffer.
#pragma once
#include <hip/hip_runtime.h>

#include <utility>

namespace pw {

template <typename T> struct cx;
template <> struct cx<double> { typedef double2 t; };
template <> struct cx<float> { typedef float2 t; };

template <typename C> __device__ __forceinline__ C cadd(C a, C b) { return C{a.x + b.x, a.y + b.y}; }
template <typename C> __device__ __forceinline__ C csub(C a, C b) { return C{a.x - b.x, a.y - b.y}; }
template <typename C> __device__ __forceinline__ C cmul(C a, C b) {
  return C{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// a * conj(b)
template <typename C> __device__ __forceinline__ C cmulc(C a, C b) {
  return C{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename C, typename T> __device__ __forceinline__ C cscale(C a, T s) { return C{a.x * s, a.y * s}; }
// multiply by (sg * i)
template <typename C, typename T> __device__ __forceinline__ C cmul_si(C a, T sg) { return C{-sg * a.y, sg * a.x}; }

// In-LDS radix-R butterfly with sign sg (-1 forward, +1 backward).
template <typename T, int R>
__device__ __forceinline__ void dft(typename cx<T>::t* v, T sg) {
  typedef typename cx<T>::t C;
  if constexpr (R == 2) {
    C a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if constexpr (R == 4) {
    C t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    C t2 = cadd(v[1], v[3]), t3 = cmul_si(csub(v[1], v[3]), sg);
    v[0] = cadd(t0, t2); v[1] = cadd(t1, t3); v[2] = csub(t0, t2); v[3] = csub(t1, t3);
  } else if constexpr (R == 3) {
    const T h = (T)0.86602540378443864676372317075294;
    C a = v[0], b = v[1], c = v[2];
    C bc = cadd(b, c);
    C t = C{a.x - (T)0.5 * bc.x, a.y - (T)0.5 * bc.y};
    C u = cmul_si(cscale(csub(b, c), h), sg);
    v[0] = cadd(a, bc); v[1] = cadd(t, u); v[2] = csub(t, u);
  } else if constexpr (R == 5) {
    const T c1 = (T)0.30901699437494742410229341718282;
    const T c2 = (T)-0.80901699437494742410229341718282;
    const T s1 = (T)0.95105651629515357211643933337938;
    const T s2 = (T)0.58778525229247312916870595463907;
    C a = v[0];
    C t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]), t3 = csub(v[1], v[4]), t4 = csub(v[2], v[3]);
    C m1 = C{a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y};
    C m2 = C{a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y};
    C n1 = cmul_si(C{s1 * t3.x + s2 * t4.x, s1 * t3.y + s2 * t4.y}, sg);
    C n2 = cmul_si(C{s2 * t3.x - s1 * t4.x, s2 * t3.y - s1 * t4.y}, sg);
    v[0] = cadd(a, cadd(t1, t2)); v[1] = cadd(m1, n1); v[4] = csub(m1, n1);
    v[2] = cadd(m2, n2); v[3] = csub(m2, n2);
  }
}

// One Stockham pass of radix R over P pencils of length n (stride ld) held in
// LDS: src -> dst.  tw[m] = exp(+2 pi i m / n), conjugated for sg < 0.
template <typename T, int R>
__device__ __forceinline__ void stockham_pass(const typename cx<T>::t* __restrict__ src,
                                              typename cx<T>::t* __restrict__ dst, int n, int P,
                                              int ld, int Ns, const typename cx<T>::t* tw, T sg) {
  typedef typename cx<T>::t C;
  const int m = n / R;
  const int tstep = n / (Ns * R);
  const int total = P * m;
  for (int t = threadIdx.x; t < total; t += blockDim.x) {
    const int p = t / m;
    const int j = t - p * m;
    const int k = j % Ns;
    const C* s = src + p * ld + j;
    C v[R];
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = s[r * m];
#pragma unroll
    for (int r = 1; r < R; r++) {
      C w = tw[r * k * tstep];  // r*k*tstep < n
      w.y *= sg;
      v[r] = cmul(v[r], w);
    }
    dft<T, R>(v, sg);
    C* d = dst + p * ld + (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; r++) d[r * Ns] = v[r];
  }
}

// Full batched 1D FFT in LDS (ping-pong a <-> b); returns the buffer holding
// the result.  Radices are taken 4 first, then 2, 3, 5 (the order of
// radix_plan on the host), derived from n so no table is indexed at run
// time.  Ends with a barrier.
template <typename T>
__device__ typename cx<T>::t* lds_fft(typename cx<T>::t* a, typename cx<T>::t* b, int n, int P,
                                      int ld, const typename cx<T>::t* tw, T sg) {
  typedef typename cx<T>::t C;
  C* src = a;
  C* dst = b;
  int Ns = 1;
  int rem = n;
  while (rem > 1) {
    int R;
    if (rem % 4 == 0) {
      R = 4;
      stockham_pass<T, 4>(src, dst, n, P, ld, Ns, tw, sg);
    } else if (rem % 2 == 0) {
      R = 2;
      stockham_pass<T, 2>(src, dst, n, P, ld, Ns, tw, sg);
    } else if (rem % 3 == 0) {
      R = 3;
      stockham_pass<T, 3>(src, dst, n, P, ld, Ns, tw, sg);
    } else {
      R = 5;
      stockham_pass<T, 5>(src, dst, n, P, ld, Ns, tw, sg);
    }
    __syncthreads();
    C* tmp = src;
    src = dst;
    dst = tmp;
    Ns *= R;
    rem /= R;
  }
  return src;
}

// Compile-time-length version: the radix sequence (4s, then 2, 3s, 5s) is
// unrolled at compile time, so every stride, twiddle step and index division
// inside the passes is a constant.
template <int REM>
struct first_radix {
  static constexpr int value = (REM % 4 == 0) ? 4 : (REM % 2 == 0) ? 2 : (REM % 3 == 0) ? 3 : 5;
};

template <typename T, int N, int NS, int REM>
__device__ __forceinline__ typename cx<T>::t* lds_fft_ct_rec(typename cx<T>::t* src, typename cx<T>::t* dst, int P,
                                                              int ld, const typename cx<T>::t* tw, T sg) {
  if constexpr (REM <= 1) {
    (void)dst; (void)P; (void)ld; (void)tw; (void)sg;
    return src;
  } else {
    constexpr int R = first_radix<REM>::value;
    stockham_pass<T, R>(src, dst, N, P, ld, NS, tw, sg);
    __syncthreads();
    return lds_fft_ct_rec<T, N, NS * R, REM / R>(dst, src, P, ld, tw, sg);
  }
}

// N > 0: compile-time length; N == 0: the runtime engine with length n.
template <typename T, int N>
__device__ __forceinline__ typename cx<T>::t* lds_fft_n(typename cx<T>::t* a, typename cx<T>::t* b, int n, int P,
                                                         int ld, const typename cx<T>::t* tw, T sg) {
  if constexpr (N > 0) {
    (void)n;
    return lds_fft_ct_rec<T, N, 1, N>(a, b, P, ld, tw, sg);
  } else {
    return lds_fft<T>(a, b, n, P, ld, tw, sg);
  }
}

// y transforms of the nk columns held in Y (column stride n2).  cg == 0:
// one batched transform ping-ponging with Yalt, result in either buffer.
// cg > 0: groups of cg columns through the scratch S (cg * n2 entries),
// copied back, so the result is always in Y.
template <typename T, int N2>
__device__ __forceinline__ typename cx<T>::t* slab_y_fft(typename cx<T>::t* Y, typename cx<T>::t* Yalt,
                                                         typename cx<T>::t* S, int n2, int nk, int cg,
                                                         const typename cx<T>::t* tw2, T sg) {
  typedef typename cx<T>::t C;
  if (cg == 0) return lds_fft_n<T, N2>(Y, Yalt, n2, nk, n2, tw2, sg);
  for (int g0 = 0; g0 < nk; g0 += cg) {
    const int nc = min(cg, nk - g0);
    C* Yg = Y + g0 * n2;
    C* R = lds_fft_n<T, N2>(Yg, S, n2, nc, n2, tw2, sg);
    if (R != Yg) {
      for (int t = threadIdx.x; t < nc * n2; t += blockDim.x) Yg[t] = R[t];
      __syncthreads();
    }
  }
  return Y;
}

// x-pass helpers of the slab transforms (N1 > 0: compile-time row length
// N1 = n1; N1 = 0: runtime)
// Rows y1 = 2q, y2 = 2q+1 (q in [q0, q0+nq)) of the y-transformed spectrum
// Y[kx * n2 + y] (kx <= kxmax) -> real rows packed as Z = u1 + i u2.
template <typename T, int N1 = 0>
__device__ typename cx<T>::t* x_c2r_rows(const typename cx<T>::t* Y, int kxmax, int n1_, int n2, int q0, int nq,
                                         typename cx<T>::t* Z0, typename cx<T>::t* Z1,
                                         const typename cx<T>::t* tw1) {
  typedef typename cx<T>::t C;
  const int n1 = N1 > 0 ? N1 : n1_;
  const int h = n1 / 2;
  for (int t = threadIdx.x; t < nq * n1; t += blockDim.x) {
    const int ql = t / n1, k = t - ql * n1;
    const int y1 = 2 * (q0 + ql), y2 = y1 + 1;
    C a1 = C{0, 0}, a2 = C{0, 0};
    int kk = k;
    bool cj = false;
    if (k > h) { kk = n1 - k; cj = true; }
    if (kk <= kxmax) {
      a1 = Y[kk * n2 + y1];
      if (y2 < n2) a2 = Y[kk * n2 + y2];
      if (cj) { a1.y = -a1.y; a2.y = -a2.y; }
      if (k == 0 || (2 * k == n1)) { a1.y = 0; a2.y = 0; }
    }
    Z0[ql * n1 + k] = C{a1.x - a2.y, a1.y + a2.x};
  }
  __syncthreads();
  return lds_fft_n<T, N1>(Z0, Z1, n1, nq, n1, tw1, (T)1);
}

// Real rows packed in Zr -> forward x FFT -> spectra of both rows for
// kx <= kxmax written to Y[kx * n2 + y].
template <typename T, int N1 = 0>
__device__ void x_r2c_rows(typename cx<T>::t* Zr, typename cx<T>::t* Zo, typename cx<T>::t* Y, int kxmax, int n1_,
                           int n2, int q0, int nq, const typename cx<T>::t* tw1) {
  typedef typename cx<T>::t C;
  const int n1 = N1 > 0 ? N1 : n1_;
  const int nk = kxmax + 1;
  C* Zf = lds_fft_n<T, N1>(Zr, Zo, n1, nq, n1, tw1, (T)-1);
  for (int t = threadIdx.x; t < nq * nk; t += blockDim.x) {
    const int ql = t / nk, k = t - ql * nk;
    const int y1 = 2 * (q0 + ql), y2 = y1 + 1;
    const C w = Zf[ql * n1 + k];
    const C wc = Zf[ql * n1 + (k == 0 ? 0 : n1 - k)];
    Y[k * n2 + y1] = C{(T)0.5 * (w.x + wc.x), (T)0.5 * (w.y - wc.y)};
    if (y2 < n2) Y[k * n2 + y2] = C{(T)0.5 * (w.y + wc.y), (T)0.5 * (wc.x - w.x)};
  }
  __syncthreads();
}

}  // namespace pw

namespace pw {

// Compile-time twiddles: cos and sin of 2 pi m / n.  The quadrant is taken
// exactly in integers (rotations by pi/2 are exact), the remaining angle
// a in [0, pi/2) goes through Taylor series to x^33 (truncation < 1e-20),
// so the constants are correctly rounded to within about an ulp.
namespace ctw {
constexpr double kHalfPi = 1.57079632679489661923132169163975144;
constexpr double sin_t(double x) {
  double x2 = x * x, term = x, sum = x;
  for (int k = 1; k <= 16; k++) {
    term *= -x2 / ((2.0 * k) * (2.0 * k + 1.0));
    sum += term;
  }
  return sum;
}
constexpr double cos_t(double x) {
  double x2 = x * x, term = 1.0, sum = 1.0;
  for (int k = 1; k <= 16; k++) {
    term *= -x2 / ((2.0 * k - 1.0) * (2.0 * k));
    sum += term;
  }
  return sum;
}
constexpr void cs2pi(long m, long n, double& c, double& s) {
  m %= n;
  if (m < 0) m += n;
  const long q = (4 * m) / n, rem = 4 * m - q * n;
  const double a = (double)rem / (double)n * kHalfPi;
  const double ca = cos_t(a), sa = sin_t(a);
  if (q == 0) { c = ca; s = sa; }
  else if (q == 1) { c = -sa; s = ca; }
  else if (q == 2) { c = -ca; s = -sa; }
  else { c = sa; s = -ca; }
}
// exp(2 pi i r STEP / N) for r = 0 .. R-1
template <int N, int R, int STEP>
struct TwR {
  double c[R], s[R];
  constexpr TwR() : c(), s() {
    for (int r = 0; r < R; r++) cs2pi((long)r * STEP, N, c[r], s[r]);
  }
};
}  // namespace ctw

// In-register FFT of compile-time length N held by one thread: the Stockham
// passes of lds_fft (radix 4, 2, 3, 5 order) with the butterfly index J, the
// stride and the twiddles all compile-time constants, so the data stays in
// registers and every twiddle is an immediate.  SG = +1 backward (exp(+i)),
// -1 forward.
template <typename T, int N, int NS, int R, int SG, int J>
__device__ __forceinline__ void reg_bfly(const typename cx<T>::t (&a)[N], typename cx<T>::t (&b)[N]) {
  typedef typename cx<T>::t C;
  constexpr int M = N / R, K = J % NS, TSTEP = N / (NS * R);
  C v[R];
#pragma unroll
  for (int r = 0; r < R; r++) v[r] = a[J + r * M];
  if constexpr (K != 0) {
    constexpr ctw::TwR<N, R, K * TSTEP> W{};
#pragma unroll
    for (int r = 1; r < R; r++) v[r] = cmul(v[r], C{(T)W.c[r], (T)(SG * W.s[r])});
  }
  dft<T, R>(v, (T)SG);
#pragma unroll
  for (int r = 0; r < R; r++) b[(J / NS) * NS * R + K + r * NS] = v[r];
}

template <typename T, int N, int NS, int R, int SG, int... J>
__device__ __forceinline__ void reg_pass(const typename cx<T>::t (&a)[N], typename cx<T>::t (&b)[N],
                                         std::integer_sequence<int, J...>) {
  (reg_bfly<T, N, NS, R, SG, J>(a, b), ...);
}

template <typename T, int N, int NS, int REM, int SG>
__device__ __forceinline__ void reg_fft_rec(typename cx<T>::t (&a)[N]) {
  if constexpr (REM > 1) {
    typedef typename cx<T>::t C;
    constexpr int R = first_radix<REM>::value;
    C b[N];
    reg_pass<T, N, NS, R, SG>(a, b, std::make_integer_sequence<int, N / R>{});
#pragma unroll
    for (int i = 0; i < N; i++) a[i] = b[i];
    reg_fft_rec<T, N, NS * R, REM / R, SG>(a);
  }
}

template <typename T, int N, int SG>
__device__ __forceinline__ void reg_fft(typename cx<T>::t (&a)[N]) {
  reg_fft_rec<T, N, 1, N, SG>(a);
}

// Value held by the lane whose index differs from this lane's in bit 0
// (X = 1: DPP quad_perm [1,0,3,2]) or in bit 1 (X = 2: quad_perm [2,3,0,1]).
template <int X>
__device__ __forceinline__ float lane_xor(float v) {
  static_assert(X == 1 || X == 2, "lane bit 0 or 1");
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), X == 1 ? 0xB1 : 0x4E, 0xf, 0xf, true));
}
template <int X>
__device__ __forceinline__ double lane_xor(double v) {
  static_assert(X == 1 || X == 2, "lane bit 0 or 1");
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), X == 1 ? 0xB1 : 0x4E, 0xf, 0xf, true);
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), X == 1 ? 0xB1 : 0x4E, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}

// v *= w^(h K), w = exp(SG 2 pi i / N), h = 0 .. 3 (a runtime lane index;
// the three candidate twiddles are compile-time constants).
template <typename T, int N, int SG, int K>
__device__ __forceinline__ void lane_twiddle(typename cx<T>::t& v, int h) {
  typedef typename cx<T>::t C;
  if constexpr (K == 0) return;
  constexpr ctw::TwR<N, 4, K> W{};
  const T c = h == 1 ? (T)W.c[1] : (h == 2 ? (T)W.c[2] : (h == 3 ? (T)W.c[3] : (T)1));
  const T s = h == 1 ? (T)(SG * W.s[1]) : (h == 2 ? (T)(SG * W.s[2]) : (h == 3 ? (T)(SG * W.s[3]) : (T)0));
  v = cmul(v, C{c, s});
}
template <typename T, int N, int SG, int M, int... K>
__device__ __forceinline__ void lane_twiddles(typename cx<T>::t (&a)[M], int h, std::integer_sequence<int, K...>) {
  (lane_twiddle<T, N, SG, K>(a[K], h), ...);
}

// Length-N FFT of one LDS column (N contiguous complex values at col) by L
// adjacent lanes, L = 1, 2 or 4: lanes with the same t / L share col and
// must be active together; h = t % L.  L = 1: the whole transform in one
// thread's registers.  L > 1, decimation in time: lane h transforms the
// samples m L + h (length M = N / L) and multiplies its output k by
// w^(h k), w = exp(SG 2 pi i / N); the L partial spectra then combine by
// the L-point DFT across the lanes, one radix-2 stage per lane bit (bit 1,
// then bit 0) with a DPP exchange, the second stage's input on lane 3
// multiplied by exp(SG 2 pi i / 4).  Lane h = 2 b1 + b0 ends with
// X[q M + k], k < M, q = b0 (L = 2) or b1 + 2 b0 (L = 4).
//
// lanes_fft is the register part: on entry lane h holds a[m] = x[m L + h],
// on exit a[k] = X[q M + k], q returned.
template <typename T, int N, int SG, int L>
__device__ __forceinline__ int lanes_fft(typename cx<T>::t (&a)[N / L], int h) {
  typedef typename cx<T>::t C;
  static_assert(L == 1 || L == 2 || L == 4, "1, 2 or 4 lanes per column");
  static_assert(N % L == 0, "L divides N");
  constexpr int M = N / L;
  reg_fft<T, M, SG>(a);
  int q = 0;
  if constexpr (L > 1) {
    lane_twiddles<T, N, SG, M>(a, h, std::make_integer_sequence<int, M>{});
    if constexpr (L == 4) {
      const bool up = h & 2;
#pragma unroll
      for (int k = 0; k < M; k++) {
        const C o{lane_xor<2>(a[k].x), lane_xor<2>(a[k].y)};
        a[k] = up ? csub(o, a[k]) : cadd(a[k], o);
        if (h == 3) a[k] = cmul_si(a[k], (T)SG);
      }
    }
    const bool odd = h & 1;
#pragma unroll
    for (int k = 0; k < M; k++) {
      const C o{lane_xor<1>(a[k].x), lane_xor<1>(a[k].y)};
      a[k] = odd ? csub(o, a[k]) : cadd(a[k], o);
    }
    q = L == 2 ? (h & 1) : ((h >> 1) + 2 * (h & 1));
  }
  return q;
}

template <typename T, int N, int SG, int L>
__device__ __forceinline__ void col_fft_lanes(typename cx<T>::t* col, int h) {
  typedef typename cx<T>::t C;
  constexpr int M = N / L;
  C a[M];
#pragma unroll
  for (int m = 0; m < M; m++) a[m] = col[m * L + h];
  const int q = lanes_fft<T, N, SG, L>(a, h);
#pragma unroll
  for (int k = 0; k < M; k++) col[q * M + k] = a[k];
}

// Ordering point for a register-held array: every element must be final
// here (each is an operand of an empty volatile asm), and no memory access
// moves across.  Without it the scheduler hoists the loads that follow an
// in-register FFT (e.g. the potential it multiplies by) above the FFT and
// keeps both live at once, which spills at two waves per SIMD.
template <typename C, int N>
__device__ __forceinline__ void pin_regs(C (&a)[N]) {
#pragma unroll
  for (int i = 0; i < N; i++) asm volatile("" : "+v"(a[i].x), "+v"(a[i].y));
  asm volatile("" ::: "memory");
}

}  // namespace pw
