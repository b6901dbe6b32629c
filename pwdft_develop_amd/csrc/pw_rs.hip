// Real-space dual-grid Gamma-point H*psi (Use_Real_Space = 1, FFT_Number_Type
// "even"/"power"): KohnSham::MultSpinor(Spinor&, NumTns<Real>&, Fourier&)
// (src/SCF/hamiltonian.cpp:3058-3125) with the real-a3 Spinor::
// AddMultSpinorFineR2C (src/SCF/spinor.cpp:376-525), IP_c2f / IP_f2c
// (src/utility.cpp:3150-3276) and the sparse real-space projectors of
// PeriodTable::CalculateNonlocalPP (src/SCF/periodtable.cpp:1042-1589).
//
// Per batch of bands (all on the handle's stream), when every grid length
// factors into 2, 3, 5 and at most one prime 7..31 (the reference's decks:
// 62 = 2*31, 124 = 4*31, 88 = 8*11, 102 = 2*3*17):
//   1. coarse R2C of psi: k_rs_x_r2c (rows), k_rs_pencil along y and z
//   2. fine C2R restricted to the coarse band: k_rs_zinv_c2f (inverse z of
//      the live columns, gathered from the coarse spectrum by IP_c2f's
//      rules), k_rs_pencil along y, k_rs_x_c2r (writing psi and V psi)
//   3. k_rs_weights: w[p, b] = vol/Nf * wgt_p * sum_i val_p[i] psi_b[idx[i]]
//   4. k_rs_nl_apply once per colour of a colouring of the atoms by shared
//      points (V psi += projector terms); k_rs_vmul + k_rs_local (V psi and
//      the projector terms through a point-major index of the entries) when
//      an atom lists a point twice
//   5. fine R2C keeping the coarse band: k_rs_x_r2c, k_rs_pencil along y and
//      z (live rows and planes only)
//   6. k_rs_f2c_ekin_p: IP_f2c as a gather on that pruned spectrum plus the
//      kinetic term
//   7. coarse C2R: k_rs_pencil along z and y, k_rs_x_c2r straight into a3.
// Other lengths are refused (PW_ERR_UNSUPPORTED).  FFTWExecute's vol/N and 1/vol factors
// (src/SCF/fourier.cpp:695-760) cancel pairwise and are folded into the
// gathers as 1/Nc and 1/Nf.

#include <algorithm>
#include <map>
#include <mutex>

#include "pw_fft.h"
#include "pw_internal.h"

using namespace pw;

namespace {

// Per-atom projector table: points, first global projector, projector count,
// offset of the atom's npts x nproj value block.
struct AtomRec {
  int npts, pg0, np, vo;
};

constexpr int kRowThreads = 64;
constexpr int kNlBands = 16;  // bands per k_rs_nl_apply workgroup

}  // namespace

struct pw_rs {
  int nc[3], nf[3];
  int hc0, hf0;             // n0/2 + 1 of the coarse and the fine grid
  long Nc, Nf, Hc, Hf;      // real and half-spectrum sizes per band
  double vol;
  int chunk;
  hipStream_t stream;
  int device;
  double* d_ekin = nullptr;
  double* d_teter = nullptr;    // b / (b + 16 a^4), a = 2 ekin (CalculateEkin)
  double* d_vtot = nullptr;
  // projectors
  int natom = 0, nproj_tot = 0, npmax = 0;
  AtomRec* d_atab = nullptr;
  int* d_apt_off = nullptr;   // natom + 1 offsets into d_idx
  int* d_idx = nullptr;
  double* d_val = nullptr;
  double* d_wgt = nullptr;
  int* d_rowptr = nullptr;    // Nf + 1
  int2* d_ent = nullptr;      // (atom, local point) sorted by point, then atom
  // atoms in colour order: no two atoms of one colour share a grid point
  // and no atom lists a point twice (else ncolor = 0 and k_rs_local runs)
  int ncolor = 0;
  std::vector<int> color_off;  // ncolor + 1 offsets into d_clist
  int* d_clist = nullptr;
  // chunk buffers
  double2* d_xc = nullptr;
  double* d_pf = nullptr;
  double* d_coef = nullptr;
  void* d_work = nullptr;
  size_t work_bytes = 0;
  // dual-grid transforms (every length factors into 2, 3, 5, 7 ... 31)
  int ny = 0, nz = 0;           // live fine rows / planes
  double2* d_tz = nullptr;      // [z][l][i]  f2 x ny x hc
  double2* d_tf = nullptr;      // [z][y][i]  f2 x f1 x hc
  double2* d_u0 = nullptr;      // [m][l][i]  nz x ny x hc
  double2* d_cv = nullptr;      // coarse [z][y][i] scratch
  double2* d_cw = nullptr;
  double* d_pv = nullptr;       // V psi on the fine grid (fused local term)
  double* d_f = nullptr;        // density weights of a batch, then sums
  double* d_part = nullptr;     // partial sums (kSumBlocks)
  int* d_ymap = nullptr;        // fine y -> live row or -1 (f1)
  int* d_zmap = nullptr;        // fine z -> live plane or -1 (f2)
  int* d_lmap = nullptr;        // live row -> fine y (ny)
  int* d_mmap = nullptr;        // live plane -> fine z (nz)
  double* d_cs = nullptr;       // prime-radix cos / sin table
  std::map<int, double2*> d_tw; // exp(+2 pi i m / n), per length
  pw_ctx* la = nullptr;         // linear algebra of the solvers (rs_linalg_ctx)
};

// ---------------------------------------------------------------- kernels

// Step 4: projector weights of one atom for BB bands.  Grid (natom, band
// groups), 256 threads striding over the atom's points; NPMAX x BB partial
// sums per thread, reduced across each wavefront with DPP and across the
// four wavefronts through LDS.  coef[p * nbc + b].
template <int NPMAX, int BB>
__global__ __launch_bounds__(256) void k_rs_weights(const double* __restrict__ pf,
                                                   const AtomRec* __restrict__ atab,
                                                   const int* __restrict__ apt_off,
                                                   const int* __restrict__ idx,
                                                   const double* __restrict__ val,
                                                   const double* __restrict__ wgt,
                                                   double* __restrict__ coef, int nbc, long Nf,
                                                   double scale) {
  __shared__ double part[4][NPMAX * BB];
  const int a = blockIdx.x;
  const int b0 = blockIdx.y * BB;
  const int nbb = min(BB, nbc - b0);
  const AtomRec rec = atab[a];
  const int* ia = idx + apt_off[a];
  const double* va = val + rec.vo;
  double acc[NPMAX][BB];
#pragma unroll
  for (int p = 0; p < NPMAX; p++)
#pragma unroll
    for (int q = 0; q < BB; q++) acc[p][q] = 0.0;
  for (int i = threadIdx.x; i < rec.npts; i += 256) {
    const long r = ia[i];
    double ps[BB];
#pragma unroll
    for (int q = 0; q < BB; q++) ps[q] = (q < nbb) ? pf[(long)(b0 + q) * Nf + r] : 0.0;
#pragma unroll
    for (int p = 0; p < NPMAX; p++) {
      if (p < rec.np) {
        const double v = va[(long)p * rec.npts + i];
#pragma unroll
        for (int q = 0; q < BB; q++) acc[p][q] = fma(v, ps[q], acc[p][q]);
      }
    }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int p = 0; p < NPMAX; p++)
#pragma unroll
    for (int q = 0; q < BB; q++) {
      const double s = wave_sum(acc[p][q]);
      if (lane == 0) part[wv][p * BB + q] = s;
    }
  __syncthreads();
  for (int t = threadIdx.x; t < rec.np * BB; t += 256) {
    const int p = t / BB, q = t % BB;
    if (q < nbb) {
      const double s = (part[0][t] + part[1][t]) + (part[2][t] + part[3][t]);
      coef[(long)(rec.pg0 + p) * nbc + b0 + q] = s * (scale * wgt[rec.pg0 + p]);
    }
  }
}

// Step 5 (general form, used when some atom lists a grid point twice): pf <-
// vtot .* pf + sum over the (atom, point) entries of each grid
// point of val_p * coef[p, b], in the reference's order (local term first,
// then atom by atom, projector by projector; spinor.cpp:437-467).  One thread
// per fine point, bands in groups of 8.
__global__ __launch_bounds__(256) void k_rs_local(double* __restrict__ pf,
                                                 const double* __restrict__ vtot,
                                                 const int* __restrict__ rowptr,
                                                 const int2* __restrict__ ent,
                                                 const AtomRec* __restrict__ atab,
                                                 const double* __restrict__ val,
                                                 const double* __restrict__ coef, int nbc, long Nf,
                                                 int has_nl) {
  const long r = (long)blockIdx.x * 256 + threadIdx.x;
  if (r >= Nf) return;
  const double v = vtot[r];
  const int e0 = has_nl ? rowptr[r] : 0, e1 = has_nl ? rowptr[r + 1] : 0;
  constexpr int G = 8;
  for (int b0 = 0; b0 < nbc; b0 += G) {
    const int nbb = min(G, nbc - b0);
    double u[G];
#pragma unroll
    for (int q = 0; q < G; q++) u[q] = (q < nbb) ? v * pf[(long)(b0 + q) * Nf + r] : 0.0;
    for (int e = e0; e < e1; e++) {
      const int2 en = ent[e];
      const AtomRec rec = atab[en.x];
      const double* vp = val + rec.vo + en.y;
      for (int p = 0; p < rec.np; p++) {
        const double dv = vp[(long)p * rec.npts];
        const double* cp = coef + (long)(rec.pg0 + p) * nbc + b0;
#pragma unroll
        for (int q = 0; q < G; q++)
          if (q < nbb) u[q] = fma(dv, cp[q], u[q]);
      }
    }
#pragma unroll
    for (int q = 0; q < G; q++)
      if (q < nbb) pf[(long)(b0 + q) * Nf + r] = u[q];
  }
}

// Step 5, split for speed when the atoms colour cleanly: pf <- vtot .* pf
// over the whole fine grid (two points and eight bands per thread) ...
__global__ __launch_bounds__(256) void k_rs_vmul(double* __restrict__ pf, const double* __restrict__ vtot,
                                                int nbc, long Nf) {
  const long r = 2 * ((long)blockIdx.x * 256 + threadIdx.x);
  if (r >= Nf) return;
  const int b0 = blockIdx.y * 8;
  const int nbb = min(8, nbc - b0);
  if (r + 1 < Nf) {
    const double2 v = *reinterpret_cast<const double2*>(vtot + r);
    double2 x[8];
#pragma unroll
    for (int q = 0; q < 8; q++)
      if (q < nbb) x[q] = *reinterpret_cast<const double2*>(pf + (long)(b0 + q) * Nf + r);
#pragma unroll
    for (int q = 0; q < 8; q++)
      if (q < nbb) {
        x[q].x *= v.x;
        x[q].y *= v.y;
        *reinterpret_cast<double2*>(pf + (long)(b0 + q) * Nf + r) = x[q];
      }
  } else {
    for (int q = 0; q < nbb; q++) pf[(long)(b0 + q) * Nf + r] *= vtot[r];
  }
}

// ... then, one launch per colour, each atom's projectors added at its own
// points: pf[b, idx_i] += sum_p val[i, p] coef[p, b].  Atoms of one colour
// share no point, so the read-modify-writes never meet.  Grid (atoms of the
// colour, groups of kNlBands bands); the atom's coefficient block sits in LDS.
template <int NPMAX>
__global__ __launch_bounds__(256) void k_rs_nl_apply(double* __restrict__ pf, const int* __restrict__ clist,
                                                    const AtomRec* __restrict__ atab,
                                                    const int* __restrict__ apt_off,
                                                    const int* __restrict__ idx,
                                                    const double* __restrict__ val,
                                                    const double* __restrict__ coef, int nbc, long Nf) {
  constexpr int BG = kNlBands;
  __shared__ double cl[NPMAX * BG];
  const int a = clist[blockIdx.x];
  const AtomRec rec = atab[a];
  const int b0 = blockIdx.y * BG;
  const int nbb = min(BG, nbc - b0);
  for (int t = threadIdx.x; t < rec.np * BG; t += 256) {
    const int p = t / BG, q = t % BG;
    cl[t] = (q < nbb) ? coef[(long)(rec.pg0 + p) * nbc + b0 + q] : 0.0;
  }
  __syncthreads();
  const int* ia = idx + apt_off[a];
  const double* va = val + rec.vo;
  for (int i = threadIdx.x; i < rec.npts; i += 256) {
    const long r = ia[i];
    double v[NPMAX];
#pragma unroll
    for (int p = 0; p < NPMAX; p++) v[p] = (p < rec.np) ? va[(long)p * rec.npts + i] : 0.0;
    // eight bands at a time: all loads issued before the first store (a
    // store may alias a later load of the same pointer, so a load-add-store
    // loop would run one round trip at a time)
    for (int q0 = 0; q0 < nbb; q0 += 8) {
      double x[8];
#pragma unroll
      for (int q = 0; q < 8; q++) x[q] = (q0 + q < nbb) ? pf[(long)(b0 + q0 + q) * Nf + r] : 0.0;
#pragma unroll
      for (int q = 0; q < 8; q++) {
        double u = 0.0;
#pragma unroll
        for (int p = 0; p < NPMAX; p++)
          if (p < rec.np) u = fma(v[p], cl[p * BG + q0 + q], u);
        x[q] += u;
      }
#pragma unroll
      for (int q = 0; q < 8; q++)
        if (q0 + q < nbb) pf[(long)(b0 + q0 + q) * Nf + r] = x[q];
    }
  }
}

// ------------------------------------------------- dual-grid transforms
//
// Batched 1D transforms in LDS along one axis of a [band][a2][a1][a0]
// array, P pencils per workgroup (row stride n), radix 4, 2, 3, 5 from
// pw_fft.h and unrolled odd-prime radices 7..31 (the decks' 62 = 2*31 and
// 124 = 4*31 points).  The fine transforms are pruned to the coarse band:
// on the way in only the kx <= Nc0/2 half-columns, the coarse rows (plus the
// Nyquist duplicate row) and planes are nonzero; on the way out only those
// are kept.  Live fine rows l = 0..ny-1 are the coarse rows j = l (fine
// index j or Nf1 - Nc1 + j) and, when Nf1 > Nc1, l = Nc1 for cjF = Nf1 -
// Nc1/2; live planes likewise.

__host__ __device__ constexpr int cs_off(int R) {
  return R == 7 ? 0 : R == 11 ? 14 : R == 13 ? 36 : R == 17 ? 62 : R == 19 ? 96 : R == 23 ? 134 : R == 29 ? 180 : 238;
}
constexpr int kCsLen = 300;   // cos, sin of 2 pi m / R, m < R, for the eight primes

// Output pairs q = Q0..Q1 of an odd-prime DFT of length R: v[r] holds
// v_r + v_{R-r} and v[R-r] holds v_r - v_{R-r} (r = 1..(R-1)/2), x0 = v_0.
// X[q] and X[R-q] share the cos and sin sums, which halves the multiplies.
// cs holds cos, sin of 2 pi m / R; the indices are compile-time constants,
// so the table is read through scalar loads.
template <int R, int Q0, int Q1>
__device__ __forceinline__ void rs_prime_pairs(const double2* v, double2 x0, double sg,
                                               const double* __restrict__ cs, double2* d, int ns) {
  constexpr int h = (R - 1) / 2;
#pragma unroll
  for (int q = Q0; q <= Q1; q++) {
    double ax = x0.x, ay = x0.y, bx = 0.0, by = 0.0;
#pragma unroll
    for (int r = 1; r <= h; r++) {
      const int m = (r * q) % R;
      const double c = cs[2 * m], sn = cs[2 * m + 1];
      ax = fma(c, v[r].x, ax);
      ay = fma(c, v[r].y, ay);
      bx = fma(sn, v[R - r].x, bx);
      by = fma(sn, v[R - r].y, by);
    }
    d[q * ns] = make_double2(ax - sg * by, ay + sg * bx);
    d[(R - q) * ns] = make_double2(ax + sg * by, ay - sg * bx);
  }
}

// Stockham pass of odd prime radix R >= 7: every butterfly is shared by the
// four wavefronts of the (256-thread) workgroup, wavefront w computing the
// output pairs of its quarter of q = 1..(R-1)/2 (lane = butterfly), so all
// four wavefronts work even when a workgroup holds only ~64 butterflies.
template <int R>
__device__ __forceinline__ void rs_pass_prime(const double2* __restrict__ src, double2* __restrict__ dst, int n,
                                              int P, int ld, int Ns, const double2* tw, double sg,
                                              const double* __restrict__ cs) {
  constexpr int h = (R - 1) / 2;
  constexpr int QS = (h + 3) / 4;
  const int m = n / R;
  const int tstep = n / (Ns * R);
  const int total = P * m;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  for (int u = threadIdx.x & 63; u < total; u += 64) {
    const int p = u / m;
    const int j = u - p * m;
    const int k = j % Ns;
    const double2* s = src + p * ld + j;
    double2 v[R];
#pragma unroll
    for (int r = 0; r < R; r++) v[r] = s[r * m];
#pragma unroll
    for (int r = 1; r < R; r++) {
      double2 t = tw[r * k * tstep];
      t.y *= sg;
      v[r] = cmul(v[r], t);
    }
    const double2 x0 = v[0];
    double2 sum = v[0];
#pragma unroll
    for (int r = 1; r <= h; r++) {
      const double2 a = v[r], b = v[R - r];
      v[r] = make_double2(a.x + b.x, a.y + b.y);
      v[R - r] = make_double2(a.x - b.x, a.y - b.y);
      sum.x += v[r].x;
      sum.y += v[r].y;
    }
    double2* d = dst + p * ld + (j / Ns) * Ns * R + k;
    switch (w) {
      case 0:
        d[0] = sum;
        rs_prime_pairs<R, 1, (QS < h ? QS : h)>(v, x0, sg, cs, d, Ns);
        break;
      case 1:
        rs_prime_pairs<R, QS + 1, (2 * QS < h ? 2 * QS : h)>(v, x0, sg, cs, d, Ns);
        break;
      case 2:
        rs_prime_pairs<R, 2 * QS + 1, (3 * QS < h ? 3 * QS : h)>(v, x0, sg, cs, d, Ns);
        break;
      default:
        rs_prime_pairs<R, 3 * QS + 1, h>(v, x0, sg, cs, d, Ns);
        break;
    }
  }
}

// One Stockham pass of radix R (2, 3, 4, 5) over P pencils of length n (row
// stride ld); odd primes go to rs_pass_prime.
template <int R>
__device__ __forceinline__ void rs_pass(const double2* __restrict__ src, double2* __restrict__ dst, int n, int P,
                                        int ld, int Ns, const double2* tw, double sg,
                                        const double* __restrict__ cs) {
  if constexpr (R > 5) {
    rs_pass_prime<R>(src, dst, n, P, ld, Ns, tw, sg, cs + cs_off(R));
  } else {
    const int m = n / R;
    const int tstep = n / (Ns * R);
    const int total = P * m;
    for (int t = threadIdx.x; t < total; t += blockDim.x) {
      const int p = t / m;
      const int j = t - p * m;
      const int k = j % Ns;
      const double2* s = src + p * ld + j;
      double2 v[R];
#pragma unroll
      for (int r = 0; r < R; r++) v[r] = s[r * m];
#pragma unroll
      for (int r = 1; r < R; r++) {
        double2 t2 = tw[r * k * tstep];
        t2.y *= sg;
        v[r] = cmul(v[r], t2);
      }
      dft<double, R>(v, sg);
      double2* d = dst + p * ld + (j / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; r++) d[r * Ns] = v[r];
    }
  }
}

// Full transform (ping-pong a <-> b) with radices 4, 2, 3, 5 and the odd
// prime PR (0: none); returns the buffer holding the result.  Ends with a
// barrier.  One prime per instantiation keeps each kernel's registers to
// what its largest butterfly needs.
template <int PR>
__device__ __forceinline__ double2* lds_fft_any(double2* a, double2* b, int n, int P, int ld, const double2* tw,
                                                double sg, const double* __restrict__ cs) {
  double2* src = a;
  double2* dst = b;
  int Ns = 1, rem = n;
  while (rem > 1) {
    int R;
    if (rem % 4 == 0) {
      R = 4;
      rs_pass<4>(src, dst, n, P, ld, Ns, tw, sg, cs);
    } else if (rem % 2 == 0) {
      R = 2;
      rs_pass<2>(src, dst, n, P, ld, Ns, tw, sg, cs);
    } else if (rem % 3 == 0) {
      R = 3;
      rs_pass<3>(src, dst, n, P, ld, Ns, tw, sg, cs);
    } else if (rem % 5 == 0) {
      R = 5;
      rs_pass<5>(src, dst, n, P, ld, Ns, tw, sg, cs);
    } else {
      if constexpr (PR > 0) {
        R = PR;
        rs_pass<PR>(src, dst, n, P, ld, Ns, tw, sg, cs);
      } else {
        R = rem;  // lengths are checked on the host: never reached
      }
    }
    __syncthreads();
    double2* tmp = src;
    src = dst;
    dst = tmp;
    Ns *= R;
    rem /= R;
  }
  return src;
}

__device__ __forceinline__ void load_tw(double2* tw, const double2* __restrict__ g, int n) {
  for (int t = threadIdx.x; t < n; t += blockDim.x) tw[t] = g[t];
}

// Pass A: inverse z transform of the fine spectrum's live columns, the
// columns gathered from the coarse spectrum by IP_c2f's rules (utility.cpp:
// 3150-3212) and scaled by 1/Nc.  Pencil g = l * hc + i; out tz[z][l][i].
template <int PR>
__global__ __launch_bounds__(256) void k_rs_zinv_c2f(const double2* __restrict__ xc, double2* __restrict__ tz,
                                                    const double2* __restrict__ twg,
                                                    const double* __restrict__ cs, int c0, int c1, int c2,
                                                    int f2, int hc, int ny, int P, long Hc, long Tzb,
                                                    double scale) {
  extern __shared__ double2 sm[];
  const int n = f2, ld = n;
  double2* tw = sm;
  double2* A = sm + n;
  double2* B = A + P * ld;
  load_tw(tw, twg, n);
  const long b = blockIdx.y;
  const int g0 = blockIdx.x * P;
  const int npen = ny * hc;
  const int In = c0 / 2, Jn = c1 / 2, Kn = c2 / 2;
  // all global loads of the phase issued before the first LDS store
  constexpr int EPT = 8;  // P * n < 2048 by the choice of P
  double2 vt[EPT];
#pragma unroll
  for (int it = 0; it < EPT; it++) {
    const int t = threadIdx.x + it * 256;
    const int p = t % P, e = t / P, g = g0 + p;
    double2 v = make_double2(0.0, 0.0);
    if (t < P * n && g < npen) {
      const int l = g / hc, i = g - l * hc;
      const int j = l < c1 ? l : Jn;
      const int dj = l < c1 ? 0 : 1;
      int k, dk;
      if (e <= Kn) { k = e; dk = 0; }
      else if (e > f2 - Kn) { k = e - f2 + c2; dk = 0; }
      else if (e == f2 - Kn) { k = Kn; dk = 1; }
      else { k = 0; dk = -1; }
      if (dk >= 0) {
        const bool ni = (i == In), nj = (j == Jn), nk = (k == Kn);
        double f;
        if (dj == 0 && dk == 0) f = (ni || nj || nk) ? 0.5 : 1.0;
        else if (dj == 1 && dk == 0) f = (!ni && !nk) ? 0.5 : 0.0;
        else if (dj == 0) f = (!ni && !nj) ? 0.5 : 0.0;
        else f = !ni ? 0.5 : 0.0;
        if (f != 0.0) {
          const double2 x = xc[b * Hc + ((long)k * c1 + j) * hc + i];
          v = make_double2(x.x * (f * scale), x.y * (f * scale));
        }
      }
    }
    vt[it] = v;
  }
#pragma unroll
  for (int it = 0; it < EPT; it++) {
    const int t = threadIdx.x + it * 256;
    if (t < P * n) A[(t % P) * ld + t / P] = vt[it];
  }
  __syncthreads();
  const double2* res = lds_fft_any<PR>(A, B, n, P, ld, tw, 1.0, cs);
  for (int t = threadIdx.x; t < P * n; t += blockDim.x) {
    const int p = t % P, e = t / P, g = g0 + p;
    if (g < npen) tz[b * Tzb + (long)e * npen + g] = res[p * ld + e];
  }
}

// A complex pass along one axis: pencil g = u * hc + i, element e at
// in[b*in_bs + u*in_su + in_map[e]*in_se + i] (in_map NULL: e; -1: zero),
// output element o <- transform element out_map[o] (NULL: o).
struct PenArgs {
  const double2* in;
  long in_bs, in_su, in_se;
  const int* in_map;
  double2* out;
  long out_bs, out_su, out_se;
  const int* out_map;
  int nout, U, hc, n, P;
  double sg;
  const double2* tw;
};

template <int PR>
__global__ __launch_bounds__(256) void k_rs_pencil(PenArgs a, const double* __restrict__ cs) {
  extern __shared__ double2 sm[];
  const int n = a.n, ld = n, P = a.P;
  double2* tw = sm;
  double2* A = sm + n;
  double2* B = A + P * ld;
  load_tw(tw, a.tw, n);
  const long b = blockIdx.y;
  const int g0 = blockIdx.x * P;
  const int npen = a.U * a.hc;
  constexpr int EPT = 8;  // P * n < 2048 by the choice of P
  double2 vt[EPT];
#pragma unroll
  for (int it = 0; it < EPT; it++) {
    const int t = threadIdx.x + it * 256;
    const int p = t % P, e = t / P, g = g0 + p;
    double2 v = make_double2(0.0, 0.0);
    if (t < P * n && g < npen) {
      const int u = g / a.hc, i = g - u * a.hc;
      const int s = a.in_map ? a.in_map[e] : e;
      if (s >= 0) v = a.in[b * a.in_bs + u * a.in_su + s * a.in_se + i];
    }
    vt[it] = v;
  }
#pragma unroll
  for (int it = 0; it < EPT; it++) {
    const int t = threadIdx.x + it * 256;
    if (t < P * n) A[(t % P) * ld + t / P] = vt[it];
  }
  __syncthreads();
  const double2* res = lds_fft_any<PR>(A, B, n, P, ld, tw, a.sg, cs);
  for (int t = threadIdx.x; t < P * a.nout; t += blockDim.x) {
    const int p = t % P, o = t / P, g = g0 + p;
    if (g < npen) {
      const int u = g / a.hc, i = g - u * a.hc;
      const int e = a.out_map ? a.out_map[o] : o;
      a.out[b * a.out_bs + u * a.out_su + o * a.out_se + i] = res[p * ld + e];
    }
  }
}

// Real rows along x, two rows per complex transform (row r0 in the real
// part, r1 = r0 + 1 in the imaginary part).  C2R: input rows hold hw
// half-spectrum entries (zero beyond), Hermitian-extended to length n.
template <int PR>
__global__ __launch_bounds__(256) void k_rs_x_c2r(const double2* __restrict__ in, double* __restrict__ out,
                                                 const double2* __restrict__ twg, const double* __restrict__ cs,
                                                 int n, int hw, int npair, long in_bs, long out_bs, int P,
                                                 const double* __restrict__ vloc, double* __restrict__ out2) {
  extern __shared__ double2 sm[];
  const int ld = n;
  double2* tw = sm;
  double2* A = sm + n;
  double2* B = A + P * ld;
  load_tw(tw, twg, n);
  const long b = blockIdx.y;
  const int g0 = blockIdx.x * P;
  constexpr int EPT = 8;  // P * n < 2048 by the choice of P
  double2 vt[EPT];
#pragma unroll
  for (int it = 0; it < EPT; it++) {
    const int t = threadIdx.x + it * 256;
    const int e = t % n, p = t / n, g = g0 + p;
    double2 z = make_double2(0.0, 0.0);
    if (t < P * n && g < npair) {
      const double2* r0 = in + b * in_bs + (long)(2 * g) * hw;
      const double2* r1 = r0 + hw;
      // one load path: entry k of each half row, conjugated above n/2
      const int k = e < hw ? e : ((2 * e > n && n - e < hw) ? n - e : -1);
      if (k >= 0) {
        const double sg = e < hw ? 1.0 : -1.0;
        const double2 a0 = r0[k], a1 = r1[k];
        z = make_double2(a0.x - sg * a1.y, sg * a0.y + a1.x);
      }
    }
    vt[it] = z;
  }
#pragma unroll
  for (int it = 0; it < EPT; it++) {
    const int t = threadIdx.x + it * 256;
    if (t < P * n) A[(t / n) * ld + t % n] = vt[it];
  }
  __syncthreads();
  const double2* res = lds_fft_any<PR>(A, B, n, P, ld, tw, 1.0, cs);
  // two neighbouring x per lane: 16-byte stores (n even, so rows are
  // 16-byte aligned)
  const int h = n / 2;
  for (int t = threadIdx.x; t < P * h; t += blockDim.x) {
    const int e = 2 * (t % h), p = t / h, g = g0 + p;
    if (g < npair) {
      const double2 z0 = res[p * ld + e], z1 = res[p * ld + e + 1];
      const long off = (long)(2 * g) * n + e;
      double2* o = reinterpret_cast<double2*>(out + b * out_bs + off);
      o[0] = make_double2(z0.x, z1.x);
      o[h] = make_double2(z0.y, z1.y);
      if (vloc) {
        // V psi beside psi (the local-potential term of the fine grid)
        const double2 v0 = *reinterpret_cast<const double2*>(vloc + off);
        const double2 v1 = *reinterpret_cast<const double2*>(vloc + off + n);
        double2* o2 = reinterpret_cast<double2*>(out2 + b * out_bs + off);
        o2[0] = make_double2(v0.x * z0.x, v0.y * z1.x);
        o2[h] = make_double2(v1.x * z0.y, v1.y * z1.y);
      }
    }
  }
}

// R2C of real rows along x, two rows per complex transform, keeping the
// first hw entries of each row's half spectrum.
template <int PR>
__global__ __launch_bounds__(256) void k_rs_x_r2c(const double* __restrict__ in, double2* __restrict__ out,
                                                 const double2* __restrict__ twg, const double* __restrict__ cs,
                                                 int n, int hw, int npair, long in_bs, long out_bs, int P) {
  extern __shared__ double2 sm[];
  const int ld = n;
  double2* tw = sm;
  double2* A = sm + n;
  double2* B = A + P * ld;
  load_tw(tw, twg, n);
  const long b = blockIdx.y;
  const int g0 = blockIdx.x * P;
  constexpr int EPT = 8;  // P * n < 2048 by the choice of P
  double2 vt[EPT];
#pragma unroll
  for (int it = 0; it < EPT; it++) {
    const int t = threadIdx.x + it * 256;
    const int e = t % n, p = t / n, g = g0 + p;
    double2 z = make_double2(0.0, 0.0);
    if (t < P * n && g < npair) {
      const double* r0 = in + b * in_bs + (long)(2 * g) * n;
      z = make_double2(__builtin_nontemporal_load(r0 + e), __builtin_nontemporal_load(r0 + n + e));
    }
    vt[it] = z;
  }
#pragma unroll
  for (int it = 0; it < EPT; it++) {
    const int t = threadIdx.x + it * 256;
    if (t < P * n) A[(t / n) * ld + t % n] = vt[it];
  }
  __syncthreads();
  const double2* res = lds_fft_any<PR>(A, B, n, P, ld, tw, -1.0, cs);
  for (int t = threadIdx.x; t < P * hw; t += blockDim.x) {
    const int k = t % hw, p = t / hw, g = g0 + p;
    if (g < npair) {
      const double2 zk = res[p * ld + k];
      const double2 zn = res[p * ld + (k == 0 ? 0 : n - k)];
      double2* o = out + b * out_bs + (long)(2 * g) * hw;
      o[k] = make_double2(0.5 * (zk.x + zn.x), 0.5 * (zk.y - zn.y));
      o[hw + k] = make_double2(0.5 * (zk.y + zn.y), -0.5 * (zk.x - zn.x));
    }
  }
}

// Pass G: k_rs_f2c_ekin on the pruned fine spectrum u0[m][l][i] (live
// planes m, live rows l); ymap / zmap send fine indices to live ones.
__global__ __launch_bounds__(kRowThreads) void k_rs_f2c_ekin_p(double2* __restrict__ xc,
                                                              const double2* __restrict__ u0,
                                                              const double* __restrict__ ekin,
                                                              const int* __restrict__ ymap,
                                                              const int* __restrict__ zmap, int c0, int c1,
                                                              int c2, int f1, int f2, int hc, int ny, long Hc,
                                                              long U0b, double sc, double sf) {
  const int row = blockIdx.x;
  const int j = row % c1, k = row / c1;
  const long b = blockIdx.y;
  const int Jn = c1 / 2, Kn = c2 / 2, In = c0 / 2;
  const int jF = (j <= Jn) ? j : f1 - c1 + j;
  const int kF = (k <= Kn) ? k : f2 - c2 + k;
  const int cjF = f1 - Jn, ckF = f2 - Kn;
  const int mjF = (f1 - jF) % f1, mkF = (f2 - kF) % f2;
  const bool nj = (j == Jn), nk = (k == Kn);
  const double2* fb = u0 + b * U0b;
  auto rowp = [&](int kk, int jj) -> const double2* {
    const int m = zmap[kk], l = ymap[jj];
    return (m >= 0 && l >= 0) ? fb + ((long)m * ny + l) * hc : nullptr;
  };
  const double2* f1row = rowp(kF, jF);
  const double2* f2row = nullptr;
  if (nj && nk) f2row = rowp(ckF, cjF);
  else if (nj) f2row = rowp(kF, cjF);
  else if (nk) f2row = rowp(ckF, jF);
  const double2* g2row;
  if (nj) g2row = rowp(mkF, jF);
  else if (nk) g2row = rowp(kF, mjF);
  else g2row = rowp(mkF, mjF);
  const double2 zero = make_double2(0.0, 0.0);
  double2* crow = xc + b * Hc + (long)row * hc;
  const double* erow = ekin + (long)row * hc;
  for (int i = threadIdx.x; i < hc; i += kRowThreads) {
    const double2 f1v = f1row ? f1row[i] : zero;
    double2 acc;
    if (i == In) {
      if (nj && nk) {
        acc = make_double2(f1v.x, 0.0);
      } else {
        const double2 f2v = g2row ? g2row[i] : zero;
        acc = make_double2(0.5 * (f1v.x + f2v.x), 0.5 * (f1v.y - f2v.y));
      }
    } else if (nj || nk) {
      const double2 f2v = f2row ? f2row[i] : zero;
      acc = make_double2(0.5 * (f1v.x + f2v.x), 0.5 * (f1v.y + f2v.y));
    } else {
      acc = f1v;
    }
    const double2 x = crow[i];
    const double e = erow[i] * sc;
    crow[i] = make_double2(fma(e, x.x, acc.x * sf), fma(e, x.y, acc.y * sf));
  }
}

// Wavefunction ecut filter on the coarse spectrum (hamiltonian.cpp:3104-3109
// and :3254-3259): entries with |G|^2/2 (ekin) above the cutoff are zeroed,
// the rest scaled by 1/Nc (the forward vol/N and backward 1/vol factors).
__global__ __launch_bounds__(256) void k_rs_cutmask(double2* __restrict__ xc, const double* __restrict__ ekin,
                                                   double cutoff, long Hc, double scale) {
  const long g = (long)blockIdx.x * 256 + threadIdx.x;
  if (g >= Hc) return;
  const long b = blockIdx.y;
  const double f = ekin[g] > cutoff ? 0.0 : scale;
  double2 v = xc[b * Hc + g];
  v.x *= f;
  v.y *= f;
  xc[b * Hc + g] = v;
}

// Density on the fine grid (KohnSham::CalculateDensity, real-space branch,
// hamiltonian.cpp:2453-2627): rho (+)= sum_b f_b psiF_b^2, bands in order.
__global__ __launch_bounds__(256) void k_rs_rho_acc(const double* __restrict__ pf, const double* __restrict__ f,
                                                   double* __restrict__ rho, int nbc, long Nf, int first) {
  const long r = (long)blockIdx.x * 256 + threadIdx.x;
  if (r >= Nf) return;
  double acc = first ? 0.0 : rho[r];
  for (int b = 0; b < nbc; b++) {
    const double v = pf[(long)b * Nf + r];
    acc += v * v * f[b];
  }
  rho[r] = acc;
}

// Deterministic sum of n doubles: each of the 1024 workgroups sums its
// stride-1024 share (fixed order, DPP wavefront sums), then one workgroup
// adds the partials in order.  scale != 0: x *= scale[0] first (in place).
constexpr int kSumBlocks = 1024;
__global__ __launch_bounds__(256) void k_rs_sum_partial(double* __restrict__ x, long n,
                                                       const double* __restrict__ scale, double* __restrict__ part) {
  __shared__ double sw[4];
  const double sc = scale ? scale[0] : 1.0;
  double acc = 0.0;
  for (long i = (long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long)kSumBlocks * 256) {
    double v = x[i];
    if (scale) {
      v *= sc;
      x[i] = v;
    }
    acc += v;
  }
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = (sw[0] + sw[1]) + (sw[2] + sw[3]);
}

// out[0] = sum of the partials; with want_scale: out[1] = c / out[0] (the
// renormalisation factor), else out[1] = out[0] * c.
__global__ __launch_bounds__(256) void k_rs_sum_final(const double* __restrict__ part, double c, int want_scale,
                                                     double* __restrict__ out) {
  __shared__ double sw[4];
  double acc = 0.0;
  for (int i = threadIdx.x; i < kSumBlocks; i += 256) acc += part[i];
  acc = wave_sum(acc);
  if ((threadIdx.x & 63) == 0) sw[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    const double s = (sw[0] + sw[1]) + (sw[2] + sw[3]);
    out[0] = s;
    out[1] = want_scale ? c / s : s * c;
  }
}

// Teter preconditioner on the coarse spectrum (Spinor::AddTeterPrecond,
// real version, spinor.cpp:309-345): xc *= teter / Nc.
__global__ __launch_bounds__(256) void k_rs_teter_spec(double2* __restrict__ xc, const double* __restrict__ teter,
                                                      long Hc, double scale) {
  const long g = (long)blockIdx.x * 256 + threadIdx.x;
  if (g >= Hc) return;
  const long b = blockIdx.y;
  const double f = teter[g] * scale;
  double2 v = xc[b * Hc + g];
  v.x *= f;
  v.y *= f;
  xc[b * Hc + g] = v;
}

// y += x over n doubles (the a3 accumulation of AddTeterPrecond)
__global__ __launch_bounds__(256) void k_rs_axpy(double* __restrict__ y, const double* __restrict__ x, long n) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i < n) y[i] += x[i];
}

// ---------------------------------------------------------------- host

namespace {

template <typename X>
int upload_rs(X** d, const X* h, size_t n) {
  PW_HIP(hipMalloc((void**)d, sizeof(X) * (n > 0 ? n : 1)));
  if (n) PW_HIP(hipMemcpy(*d, h, sizeof(X) * n, hipMemcpyHostToDevice));
  return PW_OK;
}

void free_nonlocal(pw_rs* rs) {
  hipFree(rs->d_atab);
  hipFree(rs->d_apt_off);
  hipFree(rs->d_idx);
  hipFree(rs->d_val);
  hipFree(rs->d_wgt);
  hipFree(rs->d_rowptr);
  hipFree(rs->d_ent);
  hipFree(rs->d_clist);
  hipFree(rs->d_coef);
  rs->d_atab = nullptr;
  rs->d_apt_off = rs->d_idx = rs->d_rowptr = nullptr;
  rs->d_val = rs->d_wgt = rs->d_coef = nullptr;
  rs->d_ent = nullptr;
  rs->d_clist = nullptr;
  rs->ncolor = 0;
  rs->color_off.clear();
  rs->natom = rs->nproj_tot = rs->npmax = 0;
}

// The odd prime >= 7 in n (0 if n is 2, 3, 5-smooth), or -1 if n has a
// prime factor above 31 or two different ones >= 7.
int big_prime(int n) {
  for (int f : {2, 3, 5})
    while (n % f == 0) n /= f;
  int pr = 0;
  for (int f : {7, 11, 13, 17, 19, 23, 29, 31})
    if (n % f == 0) {
      if (pr) return -1;
      pr = f;
      while (n % f == 0) n /= f;
    }
  return n == 1 ? pr : -1;
}

bool own_length(int n) { return big_prime(n) >= 0; }

#define RS_PR_SWITCH(PRV, CALL)                                                              \
  switch (PRV) {                                                                             \
    case 0: CALL(0); break;                                                                  \
    case 7: CALL(7); break;                                                                  \
    case 11: CALL(11); break;                                                                \
    case 13: CALL(13); break;                                                                \
    case 17: CALL(17); break;                                                                \
    case 19: CALL(19); break;                                                                \
    case 23: CALL(23); break;                                                                \
    case 29: CALL(29); break;                                                                \
    case 31: CALL(31); break;                                                                \
    default: return fail(PW_ERR_UNSUPPORTED, "dual-grid transform: length without a kernel"); \
  }

// pencils per workgroup: two P x n buffers (unpadded rows, so that a
// workgroup's pencils cover whole 256-byte runs of the half rows) and the
// twiddles in 64 KiB
int pencils(int n) { return std::max(1, std::min(64, (4096 - n) / (2 * n))); }

size_t pencil_lds(int n, int P) { return sizeof(double2) * ((size_t)n + 2 * (size_t)P * n); }

void live_map(int c, int f, std::vector<int>& live, std::vector<int>& fmap) {
  // live index -> fine index, and fine index -> live index (or -1)
  const int h = c / 2;
  live.clear();
  for (int l = 0; l < c; l++) live.push_back(l <= h ? l : f - c + l);
  if (f > c) live.push_back(f - h);
  fmap.assign((size_t)f, -1);
  for (size_t l = 0; l < live.size(); l++) fmap[(size_t)live[l]] = (int)l;
}

int setup_own_fft(pw_rs* rs) {
  const int hc = rs->hc0;
  std::vector<int> ly, yf, lz, zf;
  live_map(rs->nc[1], rs->nf[1], ly, yf);
  live_map(rs->nc[2], rs->nf[2], lz, zf);
  rs->ny = (int)ly.size();
  rs->nz = (int)lz.size();
  int rc;
  if ((rc = upload_rs(&rs->d_ymap, yf.data(), yf.size())) != PW_OK ||
      (rc = upload_rs(&rs->d_zmap, zf.data(), zf.size())) != PW_OK ||
      (rc = upload_rs(&rs->d_lmap, ly.data(), ly.size())) != PW_OK ||
      (rc = upload_rs(&rs->d_mmap, lz.data(), lz.size())) != PW_OK)
    return rc;
  std::vector<double> cs(kCsLen);
  for (int R : {7, 11, 13, 17, 19, 23, 29, 31})
    for (int m = 0; m < R; m++) {
      const long double a = 2.0L * 3.14159265358979323846264338327950288L * m / R;
      cs[cs_off(R) + 2 * m] = (double)cosl(a);
      cs[cs_off(R) + 2 * m + 1] = (double)sinl(a);
    }
  if ((rc = upload_rs(&rs->d_cs, cs.data(), cs.size())) != PW_OK) return rc;
  for (int n : {rs->nc[0], rs->nc[1], rs->nc[2], rs->nf[0], rs->nf[1], rs->nf[2]}) {
    if (rs->d_tw.count(n)) continue;
    std::vector<double2> tw((size_t)n);
    for (int m = 0; m < n; m++) {
      const long double a = 2.0L * 3.14159265358979323846264338327950288L * m / n;
      tw[(size_t)m] = make_double2((double)cosl(a), (double)sinl(a));
    }
    double2* d = nullptr;
    if ((rc = upload_rs(&d, tw.data(), tw.size())) != PW_OK) return rc;
    rs->d_tw[n] = d;
  }
  const size_t ch = (size_t)rs->chunk;
  const size_t tz = (size_t)rs->nf[2] * rs->ny * hc, tf = (size_t)rs->nf[2] * rs->nf[1] * hc;
  const size_t u0 = (size_t)rs->nz * rs->ny * hc;
  if (hipMalloc((void**)&rs->d_tz, sizeof(double2) * tz * ch) != hipSuccess ||
      hipMalloc((void**)&rs->d_tf, sizeof(double2) * tf * ch) != hipSuccess ||
      hipMalloc((void**)&rs->d_u0, sizeof(double2) * u0 * ch) != hipSuccess ||
      hipMalloc((void**)&rs->d_cv, sizeof(double2) * rs->Hc * ch) != hipSuccess ||
      hipMalloc((void**)&rs->d_cw, sizeof(double2) * rs->Hc * ch) != hipSuccess ||
      hipMalloc((void**)&rs->d_pv, sizeof(double) * rs->Nf * ch) != hipSuccess)
    return fail(PW_ERR_NOMEM, "pw_rs_create: cannot allocate %d-band transform buffers", rs->chunk);
  return PW_OK;
}

bool even_grid(const int* n) { return n[0] > 0 && n[1] > 0 && n[2] > 0 && n[0] % 2 == 0 && n[1] % 2 == 0 && n[2] % 2 == 0; }

}  // namespace

extern "C" int pw_rs_create(pw_rs** out, const int grid[3], const int grid_fine[3], double vol,
                            const double* ekin, int max_bands, void* stream) {
  if (!out || !grid || !grid_fine || !ekin || max_bands <= 0 || !(vol > 0.0))
    return fail(PW_ERR_ARG, "pw_rs_create: bad argument");
  *out = nullptr;
  if (!even_grid(grid) || !even_grid(grid_fine))
    return fail(PW_ERR_UNSUPPORTED,
                "pw_rs_create: grids must be even in every dimension (FFT_Number_Type even/power)");
  for (int d = 0; d < 3; d++)
    if (grid_fine[d] < grid[d]) return fail(PW_ERR_ARG, "pw_rs_create: fine grid smaller than the coarse grid");
  for (int d = 0; d < 3; d++)
    if (!own_length(grid[d]) || !own_length(grid_fine[d]) || grid_fine[d] > 1024)
      return fail(PW_ERR_UNSUPPORTED,
                  "pw_rs_create: grid %d x %d x %d / %d x %d x %d: every length must factor into 2, 3, 5 and at "
                  "most one prime 7..31 (fine lengths up to 1024)",
                  grid[0], grid[1], grid[2], grid_fine[0], grid_fine[1], grid_fine[2]);
  pw_rs* rs = new pw_rs();
  for (int d = 0; d < 3; d++) {
    rs->nc[d] = grid[d];
    rs->nf[d] = grid_fine[d];
  }
  rs->hc0 = grid[0] / 2 + 1;
  rs->hf0 = grid_fine[0] / 2 + 1;
  rs->Nc = (long)grid[0] * grid[1] * grid[2];
  rs->Nf = (long)grid_fine[0] * grid_fine[1] * grid_fine[2];
  rs->Hc = (long)rs->hc0 * grid[1] * grid[2];
  rs->Hf = (long)rs->hf0 * grid_fine[1] * grid_fine[2];
  if (rs->Nf >= (1L << 31) - 1) {
    delete rs;
    return fail(PW_ERR_UNSUPPORTED, "pw_rs_create: fine grid too large");
  }
  rs->vol = vol;
  rs->stream = (hipStream_t)stream;
  // bands per chunk: chunk buffers within about 8 GiB of the 288 GB HBM
  const double hc = rs->hc0, ny = grid[1] + 1, nz = grid[2] + 1;
  const double per_band = 48.0 * rs->Hc + 16.0 * rs->Nf +
                          16.0 * hc * ((double)grid_fine[2] * ny + (double)grid_fine[2] * grid_fine[1] + nz * ny);
  rs->chunk = (int)std::max(1.0, std::min((double)max_bands, 8589934592.0 / per_band));
  int rc = PW_OK;
  auto bail = [&](int code) {
    pw_rs_destroy(rs);
    return code;
  };
  if (hipGetDevice(&rs->device) != hipSuccess) return bail(fail(PW_ERR_HIP, "pw_rs_create: hipGetDevice failed"));
  if ((rc = upload_rs(&rs->d_ekin, ekin, (size_t)rs->Hc)) != PW_OK) return bail(rc);
  {
    // KohnSham::CalculateEkin's teter_ from the same ekin (hamiltonian.cpp:2659-2663)
    std::vector<double> t((size_t)rs->Hc);
    for (long g = 0; g < rs->Hc; g++) {
      const double a = ekin[g] * 2.0, b = 27.0 + a * (18.0 + a * (12.0 + a * 8.0));
      t[(size_t)g] = b / (b + 16.0 * pow(a, 4.0));
    }
    if ((rc = upload_rs(&rs->d_teter, t.data(), t.size())) != PW_OK) return bail(rc);
  }
  if (hipMalloc((void**)&rs->d_xc, sizeof(double2) * rs->Hc * rs->chunk) != hipSuccess ||
      hipMalloc((void**)&rs->d_pf, sizeof(double) * rs->Nf * rs->chunk) != hipSuccess)
    return bail(fail(PW_ERR_NOMEM, "pw_rs_create: cannot allocate %d-band chunk buffers", rs->chunk));
  if ((rc = setup_own_fft(rs)) != PW_OK) return bail(rc);
  *out = rs;
  return PW_OK;
}

extern "C" int pw_rs_destroy(pw_rs* rs) {
  if (!rs) return PW_OK;
  hipStreamSynchronize(rs->stream);
  if (rs->la) pw_ctx_destroy(rs->la);
  free_nonlocal(rs);
  hipFree(rs->d_ekin);
  hipFree(rs->d_teter);
  hipFree(rs->d_vtot);
  hipFree(rs->d_xc);
  hipFree(rs->d_pf);
  hipFree(rs->d_work);
  hipFree(rs->d_tz);
  hipFree(rs->d_tf);
  hipFree(rs->d_u0);
  hipFree(rs->d_cv);
  hipFree(rs->d_cw);
  hipFree(rs->d_pv);
  hipFree(rs->d_f);
  hipFree(rs->d_part);
  hipFree(rs->d_ymap);
  hipFree(rs->d_zmap);
  hipFree(rs->d_lmap);
  hipFree(rs->d_mmap);
  hipFree(rs->d_cs);
  for (auto& kv : rs->d_tw) hipFree(kv.second);
  delete rs;
  return PW_OK;
}

extern "C" int pw_rs_set_stream(pw_rs* rs, void* stream) {
  if (!rs) return fail(PW_ERR_ARG, "pw_rs_set_stream: null handle");
  PW_HIP(hipStreamSynchronize(rs->stream));
  rs->stream = (hipStream_t)stream;
  if (rs->la) PW_TRY(pw_ctx_set_stream(rs->la, stream));
  return PW_OK;
}

extern "C" int pw_rs_info(const pw_rs* rs, long* ntot, long* ntot_fine, int* chunk) {
  if (!rs) return fail(PW_ERR_ARG, "pw_rs_info: null handle");
  if (ntot) *ntot = rs->Nc;
  if (ntot_fine) *ntot_fine = rs->Nf;
  if (chunk) *chunk = rs->chunk;
  return PW_OK;
}

extern "C" int pw_rs_set_vtot(pw_rs* rs, const double* vtot) {
  if (!rs || !vtot) return fail(PW_ERR_ARG, "pw_rs_set_vtot: bad argument");
  PW_HIP(hipStreamSynchronize(rs->stream));
  if (!rs->d_vtot) PW_HIP(hipMalloc((void**)&rs->d_vtot, sizeof(double) * rs->Nf));
  PW_HIP(hipMemcpy(rs->d_vtot, vtot, sizeof(double) * rs->Nf, hipMemcpyHostToDevice));
  return PW_OK;
}

extern "C" int pw_rs_set_nonlocal(pw_rs* rs, int natom, const int* npts, const int* nproj,
                                  const int* idx, const double* val, const double* wgt) {
  if (!rs || natom < 0) return fail(PW_ERR_ARG, "pw_rs_set_nonlocal: bad argument");
  if (natom > 0 && (!npts || !nproj)) return fail(PW_ERR_ARG, "pw_rs_set_nonlocal: null counts");
  std::vector<AtomRec> atab((size_t)natom);
  std::vector<int> apt_off((size_t)natom + 1, 0);
  long tot_pts = 0, tot_val = 0;
  int tot_proj = 0, npmax = 0;
  for (int a = 0; a < natom; a++) {
    if (npts[a] < 0 || nproj[a] < 0) return fail(PW_ERR_ARG, "pw_rs_set_nonlocal: negative count for atom %d", a);
    if (nproj[a] > 32)
      return fail(PW_ERR_UNSUPPORTED, "pw_rs_set_nonlocal: atom %d has %d projectors (at most 32)", a, nproj[a]);
    atab[a] = AtomRec{npts[a], tot_proj, nproj[a], (int)tot_val};
    apt_off[a] = (int)tot_pts;
    tot_pts += npts[a];
    tot_val += (long)npts[a] * nproj[a];
    tot_proj += nproj[a];
    npmax = std::max(npmax, nproj[a]);
    if (tot_pts >= (1L << 31) || tot_val >= (1L << 31))
      return fail(PW_ERR_UNSUPPORTED, "pw_rs_set_nonlocal: too many projector entries");
  }
  apt_off[natom] = (int)tot_pts;
  if ((tot_pts > 0 && !idx) || (tot_val > 0 && !val) || (tot_proj > 0 && !wgt))
    return fail(PW_ERR_ARG, "pw_rs_set_nonlocal: null arrays");
  for (long i = 0; i < tot_pts; i++)
    if (idx[i] < 0 || idx[i] >= rs->Nf)
      return fail(PW_ERR_DOMAIN, "pw_rs_set_nonlocal: grid index %d out of the fine grid", idx[i]);
  // point-major index of (atom, local point), atoms ascending per point
  std::vector<int> rowptr((size_t)rs->Nf + 1, 0);
  for (long i = 0; i < tot_pts; i++) rowptr[(size_t)idx[i] + 1]++;
  for (long r = 0; r < rs->Nf; r++) rowptr[r + 1] += rowptr[r];
  std::vector<int2> ent((size_t)tot_pts);
  {
    std::vector<int> fill(rowptr.begin(), rowptr.end() - 1);
    for (int a = 0; a < natom; a++)
      for (int i = 0; i < npts[a]; i++) {
        const int r = idx[apt_off[a] + i];
        ent[(size_t)fill[r]++] = make_int2(a, i);
      }
  }
  // greedy colouring of the atoms by shared grid points (bit c of mask[r]:
  // an atom of colour c owns r); an atom that lists a point twice, or more
  // than 32 colours, leaves the point-major kernel in charge
  int ncolor = 0;
  std::vector<int> color_of((size_t)natom, 0);
  {
    std::vector<uint32_t> mask((size_t)rs->Nf, 0u);
    std::vector<int> stamp((size_t)rs->Nf, -1);
    bool ok = true;
    for (int a = 0; a < natom && ok; a++) {
      uint32_t used = 0;
      for (int i = 0; i < npts[a]; i++) {
        const int r = idx[apt_off[a] + i];
        if (stamp[r] == a) { ok = false; break; }
        stamp[r] = a;
        used |= mask[r];
      }
      if (!ok || used == 0xffffffffu) { ok = false; break; }
      const int c = __builtin_ctz(~used);
      color_of[a] = c;
      ncolor = std::max(ncolor, c + 1);
      for (int i = 0; i < npts[a]; i++) mask[idx[apt_off[a] + i]] |= 1u << c;
    }
    if (!ok) ncolor = 0;
  }
  std::vector<int> clist, color_off;
  if (ncolor > 0) {
    color_off.assign((size_t)ncolor + 1, 0);
    for (int a = 0; a < natom; a++) color_off[color_of[a] + 1]++;
    for (int c = 0; c < ncolor; c++) color_off[c + 1] += color_off[c];
    clist.resize((size_t)natom);
    std::vector<int> fill(color_off.begin(), color_off.end() - 1);
    for (int a = 0; a < natom; a++) clist[(size_t)fill[color_of[a]]++] = a;
  }
  PW_HIP(hipStreamSynchronize(rs->stream));
  free_nonlocal(rs);
  int rc;
  if ((rc = upload_rs(&rs->d_atab, atab.data(), atab.size())) != PW_OK ||
      (rc = upload_rs(&rs->d_apt_off, apt_off.data(), apt_off.size())) != PW_OK ||
      (rc = upload_rs(&rs->d_idx, idx, (size_t)tot_pts)) != PW_OK ||
      (rc = upload_rs(&rs->d_val, val, (size_t)tot_val)) != PW_OK ||
      (rc = upload_rs(&rs->d_wgt, wgt, (size_t)tot_proj)) != PW_OK ||
      (rc = upload_rs(&rs->d_rowptr, rowptr.data(), rowptr.size())) != PW_OK ||
      (rc = upload_rs(&rs->d_ent, ent.data(), ent.size())) != PW_OK ||
      (rc = upload_rs(&rs->d_clist, clist.data(), clist.size())) != PW_OK) {
    free_nonlocal(rs);
    return rc;
  }
  if (tot_proj > 0 &&
      hipMalloc((void**)&rs->d_coef, sizeof(double) * (size_t)tot_proj * rs->chunk) != hipSuccess) {
    free_nonlocal(rs);
    return fail(PW_ERR_NOMEM, "pw_rs_set_nonlocal: cannot allocate projector weights");
  }
  rs->natom = natom;
  rs->nproj_tot = tot_proj;
  rs->npmax = npmax;
  rs->ncolor = ncolor;
  rs->color_off = color_off;
  return PW_OK;
}

namespace {

int launch_pencil(pw_rs* rs, PenArgs a, int nbc) {
  a.P = pencils(a.n);
  a.tw = rs->d_tw[a.n];
  const int npen = a.U * a.hc;
  const dim3 g((npen + a.P - 1) / a.P, nbc);
#define RS_LAUNCH(PRV) \
  hipLaunchKernelGGL(k_rs_pencil<PRV>, g, dim3(256), pencil_lds(a.n, a.P), rs->stream, a, (const double*)rs->d_cs)
  RS_PR_SWITCH(big_prime(a.n), RS_LAUNCH)
#undef RS_LAUNCH
  PW_HIP(hipGetLastError());
  return PW_OK;
}

int launch_x(pw_rs* rs, bool c2r, const void* in, void* out, int n, int hw, int nrows, long in_bs, long out_bs,
             int nbc, const double* vt = nullptr, double* out2 = nullptr) {
  const int P = pencils(n);
  const int npair = nrows / 2;
  const dim3 g((npair + P - 1) / P, nbc);
  const double2* tw = rs->d_tw[n];
  const double* cs = rs->d_cs;
  if (c2r) {
#define RS_LAUNCH(PRV)                                                                                    \
  hipLaunchKernelGGL(k_rs_x_c2r<PRV>, g, dim3(256), pencil_lds(n, P), rs->stream, (const double2*)in, \
                     (double*)out, tw, cs, n, hw, npair, in_bs, out_bs, P, vt, out2)
    RS_PR_SWITCH(big_prime(n), RS_LAUNCH)
#undef RS_LAUNCH
  } else {
#define RS_LAUNCH(PRV)                                                                                   \
  hipLaunchKernelGGL(k_rs_x_r2c<PRV>, g, dim3(256), pencil_lds(n, P), rs->stream, (const double*)in, \
                     (double2*)out, tw, cs, n, hw, npair, in_bs, out_bs, P)
    RS_PR_SWITCH(big_prime(n), RS_LAUNCH)
#undef RS_LAUNCH
  }
  PW_HIP(hipGetLastError());
  return PW_OK;
}

// coarse psi -> xc (unscaled coarse spectrum)
int own_coarse_fwd(pw_rs* rs, const double* psi, int nbc) {
  const int* nc = rs->nc;
  const int hc = rs->hc0;
  const long Hc = rs->Hc;
  PW_TRY(launch_x(rs, false, psi, rs->d_cv, nc[0], hc, nc[1] * nc[2], rs->Nc, Hc, nbc));
  PenArgs y{rs->d_cv, Hc, (long)nc[1] * hc, hc, nullptr, rs->d_cw, Hc, (long)nc[1] * hc, hc, nullptr,
            nc[1], nc[2], hc, nc[1], 0, -1.0, nullptr};
  PW_TRY(launch_pencil(rs, y, nbc));
  PenArgs z{rs->d_cw, Hc, hc, (long)nc[1] * hc, nullptr, rs->d_xc, Hc, hc, (long)nc[1] * hc, nullptr,
            nc[2], nc[1], hc, nc[2], 0, -1.0, nullptr};
  PW_TRY(launch_pencil(rs, z, nbc));
  return PW_OK;
}

// xc -> real coarse bands (unscaled inverse)
int own_coarse_inv(pw_rs* rs, double* a3, int nbc) {
  const int* nc = rs->nc;
  const int hc = rs->hc0;
  const long Hc = rs->Hc;
  PenArgs z{rs->d_xc, Hc, hc, (long)nc[1] * hc, nullptr, rs->d_cv, Hc, hc, (long)nc[1] * hc, nullptr,
            nc[2], nc[1], hc, nc[2], 0, 1.0, nullptr};
  PW_TRY(launch_pencil(rs, z, nbc));
  PenArgs y{rs->d_cv, Hc, (long)nc[1] * hc, hc, nullptr, rs->d_cw, Hc, (long)nc[1] * hc, hc, nullptr,
            nc[1], nc[2], hc, nc[1], 0, 1.0, nullptr};
  PW_TRY(launch_pencil(rs, y, nbc));
  PW_TRY(launch_x(rs, true, rs->d_cw, a3, nc[0], hc, nc[1] * nc[2], Hc, rs->Nc, nbc));
  return PW_OK;
}

// coarse psi -> xc (coarse spectrum) -> pf (psi on the fine grid)
int own_to_fine(pw_rs* rs, const double* psi, int nbc, bool fuse_v) {
  const int *nc = rs->nc, *nf = rs->nf;
  const int hc = rs->hc0;
  const long Hc = rs->Hc;
  PW_TRY(own_coarse_fwd(rs, psi, nbc));
  // fine inverse: z (with the coarse-to-fine gather), y, x
  const long Tzb = (long)nf[2] * rs->ny * hc, Tfb = (long)nf[2] * nf[1] * hc;
  {
    const int P = pencils(nf[2]);
    const dim3 g((rs->ny * hc + P - 1) / P, nbc);
    const double2* tw = rs->d_tw[nf[2]];
    const double* cs = rs->d_cs;
#define RS_LAUNCH(PRV)                                                                                     \
  hipLaunchKernelGGL(k_rs_zinv_c2f<PRV>, g, dim3(256), pencil_lds(nf[2], P), rs->stream, rs->d_xc, rs->d_tz, \
                     tw, cs, nc[0], nc[1], nc[2], nf[2], hc, rs->ny, P, Hc, Tzb, 1.0 / (double)rs->Nc)
    RS_PR_SWITCH(big_prime(nf[2]), RS_LAUNCH)
#undef RS_LAUNCH
    PW_HIP(hipGetLastError());
  }
  PenArgs yi{rs->d_tz, Tzb, (long)rs->ny * hc, hc, rs->d_ymap, rs->d_tf, Tfb, (long)nf[1] * hc, hc, nullptr,
             nf[1], nf[2], hc, nf[1], 0, 1.0, nullptr};
  PW_TRY(launch_pencil(rs, yi, nbc));
  PW_TRY(launch_x(rs, true, rs->d_tf, rs->d_pf, nf[0], hc, nf[1] * nf[2], Tfb, rs->Nf, nbc,
                  fuse_v ? rs->d_vtot : nullptr, rs->d_pv));
  return PW_OK;
}

// pf (H psi terms on the fine grid) -> restriction + kinetic -> a3
int own_to_coarse(pw_rs* rs, const double* fine, double* a3, int nbc) {
  const int *nc = rs->nc, *nf = rs->nf;
  const int hc = rs->hc0;
  const long Hc = rs->Hc;
  const long Tzb = (long)nf[2] * rs->ny * hc, Tfb = (long)nf[2] * nf[1] * hc;
  const long U0b = (long)rs->nz * rs->ny * hc;
  PW_TRY(launch_x(rs, false, fine, rs->d_tf, nf[0], hc, nf[1] * nf[2], rs->Nf, Tfb, nbc));
  PenArgs yf{rs->d_tf, Tfb, (long)nf[1] * hc, hc, nullptr, rs->d_tz, Tzb, (long)rs->ny * hc, hc, rs->d_lmap,
             rs->ny, nf[2], hc, nf[1], 0, -1.0, nullptr};
  PW_TRY(launch_pencil(rs, yf, nbc));
  PenArgs zf{rs->d_tz, Tzb, hc, (long)rs->ny * hc, nullptr, rs->d_u0, U0b, hc, (long)rs->ny * hc, rs->d_mmap,
             rs->nz, rs->ny, hc, nf[2], 0, -1.0, nullptr};
  PW_TRY(launch_pencil(rs, zf, nbc));
  hipLaunchKernelGGL(k_rs_f2c_ekin_p, dim3(nc[1] * nc[2], nbc), dim3(kRowThreads), 0, rs->stream, rs->d_xc,
                     rs->d_u0, rs->d_ekin, rs->d_ymap, rs->d_zmap, nc[0], nc[1], nc[2], nf[1], nf[2], hc, rs->ny,
                     Hc, U0b, 1.0 / (double)rs->Nc, 1.0 / (double)rs->Nf);
  PW_HIP(hipGetLastError());
  return own_coarse_inv(rs, a3, nbc);
}

// Projector weights (from psi on the fine grid, pf), then V psi and the
// projector terms.  fused: V psi is already in d_pv (written by the C2R
// row pass) and the projector terms go there; else in place on pf.
int local_terms(pw_rs* rs, int nbc, bool fused) {
  const bool nl = rs->natom > 0 && rs->nproj_tot > 0;
  double* u = fused ? rs->d_pv : rs->d_pf;
  if (nl) {
    const double scale = rs->vol / (double)rs->Nf;
    if (rs->npmax <= 8)
      hipLaunchKernelGGL((k_rs_weights<8, 8>), dim3(rs->natom, (nbc + 7) / 8), dim3(256), 0, rs->stream,
                         rs->d_pf, rs->d_atab, rs->d_apt_off, rs->d_idx, rs->d_val, rs->d_wgt, rs->d_coef,
                         nbc, rs->Nf, scale);
    else if (rs->npmax <= 16)
      hipLaunchKernelGGL((k_rs_weights<16, 4>), dim3(rs->natom, (nbc + 3) / 4), dim3(256), 0, rs->stream,
                         rs->d_pf, rs->d_atab, rs->d_apt_off, rs->d_idx, rs->d_val, rs->d_wgt, rs->d_coef,
                         nbc, rs->Nf, scale);
    else
      hipLaunchKernelGGL((k_rs_weights<32, 2>), dim3(rs->natom, (nbc + 1) / 2), dim3(256), 0, rs->stream,
                         rs->d_pf, rs->d_atab, rs->d_apt_off, rs->d_idx, rs->d_val, rs->d_wgt, rs->d_coef,
                         nbc, rs->Nf, scale);
    PW_HIP(hipGetLastError());
  }
  if (!nl || rs->ncolor > 0) {
    if (!fused) {
      // Nf is even (even grids), so every band row starts 16-byte aligned
      const unsigned nblk = (unsigned)((rs->Nf / 2 + 255) / 256);
      hipLaunchKernelGGL(k_rs_vmul, dim3(nblk, (nbc + 7) / 8), dim3(256), 0, rs->stream, rs->d_pf, rs->d_vtot,
                         nbc, rs->Nf);
      PW_HIP(hipGetLastError());
    }
    for (int c = 0; nl && c < rs->ncolor; c++) {
      const int na = rs->color_off[c + 1] - rs->color_off[c];
      const int* cl = rs->d_clist + rs->color_off[c];
      const dim3 g(na, (nbc + kNlBands - 1) / kNlBands);
      if (rs->npmax <= 8)
        hipLaunchKernelGGL(k_rs_nl_apply<8>, g, dim3(256), 0, rs->stream, u, cl, rs->d_atab, rs->d_apt_off,
                           rs->d_idx, rs->d_val, rs->d_coef, nbc, rs->Nf);
      else if (rs->npmax <= 16)
        hipLaunchKernelGGL(k_rs_nl_apply<16>, g, dim3(256), 0, rs->stream, u, cl, rs->d_atab,
                           rs->d_apt_off, rs->d_idx, rs->d_val, rs->d_coef, nbc, rs->Nf);
      else
        hipLaunchKernelGGL(k_rs_nl_apply<32>, g, dim3(256), 0, rs->stream, u, cl, rs->d_atab,
                           rs->d_apt_off, rs->d_idx, rs->d_val, rs->d_coef, nbc, rs->Nf);
      PW_HIP(hipGetLastError());
    }
  } else {
    hipLaunchKernelGGL(k_rs_local, dim3((unsigned)((rs->Nf + 255) / 256)), dim3(256), 0, rs->stream, rs->d_pf,
                       rs->d_vtot, rs->d_rowptr, rs->d_ent, rs->d_atab, rs->d_val, rs->d_coef, nbc, rs->Nf, 1);
    PW_HIP(hipGetLastError());
  }
  return PW_OK;
}

}  // namespace

extern "C" int pw_rs_mult_spinor(pw_rs* rs, int nb, const double* psi, double* a3) {
  if (!rs || nb < 0 || (nb > 0 && (!psi || !a3))) return fail(PW_ERR_ARG, "pw_rs_mult_spinor: bad argument");
  if (!rs->d_vtot) return fail(PW_ERR_STATE, "pw_rs_mult_spinor: local potential not set");
  int dev;
  PW_HIP(hipGetDevice(&dev));
  if (dev != rs->device) return fail(PW_ERR_STATE, "pw_rs_mult_spinor: handle belongs to device %d", rs->device);
  const int* nc = rs->nc;
  const int* nf = rs->nf;
  for (int b0 = 0; b0 < nb; b0 += rs->chunk) {
    const int nbc = std::min(rs->chunk, nb - b0);
    const bool nl = rs->natom > 0 && rs->nproj_tot > 0;
    const bool fuse_v = !nl || rs->ncolor > 0;
    PW_TRY(own_to_fine(rs, psi + (long)b0 * rs->Nc, nbc, fuse_v));
    PW_TRY(local_terms(rs, nbc, fuse_v));
    PW_TRY(own_to_coarse(rs, fuse_v ? rs->d_pv : rs->d_pf, a3 + (long)b0 * rs->Nc, nbc));
  }
  return PW_OK;
}

extern "C" int pw_rs_filter(pw_rs* rs, double cutoff, int nb, double* x) {
  if (!rs || nb < 0 || (nb > 0 && !x)) return fail(PW_ERR_ARG, "pw_rs_filter: bad argument");
  int dev;
  PW_HIP(hipGetDevice(&dev));
  if (dev != rs->device) return fail(PW_ERR_STATE, "pw_rs_filter: handle belongs to device %d", rs->device);
  for (int b0 = 0; b0 < nb; b0 += rs->chunk) {
    const int nbc = std::min(rs->chunk, nb - b0);
    double* xb = x + (long)b0 * rs->Nc;
    PW_TRY(own_coarse_fwd(rs, xb, nbc));
    hipLaunchKernelGGL(k_rs_cutmask, dim3((unsigned)((rs->Hc + 255) / 256), nbc), dim3(256), 0, rs->stream,
                       rs->d_xc, rs->d_ekin, cutoff, rs->Hc, 1.0 / (double)rs->Nc);
    PW_HIP(hipGetLastError());
    PW_TRY(own_coarse_inv(rs, xb, nbc));
  }
  return PW_OK;
}

extern "C" int pw_rs_calculate_density(pw_rs* rs, int nb, const double* psi, const double* occ, int num_spin,
                                       int num_occupied, double* rho, double* val) {
  if (!rs || nb < 1 || !psi || !occ || !rho || num_spin < 1 || num_occupied < 1)
    return fail(PW_ERR_ARG, "pw_rs_calculate_density: bad argument");
  int dev;
  PW_HIP(hipGetDevice(&dev));
  if (dev != rs->device) return fail(PW_ERR_STATE, "pw_rs_calculate_density: handle belongs to device %d", rs->device);
  if (!rs->d_f) PW_HIP(hipMalloc((void**)&rs->d_f, sizeof(double) * (size_t)std::max(rs->chunk, 4)));
  if (!rs->d_part) PW_HIP(hipMalloc((void**)&rs->d_part, sizeof(double) * kSumBlocks));
  const int* nc = rs->nc;
  const int* nf = rs->nf;
  std::vector<double> f((size_t)rs->chunk);
  for (int b0 = 0; b0 < nb; b0 += rs->chunk) {
    const int nbc = std::min(rs->chunk, nb - b0);
    const double* xb = psi + (long)b0 * rs->Nc;
    PW_TRY(own_to_fine(rs, xb, nbc, false));
    for (int b = 0; b < nbc; b++) f[(size_t)b] = num_spin * occ[b0 + b];
    // the previous batch's accumulation reads d_f: copy in stream order
    PW_HIP(hipMemcpyAsync(rs->d_f, f.data(), sizeof(double) * nbc, hipMemcpyHostToDevice, rs->stream));
    PW_HIP(hipStreamSynchronize(rs->stream));
    hipLaunchKernelGGL(k_rs_rho_acc, dim3((unsigned)((rs->Nf + 255) / 256)), dim3(256), 0, rs->stream, rs->d_pf,
                       rs->d_f, rho, nbc, rs->Nf, b0 == 0 ? 1 : 0);
    PW_HIP(hipGetLastError());
  }
  // rho *= num_spin num_occ Nf / (vol sum rho) (hamiltonian.cpp:2599-2600),
  // then val = sum rho vol / Nf (:2607-2610)
  const double c = (double)num_spin * (double)num_occupied * (double)rs->Nf / rs->vol;
  double* sums = rs->d_f;  // [sum, scale, sum2, val]
  hipLaunchKernelGGL(k_rs_sum_partial, dim3(kSumBlocks), dim3(256), 0, rs->stream, rho, rs->Nf,
                     (const double*)nullptr, rs->d_part);
  hipLaunchKernelGGL(k_rs_sum_final, dim3(1), dim3(256), 0, rs->stream, (const double*)rs->d_part, c, 1, sums);
  hipLaunchKernelGGL(k_rs_sum_partial, dim3(kSumBlocks), dim3(256), 0, rs->stream, rho, rs->Nf,
                     (const double*)(sums + 1), rs->d_part);
  hipLaunchKernelGGL(k_rs_sum_final, dim3(1), dim3(256), 0, rs->stream, (const double*)rs->d_part,
                     rs->vol / (double)rs->Nf, 0, sums + 2);
  PW_HIP(hipGetLastError());
  double h[4];
  PW_HIP(hipMemcpyAsync(h, sums, sizeof(double) * 4, hipMemcpyDeviceToHost, rs->stream));
  PW_HIP(hipStreamSynchronize(rs->stream));
  if (val) *val = h[3];
  return PW_OK;
}

extern "C" int pw_rs_teter_precond(pw_rs* rs, int nb, const double* psi, double* a3) {
  if (!rs || nb < 0 || (nb > 0 && (!psi || !a3))) return fail(PW_ERR_ARG, "pw_rs_teter_precond: bad argument");
  int dev;
  PW_HIP(hipGetDevice(&dev));
  if (dev != rs->device) return fail(PW_ERR_STATE, "pw_rs_teter_precond: handle belongs to device %d", rs->device);
  for (int b0 = 0; b0 < nb; b0 += rs->chunk) {
    const int nbc = std::min(rs->chunk, nb - b0);
    const double* xb = psi + (long)b0 * rs->Nc;
    PW_TRY(own_coarse_fwd(rs, xb, nbc));
    hipLaunchKernelGGL(k_rs_teter_spec, dim3((unsigned)((rs->Hc + 255) / 256), nbc), dim3(256), 0, rs->stream,
                       rs->d_xc, rs->d_teter, rs->Hc, 1.0 / (double)rs->Nc);
    PW_HIP(hipGetLastError());
    // the fine-grid scratch (Nf >= Nc per band) takes T psi before the add
    PW_TRY(own_coarse_inv(rs, rs->d_pf, nbc));
    const long n = (long)nbc * rs->Nc;
    hipLaunchKernelGGL(k_rs_axpy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, rs->stream,
                       a3 + (long)b0 * rs->Nc, (const double*)rs->d_pf, n);
    PW_HIP(hipGetLastError());
  }
  return PW_OK;
}

int pw::rs_linalg_ctx(pw_rs* rs, int max_bands, pw_ctx** out) {
  if (!rs->la) {
    pw_ctx* c = new pw_ctx();
    c->device = rs->device;
    c->stream = rs->stream;
    if (rocblas_create_handle(&c->blas) != rocblas_status_success) {
      delete c;
      return fail(PW_ERR_HIP, "rocblas_create_handle failed");
    }
    rocblas_set_stream(c->blas, c->stream);
    rs->la = c;
  }
  rs->la->max_bands = std::max(rs->la->max_bands, max_bands);
  *out = rs->la;
  return PW_OK;
}
