// Internal declarations of libpwdft_hip.so (not part of the C-ABI).
#pragma once
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>

#include <cstdarg>
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/pwdft_hip.h"

namespace pw {

// Workgroup barrier for LDS traffic only: waits for this wavefront's LDS
// operations but leaves its global loads and stores in flight (a full
// __syncthreads also drains those).  Only for phases whose cross-wavefront
// data goes through LDS.
__device__ __forceinline__ void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" ::: "memory"); }

// Work item of block b in a one-dimensional grid of n blocks.  Blocks are
// dispatched round-robin over the 8 XCDs (block b on XCD b % 8), each with
// its own L2; this order gives XCD x the contiguous item range
// [x q + min(x, r), ..) (q = n / 8, r = n % 8), so that neighbouring items,
// which share cache lines, share an L2.  A bijection of 0 .. n-1.
__device__ __forceinline__ int xcd_item(int b, int n) {
  const int q = n >> 3, r = n & 7, x = b & 7;
  return x * q + min(x, r) + (b >> 3);
}

// Data exchanged through LDS between lanes of one wavefront: this
// wavefront's LDS operations issued so far have completed, and the compiler
// moves no memory access across (without it, a lane's load of a word another
// lane stored may be hoisted above the store, since the compiler models
// lanes as independent threads).
__device__ __forceinline__ void wave_lds_sync() { asm volatile("s_waitcnt lgkmcnt(0)" ::: "memory"); }

// Sum of a double over the 64 lanes of a wavefront, returned to every lane,
// with DPP lane moves (no LDS round trips): row_shr 1, 2, 4, 8 leave each
// 16-lane row's sum in its lane 15, row_bcast 15 and 31 carry the row sums
// up to lane 63, which is read back as a scalar.  Every lane must be active.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_f64(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xf, true);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, 0xf, true);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double wave_sum(double v) {
  v += dpp_f64<0x111, 0xf>(v);  // row_shr:1
  v += dpp_f64<0x112, 0xf>(v);  // row_shr:2
  v += dpp_f64<0x114, 0xf>(v);  // row_shr:4
  v += dpp_f64<0x118, 0xf>(v);  // row_shr:8
  v += dpp_f64<0x142, 0xa>(v);  // row_bcast:15 into rows 1 and 3
  v += dpp_f64<0x143, 0xc>(v);  // row_bcast:31 into rows 2 and 3
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), 63);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), 63);
  return __hiloint2double(hi, lo);
}

// ---- error plumbing: a thread-local message, integer status codes ----
void set_error(const char* fmt, ...);
int fail(int code, const char* fmt, ...);

#define PW_HIP(call)                                                              \
  do {                                                                            \
    hipError_t e_ = (call);                                                       \
    if (e_ != hipSuccess)                                                         \
      return ::pw::fail(PW_ERR_HIP, "%s failed: %s (%s:%d)", #call,              \
                        hipGetErrorString(e_), __FILE__, __LINE__);               \
  } while (0)

#define PW_BLAS(call)                                                             \
  do {                                                                            \
    rocblas_status s_ = (call);                                                   \
    if (s_ != rocblas_status_success)                                             \
      return ::pw::fail(PW_ERR_HIP, "%s failed: rocblas status %d (%s:%d)", #call, \
                        (int)s_, __FILE__, __LINE__);                             \
  } while (0)

#define PW_TRY(call)        \
  do {                      \
    int rc_ = (call);       \
    if (rc_ != PW_OK) return rc_; \
  } while (0)

// A linear-algebra-only context (stream, rocBLAS handle, GEMM / eigensolver
// workspaces) bound to a real-space handle, for the solvers that run on its
// bands; created on first use, freed with the handle.
int rs_linalg_ctx(pw_rs* rs, int max_bands, pw_ctx** out);

// ---- grid helpers shared by host setup and context creation ----
inline int kgrid(int i, int n) { return i <= n / 2 ? i : i - n; }
void recip_cell(const double* supercell, double* recip);
double cell_volume(const double* supercell);
bool smooth235(int n);
// Radix plan of a 2/3/5-smooth length: factors (4, 2, 3, 5 order).
std::vector<int> radix_plan(int n);
// Raise a kernel's dynamic LDS limit (above the 64 KiB default, at most
// 160 KiB).  Made once per kernel and size: the call is not stream ordered,
// so a graph capture (which always follows an eager call) never makes it.
int ensure_lds(const void* kern, size_t bytes);

// A device buffer owned by the context.
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
};

// Stick decomposition of the half sphere: the distinct (kx, ky) columns of
// the sphere; each column is transformed along z, each z slab along y and x.
struct StickPlan {
  int nstick = 0, kxmax = 0;
  int sc = 16;        // sticks per chunk in the z kernels
  int nchunk = 0;
  std::vector<int> kx, ky;        // per stick (ky = storage index 0..n2-1)
  // per chunk point list, (z, stick) ordered: packed = sphere_idx, and
  // (local stick, z) in a second array; mirrors (conj entries of the (0,0)
  // stick and nothing else) flagged by negative sphere index -(i+1).
  std::vector<int> chunk_off;     // nchunk+1
  std::vector<int> pt_idx;        // sphere index, or -(i+1) for conj(c_i)
  std::vector<int> pt_sz;         // local stick | (z << 16)
  // Gather list for the forward direction (no mirrors): same layout.
  std::vector<int> gchunk_off;
  std::vector<int> gpt_idx;
  std::vector<int> gpt_sz;
  // slab positions: kx*n2 + ky, and the mirror position for kx=0,ky>0
  std::vector<int> pos, mpos;     // mpos = -1 when no mirror
  // per (z, stick), z-major: the sphere index i of that point, -(i+1) for
  // the conj mirror of c_i (kx = ky = 0 column, kz < 0), kZsAbsent outside
  // the sphere
  std::vector<int> zs;
};
constexpr int kZsAbsent = (int)0x80000000;

}  // namespace pw

struct pw_ctx;

namespace pw {
// MFMA GEMMs (pw_gemm.hip).  C = A^T B (+ beta C), rows scaled by rowscale
// when non-null; Y = alpha A B + beta Y.  Column-major, leading dims given.
// split = 2 or 3: FP32 inputs as products of their 2 or 3 split-BF16 terms
// on v_mfma_f32_16x16x32_bf16 (FP32 accumulation); 0: FP32 MFMA.  Ignored
// for FP64.
template <typename T>
int gemm_tn(pw_ctx* c, int K, int n1, int n2, const T* A, int lda, const T* B, int ldb, T* C, int ldc,
            const T* rowscale, T beta, int split = 0);
// nbatch products C_b = A_b^T B_b of one shape, operands from device
// pointer tables (FP64 instantiated): two launches for the whole batch.
template <typename T>
int gemm_tn_batched(pw_ctx* c, int K, int n1, int n2, const T* const* dA, int lda, const T* const* dB, int ldb,
                    T* const* dC, int ldc, int nbatch);
// gemm_tn with the split-K partials in a caller's buffer (ws, ws_bytes): no
// allocation; fewer K splits when the usual count does not fit, and
// PW_ERR_UNSUPPORTED when one split does not.  ws == null: the context's
// buffer, grown on demand (gemm_tn).
template <typename T>
int gemm_tn_ws(pw_ctx* c, int K, int n1, int n2, const T* A, int lda, const T* B, int ldb, T* C, int ldc,
               const T* rowscale, T beta, int split, void* ws, size_t ws_bytes);
template <typename T>
int gemm_nn(pw_ctx* c, int m, int n, int k, T alpha, const T* A, int lda, const T* B, int ldb, T beta, T* Y,
            int ldy);
// out = H + A B (cheb false), or out = s1 (H + A B - c Yc) - s2 Yp with (c,
// s1, s2) = coef[3 step ..] on the device and Yp == null meaning s2 = 0
// (cheb true): the projector term of H Y fused with a Chebyshev step.  FP64
// on FP64 MFMA; FP32 on FP32 MFMA, or on split-BF16 MFMA with split = 2 or 3
// terms per operand; k = 0 leaves acc = 0.
template <typename T>
int gemm_nn_ep(pw_ctx* c, int m, int n, int k, const T* A, int lda, const T* B, int ldb, const T* H, const T* Yc,
               const T* Yp, T* out, int ld, const T* coef, int step, bool cheb, int split = 0);
// The FP32 projector apply on products split ahead of time (pw_gemm.hip):
// split_a_operand allocates *out and splits A (m x k) into BF16 terms in the
// operand-tile layout; gemm_nn_pre splits B (k x n) into Bws
// (split_b_bytes(n, k) bytes) and forms out = epilogue(A B) like gemm_nn_ep
// with split (2 or 3) terms.
int split_a_operand(pw_ctx* c, const float* A, int lda, int m, int k, unsigned short** out);
size_t split_b_bytes(int n, int k);
int gemm_nn_pre(pw_ctx* c, int m, int n, int k, const unsigned short* Ab, const float* B, int ldb,
                unsigned short* Bws, const float* H, const float* Yc, const float* Yp, float* out, int ld,
                const float* coef, int step, bool cheb, int split);
// Dense symmetric eigensolvers (pw_eig.hip, pw_solver.hip), in place on the
// upper triangle of A (A <- eigenvectors, w (device) <- eigenvalues
// ascending); sym_eig adds the sign convention of the oracle and takes
// rocSOLVER beyond eig_sym's n <= 1024.
int eig_sym(pw_ctx* c, int n, double* A, int lda, double* w);
// S = U^T U (U into S's upper triangle) and the inverses of U's 32 x 32
// diagonal blocks into Dinv (chol_dinv_doubles(n)); device info word 0 or
// the failing column; then X (m x n) <- X U^{-1}.  n <= chol_inv_max_n().
int chol_upper(pw_ctx* c, int n, double* S, int ld, double* Dinv, int* info);
int trsm_right_upper(pw_ctx* c, int m, int n, double* X, int ldx, const double* U, int ldu, const double* Dinv);
int chol_inv_max_n();
size_t chol_dinv_doubles(int n);
int sym_eig(pw_ctx* c, int n, double* A, int lda, double* w, double* work, int* info);
// batched forms (problem b at the given strides)
int chol_upper_batched(pw_ctx* c, int n, double* S, int ld, long sS, double* Dinv, long sD, int* info, int nbatch);
int trsm_right_upper_batched(pw_ctx* c, int m, int n, double* X, int ldx, long sX, const double* U, int ldu,
                             long sU, const double* Dinv, long sD, int nbatch);
// A x = lambda B x, n <= 256 (Sygvd itype 1 'V' 'U'): eigenvectors into A,
// U into B's upper triangle, eigenvalues ascending into w, info[b] 0 or the
// failing pivot column of B
int sygvd_batched(pw_ctx* c, int n, double* A, int lda, long sA, double* B, int ldb, long sB, double* w, long sw,
                  int* info, int nbatch);
// Scaled Chebyshev recurrence coefficients of a degree-m filter damping
// [a, b] and scaled at aL: coef[3 (k-1) ..] = (c, s1, s2) of step k, with
// Y_1 = s1 (H - c) X and Y_k = s1 (H - c) Y_{k-1} - s2 Y_{k-2}.
void cheb_coef(double a, double b, double aL, int m, double* coef);
// Stick transforms of the wavefunction sphere (pw_hpsi.hip).
template <typename T>
int sphere_to_sticks(pw_ctx* c, int nb, const void* psi, void* t1);
template <typename T>
int sticks_to_sphere_add(pw_ctx* c, int nb, const void* t2, void* out);
// Slab-kernel facts for the autotuner (pw_hpsi.hip): whether the
// register-FFT kernel exists for the context's planes in dtype with ly
// lanes per y column transform and sb bands per plane (<= 0: the defaults;
// sb = 16: the wide FP32 variant, where its LDS fits; persist: the
// persistent-grid variant), and the LDS bytes of
// the LDS kernel at (rp, cg) for complex elements of esz bytes.
bool has_reg_slab(const pw_ctx* c, int dtype, int ly = 0, int sb = 0, bool persist = false);
// the context's second stream and fork / join events (created on first use)
int side_stream(pw_ctx* c);
size_t slab_lds_bytes(const pw_ctx* c, int esz, int rp, int cg);
// One Chebyshev step Ynew = s1 (H Y - c Y) - s2 Yp (pw_hpsi.hip), the
// step's (c, s1, s2) read from the context's device coefficient table
// (d_coef / d_coef_f) at 3 step; Yp == null for the first step.  HY is
// scratch (the local, kinetic and exchange terms).
int cheb_step_apply(pw_ctx* c, int dtype, int nb, const void* Y, void* HY, const void* Yp, void* Ynew, int step);
// Exact exchange (pw_exx.hip): a3 += V_x psi when the context holds an
// exchange operator; state freed with the context.
struct ExxState;
template <typename T>
int exx_add(pw_ctx* c, int nb, const void* psi, void* a3);
// The exchange term of MultSpinor: the compressed operator when built,
// else exx_add.
template <typename T>
int exchange_add(pw_ctx* c, int nb, const void* psi, void* a3);
void exx_free(pw_ctx* c);
}  // namespace pw

struct pw_ctx {
  pw::ExxState* exx = nullptr;  // screened exchange operator, when set
  void* d_gemm_ws = nullptr;  // split-K partial tiles
  size_t gemm_ws_bytes = 0;
  int device = 0;
  hipStream_t stream = nullptr;
  // second stream (default priority) and fork / join events for work that
  // overlaps a single-workgroup factorisation; created on first use
  hipStream_t side = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  rocblas_handle blas = nullptr;
  int n1, n2, n3, h1;
  long N, Nh;
  double vol;
  double supercell[9], recip[9];
  int npw, npad, max_bands;
  pw::StickPlan sp;
  std::vector<int> h_idx;       // idxCoarseCut, host copy
  std::vector<double> h_wgt;    // projector weights, host copy
  std::vector<int> f1, f2, f3;  // radix plans
  // device arrays
  int *d_pt_idx = nullptr, *d_pt_sz = nullptr, *d_chunk_off = nullptr;
  int *d_gpt_idx = nullptr, *d_gpt_sz = nullptr, *d_gchunk_off = nullptr;
  int* d_zs = nullptr;  // StickPlan::zs
  int *d_pos = nullptr, *d_mpos = nullptr;
  double *d_ekin = nullptr, *d_teter = nullptr;
  float *d_ekin_f = nullptr, *d_teter_f = nullptr;
  double* d_tw64 = nullptr;  // roots for n1 | n2 | n3, interleaved complex
  float* d_tw32 = nullptr;
  double* d_vtot = nullptr;  // V / N
  float* d_vtot_f = nullptr;
  bool have_vtot = false;
  int nbeta = 0;
  double *d_beta = nullptr, *d_betaw = nullptr, *d_W = nullptr;
  float *d_beta_f = nullptr, *d_betaw_f = nullptr, *d_W_f = nullptr;
  // betaw (FP32) split into BF16 terms ahead of time, and the per-call split
  // of W (max_bands columns), for the projector apply on split products
  unsigned short *d_betaw_bf = nullptr, *d_W_bf = nullptr;
  void* d_t1 = nullptr;  // stick buffers: max_bands x n3 x nstick complex double
  void* d_t2 = nullptr;
  size_t tbytes = 0;
  // eigensolver workspace (allocated on first use): block buffers of
  // max_bands x npw complex double, small nb x nb FP64 matrices
  rocblas_handle solver = nullptr;  // rocSOLVER shares the rocBLAS handle type
  void* d_ws[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  double *d_S = nullptr, *d_S2 = nullptr, *d_D = nullptr, *d_E = nullptr, *d_red = nullptr;
  double* d_lz = nullptr;  // Lanczos vectors v, v_prev, f (3 x 2 npw)
  void* d_eig_ws = nullptr;  // dense eigensolver scratch, grown on demand
  size_t eig_ws_bytes = 0;
  void* d_geig_ws = nullptr;  // batched generalized eigensolver scratch, grown on demand
  size_t geig_ws_bytes = 0;
  int* d_info = nullptr;
  double* d_chol = nullptr;  // inverses of the Cholesky factor's diagonal blocks
  int ws_nb = 0;
  // Chebyshev recurrence coefficients (3 per step, degree <= 64) and the
  // captured filter graph, replayed on a private stream
  double h_coef[192];
  float h_coef_f[192];
  double* d_coef = nullptr;
  float* d_coef_f = nullptr;
  bool use_graphs = true;
  // FP32 projector products inside the Chebyshev filter as products of split
  // BF16 terms on MFMA (FP32 accumulation): 3 terms per operand (FP32's 24
  // bits, six term products) by default, PW_BF16=2 for 2 terms (16 bits,
  // three products), PW_BF16=0 for FP32 MFMA
  int bf16_split = 3;
  bool zreg = false;  // PW_ZREG=1: the register z kernels where the z length has one
  // Plans per precision ([PW_F64], [PW_F32]), set by PW_SLAB or pw_autotune:
  //   reg: the register-FFT slab kernel where compiled (PW_SLAB=lds: the LDS
  //   kernel); cg / rp / threads: the LDS kernel's y column groups (0 = two
  //   y buffers), x row pairs per step, threads per workgroup (0 = the
  //   built-in rule; PW_SLAB=lds:<cg>:<rp>:<threads>); tn_tm / tn_cap: the
  //   split-K TN product's tile (1: 64 x 64, 2: 128 x 128) and the number of
  //   partial workgroups aimed at (0 = the built-in rule)
  struct Plan {
    bool reg = true;
    int ly = 0;  // register kernel: lanes per y column transform (0: default)
    int sb = 0;  // register kernel: bands per z plane (0: default; 16: wide FP32 variant)
    bool persist = false;          // register kernel on a persistent grid (next item prefetched)
    const void* pkern = nullptr;   // persistent grid size, for this kernel and LDS size
    size_t plds = 0;
    int pgrid = 0;
    int cg = 0, rp = 0, threads = 0;
    int tn_tm = 0, tn_cap = 0;
  } plan[2];
  hipStream_t gstream = nullptr;
  hipEvent_t gev[2] = {nullptr, nullptr};
  hipGraphExec_t fg_exec = nullptr;
  long fg_key = -1, fg_warm = -1;
  void* fg_result = nullptr;
  hipGraphExec_t lz_exec = nullptr;  // Lanczos steps, keyed by their count
  long lz_key = -1, lz_warm = -1;
  // timing
  static constexpr int kTimedCalls = 1024;
  bool timing = false;
  std::vector<hipEvent_t> ev_pool;  // 5 events per timed call
  int ev_calls = 0, ev_total = 0;
  double stage_ms[4] = {0, 0, 0, 0};
};

namespace pw {
// A captured filter graph holds raw device pointers and the shape of the
// operator (potential, projectors, exchange): it is dropped whenever any of
// them is reallocated or switched on or off.
inline void invalidate_graph(pw_ctx* c) { c->fg_key = c->fg_warm = c->lz_key = c->lz_warm = -1; }
}  // namespace pw
