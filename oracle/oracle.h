/*
 * oracle.h -- FP64 CPU restatement of the PWDFT sphere-cut Gamma-point
 * H*psi path and the eigensolvers around it.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: it is loaded by
 * tests/, by __graft_entry__.smoke() and by bench.py's cpu_baseline leg, and
 * by nothing in the product (pwdft_develop_amd/).  It follows the reference
 * source line by line where the reference defines the behaviour, and the
 * published method where the reference has none (ChebFSI).
 *
 * Conventions (all arrays column-major, like NumMat/NumTns in the reference,
 * include/numtns_impl.hpp:116-134):
 *   - real grid index  r = x + n1*(y + n2*z)
 *   - R2C index        g = kx + h1*(ky + n2*kz), h1 = n1/2+1 (fourier.cpp:169-184)
 *   - a band block is npw x nb complex (interleaved re,im), G=0 first, the
 *     other half-sphere coefficients carrying a sqrt(2) factor
 *     (spinor.cpp:573-595).  Every inner product is Re(a^H b), so the block
 *     is handled as a real (2*npw) x nb matrix by the dense routines.
 *   - FFTs are unnormalised, forward = exp(-i), backward = exp(+i) (FFTW).
 *
 * Every function returns 0 on success and a negative code on failure; the
 * failure texts mirror the reference's ErrorHandling messages.
 */
#ifndef PW_ORACLE_H
#define PW_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---------- grid and sphere (fourier.cpp:97-244, 541-672) ---------- */

/* recip = 2*pi*inv(supercell)^T, rows are reciprocal vectors. */
void or_recip_cell(const double* supercell, double* recip);
double or_volume(const double* supercell);

/* Half-sphere index (idxCoarseCut), |G|^2/2 on it, and the Hermitian
 * padding pairs (idxCoarsePadding.first = src, .second = dst).  Caller
 * passes capacity-sized arrays (capacity >= h1*n2*n3 is always enough).
 * Returns npw (>0) and writes *npad, or <0 on error. */
int or_sphere_build(const int* n, const double* supercell, double ecut,
                    int* idx, double* gkk, int* npad, int* pad_src,
                    int* pad_dst);

/* Fock sphere: R2C indices with |G|^2/2 < 4*ecut, ascending
 * (fourier.cpp:642-670).  Returns the count. */
int or_fock_sphere_build(const int* n, const double* supercell, double ecut,
                         int* idx_fock, double* gkk_fock);

/* ekin = |G|^2/2, Teter b/(b+16a^4) (hamiltonian.cpp:2629-2667). */
void or_ekin_teter(int npw, const double* gkk, double* ekin, double* teter);

/* ---------- FFT ---------- */
/* 1D complex transform of length n (any n; radix 2/3/4/5 fast paths). */
int or_fft1d(int n, int sign, double* x /* 2n */);
/* Unnormalised 3D c2r: in h1*n2*n3 complex -> out n1*n2*n3 real. */
int or_fft_c2r(const int* n, const double* in, double* out);
/* Unnormalised 3D r2c: in n1*n2*n3 real -> out h1*n2*n3 complex. */
int or_fft_r2c(const int* n, const double* in, double* out);

/* ---------- real-space dual-grid H*psi (or_rs.c) ---------- */
void or_ip_c2f(const int* Nc, const int* Nf, const double* tilde, double* tildef);
void or_ip_f2c(const int* Nc, const int* Nf, double* tilde, const double* tildef);
int or_rs_mult_spinor(const int* Nc, const int* Nf, double vol, const double* ekin,
                      const double* vtot, int natom, const int* npts, const int* nproj,
                      const int* idx, const double* val, const double* wgt, int nb,
                      const double* psi, double* a3);
int or_rs_filter(const int* Nc, double vol, const double* gkk, double cutoff, int nb, double* x);
int or_rs_teter_precond(const int* Nc, double vol, const double* teter, int nb, const double* psi, double* a3);
int or_rs_density(const int* Nc, const int* Nf, double vol, int nb, const double* psi, const double* occ,
                  int num_spin, int num_occ, double* rho, double* val);

/* ---------- Hamiltonian ---------- */
typedef struct or_ham or_ham;

or_ham* or_ham_create(const int* n, const double* supercell, int npw,
                      const int* idx, int npad, const int* pad_src,
                      const int* pad_dst, const double* ekin,
                      const double* teter, const double* vtot);
void or_ham_destroy(or_ham* H);
int or_ham_set_nonlocal(or_ham* H, int nbeta, const double* beta /* npw x nbeta complex */,
                        const double* wgt);
/* Screened exact exchange (hamiltonian.cpp:3454-3551, 3553-3641). */
int or_ham_set_exx(or_ham* H, double ecut, double screen_mu, double exx_fraction,
                   int divergence_type, int nocc, const double* psi_occ /* npw x nocc */,
                   const double* occ, double* exx_div_out);
int or_ham_set_threads(or_ham* H, int nthreads);

/* Adaptively compressed exchange from the block psi (npw x nb):
 * KohnSham::CalculateVexxACE sphere branch (hamiltonian.cpp:3685-3924).
 * After it or_mult_spinor applies -xi Re(xi^H psi) in place of V_x
 * (hamiltonian.cpp:3347-3430).  nb = 0 drops it; or_ham_set_exx drops it. */
int or_ham_build_ace(or_ham* H, int nb, const double* psi);

/* KohnSham::MultSpinor (hamiltonian.cpp:3270-3452): a3 = (T + V_loc [+ V_x]) psi */
int or_mult_spinor(or_ham* H, int nb, const double* psi, double* a3);
/* Spinor::AddMultSpinorEXX (spinor.cpp:783-945): a3 += V_x psi */
int or_add_exx(or_ham* H, int nb, const double* psi, double* a3);
/* EigenSolver::NonlocalMultX (eigensolver.cpp:471-515): VnlX = beta*(w .* Re(beta^H X)) */
int or_nonlocal_multx(or_ham* H, int nb, const double* X, double* VnlX);
/* MultSpinor + NonlocalMultX, as the solvers assemble A*X (lobpcg.cpp:203-212). */
int or_apply_h(or_ham* H, int nb, const double* X, double* HX);
/* Spinor::AddTeterPrecond (spinor.cpp:348-374): a3 += teter .* psi */
int or_teter_precond(or_ham* H, int nb, const double* psi, double* a3);

/* ---------- dense FP64 linear algebra on the real (2npw) view ---------- */
/* C(n1 x n2) = A^T B, A: m x n1, B: m x n2 (ld = m). */
void or_gemm_tn(int m, int n1, int n2, const double* A, const double* B, double* C);
/* C(m x n) = alpha*A(m x k)*B(k x n) + beta*C. */
void or_gemm_nn(int m, int n, int k, double alpha, const double* A, int lda,
                const double* B, int ldb, double beta, double* C, int ldc);
/* Upper Cholesky A = U^T U, in place (lapack Potrf 'U').  Returns info. */
int or_potrf_upper(int n, double* A, int lda);
/* X(m x n) <- X * U^{-1}  (blas Trsm 'R','U','N','N'). */
void or_trsm_right_upper(int m, int n, const double* U, int ldu, double* X, int ldx);
/* Symmetric eigensolver (upper triangle used; vectors overwrite A),
 * eigenvalues ascending (lapack Syevd 'V','U'). */
int or_syev(int n, double* A, int lda, double* w);
/* lapack::Sygvd(1, 'V', 'U') (include/lapack.hpp): A <- B-orthonormal
 * eigenvectors (largest-magnitude component positive), w ascending,
 * B <- its Cholesky factor U.  -4 when B is not positive definite. */
int or_sygvd(int n, double* A, int lda, double* B, int ldb, double* w);
/* EigenSolver::SolveGenEig sequential branch (eigensolver.cpp:224-297). */
int or_solve_gen_eig(int lda, int numCol, int width, double* A, double* B, double* eig);
/* EigenSolver::Orthogonalize complex-X branch (eigensolver.cpp:386-469),
 * non-ScaLAPACK path: S = Re(X^H X), Potrf('U'), X <- X U^{-1}. */
int or_orthogonalize(int npw, int nb, double* X);

/* ---------- eigensolvers ---------- */
/* EigenSolver::LOBPCGSolveReal sphere branch (lobpcg.cpp:119-576). */
int or_lobpcg(or_ham* H, int nb, int numEig, int scfIter, int eigMaxIter,
              double eigMinTolerance, double eigTolerance, double* psi,
              double* eigval, double* resval, int* iters_out);

/* EigenSolver::PPCGSolveReal sphere branch (ppcg.cpp:119-576). */
typedef void (*or_op)(void* ctx, int nb, const double* x, double* y);
int or_ppcg_op(int m, or_op hop, or_op pop, void* ctx, int nb, int numEig, int scfIter, int eigMaxIter,
               double eigTolerance, int sbSize, double* psi, double* eigval, double* resval, int* iters_out);
int or_rs_ppcg(const int* Nc, const int* Nf, double vol, const double* ekin, const double* vtot, int natom,
               const int* npts, const int* nproj, const int* idx, const double* val, const double* wgt,
               int nb, int numEig, int scfIter, int eigMaxIter, double eigTolerance, int sbSize, double* psi,
               double* eigval, double* resval, int* iters);
int or_ppcg(or_ham* H, int nb, int numEig, int scfIter, int eigMaxIter,
            double eigMinTolerance, double eigTolerance, int sbSize, double* psi,
            double* eigval, double* resval, int* iters_out);

/* k-step Lanczos on v0 (normalised internally): Ritz values of T_k and the
 * upper bound b_up = theta_k + |beta_k| (Zhou & Li 2011). */
int or_lanczos(or_ham* H, int k, const double* v0, double* theta /* k */,
               double* b_up);
/* Scaled Chebyshev filter (Zhou, Saad, Tiago, Chelikowsky, JCP 219 (2006)
 * 172, Alg. "Chebyshev_filter_scaled"): X <- p_m(H) X, suppressing [a,b],
 * scaled at aL. */
int or_cheb_filter(or_ham* H, int nb, int m, double a, double b, double aL,
                   double* X);
/* Cholesky QR + Rayleigh-Ritz: X <- orth(X) rotated to Ritz vectors,
 * eig = Ritz values ascending, res = ||H x_j - theta_j x_j||. */
int or_rayleigh_ritz(or_ham* H, int nb, double* X, double* eig, double* res);
/* One ChebFSI sweep: Lanczos(k) bounds, degree-m filter, Cholesky QR, RR.
 * If prev_eig != NULL, a = prev_eig[nb-1] and aL = prev_eig[0]
 * (GeneralChebyStep); else aL = theta_1 and a = theta_1 + cut*(b_up - theta_1)
 * from the Lanczos values (first step, cut = 0.1). */
int or_chebfsi_sweep(or_ham* H, int nb, int m, int k, const double* v0,
                     const double* prev_eig, double* X, double* eig,
                     double* res, double* bounds_out /* a, b, aL */);

/* ---------- density and forces (adjoint checks) ---------- */
/* KohnSham::CalculateDensity sphere branch, before renormalisation
 * (hamiltonian.cpp:2509-2573): rho = sum_i numSpin*occ_i*(C2R(c_i)/sqrt(N))^2 */
int or_density(or_ham* H, int nb, const double* psi, const double* occ,
               double num_spin, double* rho);
/* Band energy E = sum_i f_i Re<psi_i|H psi_i>. */
double or_band_energy(or_ham* H, int nb, const double* psi, const double* f);
/* Nonlocal force F_a = -dE_nl/dR_a for Gaussian projectors laid out
 * atom-major with nproj_per_atom columns per atom (vnlc_ layout,
 * hamiltonian.cpp:693-797), using d beta/dR = -iG beta. */
int or_nonlocal_force(or_ham* H, int nb, const double* psi, const double* f,
                      int natom, int nproj_per_atom, double* force /* 3*natom */);

/* ---------- Harris energy (or_scf.c) ---------- */
int or_harris_energy(int ntot, int nspin, double vol, const double* density, const double* spindensity,
                     const double* vxc, const double* vhart, const double* pseudo, int nstate, const double* eig,
                     const double* occ, int num_spin, int num_occ, const double* terms, double* efree);

#ifdef __cplusplus
}
#endif
#endif
