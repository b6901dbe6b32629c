/*
 * pwdft_hip.h -- C-ABI of libpwdft_hip.so, the MI355X (gfx950) implementation
 * of PWDFT's sphere-cut Gamma-point Hamiltonian application H*psi and the
 * eigensolver filter around it.
 *
 * Every entry point replaces one reference interface; the citation in front
 * of each declaration names it (paths relative to the reference tree).
 * The C++ adapter classes in pwdft_develop_amd/host/ (Fourier, Spinor,
 * KohnSham, EigenSolver) keep the reference's signatures and call these.
 *
 * Conventions
 *   - Grids: real index r = x + n1*(y + n2*z); R2C index
 *     g = kx + h1*(ky + n2*kz), h1 = n1/2+1 (src/SCF/fourier.cpp:169-184).
 *   - A band block is npw x nb complex, column-major, interleaved (re, im),
 *     G = 0 first and sqrt(2) on the other half-sphere coefficients
 *     (src/SCF/spinor.cpp:573-595).  Inner products are Re(a^H b).
 *   - dtype PW_F64: complex double blocks; PW_F32: complex float blocks.
 *   - Block pointers passed to compute entry points are DEVICE pointers,
 *     caller-owned, leading dimension npw.  Setup arrays are HOST pointers.
 *   - Every function returns PW_OK (0) or a negative PW_ERR_* code and sets
 *     a thread-local message readable with pw_last_error().  The C++ adapter
 *     turns non-zero codes into the reference's ErrorHandling exception
 *     (src/global.cpp:71-91).
 *   - A context is bound to one device and one stream; it is not re-entrant
 *     (like Fourier's shared scratch, include/fourier.hpp:120-125).
 */
#ifndef PWDFT_HIP_H
#define PWDFT_HIP_H

#ifdef __cplusplus
extern "C" {
#endif

#define PW_OK 0
#define PW_ERR_ARG (-1)      /* invalid argument */
#define PW_ERR_HIP (-2)      /* HIP / device error */
#define PW_ERR_DOMAIN (-3)   /* "Domain size does not match" (spinor.cpp:549-551) */
#define PW_ERR_STATE (-4)    /* "Fourier is not prepared." / missing setup */
#define PW_ERR_LINALG (-5)   /* non-SPD Gram, "not enough number of columns" */
#define PW_ERR_NOMEM (-6)
#define PW_ERR_UNSUPPORTED (-7)

#define PW_F64 0
#define PW_F32 1

typedef struct pw_ctx pw_ctx;

/* Thread-local text of the last error (never NULL). */
const char* pw_last_error(void);

/* ------------------------------------------------------------------ */
/* Host-side setup: replaces Fourier::Initialize (src/SCF/fourier.cpp:97-244)
 * and Fourier::InitializeSphere real branch (src/SCF/fourier.cpp:541-612).
 * supercell[3*j+i] = component i of lattice vector j (bohr).
 * First call with idx == NULL to get *npw and *npad, then with arrays. */
int pw_sphere_build(const int grid[3], const double supercell[9], double ecut,
                    int* npw, int* npad, int* idx, double* gkk, int* pad_src,
                    int* pad_dst);

/* Fock sphere idxFineFock (src/SCF/fourier.cpp:642-670): R2C indices with
 * |G|^2/2 < 4*ecut, ascending.  idx == NULL returns the count only. */
int pw_fock_sphere_build(const int grid[3], const double supercell[9],
                         double ecut, int* nfock, int* idx, double* gkk);

/* KohnSham::CalculateEkin (src/SCF/hamiltonian.cpp:2629-2667). */
int pw_ekin_teter(int npw, const double* gkk, double* ekin, double* teter);

/* Projector block vnlc_ (src/SCF/hamiltonian.cpp:693-797) for Gaussian
 * projectors r^l exp(-r^2/(2 sigma^2)) Y_lm, l = 0..lmax, taken in G space by
 * the closed form of the radial integral that PeriodTable::CalculateNonlocalPP
 * tabulates (src/SCF/periodtable.cpp:1599-1849); m order s | pz px py.
 * wgt_channel[a*(lmax+1)+l] is the coupling of channel l of atom a; the
 * weight written is wgt*(4 pi)^2/Vol (periodtable.cpp:1661-1668).
 * beta: npw x nbeta complex, nbeta = natom*(lmax+1)^2; lmax <= 1. */
int pw_gaussian_projectors(const int grid[3], const double supercell[9],
                           int npw, const int* idx, int natom,
                           const double* pos, double sigma, int lmax,
                           const double* wgt_channel, double* beta, double* w);

/* ------------------------------------------------------------------ */
/* Device context.  Replaces the Fourier object plus the Hamiltonian state the
 * hot path reads (vtot_, ekin_, teter_, vnlc_).  Inputs are the reference's
 * own arrays: idxCoarseCut (npw), idxCoarsePadding first/second (npad), ekin,
 * teter (npw).  max_bands bounds the nb of every later call.
 * stream: a hipStream_t (NULL = default stream). */
int pw_ctx_create(pw_ctx** ctx, const int grid[3], const double supercell[9],
                  int npw, const int* idx, int npad, const int* pad_src,
                  const int* pad_dst, const double* ekin, const double* teter,
                  int max_bands, void* stream);
int pw_ctx_destroy(pw_ctx* ctx);
int pw_ctx_set_stream(pw_ctx* ctx, void* stream);
/* npw, nstick, kxmax, N; informational. */
int pw_ctx_info(const pw_ctx* ctx, int* npw, int* nstick, int* kxmax, long* ngrid);

/* Local potential vtot_[:, spinswitch_] (src/SCF/hamiltonian.cpp:3333-3334),
 * host array of N doubles. */
int pw_set_vtot(pw_ctx* ctx, const double* vtot);

/* vnlc_[0] = (beta, w) (src/SCF/hamiltonian.cpp:693-797), host arrays. */
int pw_set_nonlocal(pw_ctx* ctx, int nbeta, const double* beta, const double* w);

/* KohnSham::MultSpinor(Spinor&, NumTns<Complex>&, Fourier&)
 * (src/SCF/hamiltonian.cpp:3270-3452) with Spinor::AddMultSpinorFineR2C
 * (src/SCF/spinor.cpp:527-680): a3 = (-1/2 Laplacian + V_loc) psi, a3
 * overwritten.  No nonlocal term, as in the reference. */
int pw_mult_spinor(pw_ctx* ctx, int dtype, int nb, const void* psi, void* a3);

/* EigenSolver::NonlocalMultX (src/EigenSolver/eigensolver.cpp:471-515):
 * VnlX = beta * (w .* Re(beta^H X)), VnlX overwritten. */
int pw_nonlocal_multx(pw_ctx* ctx, int dtype, int nb, const void* X, void* VnlX);

/* pw_nonlocal_multx without any allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * projections W and the split-K partial products live in a fixed 32 MiB
 * device array per GPU (held until the process ends): fewer K splits are
 * used when the usual count does not fit, and shapes whose nbeta x nb
 * projections alone fill it return PW_ERR_UNSUPPORTED.  Every call on a GPU
 * shares that GPU's array: calls that overlap on one GPU race.  Same result
 * as pw_nonlocal_multx (bit for bit when the usual split count fits). */
int pw_nonlocal_multx_noalloc(pw_ctx* ctx, int dtype, int nb, const void* X, void* VnlX);

/* MultSpinor followed by NonlocalMultX and the add, i.e. A*X as LOBPCG
 * assembles it (src/EigenSolver/lobpcg.cpp:203-212, 312-321). */
int pw_apply_h(pw_ctx* ctx, int dtype, int nb, const void* X, void* HX);

/* Spinor::AddTeterPrecond (src/SCF/spinor.cpp:348-374): a3 += teter .* psi */
int pw_teter_precond(pw_ctx* ctx, int dtype, int nb, const void* psi, void* a3);

/* ------------------------------------------------------------------ */
/* Eigensolver layer.  X blocks are device FP64 (PW_F64 layout) unless a
 * dtype argument says otherwise.  Scratch (6 blocks of max_bands bands plus
 * small matrices) is allocated on first use and kept until
 * pw_free_workspace or pw_ctx_destroy. */

/* EigenSolver::Orthogonalize, complex X / real Gram, non-ScaLAPACK branch
 * (src/EigenSolver/eigensolver.cpp:386-469): S = Re(X^H X), Potrf('U'),
 * X <- X U^{-1}.  PW_ERR_LINALG when S is not positive definite. */
int pw_orthogonalize(pw_ctx* ctx, int nb, void* X);

/* lapack::Syevd('V', 'U', n, A, lda, w) (include/lapack.hpp:107-108) on the
 * device: the dense symmetric eigensolver every Rayleigh-Ritz and
 * SolveGenEig call uses.  A (device FP64, upper triangle read) <- the
 * eigenvectors, each with its largest-magnitude component positive (first
 * index on ties); w (device, n) <- eigenvalues ascending.  Householder
 * tridiagonalisation, multisection and inverse iteration for n <= 512
 * (environment PW_EIGEN=own: n <= 1024), rocSOLVER's dsyevd beyond. */
int pw_syevd(pw_ctx* ctx, int n, double* A, int lda, double* w);

/* lapack::Sygvd(itype 1, 'V', 'U') (include/lapack.hpp:118-124) on nbatch
 * independent problems A x = lambda B x of order 2 <= n <= 256, as PPCGSolveReal
 * solves one per sub-block (src/EigenSolver/ppcg.cpp:438-460): problem b
 * has A at A + b strideA (ld lda) and B at B + b strideB (ld ldb), upper
 * triangles read.  A <- the eigenvectors (x^T B x = 1), B <- the Cholesky
 * factor U in its upper triangle, w + b strideW <- eigenvalues ascending,
 * info[b] <- 0 or the failing pivot column of B (device memory
 * throughout).  Cholesky, C = U^-T A U^-1, the eigensolver of pw_syevd and
 * x = U^-1 y, each stage one launch for the whole batch. */
int pw_sygvd_batched(pw_ctx* ctx, int n, double* A, int lda, long strideA, double* B, int ldb, long strideB,
                     double* w, long strideW, int* info, int nbatch);

/* k-step Lanczos on the device vector v0 (one band, FP64): Ritz values of
 * T_k ascending (host theta[k]) and the upper bound theta_k + |beta_k|.
 * Spectral-bound estimator of the ChebFSI steps the reference only names
 * (src/SCF/scf.cpp:1306-1316, src/esdf.cpp:2714-2719). */
int pw_lanczos(pw_ctx* ctx, int k, const void* v0, double* theta, double* b_up);

/* Scaled Chebyshev filter of degree m (Zhou, Saad, Tiago, Chelikowsky 2006)
 * applied in place to X (dtype blocks): damps [a, b], scaled at aL. */
int pw_cheb_filter(pw_ctx* ctx, int dtype, int nb, int m, double a, double b,
                   double aL, void* X);

/* Cholesky QR + Rayleigh-Ritz on the FP64 block X: X <- Ritz vectors,
 * eig = Ritz values ascending, res = ||H x_j - theta_j x_j|| (host arrays).
 * hdtype: precision of the H X product (PW_F64 or PW_F32). */
int pw_rayleigh_ritz(pw_ctx* ctx, int hdtype, int nb, void* X, double* eig,
                     double* res);

/* One ChebFSI sweep -- the body of the reference's GeneralChebyStep(numEig,
 * filterOrder) call site (src/SCF/scf.cpp:1321-1332): Lanczos(k) bounds from
 * v0 and the degree-m filter in mode precision (PW_F64, or PW_F32 = mixed:
 * FP32 Lanczos and filter, FP64 Cholesky QR and Rayleigh-Ritz), then RR.
 * With prev_eig the
 * cut is prev_eig[nb-1] and the scale point prev_eig[0]; without, the first
 * step uses aL = theta_1, a = theta_1 + 0.1 (b_up - theta_1).  bounds (may
 * be NULL) receives {a, b_up, aL}. */
int pw_chebfsi_sweep(pw_ctx* ctx, int mode, int nb, int m, int k,
                     const void* v0, const double* prev_eig, void* X,
                     double* eig, double* res, double* bounds);

int pw_free_workspace(pw_ctx* ctx);

/* Graph capture of the Chebyshev filter and of the Lanczos steps (default
 * on): the first call of a shape ((dtype, nb, degree), resp. (dtype, k))
 * runs eagerly, the second is captured into a hipGraph, and later calls
 * replay it on a private stream ordered after the context's stream.
 * enable = 0 runs everything eagerly.  No counterpart in the reference (a
 * launch-overhead device for small cells). */
int pw_set_graphs(pw_ctx* ctx, int enable);

/* EigenSolver::LOBPCGSolveReal(numEig, scfIter, eigMaxIter, eigMinTolerance,
 * eigTolerance), sphere-cut branch (src/EigenSolver/lobpcg.cpp:59-576), with
 * SolveGenEig's sequential branch (src/EigenSolver/eigensolver.cpp:224-297).
 * X: device FP64 block (nb bands), overwritten with the eigenvectors;
 * eigval/resval: host arrays of nb (eigVal_, resVal_); *iters = iterations
 * done.  Scratch (7 blocks + (3nb)^2 matrices) is allocated per call. */
int pw_lobpcg(pw_ctx* ctx, int nb, int numEig, int scfIter, int eigMaxIter,
              double eigMinTolerance, double eigTolerance, void* X,
              double* eigval, double* resval, int* iters);

/* EigenSolver::PPCGSolveReal(numEig, scfIter, eigMaxIter, eigMinTolerance,
 * eigTolerance), sphere-cut branch (src/EigenSolver/ppcg.cpp:59-576), with
 * sbSize = esdfParam.PPCGsbSize (the sub-block of the small Sygvd problems).
 * X, eigval, resval, iters as pw_lobpcg.  PW_ERR_LINALG carries the
 * reference's "sygvd returned with info" / Potrf failures. */
int pw_ppcg(pw_ctx* ctx, int nb, int numEig, int scfIter, int eigMaxIter,
            double eigMinTolerance, double eigTolerance, int sbSize, void* X,
            double* eigval, double* resval, int* iters);

/* ------------------------------------------------------------------ */
/* Screened exact exchange.  pw_set_exx replaces KohnSham::InitializeEXX
 * (src/SCF/hamiltonian.cpp:3454-3551: exxDiv, kernel
 * 4 pi/|G|^2 (1 - exp(-|G|^2/(4 mu^2))) on the Fock sphere, G = 0 term
 * -exxDiv + pi/mu^2) plus SetPhiEXX (hamiltonian.cpp:3569-3603: phi =
 * psi sqrt(1/(vol N))).  psi_occ: device FP64 block of nocc bands; occ:
 * host occupations; divtype 0 or 1 (EXX_Divergence_Type).  Once set,
 * pw_mult_spinor and pw_apply_h add V_x psi, as KohnSham::MultSpinor does
 * when isHybrid_ && isEXXActive_ (hamiltonian.cpp:3431-3448).  nocc = 0
 * removes the operator. */
int pw_set_exx(pw_ctx* ctx, double ecut, double screen_mu, double exx_fraction,
               int divtype, int nocc, const void* psi_occ, const double* occ);

/* Spinor::AddMultSpinorEXX, sphere-cut branch (src/SCF/spinor.cpp:783-945):
 * a3 += V_x psi for nb bands (dtype blocks). */
int pw_add_exx(pw_ctx* ctx, int dtype, int nb, const void* psi, void* a3);

/* Adaptively compressed exchange: replaces KohnSham::CalculateVexxACE,
 * sphere branch (src/SCF/hamiltonian.cpp:3685-3924).  From the FP64 device
 * block psi (nb bands): W = V_x psi, M = -Re(psi^H W) = L L^T, xi = W L^{-T}.
 * Afterwards pw_mult_spinor and pw_apply_h use a3 -= xi Re(xi^H psi)
 * (hamiltonian.cpp:3347-3430, the isHybridACE branch) in place of the full
 * exchange, in either dtype; pw_add_exx still applies the full V_x.  nb = 0
 * drops the compressed operator; pw_set_exx drops it too.  PW_ERR_LINALG
 * when M is not positive definite (the reference's Potrf failure). */
int pw_build_ace(pw_ctx* ctx, int nb, const void* psi);

/* exxDiv of KohnSham::InitializeEXX (hamiltonian.cpp:3473-3517), host only. */
int pw_exx_divergence(const int grid[3], const double supercell[9], double ecut,
                      double screen_mu, int divtype, double* exx_div);

/* ------------------------------------------------------------------ */
/* Adjoint of the band energy E = sum_i f_i Re<psi_i|H[V, R]|psi_i>.
 *
 * pw_density: rho(r) (+)= sum_i f_i psi_i(r)^2 / N with psi_i = C2R(c_i):
 * the accumulation of KohnSham::CalculateDensity, sphere branch
 * (src/SCF/hamiltonian.cpp:2509-2573, f_i = numSpin * occ_i, before the
 * renormalisation at :2599-2600), which is also dE/dV(r).  rho: device array
 * of N doubles; f: host.  Bands are recomputed from the sphere, never kept
 * in real space. */
int pw_density(pw_ctx* ctx, int dtype, int nb, const void* psi, const double* f,
               double* rho, int accumulate);

/* KohnSham::CalculateDensity, sphere branch, one density component
 * (hamiltonian.cpp:2453-2627): the accumulation above with
 * f_i = num_spin * occ_i, then the renormalisation
 * rho *= num_spin * num_occupied * N / (vol * sum(rho)) (:2599-2600).
 * *val (may be NULL) = sum(rho) vol / N after it, the reference's val
 * (SCF's totalCharge_).  occ: host, nb values; rho: device, N doubles. */
int pw_calculate_density(pw_ctx* ctx, int dtype, int nb, const void* psi, const double* occ,
                         int num_spin, int num_occupied, double* rho, double* val);

/* E = sum_i f_i Re<x_i|H x_i> for a dtype block (H x in that precision,
 * the dot products summed in FP64), host f and E.  The band-energy sum of
 * KohnSham::CalculateHarrisEnergy / CalculateKSEnergy
 * (src/SCF/scf.cpp:1473-1486) for a given block. */
int pw_band_energy(pw_ctx* ctx, int dtype, int nb, const void* X, const double* f, double* E);

/* Harris-Foulkes free energy of the eigenpairs (eig, occ: host, nstate
 * values) against the input density: SCF::CalculateHarrisEnergy
 * (src/SCF/scf.cpp:1987-2140, real branch).  Device arrays on the fine grid
 * of ntot points, stored as the reference's DblNumMat(ntot, ncomp):
 * density (ntot x nspin; component RHO = 0 first), spindensity (ntot x 2,
 * UP then DN; read only when nspin = 2, may be NULL otherwise), vxc
 * (ntot x nspin), vhart and pseudo_charge (ntot).  terms (host, 8 values):
 * Exc, Eself, EIonSR, EVdw, Eext, fermi, Tbeta, temperature.  EVxc and the
 * Hartree energy are FP64 device sums in a fixed order.  *efree (host) =
 * EfreeHarris_: the band sum plus the correction terms at zero temperature
 * (or when every state is occupied), else the Fermi-Dirac free energy. */
int pw_harris_energy(pw_ctx* ctx, int ntot, int nspin, double vol, const double* density,
                     const double* spindensity, const double* vxc, const double* vhart,
                     const double* pseudo_charge, int nstate, const double* eig, const double* occ,
                     int num_spin, int num_occupied, const double* terms, double* efree);

/* Nonlocal force F_a = -dE/dR_a through the projectors, d beta/dR = -iG beta:
 * F_a = -2 sum_i f_i sum_{p in a} w_p Re<beta_p|x_i> Re<beta_p|iG x_i>, the
 * G-space form of KohnSham::CalculateForce's nonlocal term
 * (src/SCF/hamiltonian.cpp:2907-3006).  X: dtype block (PW_F64 or PW_F32:
 * the projections run in that precision, the force sums in FP64).
 * Projectors laid out atom-major, npa per atom (vnlc_ layout); force: host
 * 3*natom. */
int pw_nonlocal_force(pw_ctx* ctx, int dtype, int nb, const void* X, const double* f,
                      int natom, int npa, double* force);

/* Autotuner, in place of FFTW's planner (Fourier::Initialize plans with
 * FFTW_MEASURE, src/SCF/fourier.cpp:62, 112-184: time candidate algorithms
 * per transform shape once, keep the fastest).  For the context's grid and a
 * dtype block of nb bands it times the xy-slab transform layouts (register
 * FFT kernel; LDS kernel with x row pairs per step, y column groups and
 * threads per workgroup) on pw_mult_spinor, and the split-K tilings (64 or
 * 128 square tiles, 256..2048 partial workgroups) of the projector product
 * beta^T X, applies the fastest of each to the context and appends them to
 * the text file cache_path (NULL: no file).  Shapes already in the file are
 * applied without timing.  summary (may be NULL) receives a one-line account
 * (at most len bytes).  Results do not depend on the choice beyond
 * rounding. */
int pw_autotune(pw_ctx* ctx, int dtype, int nb, const char* cache_path, char* summary, int len);

/* Stage timing of pw_mult_spinor / pw_apply_h calls, measured with HIP
 * events recorded on the context's stream around each kernel (no host
 * synchronisation).  pw_set_timing(1) clears the totals; pw_get_timing
 * waits for the recorded events and returns the summed milliseconds per
 * stage: [0] sticks -> inverse z, [1] xy slab, [2] forward z -> sphere,
 * [3] nonlocal GEMMs; *ncalls = number of timed calls (<= 1024 per read). */
int pw_set_timing(pw_ctx* ctx, int enable);
int pw_get_timing(pw_ctx* ctx, double* ms4, int* ncalls);

/* ------------------------------------------------------------------ */
/* Many independent cells in one persistent launch.  No single counterpart
 * in the reference, which runs one cell per MPI job and applies H band by
 * band (Spinor::AddMultSpinorFineR2C, src/SCF/spinor.cpp:562-677, with
 * EigenSolver::NonlocalMultX, src/EigenSolver/eigensolver.cpp:471-515): a
 * batch stands for ncell such jobs.  Each cell c has its own grid
 * (grid[3c..3c+2]), cell (supercell[9c..9c+8]), sphere (npw[c] entries of
 * idx, ekin, teter), padding pairs (npad[c] of pad_src / pad_dst), local
 * potential (N_c entries of vtot), projectors (nbeta[c] columns of 2 npw[c]
 * reals in beta, nbeta[c] weights in wgt) and band count nb[c]; every array
 * holds the cells' parts one after another.  nbeta[c] <= 256.
 * The band blocks of all cells live in one device buffer of complex floats
 * (FP32, the PW_F32 layout per cell): cell c's npw[c] x nb[c] block starts
 * at complex offset xoff[c] = sum_{c' < c} npw[c'] nb[c'].
 * One launch of a persistent grid (as many workgroups as the device holds
 * at once) runs every filter step as four stages over queues that mix all
 * cells' items -- z-inverse sticks of one band and 16 x 16 projection tiles;
 * one z plane of one band; z-forward sticks of one band; 64 rows x 16 bands
 * of the projector apply fused with the recurrence -- with a device-wide
 * barrier between stages.  The reset of the barrier and the launch are
 * captured into a hipGraph on the second call of a shape and replayed
 * afterwards.  Calls return after the launch has finished (the barrier's
 * error word is read back: PW_ERR_HIP if a barrier ever timed out). */
typedef struct pw_batch pw_batch;
int pw_batch_create(pw_batch** batch, int ncell, const int* grid, const double* supercell,
                    const int* npw, const int* idx, const int* npad, const int* pad_src,
                    const int* pad_dst, const double* ekin, const double* teter,
                    const double* vtot, const int* nbeta, const double* beta,
                    const double* wgt, const int* nb, void* stream);
int pw_batch_destroy(pw_batch* batch);
int pw_batch_set_stream(pw_batch* batch, void* stream);
/* Total complex elements of the block buffer, jobs per filter step,
 * workgroups of the persistent grid, LDS bytes per workgroup. */
int pw_batch_info(const pw_batch* batch, long* xtotal, int* njobs, int* nwg, int* lds);
/* Re-plan: x row pairs per slab job, 16-stick chunks per z job, workgroups
 * (0 = the default of each; the grid never exceeds what the device holds at
 * once).  The autotuner's knobs. */
int pw_batch_tune(pw_batch* batch, int rp, int zc, int nwg);
int pw_batch_set_graphs(pw_batch* batch, int enable);

/* Stage timing of the batched launch (as pw_set_timing / pw_get_timing for a
 * context; no counterpart in the reference): when enabled, the launch stamps
 * the device wall clock after each of its stage barriers, and
 * pw_batch_timing returns the last launch's time per stage (ms, 4 values:
 * z-inverse + projections, xy slabs, z-forward, projector apply +
 * recurrence), summed over its steps.  Enabling or disabling drops a
 * captured graph. */
int pw_batch_set_timing(pw_batch* batch, int enable);
int pw_batch_timing(pw_batch* batch, double* stage_ms);
/* The batch's autotuner (as pw_autotune): times work-per-item, row-pair
 * and z-chunk splits on pw_batch_apply_h over every cell, applies the
 * fastest and records it in cache_path (NULL: no file). */
int pw_batch_autotune(pw_batch* batch, const char* cache_path, char* summary, int len);
/* HX = H X (FP32, local + kinetic + nonlocal) for every cell in one launch;
 * X and HX: device, xtotal complex floats each, not aliased. */
int pw_batch_apply_h(pw_batch* batch, const void* X, void* HX);
/* Scaled Chebyshev filter of degree m (as pw_cheb_filter, FP32) applied in
 * place to every cell's block in one launch; bounds: host, (a, b, aL) per
 * cell. */
int pw_batch_cheb_filter(pw_batch* batch, int m, const double* bounds, void* X);
/* The same filter cell by cell through pw_cheb_filter on each cell's own
 * context: the one-by-one reference point. */
int pw_batch_cheb_filter_serial(pw_batch* batch, int m, const double* bounds, void* X);

/* ------------------------------------------------------------------ */
/* Real-space dual-grid Gamma-point path (Use_Real_Space = 1, FFT_Number_Type
 * "even" or "power"): wavefunctions are real on the coarse grid (ntot =
 * n1*n2*n3 values per band, band-major, DEVICE pointers), the local
 * potential and the projectors live on the fine grid (numGridFine).  The
 * handle plays the part of Fourier's R2C and R2CFine plans and buffers
 * (src/SCF/fourier.cpp:166-183, 246-330); grids must be even and the fine
 * grid at least the coarse one in every dimension.
 * ekin: host array on the coarse R2C grid ((n1/2+1)*n2*n3), the reference's
 * ekin_ without sphere cut (src/SCF/hamiltonian.cpp:2629-2667).
 * max_bands caps the bands transformed per batch (the library also keeps
 * its batch buffers within about 8 GiB); any nb is accepted later. */
typedef struct pw_rs pw_rs;
int pw_rs_create(pw_rs** rs, const int grid[3], const int grid_fine[3], double vol,
                 const double* ekin, int max_bands, void* stream);
int pw_rs_destroy(pw_rs* rs);
int pw_rs_set_stream(pw_rs* rs, void* stream);
/* ntot, ntotFine and the bands per batch; informational. */
int pw_rs_info(const pw_rs* rs, long* ntot, long* ntot_fine, int* chunk);

/* vtot_[:, spinswitch_] on the fine grid (src/SCF/hamiltonian.cpp:3120-3122),
 * host array of ntotFine doubles. */
int pw_rs_set_vtot(pw_rs* rs, const double* vtot);

/* pseudo_[a].vnlList (src/SCF/periodtable.cpp:1042-1589), host arrays, per
 * atom a: npts[a] fine-grid indices (the shared iv of the atom's SparseVecs),
 * nproj[a] projectors (<= 32) whose VAL columns form an npts[a] x nproj[a]
 * column-major block of val, and one weight each in wgt.  natom = 0 removes
 * the nonlocal term. */
int pw_rs_set_nonlocal(pw_rs* rs, int natom, const int* npts, const int* nproj,
                       const int* idx, const double* val, const double* wgt);

/* KohnSham::MultSpinor(Spinor&, NumTns<Real>&, Fourier&)
 * (src/SCF/hamiltonian.cpp:3058-3125) with the real Spinor::
 * AddMultSpinorFineR2C (src/SCF/spinor.cpp:376-525): a3 = H psi (kinetic on
 * the coarse grid, local and nonlocal terms on the fine grid through IP_c2f /
 * IP_f2c, src/utility.cpp:3150-3276), a3 overwritten; psi is not modified. */
int pw_rs_mult_spinor(pw_rs* rs, int nb, const double* psi, double* a3);

/* The wavefunction ecut filter of that MultSpinor (set_wfn_filter,
 * include/hamiltonian.hpp:347-352; applied to psi on the first call,
 * src/SCF/hamiltonian.cpp:3090-3116, and to a3 on every call, :3240-3265):
 * x <- C2R(mask .* R2C(x)) in place, zeroing the R2C entries whose |G|^2/2
 * (the handle's ekin, i.e. Fourier::gkkR2C) exceeds cutoff. */
int pw_rs_filter(pw_rs* rs, double cutoff, int nb, double* x);

/* Spinor::AddTeterPrecond, real version (src/SCF/spinor.cpp:309-345):
 * a3 += C2R(teter .* R2C(psi)) with the reference's scalings; teter is
 * KohnSham::CalculateEkin's b / (b + 16 a^4), a = 2 ekin, from the handle's
 * ekin (src/SCF/hamiltonian.cpp:2659-2663).  psi, a3: device, nb x ntot. */
int pw_rs_teter_precond(pw_rs* rs, int nb, const double* psi, double* a3);

/* EigenSolver::PPCGSolveReal, real-space branch (src/EigenSolver/ppcg.cpp:
 * 683-1246): the pw_ppcg iteration on the handle's real coarse-grid bands X
 * (device, nb x ntot, overwritten with the Ritz vectors) with
 * pw_rs_mult_spinor as H and pw_rs_teter_precond as the preconditioner;
 * host eigval, resval (nb), iterations in *iters. */
int pw_rs_ppcg(pw_rs* rs, int nb, int numEig, int scfIter, int eigMaxIter, double eigMinTolerance,
               double eigTolerance, int sbSize, double* X, double* eigval, double* resval, int* iters);

/* KohnSham::CalculateDensity, real-space branch, one density component
 * (src/SCF/hamiltonian.cpp:2453-2627): rho = sum_b num_spin occ_b psiF_b^2 on
 * the fine grid (psiF the fine-grid interpolant of band b), renormalised to
 * rho *= num_spin num_occupied Nf / (vol sum rho) (:2599-2600); *val (may be
 * NULL) = sum(rho) vol / Nf after it.  psi: device, nb x ntot; occ: host,
 * nb; rho: device, ntotFine doubles. */
int pw_rs_calculate_density(pw_rs* rs, int nb, const double* psi, const double* occ, int num_spin,
                            int num_occupied, double* rho, double* val);

#ifdef __cplusplus
}
#endif
#endif
