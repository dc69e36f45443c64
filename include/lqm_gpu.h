/*
 * lqm_gpu.h -- C ABI of liblqmgpu.so, the MI355X (gfx950) k-space band path of
 * LatticeQM.jl: Bloch assembly H(k) = sum_R t_R e^{2 pi i k.R}, batched Hermitian
 * eigensolve over a whole k-grid, DOS, chemical potential and the density matrix
 * rho(R) = sum_k w_k e^{-2 pi i k.R} U f(eps) U^H.
 *
 * Conventions (all entry points):
 *   - Plain C types only.  Complex numbers are interleaved (re, im) doubles, the
 *     layout of Julia's ComplexF64 and C99 double _Complex.
 *   - Arrays are column-major like Julia's:
 *       R      int32  [d x nR]       lattice vectors (keys of the Hops dict)
 *       T      cplx   [N x N x nR]   hopping blocks, block r = T + 2*r*N*N doubles
 *       ks     double [d x Nk]       fractional k-points (Structure.points(ks))
 *       bands  double [N x Nk]       ascending eigenvalues, column j = k_j
 *       evecs  cplx   [N x N x Nk]   column n of block j = eigenvector of bands[n, j]
 *       rho    cplx   [N x N x nRr]  density-matrix blocks for the keys Rr
 *   - Every function returns an lqm_status; on failure lqm_last_error() returns a
 *     message for the calling thread.  No exceptions or callbacks cross the ABI.
 *   - Host-buffer functions are synchronous: pointers are read/written only during
 *     the call.  *_dev functions take device pointers, enqueue their kernels on the
 *     context's stream and return once that stream has reached the end of the
 *     call (the device-side status word is read back then).
 *   - A context is used by one host thread at a time; several may coexist.
 */
#ifndef LQM_GPU_H
#define LQM_GPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    LQM_OK = 0,
    LQM_ERR_ARG = 1,         /* shape / argument error (DimensionMismatch, ArgumentError) */
    LQM_ERR_DEVICE = 2,      /* HIP runtime error or device allocation failure */
    LQM_ERR_UNSUPPORTED = 3, /* size or mode this build does not handle */
    LQM_ERR_NOCONV = 4,      /* eigensolver did not converge (LAPACK info > 0 analogue) */
    LQM_ERR_NONFINITE = 5    /* NaN/Inf met in the input or the results */
} lqm_status;

typedef enum {
    LQM_PREC_FP64 = 0,         /* FP64 complex everywhere: the plain path */
    LQM_PREC_FP32_REFINED = 1  /* FP32 H(k) + eigensolve, FP64 Ogita-Aishima refinement of
                                  the eigenpairs, FP64 occupations, rho_k and contraction */
} lqm_precision;

typedef struct lqm_context lqm_context;
typedef struct lqm_model lqm_model;
typedef struct lqm_sparse_model lqm_sparse_model;

/* Message of the last failing call on this thread ("" if none). */
const char *lqm_last_error(void);
/* ABI version, bumped on incompatible changes. */
int lqm_abi_version(void);

/* ---- context ------------------------------------------------------------------ */
/* Replaces the executor choice of Parallel.to_executor (Parallel/Parallel.jl:85-109):
 * a context is the GPUExec a multimode=:gpu call resolves to. */
lqm_status lqm_context_create(int device, int precision, lqm_context **out);
lqm_status lqm_context_destroy(lqm_context *ctx);
/* Run on an existing hipStream_t (NULL restores the context's own stream).  Work the
 * context places on its internal side stream is forked from and joined back to this stream
 * with events, so every call is ordered on it as a whole. */
lqm_status lqm_context_set_stream(lqm_context *ctx, void *hip_stream);
lqm_status lqm_context_synchronize(lqm_context *ctx);
/* Kernel-stage timing (no reference counterpart: the reference has no tracer).
 * When enabled every kernel launch is bracketed by HIP events on the context
 * stream.  lqm_context_stage_times returns, per stage, the accumulated device
 * milliseconds, launch counts and k-points processed (0 for stages that do not work
 * per k-point) since the last query, then resets them; any output may be NULL.
 * Stages: 0 hk_tridiag, 1 tridiag_ql, 2 replay, 3 backtransform, 4 zfz,
 * 5 rho_twosided, 6 phase_contract, 7 phase_fold, 8 dos, 9 reduce, 10 herm_hk_fp64,
 * 11 zgemm_nn_fp64, 12 zgemm_cn_fp64, 13 zgemm_nc_fp64, 14 oa_correct, 15 hf_meanfield,
 * 16 mix, 17 mu_bisect, 18 blocked_hk, 19 blocked_panel, 20 blocked_update,
 * 21 blocked_bisect, 22 blocked_invit, 23 blocked_backtransform, 24 blocked_rho,
 * 25 oa_correct_later (14 = the first refinement sweep, on every k-point; 25 = later sweeps
 * and cluster corrections, on listed k-points), 26 response_sums (the ldos / density site
 * weights and the Kubo pair sums), 27 partial_filter (the Chebyshev filter of the partial
 * spectrum) (LQM_NSTAGES = 28); lqm_stage_name gives the names. */
#define LQM_NSTAGES 28
lqm_status lqm_context_set_timing(lqm_context *ctx, int enable);
/* SCF iterations from captured HIP graphs (no reference counterpart; default on).  Where an
 * iteration needs no host round trip (per-k FP64 kernels, T > 0, the grid in one chunk),
 * lqm_hartreefock_scf runs the first iteration eagerly, captures 8 iterations in a graph and
 * replays it until a device-side step kernel reports convergence, the iteration limit or an
 * error (iterations after that leave the state unchanged); the results are bitwise those of
 * the host-driven loop.  0 keeps the host-driven loop everywhere; timing mode does too. */
lqm_status lqm_context_set_graphs(lqm_context *ctx, int enable);
/* Progress of lqm_hartreefock_scf_batch (the reference's SCF progress bar, no numerical
 * counterpart): fn(problem index, iterations (-1 on error), converged, user) is called once per
 * problem as it finishes, from the batch's worker threads, one call at a time.  NULL: none. */
typedef void (*lqm_progress_fn)(int problem, int iterations, int converged, void *user);
lqm_status lqm_context_set_progress(lqm_context *ctx, lqm_progress_fn fn, void *user);
lqm_status lqm_context_stage_times(lqm_context *ctx, double *ms, int64_t *launches,
                                   int64_t *kpoints, int n);
const char *lqm_stage_name(int stage);

/* Kernel variants (no reference counterpart; the reference's only tunable is
 * the executor).  kind 0: tridiagonalization (0 = 256-thread full tile, 1 = one
 * wave with the packed Hermitian tile, 2 = register tile on a materialised H(k), N <= 64,
 * 3 = column per lane on a materialised H(k), N <= 64, the default there: one wave up to
 * 32 orbitals, two waves above, 4 = the same with FP32 on one wave of 64 rows);
 * kind 1: rho_k (0 = V = QZ then V f V^H,
 * 1 = two-sided reflectors on Z f Z^T, the default for N <= 8); kind 2: Ogita-Aishima sweep of the mixed path
 * (0 = batched GEMM chain, 1 = one fused kernel per k-point with X and W in LDS, 2 = the
 * fused kernel in column-panel form with the products in matrix-core accumulators, two
 * blocks per CU; 1 and 2 need N <= 64, 2 is the default for N > 16);
 * kind 3: k-points per block of the lane-per-k QL (64 >> variant, 0..3); kind 4:
 * back-transformation (0 = reflector by reflector, 1 = compact WY blocks on the matrix
 * cores, N <= 64, the default for N > 16); kind 5: eigensolver layout (0 = per-k-point
 * kernels, 1 = the large-N path, the default for N > 64 in both precision modes:
 * Householder tridiagonalization -- for 64 < N <= 128 one workgroup per k-point with the
 * matrix in registers, a column per lane (kind 0 variant 0 selects the panels instead),
 * above that zlatrd panels of 32 with matrix-core trailing updates -- Sturm
 * bisection, inverse iteration with clusters closer than 1e-5 ||T|| orthogonalised,
 * compact-WY back-transformation and rho_k on the matrix cores; FP64, or FP32 followed by
 * FP64 Ogita-Aishima sweeps in the mixed mode, N <= 4096; occupations below 1e-18 leave
 * rho_k out); kind 6: eigenpairs of the
 * tridiagonal ahead of the compact-WY back-transformation (0 = QL with a rotation log, then
 * its replay; 1 = the default: bisection with Newton steps and twisted-factorisation
 * eigenvectors, O(N^2) per k-point, Gram-Schmidt inside clusters of eigenvalues closer
 * than 1e-3 ||T||, and the QL path for the k-points whose cluster vectors come out
 * dependent; eigenvalue-only passes for N <= 64 use the bisection alone).
 * variant -1 returns to the tuned/default
 * choice.  lqm_autotune times every variant of every kind on Nk device k-points of
 * the model and records the fastest per orbital count N in the context;
 * lqm_context_get_variant reports the choice in force (-1: default, not tuned). */
lqm_status lqm_context_set_variant(lqm_context *ctx, int kind, int variant);

/* Mixed precision (LQM_PREC_FP32_REFINED; no reference counterpart).  Eigenpairs
 * from the FP32 eigensolver are refined in FP64 by Ogita-Aishima sweeps
 * (R = I - X^H X, S = X^H A X, X <- X + X E) with a Rayleigh-Ritz step on
 * near-degenerate clusters; eigenvalues are the Rayleigh quotients s_ii/g_ii with their
 * second-order correction from the off-diagonal of (S, G).  The first sweep refines every
 * k-point; a k-point takes a further sweep only while its last correction rotated or
 * left a cluster, or was so large that its square -- the error the sweep leaves --
 * exceeds the threshold (default 1e-8; <= 0 gives every k-point at least two sweeps),
 * up to max(2, max_sweeps) sweeps (max_sweeps default 3).  The report of the last call:
 * out[0] sweeps used, out[1] max_k ||I - X^H X||_F, out[2] max_k delta (cluster
 * threshold), out[3] k-points still holding near-degenerate pairs, out[4] k-point
 * sweeps run (the sum over sweeps of the k-points refined), out[5] k-points whose rho_k
 * the fused sweep formed in its epilogue (out must hold 6 doubles).  Eigenvector output
 * (bandmatrix with evecs) always uses the FP64 path. */
lqm_status lqm_context_set_refinement(lqm_context *ctx, int max_sweeps);
lqm_status lqm_context_set_refinement_threshold(lqm_context *ctx, double threshold);
lqm_status lqm_context_refinement_report(lqm_context *ctx, double *out);
/* Precision the last pass over a grid actually ran in (LQM_PREC_FP64 or
 * LQM_PREC_FP32_REFINED; -1 before the first pass).  A LQM_PREC_FP32_REFINED context runs
 * the FP64 path for eigenvector output and for the large-N layout (kind 5 = 1); the
 * refinement report is cleared at the start of every pass, so it reads all zeros after a
 * pass that did no refinement.  No reference counterpart (the reference is FP64 only). */
lqm_status lqm_context_last_precision(lqm_context *ctx, int *precision);
lqm_status lqm_context_get_variant(lqm_context *ctx, int kind, int N, int *variant);
lqm_status lqm_autotune(lqm_model *m, int Nk, const double *ks_dev);
/* The autotune cache: the context's tuned variants (kind, N, variant) written to / merged
 * from a text file whose first line names the device architecture ("lqm-tuning 1 gfx950");
 * loading a file tuned on another architecture, or one with an entry out of range, is
 * LQM_ERR_ARG and changes nothing.  No reference counterpart. */
lqm_status lqm_context_save_tuning(lqm_context *ctx, const char *path);
lqm_status lqm_context_load_tuning(lqm_context *ctx, const char *path);

/* ---- model: the Hops operator, resident in HBM --------------------------------- */
/* Upload of Hops{Vector{Int},Matrix{ComplexF64}} (TightBinding/types.jl:54-60):
 * nR keys R (d x nR) and their N x N blocks T.  Key order is the summation order. */
lqm_status lqm_model_create(lqm_context *ctx, int N, int d, int nR, const int32_t *R,
                            const double *T, lqm_model **out);
/* Same, T already in device memory (copied into the model). */
lqm_status lqm_model_create_dev(lqm_context *ctx, int N, int d, int nR, const int32_t *R,
                                const double *T_dev, lqm_model **out);
lqm_status lqm_model_destroy(lqm_model *m);

/* ---- H(k) ------------------------------------------------------------------------ */
/* fouriersum! (TightBinding/types.jl:18-25) at Nk points: Hk [N x N x Nk]. */
lqm_status lqm_fouriersum(lqm_model *m, int Nk, const double *ks, double *Hk);
lqm_status lqm_fouriersum_dev(lqm_model *m, int Nk, const double *ks_dev, double *Hk_dev);

/* ---- bands ------------------------------------------------------------------------ */
/* bandmatrix(H, ks) with no projectors (Spectrum/bands.jl:164-227): eigenvalues
 * of Hermitian(H(k)) (upper triangle), ascending.  evecs may be NULL; when given,
 * the eigenvectors are returned too (the geteigen! branch, bands.jl:222-224). */
lqm_status lqm_bandmatrix(lqm_model *m, int Nk, const double *ks, double *bands, double *evecs);
lqm_status lqm_bandmatrix_dev(lqm_model *m, int Nk, const double *ks_dev, double *bands_dev,
                              double *evecs_dev);

/* bandmatrix(H, ks, projectors) (Spectrum/bands.jl:52-66, :107-112, :222-224): bands and
 * the expectation values obs[n, j, p] = real(psi_nj^H O_p(k_j) psi_nj) [N x Nk x nP] of
 * nP operators, each a model handle on the same context (a k-independent matrix is a
 * model with the single zero key; a Hops operator gives O_p(k) = sum_R O_R e^{2 pi i k.R},
 * the k-dependent ExpvalMap). */
lqm_status lqm_bandmatrix_expval(lqm_model *m, int Nk, const double *ks, int nP,
                                 lqm_model *const *projectors, double *bands, double *obs);
lqm_status lqm_bandmatrix_expval_dev(lqm_model *m, int Nk, const double *ks_dev, int nP,
                                     lqm_model *const *projectors, double *bands_dev,
                                     double *obs_dev);

/* ---- eigenvector consumers ------------------------------------------------------- */
/* ldos(H, ks, frequencies; G) (Spectrum/dos.jl:57-105): site-resolved local DOS
 *   ldos[i] = 1/(Nk pi) sum_k sum_n |psi_nk[i]|^2 (1/Nw) sum_o G / ((w_o - e_nk)^2 + G^2)
 * (with kweights: sum_k w_k ... / pi).  One frequency is the reference's result exactly; over a
 * window the reference divides its running per-task accumulator by Nw once per k-point (an
 * executor-dependent result), here the window is averaged once per state. */
lqm_status lqm_ldos(lqm_model *m, int Nk, const double *ks, const double *kweights, int Nw,
                    const double *omegas, double gamma, double *ldos);
lqm_status lqm_ldos_dev(lqm_model *m, int Nk, const double *ks_dev, const double *kweights_dev,
                        int Nw, const double *omegas_dev, double gamma, double *ldos_dev);
/* density(H, ks, mu) (Spectrum/dos.jl:20-52): n[i] = 1/Nk sum_k sum_{e_nk <= mu} |psi_nk[i]|^2
 * (with kweights: sum_k w_k ...). */
lqm_status lqm_density(lqm_model *m, int Nk, const double *ks, const double *kweights, double mu,
                       double *n);
lqm_status lqm_density_dev(lqm_model *m, int Nk, const double *ks_dev, const double *kweights_dev,
                           double mu, double *n_dev);
/* opticalconductivity(frequencies, H, J1, J2, ks; mu, G, T) (LinearResponse/
 * opticalconductivity.jl:101-127 with kubo!, kubo.jl:11-31): the Kubo current-current
 * correlator, complex chi [Nw] (interleaved re, im),
 *   chi(w) = -(i/Nk) sum_k sum_{m != n} (f_m - f_n) <m|J1(k)|n> <n|J2(k)^H|m> / (e_m - e_n - w - iG)
 * with f = fermidirac(e; mu, T) (T = 0: the step with 1/2 at e = mu); with kweights the sum
 * over k carries w_k instead of 1/Nk.  j1, j2: model handles (Hops of the currents) on the
 * context of h; j1 == j2 computes U^H J U once. */
lqm_status lqm_kubo(lqm_model *h, lqm_model *j1, lqm_model *j2, int Nk, const double *ks,
                    const double *kweights, int Nw, const double *omegas, double mu, double gamma,
                    double T, double *chi);
lqm_status lqm_kubo_dev(lqm_model *h, lqm_model *j1, lqm_model *j2, int Nk, const double *ks_dev,
                        const double *kweights_dev, int Nw, const double *omegas_dev, double mu,
                        double gamma, double T, double *chi_dev);

/* ---- partial spectrum (format = :sparse, num_bands) -------------------------------- */
/* A sparse Bloch Hamiltonian H(k) = sum_r T_r e^{2 pi i k.R_r} in one CSR for all keys: row i
 * holds the entries rowptr[i] .. rowptr[i+1]-1, entry e = (column col[e], key key[e], value
 * val[2e] + i val[2e+1]); entries of one (i, j) under several keys add.  The sparse Hops of
 * the reference's format = :sparse (TightBinding/types.jl:102). */
lqm_status lqm_sparse_model_create(lqm_context *ctx, int N, int d, int nR, const int32_t *R, int64_t nnz,
                                   const int64_t *rowptr, const int32_t *col, const int32_t *key,
                                   const double *val, lqm_sparse_model **out);
lqm_status lqm_sparse_model_destroy(lqm_sparse_model *m);
/* max_i sum_e |t_e|: a bound on ||H(k)||_2 at every k (the residual scale of the solver). */
lqm_status lqm_sparse_model_bound(lqm_sparse_model *m, double *bound);
/* eigen_sparse(H(k); num_bands, sigma, which) (Eigen/eigen_sparse_krylovkit.jl:44-88) at Nk
 * k-points: which = 0 (:LM) the nb eigenpairs nearest sigma, 1 (:SR) the smallest, 2 (:LR) the
 * largest; eigenvalues ascending, evals [nb x Nk], evecs (may be NULL) [N x nb x Nk] complex,
 * orthonormal columns.  The reference's shift-invert Lanczos is replaced by Chebyshev-filtered
 * subspace iteration (subspace of max(2 nb + 10, 20) vectors, at most 64, so nb <= 56; start
 * block seeded deterministically): each iteration filters the block with a Chebyshev
 * polynomial in B = (H - sigma)^2 (LM) or +-H (SR, LR) that damps B's spectrum above the
 * widest Ritz value, orthonormalises it (shifted CholeskyQR and two plain passes) and takes
 * Rayleigh-Ritz pairs; a k-point is done when every wanted pair has ||H x - l x|| <=
 * tol * bound.  iters (may be NULL) [Nk]: the iteration a k-point converged in, -1 if maxiter
 * ran out (its last Ritz pairs are returned). */
lqm_status lqm_eigen_partial(lqm_sparse_model *m, int Nk, const double *ks, int nb, int which, double sigma,
                             double tol, int maxiter, double *evals, double *evecs, int32_t *iters);
lqm_status lqm_eigen_partial_dev(lqm_sparse_model *m, int Nk, const double *ks_dev, int nb, int which,
                                 double sigma, double tol, int maxiter, double *evals_dev, double *evecs_dev,
                                 int32_t *iters);

/* ---- DOS -------------------------------------------------------------------------- */
/* dos_compute! (Spectrum/dos.jl:202-232): DOS(w) = sum_k wk sum_n Im[1/(w - i G - e_nk)],
 * divided by Nk when kweights == NULL (the unweighted branch, dos.jl:221). */
lqm_status lqm_dos(lqm_model *m, int Nk, const double *ks, const double *kweights, int Nw,
                   const double *omegas, double gamma, double *dos);
/* DOS from eigenvalues already on the device (bands_dev [N x Nk]). */
lqm_status lqm_dos_from_bands_dev(lqm_context *ctx, int N, int Nk, const double *bands_dev,
                                  const double *kweights_dev, int Nw, const double *omegas_dev,
                                  double gamma, double *dos_dev);

/* Allocation-free variant of lqm_dos_from_bands_dev.  No reference counterpart: it is for
 * callers that cannot allocate device memory (code under HIP graph capture, tight serving
 * loops, harnesses that forbid hipMalloc).  Same kernels and result; the DOS partials live in
 * a fixed device array of 1024 x 256 doubles (2 MiB, held by each device the library runs on
 * until the process ends) instead of the context's workspace, and the call makes no
 * hipMalloc / hipFree.  Nw above 262144 returns LQM_ERR_UNSUPPORTED.  Every call on one device
 * shares that array: calls that overlap on one device (several contexts or streams) race. */
lqm_status lqm_dos_from_bands_dev_noalloc(lqm_context *ctx, int N, int Nk, const double *bands_dev,
                                          const double *kweights_dev, int Nw,
                                          const double *omegas_dev, double gamma, double *dos_dev);

/* ---- chemical potential ----------------------------------------------------------- */
/* chemicalpotential(H, ks, filling; T) (Spectrum/chemicalpotential.jl:6-65):
 * T == 0: mean of the i-th and (i+1)-th smallest of all N*Nk eigenvalues with
 * i = 1 + floor(filling*(L-1)); T > 0: bisection of getfilling(mu) = filling on
 * [min, max] down to adjacent doubles.  kweights NULL = unweighted mean. */
lqm_status lqm_chemicalpotential(lqm_model *m, int Nk, const double *ks, const double *kweights,
                                 double filling, double T, double *mu);
/* Same from eigenvalues already on the device; *mu is a host scalar. */
lqm_status lqm_chemicalpotential_from_bands_dev(lqm_context *ctx, int N, int Nk,
                                                const double *bands_dev,
                                                const double *kweights_dev, double filling,
                                                double T, double *mu);

/* ---- density matrix --------------------------------------------------------------- */
/* getdensitymatrix!(rho, H, ks, [kweights,] mu; T) (Operators/densitymatrix.jl:17-145):
 * rho (keys Rr, d x nRr) is overwritten with sum_k w_k e^{-2 pi i k.Rr} U f U^H,
 * f = fermidirac(e - mu; T), w_k = 1/Nk when kweights == NULL.  *eband receives
 * sum_k w_k sum_n e_n f_n, the value the reference returns. */
lqm_status lqm_densitymatrix(lqm_model *m, int Nk, const double *ks, const double *kweights,
                             double mu, double T, int nRr, const int32_t *Rr, double *rho,
                             double *eband);
/* Device variant: ks_dev, kweights_dev (may be NULL), rho_dev, eband_dev (one double)
 * are device pointers; Rr stays a host array.  bands_dev, if not NULL, receives
 * the ascending eigenvalues [N x Nk] of the same pass. */
lqm_status lqm_densitymatrix_dev(lqm_model *m, int Nk, const double *ks_dev,
                                 const double *kweights_dev, double mu, double T, int nRr,
                                 const int32_t *Rr, double *rho_dev, double *eband_dev,
                                 double *bands_dev);

/* ---- spectral adjoints: grand potential and weighted density matrices -------------- */
/* No reference counterpart (the reference has no grand potential or gradients); these are
 * the pullbacks a ChainRulesCore rrule for the band path binds.
 * Grand potential Omega = sum_k w_k sum_n omega(e_nk), omega(e) = -T ln(1 + e^{-(e - mu)/T})
 * (T > 0) or min(e - mu, 0) (T = 0); w_k = 1/Nk when kweights == NULL.  grad (may be NULL)
 * receives dOmega/dt_R for every model key R, in model key order [N x N x nR]:
 * [grad_R]_ij = dOmega/d(t_R)_ij = rho(-R)_ji (Hellmann-Feynman: dOmega = sum_k w_k
 * tr(f(H_k) dH_k)), with rho(-R) the density matrix of getdensitymatrix! at (mu, T).
 * For a perturbation that keeps H Hermitian (t_R += h D, t_-R += h D^H) the directional
 * derivative is 2 Re tr(rho(-R) D) (R != 0) or tr(rho(0) D) (R = 0, D Hermitian). */
lqm_status lqm_grandpotential(lqm_model *m, int Nk, const double *ks, const double *kweights,
                              double mu, double T, double *omega, double *grad);
/* Device variant: ks_dev, kweights_dev, grad_dev on the device; *omega a host scalar. */
lqm_status lqm_grandpotential_dev(lqm_model *m, int Nk, const double *ks_dev,
                                  const double *kweights_dev, double mu, double T, double *omega,
                                  double *grad_dev);
/* Weighted density matrix rho_g(R) = sum_k w_k e^{-2 pi i k.R} U_k diag(g_k) U_k^H on the
 * keys Rr, where occ [N x Nk] holds g_k for the ascending eigenvalues at k_j.  The
 * pullback of any spectral loss L = sum_k w_k sum_n l(e_nk) is dL/d(t_R)_ij =
 * rho_{l'}(-R)_ji (occ = l'(bands)).  rho is overwritten. */
lqm_status lqm_densitymatrix_occ(lqm_model *m, int Nk, const double *ks, const double *kweights,
                                 const double *occ, int nRr, const int32_t *Rr, double *rho);
/* Device variant (Rr stays a host array; bands_dev may be NULL). */
lqm_status lqm_densitymatrix_occ_dev(lqm_model *m, int Nk, const double *ks_dev,
                                     const double *kweights_dev, const double *occ_dev, int nRr,
                                     const int32_t *Rr, double *rho_dev, double *bands_dev);

/* ---- Bogoliubov-de Gennes (Nambu) operators ----------------------------------------- */
/* m holds the Nambu model H_BdG at mu = 0 (BdGOperator, Superconductivity/types.jl:43-107):
 * N = 2d, electrons in rows 0..d-1, holes in d..2d-1, and it must have the zero key.  A
 * chemical potential mu enters as H_BdG - mu tau_z on the zero key (addchemicalpotential!
 * with -mu, Superconductivity/operators.jl:14-17), written on the device from a saved copy
 * of the block and undone before the call returns.
 *
 * getdensitymatrix!(rho::BdGOperator, H::BdGOperator, ks, [kweights,] mu; T)
 * (Superconductivity/green.jl:38-54): rho [N x N x nRr] on the keys Rr is overwritten with
 * the Nambu density matrix of H - mu tau_z; *ekin (may be NULL) receives the physical
 * <H - mu N> = E_band / 2 + (Re tr h_e(0) - d mu) / 2, h_e(0) the electron block of the zero
 * key. */
lqm_status lqm_bdg_densitymatrix(lqm_model *m, int Nk, const double *ks, const double *kweights,
                                 double mu, double T, int nRr, const int32_t *Rr, double *rho,
                                 double *ekin);
/* Spectrum.filling(H::BdGOperator, ks, mu; T) (Superconductivity/operators.jl:107-112):
 * Re tr(rho_el(0)) / d of the full Nambu density matrix at mu (one density-matrix pass). */
lqm_status lqm_bdg_filling(lqm_model *m, int Nk, const double *ks, const double *kweights,
                           double mu, double T, double *filling);
/* Spectrum.chemicalpotential(H::BdGOperator, ks, target; T) (Superconductivity/operators.jl:
 * 146-162): the root of filling(mu) - target on the bracket [lo, hi] by Roots.jl's Float64
 * bisection (bit-pattern midpoints down to adjacent doubles), one full Nambu density-matrix
 * pass per evaluation, all on the device (one scalar comes back per pass).  The caller
 * chooses the bracket as the reference does (extrema of the electron-sector bands +- 1,
 * clipped to [mu_lo, mu_hi]).  *passes (may be NULL) receives the passes run;
 * LQM_ERR_ARG when the target is not bracketed. */
lqm_status lqm_bdg_chemicalpotential(lqm_model *m, int Nk, const double *ks, const double *kweights,
                                     double target, double T, double lo, double hi, double *mu,
                                     int *passes);

/* ---- Berry phases, Chern numbers, Wilson loops ---------------------------------------- */
/* Link variables of the band topology on eigenvector grids (Spectrum/berry.jl): U holds npts
 * states, each an N x M block of M eigenvector columns (column-major, block p at
 * U + 2*p*N*M doubles); pairs [2 x nlink] names the states (a_l, b_l) of each link.
 * lqm_link_overlaps: O_l = U_a^H U_b [M x M x nlink] (the overlaps of Wilson1D, berry.jl:
 * 370-383), on the matrix cores.  lqm_link_phases: the unit phase det(O_l) / |det(O_l)|
 * [2 x nlink doubles] (L(State1, State2), berry.jl:116-119; (0, 0) for a singular overlap) by
 * LU with partial pivoting on the device.  The plaquette sums of berry / getcherns /
 * gapcherns are formed from these phases by the host wrappers.  *_dev: U, O, phase are
 * device pointers, pairs stays on the host. */
lqm_status lqm_link_overlaps(lqm_context *ctx, int N, int M, int npts, const double *U, int nlink,
                             const int32_t *pairs, double *O);
lqm_status lqm_link_phases(lqm_context *ctx, int N, int M, int npts, const double *U, int nlink,
                           const int32_t *pairs, double *phase);
lqm_status lqm_link_overlaps_dev(lqm_context *ctx, int N, int M, const double *U_dev, int nlink,
                                 const int32_t *pairs, double *O_dev);
lqm_status lqm_link_phases_dev(lqm_context *ctx, int N, int M, const double *U_dev, int nlink,
                               const int32_t *pairs, double *phase_dev);

/* ---- Hartree-Fock mean field and self-consistency ----------------------------------- */
/* Flags of the HF functional: which channels HartreeFock(h, v; hartree, fock) keeps
 * (Meanfield/hartreefock.jl:158). */
#define LQM_HF_HARTREE 1
#define LQM_HF_FOCK 2

/* The HartreeFock functional applied to rho (Meanfield/hartreefock.jl:178-221; the
 * meanfieldOperator! / meanfieldScalar! pair behind hf(rho)):
 *   H_MF[L] = h[L] - v[L] .* rho[L]                      (Fock, no conjugate, :247-276)
 *   H_MF[0][i,i] += (V0 * diag(rho[0]))[i], V0 = sum_L v[L]   (Hartree, :278-289)
 *   eps_H = 1/2 n^T V0 n, eps_F = -1/2 sum_L sum_ij v_L |rho_L|^2   (:392-410).
 * h: nRh keys Rh [d x nRh] with blocks Th; v: nRv keys Rv, blocks Tv; rho: nRr keys Rr,
 * blocks rho.  The key set of H_MF is the keys of h in order, then the keys of v that h
 * lacks, then the zero key if neither has it: *nK_out (<= nRh + nRv + 1) keys are
 * written to RK_out [d x nK] and their blocks to hmf_out [N x N x nK] (RK_out and
 * hmf_out may be NULL).  energies_out[2] = {eps_H, eps_F}.  The Hartree term needs the
 * zero key in rho (LQM_ERR_ARG otherwise); a complex eps_H or eps_F beyond
 * sqrt(eps) max(1, N) (a non-Hermitian rho) is LQM_ERR_ARG, the reference's assertion. */
lqm_status lqm_hartreefock_meanfield(lqm_context *ctx, int N, int d, int nRh, const int32_t *Rh,
                                     const double *Th, int nRv, const int32_t *Rv,
                                     const double *Tv, int nRr, const int32_t *Rr,
                                     const double *rho, int flags, int *nK_out, int32_t *RK_out,
                                     double *hmf_out, double *energies_out);

/* Self-consistent Hartree-Fock: solvehartreefock(h, v, rho_init, filling, ks; T, beta,
 * tol, iterations) (Meanfield/selfconsistent.jl:19, :366-410, :256-336 and the linear
 * branch of fixedpoint!, Meanfield/fixedpoint.jl:39-194).  Every iteration, on the device:
 *   H_MF[rho0] and its energies; mu = chemicalpotential(H_MF, ks, filling; T) from an
 *   eigenvalues pass; rho1 = getdensitymatrix(H_MF, ks, mu; T) on the keys Rr, returning
 *   eps_band; E = eps_band - (eps_H + eps_F); rho1 <- beta rho1 + (1 - beta) rho0;
 *   residual = ||rho1 - rho0|| / ||rho1|| (Frobenius over all blocks, after mixing;
 *   absolute when ||rho1|| <= eps); rho0 <- rho1; stop when residual < tol.
 * Then H_MF and the energies are refreshed at the final rho and
 *   eps_kin = eps_band - 2 (eps_H + eps_F), E_GS = eps_band - (eps_H + eps_F).
 * rho [N x N x nRr] holds the initial guess on entry (the caller has made it Hermitian,
 * sanitize!) and the final density matrix on return.  kweights (may be NULL: uniform
 * 1/Nk) weight both mu and rho.  result[8] = {E_GS, mu, eps_band, eps_H, eps_F, eps_kin,
 * residual, iterations}; *converged = residual < tol; history (NULL or 3 x max_iter
 * doubles) receives (E, residual, mu) per iteration; nK_out / RK_out / hmf_out as in
 * lqm_hartreefock_meanfield, for the final H_MF.  The context's precision selects the
 * eigensolver of both k passes; mu, occupations, mixing and energies are FP64. */
lqm_status lqm_hartreefock_scf(lqm_context *ctx, int N, int d, int nRh, const int32_t *Rh,
                               const double *Th, int nRv, const int32_t *Rv, const double *Tv,
                               int nRr, const int32_t *Rr, double *rho, int Nk, const double *ks,
                               const double *kweights, double filling, double T, double beta,
                               double tol, int max_iter, int flags, double *result,
                               int *converged, double *history, int *nK_out, int32_t *RK_out,
                               double *hmf_out);

/* Grand potential of the mean field and its pullback onto the mean-field parameters (no
 * reference counterpart; the rrule of Omega(h, v) = Omega[H_MF(rho; h, v)] at fixed rho).
 * H_MF = HartreeFock(h, v)(rho) as in lqm_hartreefock_meanfield; *omega = Omega of H_MF at
 * (mu, T) as in lqm_grandpotential.  grad_h [N x N x nRh] and grad_v [N x N x nRv] (either
 * may be NULL) receive dOmega/d(h_R)_ij = G_R,ij and
 *   dOmega/d(v_L)_ij = -G_L,ij rho_L,ij (Fock, L a key of rho) + G_0,ii n_j (Hartree),
 * with G_q = rho_MF(-q)^T the grand-potential gradient of H_MF and n = diag(rho(0)): the
 * derivative along a perturbation that keeps H_MF Hermitian is sum_ij G_ij D_ij over the
 * perturbed blocks (for the onsite Hubbard form v(0) = U (I (x) sigma_x), dOmega/dU is
 * sum_ij grad_v(0)_ij (I (x) sigma_x)_ij). */
lqm_status lqm_hartreefock_grandpotential(lqm_context *ctx, int N, int d, int nRh, const int32_t *Rh,
                                          const double *Th, int nRv, const int32_t *Rv,
                                          const double *Tv, int nRr, const int32_t *Rr,
                                          const double *rho, int flags, int Nk, const double *ks,
                                          const double *kweights, double mu, double T,
                                          double *omega, double *grad_h, double *grad_v);

/* Many independent Hartree-Fock problems of uneven size (no reference counterpart: the
 * reference runs models one at a time).  Each problem is lqm_hartreefock_scf on its own
 * model, k-grid and filling; T, beta, tol, max_iter, flags and d are shared.  The batch
 * runs on nstreams (1..16) worker contexts, each with its own stream, fed from one queue
 * in decreasing N^3 Nk order, so that small models run beside large ones.  Per problem:
 * rho is the initial guess on entry and the result on return; result / converged as in
 * lqm_hartreefock_scf; status = that problem's lqm_status.  Returns the first failure. */
typedef struct {
    int N;
    int nRh;
    const int32_t *Rh;
    const double *Th;
    int nRv;
    const int32_t *Rv;
    const double *Tv;
    int nRr;
    const int32_t *Rr;
    double *rho;
    int Nk;
    const double *ks;
    const double *kweights;
    double filling;
    double result[8];
    int converged;
    int status;
} lqm_hf_problem;
lqm_status lqm_hartreefock_scf_batch(lqm_context *ctx, int d, int M, lqm_hf_problem *problems,
                                     double T, double beta, double tol, int max_iter, int flags,
                                     int nstreams);

#ifdef __cplusplus
}
#endif

#endif /* LQM_GPU_H */
