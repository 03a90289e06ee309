/*
 * dmftgpu.h -- C ABI of the MI355X lattice Green's-function library.
 *
 * Replaces the k-space lattice Green's function path of DMFTpack
 * (src/GreenFtn.cpp, src/FourierTransform.cpp).  Plain C: pointers, sizes and
 * an opaque handle.  No HIP, torch or Eigen types cross this boundary; the HIP
 * stream is passed as a void*.
 *
 * Data layout (all complex numbers are complex128 = two doubles, re then im):
 *   every matrix is column-major, element (i,j) at offset i + j*rows, which is
 *   Eigen's MatrixXcd::data() layout; a stack of matrices is contiguous,
 *   matrix m starting at m*rows*cols.
 *     Hk        [n_k][n][n]            H(k) in the model space, correlated
 *                                      orbitals first (downfolding.cpp:69-79)
 *     sigma     [n_freq][nc][nc]       SelfE_w(n), n < N_freq (ImgTimeFtn.h:32-98)
 *     moments   [4][nc][nc]            m0..m3 of getAsymto_moments
 *     gloc      [n_freq][nc][nc]       Gw (block-diagonal per cluster)
 *   Array arguments are DEVICE pointers unless the comment says "host".
 *   nc = n_corr = NumCluster * NumHartrOrbit_per_cluster.
 *
 * Errors: every function returns a status (0 = success); the message of the
 * last failure is available from dmftgpu_last_error().  The reference asserts
 * or calls exit(1) in the same situations (GreenFtn.cpp:85,113;
 * FourierTransform.cpp:321,440,557); the C++ adapter in dmftgpu_shim.hpp
 * maps a non-zero status back to that behaviour.
 *
 * Threading: a handle is not thread-safe; calls on one handle are ordered on
 * its stream.  Functions that return host scalars synchronise that stream.
 */
#ifndef DMFTGPU_H
#define DMFTGPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DMFTGPU_OK 0
#define DMFTGPU_ERR_ARG 1         /* invalid argument or size out of range */
#define DMFTGPU_ERR_STATE 2       /* dmftgpu_prepare not called / stale */
#define DMFTGPU_ERR_HIP 3         /* HIP runtime error */
#define DMFTGPU_ERR_NUMERIC 4     /* non-finite result, failed bracketing */
#define DMFTGPU_ERR_UNSUPPORTED 5 /* size or option this build does not handle */

/* Methods for the per-(k, iw) solve. */
#define DMFTGPU_PLAIN_F64 0  /* full n x n inverse per (k,iw) in FP64 (GreenFtn.cpp:261) */
#define DMFTGPU_SCHUR_F64 1  /* Schur complement on the correlated block, FP64 */
#define DMFTGPU_SCHUR_F32R 2 /* Schur complement, FP32 solve + FP64-residual refinement */

typedef struct dmftgpu_lattice {
    int32_t n;          /* NBAND[k] (TB.h:29), uniform over k, 1..32 */
    int32_t n_corr;     /* NumCluster*NumHartrOrbit_per_cluster (tight_common.h:61-62) */
    int32_t n_cluster;  /* NumCluster; gloc is block-diagonal over clusters */
    int32_t n_k;        /* knum: k points held in Hk */
    int32_t n_k_global; /* knum_mpiGlobal: normalisation of the k average */
    int32_t n_freq;     /* N_freq (tight_common.h:67) */
    int32_t high_freq;  /* frequencies summed for densities = high_freq*N_freq (highFreq2, GreenFtn.cpp:8) */
    double beta;        /* inverse temperature (tight_common.h:64) */
} dmftgpu_lattice;

typedef struct dmftgpu_handle_s* dmftgpu_handle;

/* ---- handle ---------------------------------------------------------------- */
/* No reference counterpart: the reference holds this state in globals and in
 * GreenFtn_w's locals.  A handle owns the device workspaces, the bound H(k)
 * and its static eigen-decomposition; hip_stream NULL = the default stream. */
int dmftgpu_create(dmftgpu_handle* h, void* hip_stream);
int dmftgpu_destroy(dmftgpu_handle h);
int dmftgpu_set_stream(dmftgpu_handle h, void* hip_stream);
const char* dmftgpu_last_error(dmftgpu_handle h);
const char* dmftgpu_version(void);

/* Tuning options (no reference counterpart). */
#define DMFTGPU_OPT_REFINE_STEPS 1      /* FP64-residual refinement steps of SCHUR_F32R (1 or 2, default 2) */
#define DMFTGPU_OPT_TARGET_WORKGROUPS 2 /* work items per lattice launch, sets the k chunking (0 = default: up to
                                           32768 for the one-matrix-per-lane Schur kernels, keeping >= 64 k per
                                           chunk and >= 8192 items; 8192 for the others) */
#define DMFTGPU_OPT_SCHUR_WG 3          /* frequencies per workgroup of the Schur kernels (64 or 128; default 0 =
                                           the adjoint at 128, the FP32R forward over the first N_freq
                                           frequencies at n_corr >= 5 as two wavefronts over 64 frequencies
                                           splitting the k walk, the others at 64) */
#define DMFTGPU_OPT_SCHUR_LAYOUT 4      /* n_c <= 6: 0 = one (k,iw) matrix per lane (default), 1 = one column
                                           per lane (always used for n_c > 6) */
#define DMFTGPU_OPT_UPFOLD_CHUNK 5      /* k points per chunk of dmftgpu_upfold_density (0 = as many as 512 MiB of
                                           scratch holds, the default) */
#define DMFTGPU_OPT_DENSITY_KERNEL 6    /* dmftgpu_density_matrix: 0 = block form with the frequency sum on FP64
                                           MFMA where it applies (n <= 16, n_corr <= 6; the default), 1 = one full
                                           n x n inverse per (k, iw) always */
int dmftgpu_set_option(dmftgpu_handle h, int option, int value);

/* Instrumentation (no reference counterpart): with timing enabled, HIP events
 * are recorded on the handle's stream around the main lattice kernel of each
 * mode (0 forward, 1 density, 2 VJP); dmftgpu_kernel_time returns the last
 * such kernel's duration in milliseconds (HOST) and waits for it. */
int dmftgpu_enable_timing(dmftgpu_handle h, int enable);
int dmftgpu_kernel_time(dmftgpu_handle h, int mode, double* ms);

/* ---- lattice path ---------------------------------------------------------- */

/* Binds H(k) and the Sigma tail moments to the handle and computes the static
 * eigen-decomposition H(k) + P m0 P^H = U diag(eps) U^H for every k
 * (GreenFtn.cpp:197-215).  Hk must stay valid until the next prepare.
 * For the Schur methods it also diagonalises H_dd(k) once per k. */
int dmftgpu_prepare(dmftgpu_handle h, const dmftgpu_lattice* lat, const double* Hk, const double* moments);

/* New tail moments for the H(k) bound by dmftgpu_prepare (a DMFT iteration
 * changes Sigma, H(k) stays): redoes only the static eigen-decomposition
 * (GreenFtn.cpp:197-215); the Schur data of H(k) is kept. */
int dmftgpu_update_moments(dmftgpu_handle h, const double* moments);

/* Copies the static eigenpairs out: eps [n_k][n] ascending, U [n_k][n][n]
 * (the reference keeps them in GreenFtn_w's locals, GreenFtn.cpp:197-215). */
int dmftgpu_static_eig(dmftgpu_handle h, double* eps, double* U);

/* GreenFtn_w (GreenFtn.cpp:142-422) for one mu.
 *   gloc        [n_freq][nc][nc]  required; inter-cluster blocks are zero (:371-383)
 *   gloc_hf     [n_freq][nc][nc]  optional (NULL): GwHF, the static-part G0 (:277)
 *   num_matrix  [nc][nc]          required; the global NumMatrix (:322-356)
 *   n_total     HOST, optional (NULL): total N(mu) of the same pass, the
 *               quantity GiwToNele returns (GreenFtn.cpp:600-634)
 * Method: DMFTGPU_PLAIN_F64 or DMFTGPU_SCHUR_*. */
int dmftgpu_greenftn_w(dmftgpu_handle h, int method, const double* sigma, double mu, double* gloc, double* gloc_hf,
                       double* num_matrix, double* n_total);

/* No reference counterpart: dmftgpu_greenftn_w for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and results; it calls no
 * hipMalloc / hipFree and uses no host heap.  Its scratch (k-chunk partial
 * sums, the Sigma table, the static weights) is one fixed array of 128 MiB per
 * device, loaded with the library and held until the process ends; the k
 * chunking is reduced until the partial sums fit, and a lattice whose
 * scratch exceeds 128 MiB even as one chunk (roughly n_freq*high_freq*n_corr^2
 * * 48 B plus 16 B per k point and band) is rejected with
 * DMFTGPU_ERR_UNSUPPORTED.  dmftgpu_create takes the array's address and
 * dmftgpu_prepare (which allocates) must have run.  Every call on one device
 * shares the array: calls that overlap on a device race. */
int dmftgpu_greenftn_w_noalloc(dmftgpu_handle h, int method, const double* sigma, double mu, double* gloc,
                               double* gloc_hf, double* num_matrix, double* n_total);

/* Real-axis Green's function of retarded_GreenFtn2 (GreenFtn.cpp:425-466),
 * summed over k as qs_spectra accumulates it (qs_spectra.cpp:118-131):
 *   G^R(k, E_j) = [(E_j + mu + i eta) I - H(k) - P Sigma^R(E_j) P^H]^{-1},
 *   E_j = e0 + j de, j < n_e, eta > 0 (the reference's `infinitesimal`)
 *   sigma_r [n_e][nc][nc]  Sigma^R(E_j)
 *   gloc_r  [n_e][nc][nc]  (1/n_k_global) sum_k G^R_cc (full correlated block)
 *   trace   [n_e] complex, optional (NULL): (1/n_k_global) sum_k tr G^R over
 *           the model space (-Im trace / pi is the total spectral weight)
 * Uses the H(k) bound by dmftgpu_prepare.  The Schur methods run the
 * one-matrix-per-lane kernels for n_corr <= 6 (the column-per-lane ones with
 * DMFTGPU_OPT_SCHUR_LAYOUT = 1, and always above 6). */
int dmftgpu_greenftn_real(dmftgpu_handle h, int method, int n_e, double e0, double de, double eta,
                          const double* sigma_r, double mu, double* gloc_r, double* trace);

/* Per-k form of dmftgpu_greenftn_real, for the band spectra of qs_spectra
 * (qs_spectra.cpp:197-221: A(k, E) = -Im tr G^R(k, E) / pi and the partial
 * weights -Im G^R_ii(k, E) / pi of the correlated orbitals), H(k) bound by
 * dmftgpu_prepare being the k path:
 *   gcc_k   [n_k][n_e][nc][nc]  G^R_cc(k, E_j), or NULL
 *   trace_k [n_k][n_e] complex  tr G^R(k, E_j) over the model space, or NULL
 * (at least one of them); nothing is averaged over k. */
int dmftgpu_greenftn_real_k(dmftgpu_handle h, int method, int n_e, double e0, double de, double eta,
                            const double* sigma_r, double mu, double* gcc_k, double* trace_k);

/* Per-k density matrix densityMatDFT (GreenFtn.cpp:291-319): dens [n_k][n][n],
 * FP64.  For n <= 16 and n_corr <= 6 the frequency sum is taken in block form
 * (Schur complement of the correlated block, then the rank-n_corr frequency
 * sum as one complex product per k on the matrix cores); otherwise one full
 * n x n inverse per (k, iw).  DMFTGPU_OPT_DENSITY_KERNEL = 1 forces the latter. */
int dmftgpu_density_matrix(dmftgpu_handle h, const double* sigma, double mu, double* dens);

/* Kohn-Sham bands of the bound H(k) for the core-band term of GiwToNele
 * (GreenFtn.cpp:622-631), which counts the Fermi weight of the KS bands that
 * lie outside the model window:
 *     (1/knum_mpiGlobal) sum_k [ sum_{l<NumOrbit} f(eps_kl - mu) - sum_{l0<NBAND} f(eps_k,FromValToKS[k][l0] - mu) ]
 *   n_orb      NumOrbit (TB.h:7), 1..4096; 0 removes the bands again
 *   ks_energy  [n_k][n_orb] KS_eigenEnergy (TB.h:19), host or device
 *   val_to_ks  [n_k][n] HOST FromValToKS (TB.h:31): the KS band of each window
 *              band, distinct per k; -1 marks a padded band (NBAND[k] < n)
 * Once bound, N(mu) of dmftgpu_density, dmftgpu_adjust_mu and dmftgpu_greenftn_w
 * (n_total) is the total electron number of GiwToNele, and dN/dmu, mu_bar and
 * the loss of dmftgpu_gloc_loss_grad include the term.  dmftgpu_prepare
 * removes the binding (the bands belong to the H(k) it replaces). */
int dmftgpu_set_core_bands(dmftgpu_handle h, int n_orb, const double* ks_energy, const int* val_to_ks);

/* Total electron number N(mu) of GiwToNele (GreenFtn.cpp:513-635, via the
 * trace identity sum_l 1/(z+mu-lambda_l) = tr G) and its analytic dN/dmu,
 * with the core-band term when Kohn-Sham bands are bound (dmftgpu_set_core_bands).
 * n_out, dn_dmu_out are HOST pointers (either may be NULL). */
int dmftgpu_density(dmftgpu_handle h, int method, const double* sigma, double mu, double* n_out, double* dn_dmu_out);

/* GreenFtn_w_adjust_mu (GreenFtn.cpp:66-139): mu with N(mu) = n_target
 * (NumberOfElectron: with bound Kohn-Sham bands the total count, core bands
 * included, as in the reference).  Safeguarded Newton on the analytic dN/dmu
 * (bracket + bisection fallback), stopping when |dmu| <= tol.  mu0 initial
 * guess; mu_out, n_eval HOST. */
int dmftgpu_adjust_mu(dmftgpu_handle h, int method, const double* sigma, double mu0, double n_target, double tol,
                      double* mu_out, int* n_eval);

/* Vector-Jacobian product of (G_loc, N) with respect to (Sigma(iw_n), n<N_freq)
 * and mu, moments held fixed; G is recomputed, never stored.
 *   gloc_bar  [n_freq][nc][nc]   cotangent of the full nc x nc G_loc
 *   n_bar     cotangent of N
 *   sigma_bar [n_freq][nc][nc]   output, convention dL/dRe + i dL/dIm
 *   mu_bar    HOST output
 * No reference counterpart (the reference has no derivatives). */
int dmftgpu_gloc_vjp(dmftgpu_handle h, int method, const double* sigma, double mu, const double* gloc_bar,
                     double n_bar, double* sigma_bar, double* mu_bar);

/* Jacobian-vector product of (G_loc, N) in (Sigma(iw_n), n < N_freq) and mu,
 * moments held fixed, G recomputed -- the J v of a Newton-Krylov update of
 * the DMFT self-consistency (the reference mixes with Pulay, pulay.cpp:48-167):
 *   dsigma [n_freq][nc][nc], dmu   tangent of (Sigma, mu)
 *   dgloc  [n_freq][nc][nc]        output: dG_loc = (1/Nk) sum_k G_cc (dSigma - dmu S') G_cc,
 *                                  the full nc x nc block (as the cotangent of dmftgpu_gloc_vjp)
 *   dn     HOST output (may be NULL): dN
 * The adjoint of dmftgpu_gloc_vjp: Re<gloc_bar, dgloc> + n_bar dn =
 * Re<sigma_bar, dsigma> + mu_bar dmu.  No reference counterpart. */
int dmftgpu_gloc_jvp(dmftgpu_handle h, int method, const double* sigma, double mu, const double* dsigma, double dmu,
                     double* dgloc, double* dn);

/* Loss and gradient in one call, for fitting / Newton-Krylov updates:
 *   L = sum_{n<N_freq} ||G_loc(iw_n) - G_tgt(iw_n)||_F^2 + (N - n_target)^2
 * with G_loc as dmftgpu_greenftn_w returns it (cluster-block-diagonal).
 *   gloc_target [n_freq][nc][nc]   in;  gloc [n_freq][nc][nc], num_matrix [nc][nc] out
 *   n_total, loss, mu_bar          HOST outputs (each may be NULL)
 *   sigma_bar [n_freq][nc][nc]     out, dL/dRe + i dL/dIm (as dmftgpu_gloc_vjp)
 * Equal to dmftgpu_greenftn_w, the cotangents 2 (G_loc - G_tgt) and 2 (N - n_target),
 * and dmftgpu_gloc_vjp; on the one-matrix-per-lane Schur path the forward also forms
 * tr G^2 beyond N_freq, which the adjoint then skips.  No reference counterpart. */
int dmftgpu_gloc_loss_grad(dmftgpu_handle h, int method, const double* sigma, double mu, const double* gloc_target,
                           double n_target, double* gloc, double* num_matrix, double* n_total, double* loss,
                           double* sigma_bar, double* mu_bar);

/* ---- batched lattices ----------------------------------------------------------
 * No reference counterpart: DMFTpack handles one lattice per call.  Many
 * independent lattices (uneven n <= 32, n_corr <= 16, n_k) run as one
 * replayable step, one persistent launch per n_corr group and mode (one
 * (k, iw) matrix per lane for n_corr <= 6, one column per lane above): per
 * system GreenFtn_w (G_loc, NumMatrix, N) and, when a target is
 * given, the VJP of L = sum_w ||G_loc - gloc_target||^2 + (N - n_target)^2
 * (G_loc cluster-block-diagonal as returned, as in dmftgpu_gloc_loss_grad).
 * All pointers in dmftgpu_system are DEVICE pointers; n_total and mu_bar
 * point to one double each.  The batch owns a stream (default priority) and
 * orders itself after / before the caller's handle stream. */
typedef struct dmftgpu_system {
    dmftgpu_lattice lat;
    const double* Hk;          /* [n_k][n][n] */
    const double* sigma;       /* [n_freq][nc][nc] */
    const double* moments;     /* [4][nc][nc] */
    double mu;
    double* gloc;              /* out [n_freq][nc][nc] */
    double* num_matrix;        /* out [nc][nc] */
    double* n_total;           /* out, 1 double */
    const double* gloc_target; /* optional [n_freq][nc][nc]: enables the VJP */
    double n_target;
    double* sigma_bar;         /* out [n_freq][nc][nc] (with a target) */
    double* mu_bar;            /* out, 1 double (with a target) */
} dmftgpu_system;

typedef struct dmftgpu_batch_s* dmftgpu_batch;
#define DMFTGPU_BATCH_TUNE 1    /* autotune workgroup width and k chunking per n_corr group */
#define DMFTGPU_BATCH_GRAPH 2   /* capture the step into a hipGraph replayed by dmftgpu_batch_run */
#define DMFTGPU_BATCH_PREPARE 4 /* recompute the static eigen-decompositions in every step */
int dmftgpu_batch_create(dmftgpu_handle h, int count, const dmftgpu_system* systems, int method, int flags,
                         dmftgpu_batch* out);
int dmftgpu_batch_run(dmftgpu_batch b);
/* Tuning result of n_corr group `group` (returns the number of groups when out of range):
 * wg[2], target[2], ms[2] per mode (forward, adjoint); wg = 1 is the FP32R forward's k-pair form
 * (two wavefronts over 64 frequencies, each taking every other k of a chunk). */
int dmftgpu_batch_tuning(dmftgpu_batch b, int group, int* nc, int* systems, int* wg, double* target, double* ms);
int dmftgpu_batch_destroy(dmftgpu_batch b);

/* ---- Matsubara tails and Fourier transforms ---------------------------------- */

/* getAsymto_moments (FourierTransform.cpp:184-465): fw [n_freq][d][d] ->
 * moments [4][d][d].  green_asym != 0 selects the G-like branch (m0=0, m1=I).
 * windows (HOST, optional, 4 ints) receives the selected window centres
 * (-1 for a moment the branch does not fit).  The window choice runs on the
 * device: without windows the call only enqueues work on the handle's stream
 * (moments is ready once the stream has passed it); with windows it waits. */
int dmftgpu_moments(dmftgpu_handle h, int d, int n_freq, double beta, const double* fw, int green_asym,
                    double* moments, int* windows);

/* FourierTransform_ (FourierTransform.cpp:126-178) with FourierTransform_tau
 * (:16-118): G(tau_j), tau_j = j beta/n_tau, j = 0..n_tau, two-pole tail
 * subtraction, Hermitisation and the causality patch.
 *   gtau [n_tau+1][d][d]; patched HOST (optional) = the reference's CHECK flag. */
int dmftgpu_ft_w2tau(dmftgpu_handle h, int d, int n_freq, int n_tau, double beta, const double* fw,
                     const double* moments, double* gtau, int* patched);

/* FourierTransform_tau (FourierTransform.cpp:16-118) at n_points arbitrary
 * tau values (taus_host, HOST); out [n_points][d][d].  No causality patch,
 * as in the single-tau FourierTransform overloads (:494-512). */
int dmftgpu_ft_w2tau_points(dmftgpu_handle h, int d, int n_freq, double beta, const double* fw,
                            const double* moments, int n_points, const double* taus_host, double* out);

/* FT_t_to_w (FourierTransform.cpp:521-561): exact transform of the piecewise
 * linear interpolant of G(tau).  violations HOST (optional) counts frequencies
 * failing the causality test of :551-558 (the reference asserts). */
int dmftgpu_ft_tau2w(dmftgpu_handle h, int d, int n_freq, int n_tau, double beta, const double* gtau, double* fw,
                     int* violations);

/* ---- callers either side of the path ----------------------------------- */

/* Construct_hyb_delta (Construct_hyb_delta.cpp:47-119, off-diagonal form
 * :96-117): hybridisation function of one impurity whose dim orbitals are
 * orbit[0..dim) of the correlated block,
 *   delta(iw_n) = (iw_n + mu) I - H_imp - Sigma_imp(iw_n) - [G_imp(iw_n)]^{-1},  n < n_freq,
 * X_imp(a, b) = X(orbit[a], orbit[b]).
 *   gloc   [n_freq][nc][nc]   G_loc (dmftgpu_greenftn_w output)
 *   sigma  [n_freq][nc][nc]   SelfE_w
 *   h_imp  [nc][nc]           impurity_site_Hamiltonian
 *   orbit  [dim]              HOST, 0 <= orbit[a] < nc (impurityOrbit)
 *   delta  [n_freq][dim][dim] weiss_fieldTB block of this atom
 * dim = 0 returns without work, as the reference skips the inverse (:98).
 * Delta -> tau is dmftgpu_moments + dmftgpu_ft_w2tau with green_asym = 0
 * (write_hyb_t, Construct_hyb_delta.cpp:126-138). */
int dmftgpu_hyb_delta(dmftgpu_handle h, int nc, int n_freq, double beta, double mu, const double* gloc,
                      const double* sigma, const double* h_imp, int dim, const int* orbit, double* delta);

/* H(k) = sum_R H(R) exp(-i k.a.R) (ConstructModelHamiltonian.cpp:58-77) from
 * the hopping list of read_HR (read_TB_parm.cpp:6-76):
 *   hop_idx [n_hop][5]  HOST int: R1 R2 R3 i j (0-based orbitals i, j < n)
 *   hop     [n_hop]     HOST complex t
 *   kphase  [n_k][3]    k_x a_x, k_y a_y, k_z a_z
 *   Hk      [n_k][n][n] output; hops of one element are summed in list order. */
int dmftgpu_hk_from_hr(dmftgpu_handle h, int n, int n_hop, const int* hop_idx, const double* hop, int n_k,
                       const double* kphase, double* Hk);

/* upfolding_density (upfolding.cpp:38-139) with upfolding_k (:10-35), the
 * downfolding == 1 branch (:45-51): the per-k density matrix of
 * dmftgpu_density_matrix taken to the full DFT basis and summed to R,
 *   E_k  = (T_k^H)^{-1} K_k                       (:46)
 *   K'_k = diag f(eps_kb - mu), f = 0 / 1 where beta (eps - mu) > 100 / < -100,
 *          with K'_k(v_ki, v_kj) = (C_k^H X_k C_k)(i, j), i, j < n   (:13-32)
 *   D_k  = E_k K'_k E_k^H                         (:34)
 *   dmr[0][r] = sum_k (S_k^{-1} D_k S_k^{-1})(i0, m0) exp(+i kphase_k . R) / n_k_global
 *   dmr[1][r] = same with S_k^{-1} D_k,  dmr[2][r] = same with D_k   (:49-50, :67-86)
 * (the DM_DFT_DMFT.HWR, dual_DM_direct.HWR and direct_DM_direct.HWR values).
 *   n, n_orb   NBAND (uniform over k) and NumOrbit, 1 <= n <= n_orb <= 512
 *   dm_model   [n_k][n][n]          densityMatDFT in the model basis
 *   corr_base  [n_k][n][n]          DF_CorrBase
 *   ks_energy  [n_k][n_orb]         KS_eigenEnergy (double)
 *   val_to_ks  [n_k][n]   HOST int  FromValToKS, distinct bands per k; -1 marks a padded
 *                                   band (NBAND[k] < n), whose rows and columns of
 *                                   dm_model must be zero
 *   transform  [n_k][n_orb][n_orb]  transformMatrix_k
 *   ks_ortho   [n_k][n_orb][n_orb]  KS_eigenVectors_orthoBasis
 *   overlap    [n_k][n_orb][n_orb]  S_overlap, or NULL for S = I
 *   r_idx      [n_r][5]   HOST int  H_Rindex rows: R1 R2 R3 i0 m0 (0-based)
 *   kphase     [n_k][3]             k_x a_x, k_y a_y, k_z a_z
 *   dm_k       [n_k][n_orb][n_orb]  output D_k (densityMatDFT[k] after upfolding_k), or NULL
 *   dmr        [3][n_r] complex     output (n_r = 0: dm_k only)
 * Inverses are Gauss-Jordan with partial pivoting; a singular T_k or S_k
 * returns DMFTGPU_ERR_NUMERIC.  With k split over ranks the R sums are the
 * rank's share (the operands of the reference's MPI_Allreduce, :80-82). */
int dmftgpu_upfold_density(dmftgpu_handle h, int n, int n_orb, int n_k, int n_k_global, double beta, double mu,
                           const double* dm_model, const double* corr_base, const double* ks_energy,
                           const int* val_to_ks, const double* transform, const double* ks_ortho,
                           const double* overlap, int n_r, const int* r_idx, const double* kphase, double* dm_k,
                           double* dmr);

#ifdef __cplusplus
}
#endif
#endif /* DMFTGPU_H */
