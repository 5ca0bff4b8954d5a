/*
 * jvmc_hip.h -- C ABI of libjvmc_hip.so, the gfx950 (MI355X) implementation of
 * jVMC's VMC inner loop (one TDVP / MinSR right-hand-side evaluation).
 *
 * Conventions
 *   - Every entry point returns an int status: JV_OK (0) or a negative JV_E_*
 *     code; jv_last_error() returns the message of the last failure on the
 *     calling thread.  No C++ exception crosses this boundary.
 *   - Array arguments are DEVICE pointers owned by the caller, except where a
 *     parameter is documented as host memory.  Complex numbers are interleaved
 *     (re, im) doubles.  Matrices are row-major.
 *   - All calls are asynchronous on the given HIP stream (NULL = default
 *     stream), except jv_eigh/jv_minsr_solve/jv_tdvp_solve, which read a small
 *     status word back to the host.
 *   - The library keeps no per-job state.  Solver temporaries live in a
 *     workspace per host thread and device, grown on demand and released by
 *     jv_shutdown() or at thread exit; calls from different host threads may
 *     run concurrently on different streams.  Calls from one thread that use
 *     workspaces must be ordered on one stream.
 *
 * Each entry point names the jVMC 1.5.8 interface it replaces (file:line).
 */
#ifndef JVMC_HIP_H
#define JVMC_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define JV_OK 0
#define JV_E_ARG (-1)      /* invalid argument / unsupported configuration */
#define JV_E_HIP (-2)      /* HIP runtime error */
#define JV_E_SOLVER (-3)   /* rocSOLVER failure (info != 0) */
#define JV_E_NOMEM (-4)    /* allocation failure or LDS budget exceeded */

#define JV_NET_CPXRBM 1
#define JV_NET_RBM 2
#define JV_NET_CNN 3

#define JV_MAX_LAYERS 8

/* Arithmetic of the network kernels (jv_net.precision).  FP64: the reference's
 * precision throughout.  MIXED: the CNN window kernels (Metropolis and local
 * energy) evaluate activations in FP32 with FP64 accumulation of log psi
 * differences and of O_loc, and jv_net_gradients (layout 0) writes O in FP32
 * (jv_net_gradients_f32); kernels without an FP32 form run in FP64. */
#define JV_PREC_FP64 0
#define JV_PREC_MIXED 1

/* Network descriptor.  The caller fills the "input" fields; jv_net_init()
 * validates them and fills the "derived" fields (parameter layout).  The flat
 * real parameter vector follows jVMC's layout (vqs.py:430-466): leaves in JAX
 * tree_flatten order (dict keys sorted as strings), each complex leaf stored as
 * [Re(ravel), Im(ravel)]. */
typedef struct jv_net {
  /* input */
  int32_t kind;                       /* JV_NET_* */
  int32_t ndim;                       /* lattice rank, 1 or 2 */
  int32_t shape[2];                   /* lattice extents; shape[1] = 1 when ndim == 1 */
  int32_t num_hidden;                 /* RBMs: hidden units */
  int32_t hidden_bias;                /* RBMs: 0/1 (rbm.py:32 use_bias) */
  int32_t visible_bias;               /* CpxRBM only: 0/1 (extension, not in jVMC) */
  int32_t num_layers;                 /* CNN: number of conv layers (cnn.py:159) */
  int32_t filter[2];                  /* CNN: filter extents F (cnn.py:122) */
  int32_t channels[JV_MAX_LAYERS];    /* CNN: output channels per layer */
  int32_t layer_bias[JV_MAX_LAYERS];  /* CNN: bias per layer (cnn.py:144-145) */
  int32_t precision;                  /* JV_PREC_FP64 (0) or JV_PREC_MIXED (extension, not in jVMC) */
  /* derived (jv_net_init) */
  int32_t num_sites;                  /* shape[0] * shape[1] */
  int32_t is_complex;                 /* complex (holomorphic) parameters */
  int64_t num_params;                 /* parameters as the network sees them (complex count for CpxRBM) */
  int64_t num_real_params;            /* length of the flat real vector */
  int64_t off_hidden_bias;            /* RBMs: offset of leaf in the flat vector, -1 if absent */
  int64_t off_kernel;                 /* RBMs */
  int64_t off_visible_bias;           /* CpxRBM */
  int64_t off_conv_kernel[JV_MAX_LAYERS]; /* CNN */
  int64_t off_conv_bias[JV_MAX_LAYERS];   /* CNN, -1 if absent */
} jv_net;

/* Compiled operator table (branch_free.py:348-414).  All pointers are device
 * pointers.  Strings are padded to max_len factors with the identity on site 0;
 * factor k of string o is applied k-th (right-to-left order of the string). */
typedef struct jv_optable {
  int32_t num_ops;
  int32_t max_len;
  int32_t l_dim;             /* local dimension, must be 2 */
  int32_t num_diag;
  int32_t max_sites_changed; /* upper bound on sites one string changes (factors with a
                                non-identity map); 0 = unknown.  1 enables the single-flip
                                window kernels. */
  int32_t reserved;
  const int32_t* idx;        /* [num_ops][max_len] site of each factor */
  const int32_t* map;        /* [num_ops][max_len][l_dim] new local state */
  const double* matel;       /* [num_ops][max_len][l_dim] complex factor */
  const int32_t* diag;       /* [num_diag] indices of all-diagonal strings */
  const int32_t* fermi;      /* [num_ops][max_len] 1 for a fermionic factor (creation /
                                annihilation, branch_free.py:149-191), else 0; NULL when the
                                operator has none.  A fermionic factor on site i multiplies m
                                by (-1)^(number of occupied sites j > i) of the configuration
                                it acts on (branch_free.py:453,463-466). */
} jv_optable;

/* Symmetrisation of a network over an orbit of lattice symmetries (SymNet,
 * sym_wrapper.py:29-83).  The orbit matrices of symmetries.py:75-206 are signed
 * permutations; element g is stored as a site map: (g s)_i = s_j for
 * map[g][i] = j >= 0 and 1 - s_j for map[g][i] = -1 - j, which is
 * (o (2s - 1) + 1) // 2 of sym_wrapper.py:70-73 for a matrix row with a single
 * entry +1 (-1) in column j.  All pointers are device pointers. */
#define JV_AVG_EXP 0  /* avgFun_Coefficients_Exp: logsumexp(c, b = factor) (sym_wrapper.py:7-9) */
#define JV_AVG_LOG 1  /* avgFun_Coefficients_Log: mean(c) (sym_wrapper.py:12-13) */
#define JV_AVG_SEP 2  /* avgFun_Coefficients_Sep (sym_wrapper.py:16-19) */
typedef struct jv_orbit {
  int32_t num_elements;    /* |orbit| */
  int32_t num_sites;       /* must equal the network's num_sites */
  int32_t avg;             /* JV_AVG_* */
  int32_t complex_factor;  /* factors have a complex dtype: log psi promotes to complex
                              (jax logsumexp type promotion) */
  const int32_t* map;      /* [num_elements][num_sites] */
  const double* factor;    /* complex [num_elements] */
} jv_orbit;

/* Metropolis arguments (sampler.py:103-104, 195-358). */
typedef struct jv_mc_args {
  int32_t num_chains;
  int32_t therm_steps;        /* thermalizationSweeps * sweepSteps */
  int32_t samples_per_chain;  /* from distribute_sampling (mpi_wrapper.py:53-95) */
  int32_t sweep_steps;
  double mu;
  uint64_t seed;              /* Philox4x32-10 key */
  uint32_t call_index;        /* Philox counter word 2; advance once per call */
  uint32_t stream;            /* Philox counter word 3 */
} jv_mc_args;

/* Solver outputs copied to host memory by the solve calls. */
typedef struct jv_solve_info {
  double residual;           /* tdvp_residual (tdvp.py:209) */
  double pinv_cutoff;        /* max(cutoff, pinvCutoff) (tdvp.py:213) */
  int32_t iterations;        /* TDVP: regulariser iterations taken (tdvp.py:198);
                                MinSR: 1 when no eigenvalue lies at or below the
                                cutoff and the update came from a Cholesky
                                solve, 0 from the eigendecomposition */
  int32_t rank;              /* MinSR: eigenvalues kept by the pseudo-inverse */
  double e_var;              /* TDVP and MinSR deferred dual: sum_n |e_n|^2 of the centred,
                                weighted E_loc = Var E_loc (stats.py:248-252) */
  double tdvp_error;         /* TDVP: tdvp_error metadata (tdvp.py:102-104) */
  double e_mean[2];          /* MinSR deferred dual: weighted mean of E_loc (stats.py:197-205),
                                complex (re, im) */
} jv_solve_info;

/* ------------------------------------------------------------------ misc */
int jv_version(void);
const char* jv_last_error(void);
/* Selects the device for the calling thread.  Replaces the rank-to-device pick
 * of global_defs.py:15-19. */
int jv_set_device(int device);
/* Loads the solver libraries' kernels on `device` once per process and makes it
 * current (no reference counterpart: jax initialises its backends on import).
 * Every first solver call does the same, serialised; a server calls it before
 * it starts job threads so no job pays the load. */
int jv_init(int device);
/* Releases the calling thread's workspaces and solver handles.  Worker threads
 * release theirs when they end; the main thread's are kept until this call or
 * process exit (no HIP call is made from exit handlers). */
int jv_shutdown(void);
/* Bytes of device memory held by the calling thread's workspaces and solver
 * library handle on the current device (no reference counterpart: jax sizes
 * its buffers inside XLA).  A server compares it with its per-job HBM
 * estimate. */
int jv_workspace_bytes(int64_t* bytes);
/* Kernel timing (no reference counterpart; measurement support): when enabled,
 * HIP events are recorded on the launch stream around the named kernels
 * ("gram", "gemm", "eigh", "eloc", "mc_sweep", "net_grad", "net_forward").
 * jv_timer_read synchronises on the recorded events and returns the summed
 * duration and the launch count. */
int jv_timer_enable(int on);
int jv_timer_reset(void);
int jv_timer_read(const char* name, double* total_ms, int64_t* count);

/* ------------------------------------------------------------------ networks */
/* Validates a descriptor and computes the flat parameter layout (host only).
 * Replaces NQS.init_net's shape bookkeeping (vqs.py:187-218). */
int jv_net_init(jv_net* net);

/* log psi(s) for n configurations (int32 [n][num_sites]); out: complex [n].
 * Replaces NQS.__call__ / NQS._eval (vqs.py:223-251) with CpxRBM.__call__
 * (rbm.py:30-38), RBM.__call__ (rbm.py:80-88) or CNN.__call__ (cnn.py:131-171). */
int jv_net_forward(const jv_net* net, const double* params, const int32_t* configs,
                   int64_t n, double* logpsi, void* stream);

/* Per-sample O_k = d log psi / d theta_k.  layout 0 ("compact"): complex nets
 * write complex [n][num_params] (one derivative per complex parameter), real
 * nets write real [n][num_params].  layout 1 ("jvmc"): complex [n][num_real_params]
 * in jVMC's flat layout ([g, i*g] per leaf for holomorphic nets, vqs.py:66-69;
 * g + 0i for real nets, vqs.py:46-51).  logpsi may be NULL.
 * Replaces NQS.gradients / NQS._get_gradients (vqs.py:256-287). */
int jv_net_gradients(const jv_net* net, const double* params, const int32_t* configs,
                     int64_t n, int32_t layout, double* grad, double* logpsi, void* stream);

/* jv_net_gradients in the mixed-precision form: O (layout 0, real [n][num_params])
 * in FP32 from FP32 activations; logpsi (FP64 sum, may be NULL).  CNN only.
 * Replaces NQS.gradients (vqs.py:256-287) in the mixed form. */
int jv_net_gradients_f32(const jv_net* net, const double* params, const int32_t* configs, int64_t n, float* grad,
                         double* logpsi, void* stream);

/* ------------------------------------------------------------------ symmetrised networks */
/* log psi(s) = avg_g log psi_net(g s) (complex [n]).  EXP: log(sum_g b_g
 * exp(c_g)) with the jax logsumexp conventions: entries with b_g = 0 are
 * dropped; for a real network with real factors the result is log|sum| and NaN
 * where the sum is negative.  LOG: the mean.  SEP: 0.5 log mean exp(2 Re c) +
 * i angle(mean exp(i Im c)).
 * Replaces NQS.__call__ with SymNet.__call__ (sym_wrapper.py:66-77, vqs.py:223-251). */
int jv_sym_forward(const jv_net* net, const double* params, const jv_orbit* orbit,
                   const int32_t* configs, int64_t n, double* logpsi, void* stream);

/* Per-sample gradients of the symmetrised log psi: O(s) = sum_g w_g O_net(g s),
 * w_g the derivative of the average with respect to c_g (EXP: b_g exp(c_g) /
 * sum; LOG: 1/|orbit|; SEP: the softmax of 2 Re c, real-parameter networks
 * only; real EXP samples whose sum is negative get w = 0, as jax's where()).
 * Layouts as jv_net_gradients, except that layout 0 of a real network is
 * complex [n][num_params] when the weights are complex (EXP with complex
 * factors).  Replaces NQS.gradients for a SymNet (vqs.py:40-82, 256-287). */
int jv_sym_gradients(const jv_net* net, const double* params, const jv_orbit* orbit,
                     const int32_t* configs, int64_t n, int32_t layout, double* grad,
                     double* logpsi, void* stream);

/* O_loc(s) = sum_j m_j exp(log psi(s'_j) - logpsi_s) for a symmetrised network:
 * s' materialised per sample chunk (jv_get_s_primes), log psi(s') from
 * jv_sym_forward.  Replaces Operator.get_O_loc (base.py:162-192) for a SymNet. */
int jv_sym_eloc(const jv_net* net, const double* params, const jv_orbit* orbit,
                const jv_optable* op, const double* prefactor, const int32_t* configs,
                const double* logpsi_s, int64_t n, double* out, void* stream);

/* Metropolis sampling of a symmetrised network with single-spin-flip proposals:
 * the same arguments, Philox draws and emission order as jv_mc_sample; each
 * proposal is evaluated with jv_sym_forward on all chains at once and accepted
 * with u < exp(mu Re(log psi' - log psi)).
 * Replaces MCSampler._sweep / _get_samples_mcmc (sampler.py:264-367) for a SymNet. */
int jv_sym_mc_sample(const jv_net* net, const double* params, const jv_orbit* orbit,
                     const jv_mc_args* args, int32_t* walkers, int32_t* configs, double* logpsi,
                     int64_t* counters, void* stream);

/* ------------------------------------------------------------------ operators */
/* s' enumeration with nonzero compaction.  Outputs, per sample, all num_ops
 * slots: the first count[n] slots hold the strings with |m| > 1e-6 in string
 * order, the rest repeat the last string's s' with m = 0; sp: int32
 * [n][num_ops][num_sites], matel: complex [n][num_ops], count: int32 [n].
 * prefactor: complex [num_ops] (device), the host-evaluated f(*args).
 * Replaces BranchFreeOperator._get_s_primes (branch_free.py:443-487) and
 * Operator.get_s_primes / _find_nonzero / set_zero_to_zero (base.py:91-160). */
int jv_get_s_primes(const jv_optable* op, const double* prefactor, const int32_t* configs,
                    int64_t n, int32_t num_sites, int32_t* sp, double* matel, int32_t* count,
                    void* stream);

/* Fused local estimator: O_loc(s) = sum_j m_j exp(log psi(s'_j) - logpsi_s) with
 * the s' enumerated in-kernel, never materialised, and log psi(s') obtained by
 * incremental (RBM) or receptive-field-limited (CNN) re-evaluation.
 * logpsi_s: complex [n]; out: complex [n].
 * Replaces Operator.get_O_loc / _get_O_loc / get_O_loc_batched (base.py:162-278). */
int jv_eloc(const jv_net* net, const double* params, const jv_optable* op,
            const double* prefactor, const int32_t* configs, const double* logpsi_s,
            int64_t n, double* out, void* stream);

/* ------------------------------------------------------------------ sampling */
/* Metropolis sampling with single-spin-flip proposals.  walkers: int32
 * [num_chains][num_sites], updated in place.  configs: int32
 * [samples_per_chain * num_chains][num_sites] with sample index t*C + c;
 * logpsi: complex [samples_per_chain * num_chains]; counters: int64 [2]
 * (proposed, accepted), overwritten.
 * Replaces MCSampler._get_samples_mcmc / _get_samples / _sweep / _mc_init and
 * propose_spin_flip (sampler.py:15-19, 264-367). */
int jv_mc_sample(const jv_net* net, const double* params, const jv_mc_args* args,
                 int32_t* walkers, int32_t* configs, double* logpsi, int64_t* counters,
                 void* stream);

/* Exact-sampler weights: p_k = exp(Re(logpsi_k - last_norm)/log_prob_factor),
 * normalised; *norm_out (host) receives the unnormalised sum.
 * Replaces ExactSampler._compute_probabilities / sample (sampler.py:474-526). */
int jv_exact_weights(const double* logpsi, int64_t n, double last_norm, double log_prob_factor,
                     double* p, double* norm_out, void* stream);

/* Metropolis weights p = exp((1/logProbFactor - mu) Re logpsi), normalised
 * (sampler.py:234-235). */
int jv_mc_weights(const double* logpsi, int64_t n, double mu, double log_prob_factor,
                  double* p, void* stream);

/* ------------------------------------------------------------------ statistics */
/* SampledObs centring: mean = sum_n w_n O_n, data_n = sqrt(w_n) (O_n - mean).
 * obs: [n][p] real (is_complex 0) or complex (1); data may alias obs.
 * Replaces SampledObs._compute_data_and_mean (stats.py:197-205). */
int jv_center(const double* obs, const double* w, int64_t n, int64_t p, int32_t is_complex,
              double* mean, double* data, void* stream);

/* jv_center for FP32 observations (the mixed form): the weighted mean is
 * accumulated and returned in FP64 (mean: [p] real or complex doubles), the
 * centred data are rounded to FP32.  Replaces SampledObs._compute_data_and_mean
 * (stats.py:197-205) in the mixed form. */
int jv_center_f32(const float* obs, const double* w, int64_t n, int64_t p, int32_t is_complex, double* mean,
                  float* data, void* stream);

/* jv_center without any allocation, for callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The partial column sums live
 * in a fixed per-device array of 2^21 doubles (16 MiB, held until the process
 * ends); ceil(n / 256) * p * (is_complex ? 2 : 1) must not exceed it, else
 * JV_E_ARG.  Every call on a device shares that array: calls that overlap on
 * one device (different streams or threads) race and must be serialised by
 * the caller.  Same results as jv_center. */
int jv_center_noalloc(const double* obs, const double* w, int64_t n, int64_t p, int32_t is_complex, double* mean,
                      double* data, void* stream);

/* Column sums of squares: out[j] = sum_n |data[n][j]|^2.
 * Replaces SampledObs.var (stats.py:248-252). */
int jv_col_norm2(const double* data, int64_t n, int64_t p, int32_t is_complex, double* out, void* stream);

/* General FP64 GEMM on MFMA: C[M][N] = alpha * opA opB with opA(m, k) =
 * A[m*sam + k*sak], opB(k, n) = B[k*sbk + n*sbn] (element strides; complex
 * elements are two doubles), optional conjugation of either operand.  Used for
 * the products of stats.py:50-83 (covar, covar_data + transform) and
 * tdvp.py:175-179. */
int jv_gemm(int64_t M, int64_t N, int64_t K, const double* A, int64_t sam, int64_t sak, int32_t conjA,
            const double* B, int64_t sbk, int64_t sbn, int32_t conjB, double* C, int64_t ldc,
            int32_t is_complex, double alpha, void* stream);

/* y = alpha * op(M) x for row-major M [rows][cols]: trans 0 -> M x, trans 1 ->
 * M^T x (conj 1: M^H x).  Replaces the jnp.dot / matmul GEMVs of tdvp.py:171,211
 * and minsr.py:65,78. */
int jv_gemv(int32_t trans, const double* M, const double* x, int64_t rows, int64_t cols, int32_t is_complex,
            int32_t conj, double alpha, double* y, void* stream);

/* Tangent kernel T = D D^dagger ([n][n]) of centred data D [n][p].
 * Replaces SampledObs.tangent_kernel (stats.py:332-336) and the T build of
 * MinSR.solve (minsr.py:59-71). */
int jv_gram_nn(const double* data, int64_t n, int64_t p, int32_t is_complex, double* T,
               void* stream);

/* Covariance S = D^dagger D ([p][p]) of centred data D [n][p].
 * Replaces SampledObs.covar (stats.py:235-245) for the S matrix. */
int jv_gram_pp(const double* data, int64_t n, int64_t p, int32_t is_complex, double* S,
               void* stream);

/* Mixed-precision covariance S = D^dagger D ([p][p], FP64 output) of centred
 * data D [n][p] stored in FP32 (data_f32 = 1) or FP64 (0), complex data
 * interleaved.  Products on the BF16 matrix cores with compensated splits
 * (x = bf16(x) + bf16(x - bf16(x)), three products per multiply, FP32
 * accumulation within slices of the sample axis, slices summed in FP64):
 * relative error of each entry ~1e-5 of sum_n |D_ni| |D_nj|.
 * Replaces SampledObs.covar (stats.py:235-245) / the Gram of minsr.py:67-78
 * in the mixed-precision form. */
int jv_gram_pp_mixed(const void* data, int32_t data_f32, int64_t n, int64_t p, int32_t is_complex, double* S,
                     void* stream);

/* Symmetric / Hermitian eigendecomposition (ascending eigenvalues) in FP64:
 * for n >= 128 own Householder tridiagonalisation kernels, the tridiagonal
 * divide and conquer of jv_stedc (rocSOLVER dstedc with
 * VMC_JAX_AMD_STEDC=rocsolver) and the back-transformation; below that rocSOLVER
 * syevd / heevd.  A [n][n] is overwritten by eigenvectors stored as COLUMNS of
 * the row-major matrix (A[i][k] = component i of eigenvector k).
 * Replaces jnp.linalg.eigh (tdvp.py:153-171). */
int jv_eigh(int64_t n, int32_t is_complex, double* A, double* w, void* stream);

/* Eigenvalues and eigenvectors of the real symmetric tridiagonal matrix with
 * diagonal d[n] and off-diagonal e[n-1], by divide and conquer on the device
 * (stream ordered, no host synchronisation in steady state).  On exit d holds
 * the eigenvalues ascending and z [n][n] the eigenvectors as ROWS (z[k][i] =
 * component i of eigenvector k); e is not modified.  n <= 3276.  No
 * counterpart in the reference, where jnp.linalg.eigh (tdvp.py:153-171,
 * minsr.py:74) runs the whole decomposition: it is the tridiagonal stage of
 * jv_eigh, exposed for callers that bring their own reduction. */
int jv_stedc(int64_t n, double* d, const double* e, double* z, void* stream);

/* MinSR update (minsr.py:53-80) from centred data.  Holomorphic (is_complex
 * with holo=1): T = D D^dagger, update = -D^dagger pinv(T) e with the jVMC flat
 * layout handled by the caller.  Real-parameter nets (is_complex 0): D real
 * [n][p], e complex [n]; T = D D^T + shift I (the imaginary stacked half of
 * minsr.py:68-78 is identically zero and drops out exactly), update real [p].
 * pinv keeps |lambda| > rcond * max|lambda|.  try_cholesky != 0: when a
 * Cholesky factorisation of T - c I (complex T: of its real 2n x 2n form
 * [[Re T, -Im T], [Im T, Re T]], same eigenvalues) with c = ||T||_F (rcond +
 * 8 n eps) succeeds, no eigenvalue is cut and x = T^-1 e comes from L^-1
 * (info->iterations = 1); otherwise the eigendecomposition runs.  Centred data
 * with uniform weights leave T the null vector u0 = 1/sqrt(n): when
 * ||T u0|| <= rcond max Re T_ii / 2 the test runs on T + (max Re T_ii) u0 u0^T
 * and, if it passes, x = T'^-1 (e - u0 u0^T e) (the pseudo-inverse with u0's
 * eigenvalue cut; info->rank = n - 1).  info (host) may be NULL. */
int jv_minsr_solve(const double* data, const double* e, int64_t n, int64_t p,
                   int32_t is_complex, double rcond, double diag_shift, int32_t try_cholesky,
                   double* update, jv_solve_info* info, void* stream);

/* Same update as jv_minsr_solve from the P x P Gram S = D^dagger D (+ shift I)
 * instead of the N x N tangent kernel: the nonzero spectra of D D^dagger and
 * D^dagger D coincide, D^dagger annihilates the null space of T, so
 * -D^dagger pinv(T) e = -pinv(S) D^dagger e with the same relative cutoff.
 * Cheaper whenever N > P.  No separate reference function (exact rewrite of
 * minsr.py:53-80).  Real data with try_cholesky != 0: when a Cholesky
 * factorisation of S - c I, c = ||S||_F (rcond + 8 p eps) >= rcond
 * lambda_max, succeeds, no eigenvalue lies at or below the cutoff, pinv(S) =
 * S^-1 and the update comes from a Cholesky solve (info->iterations = 1);
 * otherwise, or with try_cholesky = 0, from the eigendecomposition.  Both give
 * the same update; try_cholesky only saves the factorisation when the caller
 * expects a cut spectrum. */
int jv_minsr_solve_dual(const double* data, const double* e, int64_t n, int64_t p,
                        int32_t is_complex, double rcond, double diag_shift, int32_t try_cholesky,
                        double* update, jv_solve_info* info, void* stream);

/* jv_minsr_solve_dual with the local estimator still in flight: e_raw is the
 * uncentred E_loc (complex [n]) with the sample weights w [n], complete once the
 * event e_ready (a hipEvent_t; NULL = ready in stream order) has fired, so
 * the producer of e_raw overlaps the Gram and its factorisation or reduction.
 * On the Cholesky route the centring of e_raw (stats.py:197-205) and
 * b = D^dagger e follow e_ready on the calling stream; on the real tridiagonal
 * path they run with Q^T b on the library's auxiliary stream, overlapping the
 * tridiagonal eigensolver.  info->e_mean / e_var return the statistics of
 * E_loc.  Same update as jv_minsr_solve_dual on
 * the centred data.  No separate reference function (minsr.py:82-165 computes
 * E_loc before the solve; the order does not change the result). */
int jv_minsr_solve_dual_deferred(const double* data, const double* e_raw, const double* w, void* e_ready,
                                 int64_t n, int64_t p, int32_t is_complex, double rcond, double diag_shift,
                                 int32_t try_cholesky, double* update, jv_solve_info* info, void* stream);

/* jv_minsr_solve_dual_deferred in the mixed-precision form: centred data D in
 * FP32 (real only, is_complex must be 0), S = D^T D from the compensated BF16
 * Gram (jv_gram_pp_mixed) and b = D^T e accumulated in FP64; the factorisation
 * or eigendecomposition, the pseudo-inverse and the E_loc statistics stay in
 * FP64.  No separate reference function (minsr.py:53-80 in FP64). */
int jv_minsr_solve_dual_deferred_mixed(const float* data, const double* e_raw, const double* w, void* e_ready,
                                       int64_t n, int64_t p, int32_t is_complex, double rcond, double diag_shift,
                                       int32_t try_cholesky, double* update, jv_solve_info* info, void* stream);

/* The dual solve in two calls, so that the caller can start the local
 * estimator after the Gram: jv_minsr_dual_gram forms S = D^dagger D (+ shift I
 * for real data) into the caller's buffer S (p x p, complex interleaved
 * p x 2p when is_complex) on the stream: FP64 data (f32 = 0) through the FP64
 * MFMA Gram, FP32 data (f32 = 1, real only) through the compensated BF16 Gram
 * of the mixed form.  jv_minsr_solve_dual_deferred_gram is then
 * jv_minsr_solve_dual_deferred (f32 = 0) or its mixed form (f32 = 1) on that
 * S, which it consumes.  Same update as the one-call forms.  No separate
 * reference function (minsr.py:82-165 computes E_loc before the solve; the
 * order does not change the result). */
int jv_minsr_dual_gram(const void* data, int32_t f32, int64_t n, int64_t p, int32_t is_complex, double diag_shift,
                       double* S, void* stream);
int jv_minsr_solve_dual_deferred_gram(const void* data, int32_t f32, double* S, const double* e_raw,
                                      const double* w, void* e_ready, int64_t n, int64_t p, int32_t is_complex,
                                      double rcond, double diag_shift, int32_t try_cholesky, double* update,
                                      jv_solve_info* info, void* stream);

/* jv_minsr_solve_dual_deferred_gram without the host synchronisation at its
 * end: info is filled by a copy enqueued on the stream after the solve, so it
 * is valid once the stream has passed the call (an event recorded after it
 * has completed); info must stay allocated until then, and must be pinned host
 * memory (hipHostMalloc) for the call not to block.  A Cholesky factorisation
 * that failed on the accepted route, an error return of the blocking form, is
 * reported as info->iterations = -1 - its info.  The host still waits inside
 * the call for the Cholesky attempt's outcome (it selects the kernels that
 * follow), not for E_loc or the rest of the solve.  Same update.  No
 * separate reference function (minsr.py:82-165 returns the update and keeps
 * the energy for later reads). */
int jv_minsr_solve_dual_deferred_gram_async(const void* data, int32_t f32, double* S, const double* e_raw,
                                            const double* w, void* e_ready, int64_t n, int64_t p,
                                            int32_t is_complex, double rcond, double diag_shift,
                                            int32_t try_cholesky, double* update, jv_solve_info* info,
                                            void* stream);

/* TDVP solve (tdvp.py:134-213).  data: centred O in jVMC's flat layout, complex
 * [n][pr]; e: centred E_loc, complex [n]; w: weights [n]; x = rhsPrefactor
 * (complex, host); make_real 0 = 'real', 1 = 'imag'.  use_snr selects the SNR
 * term of the regulariser (MCSampler) and glob_num_samples is mpi.globNumSamples.
 * Outputs (device): update real [pr], S (makeReal'd, complex [pr][pr]), F (complex
 * [pr]), ev [pr], snr [pr]; E mean / variance complex (device, 2 values each).
 * Replaces TDVP.get_tdvp_equation / _transform_to_eigenbasis / _get_snr / solve. */
int jv_tdvp_solve(const double* data, const double* e, const double* w, int64_t n, int64_t pr,
                  const double* x, int32_t make_real, double diag_shift, int32_t use_snr,
                  double glob_num_samples, double snr_tol, double pinv_tol, double pinv_cutoff,
                  double* update, double* S, double* F, double* ev, double* snr,
                  jv_solve_info* info, void* stream);

/* jv_tdvp_solve for a holomorphic network, whose flat gradient rows are
 * [g, i g] per complex leaf (vqs.py:66-69): column col_ig[c] of D equals i times
 * column col_g[c] (device int64 [pc] each, pr = 2 pc).  Then S0 = O^H O is
 * [[A, iA], [-iA, A]] in split order with A = G^H G (G = D[:, col_g]), so S0 is
 * assembled from the Pc x Pc Gram and S is diagonalised through A: 'real'
 * S = Re S0 has A's eigenvalues twice with eigenvectors [a; b], [-b; a] for
 * u = a + ib; 'imag' S = i Im S0 has eigenvalues +-lambda(A) with eigenvectors
 * [u; -iu] / sqrt2 and [conj u; i conj u] / sqrt2.  Same outputs as jv_tdvp_solve
 * (within a degenerate eigenspace the basis may differ; the update does not).
 * No separate reference function (exact rewrite of tdvp.py:134-171). */
int jv_tdvp_solve_holomorphic(const double* data, const double* e, const double* w, int64_t n, int64_t pr,
                              const int64_t* col_g, const int64_t* col_ig, int64_t pc, const double* x,
                              int32_t make_real, double diag_shift, int32_t use_snr, double glob_num_samples,
                              double snr_tol, double pinv_tol, double pinv_cutoff, double* update, double* S,
                              double* F, double* ev, double* snr, jv_solve_info* info, void* stream);

/* jv_tdvp_solve_holomorphic in the mixed-precision form: A = G^H G from the
 * compensated BF16 Gram (jv_gram_pp_mixed on the FP64 data); S, F, the
 * eigendecomposition, the SNR and the regulariser stay FP64.  No separate
 * reference function (tdvp.py:134-213 in FP64). */
int jv_tdvp_solve_holomorphic_mixed(const double* data, const double* e, const double* w, int64_t n, int64_t pr,
                                    const int64_t* col_g, const int64_t* col_ig, int64_t pc, const double* x,
                                    int32_t make_real, double diag_shift, int32_t use_snr, double glob_num_samples,
                                    double snr_tol, double pinv_tol, double pinv_cutoff, double* update, double* S,
                                    double* F, double* ev, double* snr, jv_solve_info* info, void* stream);

/* SampledObs.subset (stats.py:310-329): rows r_k = start + k step, k < count, of
 * centred data [n][p] (real or complex) with weights w [n] and mean [p]:
 * out_w = w[r] / nrm (nrm = sum_k w[r_k]), out_mean = sum_k sqrt(out_w_k)
 * data[r_k] / sqrt(nrm) + mean, out_data[k] = data[r_k] / sqrt(nrm) +
 * sqrt(out_w_k) (mean - out_mean).  Device pointers. */
int jv_subset(const double* data, const double* w, const double* mean, int64_t n, int64_t p, int32_t is_complex,
              int64_t start, int64_t step, int64_t count, double* out_data, double* out_w, double* out_mean,
              void* stream);

/* Cross-validation check of a TDVP update (tdvp.py:306-324) on a second sample
 * set with centred data D [n][pr] (complex, jVMC layout) and centred E_loc e [n]:
 * with S0 = D^dagger D, F0 = -x D^dagger e, S = makeReal(S0) (+ diag shift as
 * tdvp.py:145-146) and F = makeReal(F0), out (host, 2 doubles) receives
 * out[0] = ||S u - F|| / ||F|| and out[1] = Re(u^T S0 u) - 2 Re(u^T F0), the
 * numerator of the tdvp_error (tdvp.py:102-104).  S0 is applied as D^dagger (D u):
 * no pr x pr matrix is formed.  update u: real [pr] (device). */
int jv_tdvp_cross_validate(const double* D, const double* e, int64_t n, int64_t pr, const double* x,
                           int32_t make_real, double diag_shift, const double* update, double* out, void* stream);

/* ------------------------------------------------------------------ batched field sweep */
/* Many independent RBM states (uneven lattice sizes, hidden-unit counts,
 * operators and sample counts) advanced by one launch per phase: a
 * device-resident instance table maps each workgroup to one instance, which
 * runs the same per-chain / per-sample routines as jv_mc_sample, jv_eloc and
 * jv_net_gradients.  No counterpart in the reference, where each state is a
 * separate jVMC driver (sampler.py:301-358, operator/base.py:166-192,
 * vqs.py:256-287 per instance). */
#define JV_BATCH_SAMPLE 0  /* Metropolis pass (then the Philox call index advances) */
#define JV_BATCH_GRAD 1    /* log psi and O_k of the emitted samples */
#define JV_BATCH_ELOC 2    /* O_loc of the emitted samples */

typedef struct jv_batch_item {
  jv_net net;                /* an RBM or CpxRBM after jv_net_init */
  const double* params;      /* device flat parameters, read at every launch: update in place */
  jv_optable op;             /* device operator tables */
  const double* prefactor;   /* device complex [num_ops] */
  jv_mc_args mc;             /* call_index: the first pass's; the batch advances it on the device */
  int32_t* walkers;          /* [num_chains][L], persistent */
  int32_t* configs;          /* [samples_per_chain * num_chains][L], sample index t*C + c */
  double* logpsi;            /* complex [n] */
  int64_t* counters;         /* [2] proposed, accepted of the last pass */
  double* eloc;              /* complex [n] */
  double* grad;              /* gradient rows in grad_layout (jv_net_gradients), may be NULL */
  int32_t grad_layout;       /* 0 compact, 1 jVMC */
  int32_t reserved;
} jv_batch_item;

typedef struct jv_batch jv_batch;
typedef struct jv_graph jv_graph;

/* Validates the items and uploads the instance table (synchronous). */
int jv_batch_create(const jv_batch_item* items, int32_t count, jv_batch** out);
int jv_batch_destroy(jv_batch* batch);
/* Launch shape of a phase: waves per workgroup (1, 2, 4, 8) and, for the
 * gradient and O_loc phases, samples per wave (1..64); re-uploads the table.
 * Chosen by the package's autotuner (vmc_jax_amd/batch.py). */
int jv_batch_set_shape(jv_batch* batch, int32_t phase, int32_t waves, int32_t samples_per_wave);
/* One phase for every instance. */
int jv_batch_phase(jv_batch* batch, int32_t phase, void* stream);
/* Counters cleared, then sampling, gradients (+ log psi) and O_loc for every
 * instance: 5 launches, no allocation, host synchronisation or memset. */
int jv_batch_step(jv_batch* batch, void* stream);
/* jv_batch_step captured into a HIP graph on `stream` (not the default
 * stream) and instantiated; jv_graph_launch replays it (parameters, walkers
 * and the Philox call index are read from device memory at replay). */
int jv_graph_capture(jv_batch* batch, void* stream, jv_graph** out);
int jv_graph_launch(jv_graph* graph, void* stream);
int jv_graph_destroy(jv_graph* graph);

#ifdef __cplusplus
}
#endif

#endif /* JVMC_HIP_H */
