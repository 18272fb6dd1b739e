/*
 * edhmc -- MI355X-native Hamiltonian Monte Carlo engine, C ABI.
 *
 * This is the drop-in boundary under Edward's `ed.HMC` surface.  In Edward
 * 1.3.5 the hot path is one TensorFlow session call per sample:
 * `MonteCarlo.update` (edward/inferences/monte_carlo.py:111-150) runs the op
 * that `HMC.build_update` (edward/inferences/hmc.py:61-130) built, whose
 * arithmetic is `leapfrog` (hmc.py:195-210) over `HMC._log_joint`
 * (hmc.py:161-192).  Each entry point below replaces one of those pieces;
 * its comment cites the reference lines whose behaviour it reproduces.
 *
 * Conventions
 *   - All array arguments are DEVICE pointers on the context's GPU, owned by
 *     the caller; the library never frees them.  The library owns only the
 *     scratch behind an opaque context (data layouts, residuals, partial
 *     sums, chain state).
 *   - `stream` is a hipStream_t passed as void* (NULL = default stream).
 *     Calls enqueue work and return; the caller synchronises.
 *   - Every function returns 0 on success or a negative EDHMC_E* code;
 *     edhmc_last_error() then holds a message.  No C++ exception crosses
 *     the ABI.  A context is not thread-safe (one per thread or a lock).
 *   - Matrices are row-major.  C chains of a D-dimensional latent are [C, D].
 */
#ifndef EDHMC_H_
#define EDHMC_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct edhmc_ctx edhmc_ctx;

#define EDHMC_OK 0
#define EDHMC_EINVAL (-1)      /* bad argument (TypeError/ValueError in Edward) */
#define EDHMC_ENOMEM (-2)      /* device allocation failed */
#define EDHMC_EHIP (-3)        /* HIP runtime error */
#define EDHMC_EUNSUPPORTED (-4)
#define EDHMC_ESTATE (-5)      /* call order: e.g. transition before bind */
#define EDHMC_ENONFINITE (-6)  /* non-finite data (InvalidArgumentError in ed.dot) */

#define EDHMC_F32 0
#define EDHMC_BF16 1

#define EDHMC_LIK_BERNOULLI 0  /* Bernoulli(logits = X w)          */
#define EDHMC_LIK_NORMAL 1     /* Normal(loc = X w, scale = s)      */

/* Library version string. */
const char* edhmc_version(void);

/* Create a context on HIP device `device`.  Replaces the global session of
 * edward/util/graphs.py:13-41 (get_session). */
int edhmc_create(int device, edhmc_ctx** out);
int edhmc_destroy(edhmc_ctx* ctx);
const char* edhmc_last_error(const edhmc_ctx* ctx);

/* Bind a generalised linear model  y ~ Lik(dot(X, w)),  w ~ Normal(loc, scale).
 * Replaces the data binding of edward/inferences/inference.py:81-95 (data
 * copied once into a runtime-owned variable) and the model graph whose
 * log-joint hmc.py:161-192 sums: prior Normal.log_prob + likelihood log_prob
 * (Bernoulli(logits=ed.dot(X, w)), examples/bayesian_logistic_regression.py:41-44,
 * or Normal(loc=ed.dot(X, w), scale=s), tests/inferences/hmc_test.py:63-66).
 *   X         [N, D] in x_in_dtype (EDHMC_F32 or EDHMC_BF16)
 *   y         [N] float32 (0/1 for Bernoulli)
 *   prior_loc, prior_scale  [D] float32
 *   compute_dtype  EDHMC_F32: fp32 X, fp32 MFMA;  EDHMC_BF16: X rounded to
 *                  bf16 (RNE), bf16 MFMA with fp32 accumulation.
 * Non-finite X raises EDHMC_ENONFINITE like ed.dot's verify_tensor_all_finite
 * (edward/util/tensorflow.py:33-36).  Rebinding discards chain state.
 * Device memory for the data: X is kept in two layouts (row-major and
 * K-blocked), and with bf16, a Bernoulli likelihood, D_pad (D rounded up to
 * 64) a multiple of 128 up to 512 and N >= 8192 in a third,
 * MFMA-operand-ordered one -- 3 x N x D_pad x 2 bytes (3 GB at N = 1e6,
 * D = 512). */
int edhmc_bind_glm(edhmc_ctx* ctx, const void* X, int32_t x_in_dtype, const float* y,
                   int64_t N, int32_t D, int32_t likelihood, float lik_scale,
                   const float* prior_loc, const float* prior_scale,
                   int32_t compute_dtype, void* stream);

/* Set the C chains' positions W [C, D] and evaluate logp and its gradient
 * there (the state every following transition starts from).  This is the
 * gather `params[max(t-1, 0)]` of hmc.py:81-85 when a run starts.
 * Non-finite W -> EDHMC_ENONFINITE (ed.dot check). */
int edhmc_set_state(edhmc_ctx* ctx, const float* W, int32_t C, void* stream);

/* Copy the current positions [C, D], log joints [C] (fp64) and gradients
 * [C, D]; any output may be NULL. */
int edhmc_get_state(edhmc_ctx* ctx, float* W_out, double* logp_out, float* grad_out,
                    void* stream);

/* Log joint and gradient at W [C, D]: logp_out [C] fp64, grad_out [C, D].
 * The value of HMC._log_joint_unconstrained (hmc.py:132-159, identity
 * transform for real support) and of tf.gradients of it (hmc.py:199,206).
 * Also makes W the current chain state.  Parity hook. */
int edhmc_logp_grad(edhmc_ctx* ctx, const float* W, int32_t C, float* grad_out,
                    double* logp_out, void* stream);

/* Log joint only at W [C, D] (no gradient): logp_out [C] fp64.  The forward
 * pass the Metropolis test of hmc.py:104-105 evaluates; in bf16 precision it
 * uses the hi/lo split positions (two bf16 products) with the FP64
 * reduction.  Leaves the chain positions untouched, but a C different from
 * the current chain count discards the chain state. */
int edhmc_logp(edhmc_ctx* ctx, const float* W, int32_t C, double* logp_out, void* stream);

/* Posterior-predictive pass over the bound rows for S posterior draws
 * W [S, D] (the sess.run loops over `output_key.probs` / `output_key` /
 * `log_prob(y_true)` of edward/criticisms/evaluate.py:128-150,199-204 and the
 * replicated data sets of edward/criticisms/ppc.py:100-121), any output may
 * be NULL:
 *   mean_out   [N] fp32: mean over the S draws of E[y_n | w_s] (sigmoid(eta)
 *              for Bernoulli, eta for Normal)
 *   loglik_out [S] fp64: sum over rows of log p(y_n | w_s), y = the bound y
 *   rep_out    [S, N] fp32: y_rep ~ p(y | w_s), Philox4x32-10 keyed by seed,
 *              counter (s, n) (Bernoulli: u < sigmoid(eta) with 24-bit u;
 *              Normal: eta + s * Box-Muller normal)
 * In bf16 precision the forward uses the hi/lo split of W.  Uses the chain
 * buffers: a draw count different from the current chain count discards
 * the chain state. */
int edhmc_predict(edhmc_ctx* ctx, const float* W, int32_t S, float* mean_out,
                  double* loglik_out, float* rep_out, uint64_t seed, void* stream);

/* One HMC transition of every chain: momentum draw (hmc.py:88-91),
 * leapfrog with n_steps steps of size eps (hmc.py:94-97,195-210), ratio and
 * accept (hmc.py:100-111), selection, and the sample write + n_accept
 * increment (hmc.py:121-129).
 *   r0       [C, D] fp32 momenta, or NULL to draw them on the device
 *            (Philox4x32-10, key = seed, counter = `counter`)
 *   u        [C] fp64 accept uniforms, or NULL to draw them on the device
 *   accepted_out [C] int32 0/1 (may be NULL)
 *   sample_out   [C, D] fp32: the Empirical row params[t] (may be NULL)
 *   n_accept     [C] int32, incremented on accept (may be NULL)
 * The gradient at the current state is carried from the previous
 * transition (the reference re-evaluates it, hmc.py:199; the value is the
 * same), so one transition costs n_steps gradient evaluations. */
int edhmc_transition(edhmc_ctx* ctx, float eps, int32_t n_steps, const float* r0,
                     const double* u, uint64_t seed, uint64_t counter,
                     int32_t* accepted_out, float* sample_out, int32_t* n_accept,
                     void* stream);

/* Stochastic-gradient samplers over the bound GLM (the chain state set by
 * edhmc_set_state; the gradient at the current positions is evaluated first
 * when the last call moved them).  Each call writes the new positions to
 * sample_out [C, D] (may be NULL) and adds 1 to n_accept [C] (may be NULL),
 * as edward/inferences/sgld.py:82-89 and sghmc.py:88-97 do.
 *   edhmc_sgld_step  (sgld.py:62-80): z <- z + lr/2 grad + sqrt(lr) xi, with
 *                    lr = step_size / (t + 1)^0.55 computed by the caller
 *   edhmc_sghmc_step (sghmc.py:66-86): sample = z + v,
 *                    v <- (1 - friction/2) v + lr grad + sqrt(lr friction) xi,
 *                    lr = step_size * 0.01; v [C, D] fp32 in/out (caller owned,
 *                    zero at the start as the reference's velocity variable)
 * xi ~ N(0, I) from the momentum stream of (seed, counter). */
int edhmc_sgld_step(edhmc_ctx* ctx, float learning_rate, uint64_t seed, uint64_t counter,
                    float* sample_out, int32_t* n_accept, void* stream);
int edhmc_sghmc_step(edhmc_ctx* ctx, float learning_rate, float friction, float* v,
                     uint64_t seed, uint64_t counter, float* sample_out, int32_t* n_accept,
                     void* stream);
/* Random-walk Metropolis-Hastings over the bound GLM
 * (edward/inferences/metropolis_hastings.py:64-156) with the proposal
 * Normal(loc = z, scale = proposal_scale[D]) (device fp32): z' = z + s xi
 * (xi from the momentum stream of (seed, counter)), ratio = log p(x, z') -
 * log p(x, z) (the symmetric proposal's terms cancel), accept iff
 * log u < ratio with u [C] fp64 fed or drawn on the device.  The log joint is
 * the forward pass's (bf16: hi/lo split, FP64 reduction).  Outputs as
 * edhmc_transition. */
int edhmc_mh_step(edhmc_ctx* ctx, const float* proposal_scale, const double* u, uint64_t seed,
                  uint64_t counter, int32_t* accepted_out, float* sample_out, int32_t* n_accept,
                  void* stream);

/* Replica-exchange Monte Carlo (edward/inferences/replica_exchange_mc.py:
 * 78-237).  The C chains form n_replica replicas of C / n_replica ladders
 * (chain = replica * ladders + ladder); beta [C] (device fp32) is each
 * chain's inverse temperature.  Each chain takes the random-walk MH step of
 * edhmc_mh_step with its ratio scaled by beta (replica_exchange_mc.py:200);
 * then, per ladder, with probability exchange_freq and from a random start
 * in {0, 1}, adjacent replicas (i, i+1), (i+2, i+3), ... exchange states iff
 * log u < (logp_i - logp_{i+1}) (beta_{i+1} - beta_i) (:214-237).  Draws: the
 * MH proposal and accept as edhmc_mh_step; the exchange coin, start and
 * uniforms from the uniform stream with a = 1 + ladder.  sample_out [C, D]
 * gets every chain's state after the exchange, n_accept [C] the MH accepts,
 * swaps [C / n_replica] int32 the exchanges (any output may be NULL). */
int edhmc_replica_exchange_step(edhmc_ctx* ctx, const float* proposal_scale, const float* beta,
                                int32_t n_replica, float exchange_freq, uint64_t seed,
                                uint64_t counter, int32_t* accepted_out, float* sample_out,
                                int32_t* n_accept, int32_t* swaps, void* stream);

/* Multiplier of the likelihood term in the log joint and its gradient:
 * Edward's `scale` of the observed variable (inference.py:187-193, used by
 * sgld.py:95-121 and sghmc.py:99-135 for minibatches; HMC ignores it,
 * hmc.py:161-192, so the HMC entry points see 1; a model without observed
 * data binds one zero row with weight 0).  bind resets it to 1. */
int edhmc_set_lik_weight(edhmc_ctx* ctx, float weight);

/* Constrained priors for auto-transformed latents (edward/inferences/
 * inference.py:222-264 builds the transformations; hmc.py:132-159 adds the
 * log-Jacobian of the map from the unconstrained position u).  kind [D]
 * uint8 (device): low nibble = family of the prior in u-space, high nibble =
 * the operand the linear predictor receives:
 *   family 0 Normal(loc, scale) on u           operand 0 u
 *          1 Gamma(loc=a, scale=b) at softplus(u)       1 softplus(u)
 *          2 Beta(loc=a, scale=b) at sigmoid(u)         2 sigmoid(u)
 *          3 InverseGamma(loc=a, scale=b) at softplus(u)
 * where loc/scale are the bound prior vectors, read as the family's two
 * parameters; lnorm [D] fp64 (device) is each coordinate's normalising
 * constant (Normal: -log scale - log(2 pi)/2; Gamma and InverseGamma:
 * a log b - lgamma(a); Beta: -log B(a, b)).  kind = NULL restores plain
 * Normal priors with identity operands.  Binding a model clears it.  Not
 * available with the surrogate transition (EDHMC_EUNSUPPORTED there). */
int edhmc_set_prior_family(edhmc_ctx* ctx, const uint8_t* kind, const double* lnorm,
                           void* stream);

/* Newton terms of the bound GLM's log joint at one position w [D] (fp32,
 * device): logp_out (fp64 scalar), grad_out [D] fp64 and neg_hess_out
 * [D, D] fp64 = -Hessian = X^T diag(s) X + diag(1 / scale^2), s_n the
 * likelihood curvature in eta_n.  The mode and N(mode, neg_hess^-1) are
 * the Laplace approximation (edward/inferences/laplace.py:17-137); the
 * engine uses them for a warm start of the chains.  Normal priors and
 * likelihood weight 1 only (EDHMC_EUNSUPPORTED otherwise). */
int edhmc_newton_terms(edhmc_ctx* ctx, const float* w, double* logp_out, double* grad_out,
                       double* neg_hess_out, void* stream);

/* The run loop of Inference.run (edward/inferences/inference.py:145-147):
 * transitions t = t0 .. t0+n_iter-1, each writing params[t] of the sample
 * store params [T, C, D] and using momentum counter counter0 + t.  The chain
 * state must already equal params[max(t0-1, 0)] (edhmc_set_state). */
int edhmc_run(edhmc_ctx* ctx, float eps, int32_t n_steps, float* params, int64_t T,
              int64_t t0, int64_t n_iter, uint64_t seed, uint64_t counter0,
              int32_t* n_accept, void* stream);

/* Running FP64 per-chain moments (Welford) of every sample the transitions
 * write; the readout of Empirical.mean/stddev (edward/models/empirical.py:87-96)
 * without reading the whole sample store.  enable != 0 clears them. */
int edhmc_moments_enable(edhmc_ctx* ctx, int32_t enable);
/* mean_out, m2_out [C, D] fp64 (m2 = sum of squared deviations); count_out
 * host int64. */
int edhmc_moments_get(edhmc_ctx* ctx, double* mean_out, double* m2_out,
                      int64_t* count_out, void* stream);

/* Event profiler: with enable != 0, every kernel launch is bracketed by
 * hipEvents recorded on its stream; edhmc_profile_read synchronises and
 * returns, per kernel class, the summed duration (ms) and launch count:
 *   0 GEMM1 gradient pass (X w with residual epilogue)
 *   1 GEMM1 forward with FP64 log-likelihood epilogue only
 *   2 GEMM2 (X^T residual, split-K)
 *   3 per-chain leapfrog / accept kernels
 *   4 surrogate MLP GEMMs
 *   5 hierarchical segmented reductions
 * total_ms_out, count_out: host arrays of 6.  Enabling clears the totals. */
int edhmc_profile_enable(edhmc_ctx* ctx, int32_t enable);
int edhmc_profile_read(edhmc_ctx* ctx, double* total_ms_out, int64_t* count_out);
/* Time only the classes whose bit is set (bit k = class k above; default all
 * six).  Each timed launch costs two event records on its stream, about
 * 10 us between dependent launches: time only the class being measured
 * inside a timed region.  Classes left out report zero launches. */
int edhmc_profile_select(edhmc_ctx* ctx, uint32_t class_mask);

/* Record the (position, exact gradient) pair of every `stride`-th inner
 * leapfrog point (steps k = stride, 2 stride, ... < n_steps) of the following
 * plain transitions into W_buf / G_buf [capacity, D] (rows appended
 * chain-major per leapfrog step) until full; W_buf = NULL stops.  These pairs
 * train the gradient surrogate. */
int edhmc_record_pairs(edhmc_ctx* ctx, float* W_buf, float* G_buf, int64_t capacity,
                       int32_t stride);
/* Rows recorded by the current (or the last stopped) recording. */
int64_t edhmc_record_count(edhmc_ctx* ctx);

/* Empirical.mean / Empirical.variance (edward/models/empirical.py:87-96) of a
 * sample store params [T, K] (K = product of the event shape) over rows
 * t0 .. T-1: two-pass FP64, divisor T - t0 (population variance, as the
 * reference's stddev).  params_dtype EDHMC_F32 or 2 (= fp64);
 * mean_out, var_out [K] fp64. */
int edhmc_sample_moments(edhmc_ctx* ctx, const void* params, int32_t params_dtype, int64_t T,
                         int64_t K, int64_t t0, double* mean_out, double* var_out, void* stream);

/* Column means and standard deviations (FP64 accumulation, divisor P) of
 * X [P, D]: mean_out, std_out [D] fp32.  Deterministic. */
int edhmc_col_stats(edhmc_ctx* ctx, const float* X, int64_t P, int32_t D, float* mean_out,
                    float* std_out, void* stream);

/* The same column statistics, allocating nothing: no counterpart in the
 * reference -- for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The partial sums live in a 16 MiB file-scope device array of the library
 * (one per GPU, held until the process ends) instead of context scratch;
 * D is limited to 4096 (EDHMC_EINVAL beyond).  Results are bitwise those of
 * edhmc_col_stats.  Calls that overlap on one GPU share the array and race:
 * serialise them (one stream, or an external lock). */
int edhmc_col_stats_noalloc(edhmc_ctx* ctx, const float* X, int64_t P, int32_t D,
                            float* mean_out, float* std_out, void* stream);

/* Number of chains whose proposal was non-finite (rejected) since binding,
 * [C] int32 device array copy. */
int edhmc_divergences(edhmc_ctx* ctx, int32_t* out, void* stream);

/* bf16 models: exact = 1 makes the inner leapfrog gradients of
 * edhmc_transition use both bf16 planes of the positions (hi + lo, the
 * positions to 2^-17, as the first and last gradients always do); the
 * default 0 uses the hi plane only (the positions rounded to bf16, about a
 * third cheaper).  Either way the accept test of hmc.py:100-111 uses the
 * exact log joint, so this changes the proposal, not the target.  fp32
 * models ignore it. */
int edhmc_set_inner_gradient(edhmc_ctx* ctx, int32_t exact);

/* Parity hook: the FP64 Metropolis log ratio
 *   K(r0) - K(rL) + log p(zL) - log p(z0)
 * of every chain's last edhmc_transition / edhmc_transition_surrogate
 * (edward/inferences/hmc.py:100-107), [C] fp64 device array copy; the
 * accept test was log(u) < ratio (hmc.py:108-109). */
int edhmc_last_ratio(edhmc_ctx* ctx, double* ratio_out, void* stream);

/* Parity hook: the momenta r_L at the end of every chain's last GLM
 * transition (plain or surrogate; kept whether or not the proposal was
 * accepted), r_out [C, D] fp32 -- the r of hmc.py:100-102's kinetic term. */
int edhmc_get_momentum(edhmc_ctx* ctx, float* r_out, void* stream);

/* Device RNG parity hooks: the momenta / uniforms edhmc_transition would
 * draw for (seed, counter).  r_out [C, D] fp32, u_out [C] fp64. */
int edhmc_draw_momenta(edhmc_ctx* ctx, int32_t C, int32_t D, uint64_t seed,
                       uint64_t counter, float* r_out, void* stream);
int edhmc_draw_uniforms(edhmc_ctx* ctx, int32_t C, uint64_t seed, uint64_t counter,
                        double* u_out, void* stream);

/* Multi-chain autocovariance for effective-sample-size estimation over a
 * sample store s [T, C, D] fp32 (no reference counterpart; build-only):
 *   acov_out   [max_lag+1, D] fp64  sum over chains of the biased (/T)
 *              within-chain autocovariance at each lag
 *   chain_mean [C, D] fp64 (may be NULL to skip)                            */
int edhmc_autocov(edhmc_ctx* ctx, const float* s, int64_t T, int32_t C, int32_t D,
                  int32_t max_lag, double* acov_out, double* chain_mean_out,
                  void* stream);

/* ------------------------------------------------------------------------
 * Learned gradient surrogate (no reference counterpart; build-only).
 * An MLP D -> H -> H -> D with tanh-GELU whose output layer also reads the
 * normalised input (linear skip): out = W3 [GELU(a2) | x] + b3.  fp32 master
 * weights, bf16 MFMA GEMMs with fp32 accumulation, trained with Adam on
 * (w, grad log p) pairs.  It drives the inner leapfrog steps of
 * edhmc_transition_surrogate; the exact log joint still decides every
 * Metropolis test.
 * Weight layout (fp32, one flat buffer, PyTorch Linear convention):
 *   W1 [H, D] b1 [H] W2 [H, H] b2 [H] W3 [D, H + D] b3 [D]
 * (W3's first H columns act on the hidden layer, the last D on x).
 * Normalisation: x = (w - in_shift) * in_scale,  grad = out * out_scale + out_shift.
 * ------------------------------------------------------------------------ */
int edhmc_sur_create(edhmc_ctx* ctx, int32_t D, int32_t H, int32_t max_batch);
int64_t edhmc_sur_num_params(edhmc_ctx* ctx);
int edhmc_sur_set_params(edhmc_ctx* ctx, const float* params, void* stream);
int edhmc_sur_get_params(edhmc_ctx* ctx, float* params_out, void* stream);
int edhmc_sur_set_norm(edhmc_ctx* ctx, const float* in_shift, const float* in_scale,
                       const float* out_shift, const float* out_scale, void* stream);
/* grad_out [B, D] = surrogate gradient at W [B, D]. */
int edhmc_sur_forward(edhmc_ctx* ctx, const float* W, int32_t B, float* grad_out,
                      void* stream);
/* Loss = mean over batch and coordinates of (out - target_n)^2 with
 * target_n = (G - out_shift) / out_scale.  Writes the loss (fp64, device
 * [1]) and the gradient of the loss w.r.t. the flat parameters (fp32, device,
 * edhmc_sur_num_params entries); with param_grad_out NULL only the loss is
 * computed (no backward pass: held-out losses).  The loss is a fixed-order
 * sum of per-wave partials: the same bits on every run.  Parity hook for the
 * backward kernels. */
int edhmc_sur_loss_grad(edhmc_ctx* ctx, const float* W, const float* G, int32_t B,
                        double* loss_out, float* param_grad_out, void* stream);
/* One Adam step (Kingma & Ba 2015) on a batch: loss_grad then
 * m = b1 m + (1-b1) g, v = b2 v + (1-b2) g^2,
 * p -= lr (m / (1 - b1^t)) / (sqrt(v / (1 - b2^t)) + adam_eps). */
int edhmc_sur_train_step(edhmc_ctx* ctx, const float* W, const float* G, int32_t B,
                         float lr, float beta1, float beta2, float adam_eps, int32_t step,
                         double* loss_out, void* stream);
/* n_batches Adam steps in one call: batch b is the pairs W[perm[b*batch+i]],
 * G[perm[b*batch+i]] (i < batch; W, G [P, D], perm int64 device), trained
 * as edhmc_sur_train_step with learning rate lr_host[b] (host array) and
 * Adam step step0 + b + 1; losses_out [n_batches] fp64 (device).
 * Permutation entries outside [0, P) read as NaN rows (the batch's loss is
 * NaN).  Equal to the sequence of single steps, without a host round trip
 * per batch. */
int edhmc_sur_train_epoch(edhmc_ctx* ctx, const float* W, const float* G, int64_t P,
                          const int64_t* perm, int32_t batch, int32_t n_batches,
                          const float* lr_host, float beta1, float beta2, float adam_eps,
                          int32_t step0, double* losses_out, void* stream);

/* Rows of a device matrix: dst [B, D] = src [P, D] rows idx[0..B) (int64);
 * an index outside [0, P) gives a NaN row.  (The surrogate's held-out
 * training pairs; no reference counterpart.) */
int edhmc_gather_rows(edhmc_ctx* ctx, const float* src, int64_t P, int32_t D,
                      const int64_t* idx, int32_t B, float* dst, void* stream);

/* Gram products for the least-squares start of the surrogate's linear skip:
 * over the rows idx[0..n) of the pairs (W [P, D], G [P, D]), with the
 * normalisation set by edhmc_sur_set_norm, x = (w - in_shift) * in_scale and
 * t = (g - out_shift) / out_scale, xtx_out = sum x^T x, xtt_out = sum x^T t,
 * [D, D] fp64.  No reference counterpart. */
int edhmc_sur_skip_gram(edhmc_ctx* ctx, const float* W, const float* G, int64_t P,
                        const int64_t* idx, int32_t n, double* xtx_out, double* xtt_out,
                        void* stream);

/* out [n, D] = shift [D] + E [n, D] R [D, D] (fp32): draws of a Gaussian
 * N(shift, R^T R) from standard normal rows E (the Laplace warm start's
 * draws, edward/inferences/laplace.py:17-137 fits that Gaussian). */
int edhmc_affine_rows(edhmc_ctx* ctx, const float* E, int64_t n, int32_t D, const float* R,
                      const float* shift, float* out, void* stream);
/* HMC transition whose leapfrog forces come from the surrogate; the ratio
 * uses the exact log joint at the end point (bf16 X with a hi/lo bf16 split
 * of w, FP64 reduction), so the chain targets the exact posterior.  Same
 * arguments and semantics as edhmc_transition. */
int edhmc_transition_surrogate(edhmc_ctx* ctx, float eps, int32_t n_steps, const float* r0,
                               const double* u, uint64_t seed, uint64_t counter,
                               int32_t* accepted_out, float* sample_out, int32_t* n_accept,
                               void* stream);

/* ------------------------------------------------------------------------
 * Batched hierarchical logistic regression (G groups, rows sorted by group):
 *   a_g ~ N(mu, exp(log_sigma)^2), mu ~ N(0,1), log_sigma ~ N(0,1),
 *   w ~ N(0, I_D), y_n ~ Bernoulli(a_{g(n)} + x_n . w).
 * Latent layout per chain: [a (G), mu, log_sigma, w (D)].  In Edward:
 * Bernoulli(logits=tf.gather(a, group) + ed.dot(X, w)); the gather gradient
 * is the IndexedSlices case of tests/inferences/hmc_test.py:105-123 and
 * hmc.py:203,208.  Each gradient pass is the fp32 GLM GEMM pair with the
 * group intercept gathered into the GEMM1 epilogue, plus a ragged segmented
 * reduction of the residuals for the intercept gradients.
 * ------------------------------------------------------------------------ */
int edhmc_bind_hier(edhmc_ctx* ctx, const float* X, const float* y, const int64_t* offsets,
                    int32_t G, int64_t N, int32_t D, void* stream);
int edhmc_hier_logp_grad(edhmc_ctx* ctx, const float* theta, int32_t C, float* grad_out,
                         double* logp_out, void* stream);
/* Prior of the group-scale coordinate u of the hierarchical model (bind
 * sets u = log sigma ~ Normal(0, 1)).  kind = family | map << 4: family 0
 * Normal(p1, p2) on u, 1 Gamma(p1, p2) or 3 InverseGamma(p1, p2) on sigma
 * with the log-Jacobian of the map (edward/inferences/inference.py:222-264,
 * hmc.py:132-159); map 0: sigma = exp(u) (a log-scale latent), 1: sigma =
 * softplus(u) (a positive latent auto-transformed by inverse softplus,
 * util/random_variables.py:856-917).  lnorm is the family's normalising
 * constant.  Invalidates the chain state (set it again). */
int edhmc_hier_set_scale_prior(edhmc_ctx* ctx, int32_t kind, double p1, double p2,
                               double lnorm);

int edhmc_hier_set_state(edhmc_ctx* ctx, const float* theta, int32_t C, void* stream);
int edhmc_hier_transition(edhmc_ctx* ctx, float eps, int32_t n_steps, const float* r0,
                          const double* u, uint64_t seed, uint64_t counter,
                          int32_t* accepted_out, float* sample_out, int32_t* n_accept,
                          void* stream);

#ifdef __cplusplus
}
#endif

#endif /* EDHMC_H_ */
