/*
 * gpu_bpf.h -- C ABI of the MI355X (gfx950) bootstrap particle filter for the
 * stochastic-volatility model
 *
 *     x_t = mu + phi (x_{t-1} - mu) + sigma_eta eta_t,   y_t = exp(x_t / 2) eps_t
 *
 * (reference names: rho = phi, sigma_z = sigma_eta, h = x).
 *
 * Drop-in for the reference's plain-filter API (gpu_bpf/gpu_bpf_full.cuh and
 * gpu_bpf/gpu_bpf.cuh): the same extern "C" symbols, argument meanings and
 * BpfResult layout.  The state is opaque; callers that poked GpuBpfState
 * fields use the setters below instead.  Plain C: no HIP or CUDA types.
 *
 * Threading: one HIP stream per filter state; distinct states may be used from
 * distinct threads; a single state is not reentrant.
 * Errors: functions that can fail return NULL / a negative code / NaN and set
 * a thread-local message readable with gpu_bpf_last_error().  Nothing calls
 * exit().  A NaN observation is propagated (NaN statistics, no resampling),
 * not trapped.
 */
#ifndef GPU_BPF_H
#define GPU_BPF_H

#ifdef __cplusplus
extern "C" {
#endif

/* Per-tick output, returned by value (8-byte POD).
 * Replaces BpfResult of gpu_bpf/gpu_bpf_full.cuh:33-36 (identical layout). */
typedef struct {
    float h_mean;   /* posterior mean E[x_t | y_{0:t}], pre-resample */
    float log_lik;  /* reference semantics: max_lw + log(sum_w / N) of the
                       log-weights accumulated since the last resample
                       (ptx_assembly/bpf_kernels_full.ptx:755-775) */
} BpfResult;

/* Full per-tick statistics (build-owned extension, 64 bytes). */
typedef struct {
    double mean;     /* sum w x / sum w */
    double var;      /* sum w x^2 / sum w - mean^2 (filtered variance) */
    double ess;      /* (sum w)^2 / sum w^2 */
    double ll_incr;  /* incremental log p(y_t | y_{0:t-1}) estimate */
    double ll_cum;   /* FP64 running sum of ll_incr since create */
    double ll_ref;   /* BpfResult.log_lik before the cast to float */
    float  h_mean;   /* (float)mean */
    float  log_lik;  /* (float)ll_ref */
    int    resampled;
    int    tick;
} BpfTickStats;

/* Opaque filter state.  Replaces the public struct GpuBpfState of
 * gpu_bpf/gpu_bpf_full.cuh:80-126 (gpu_bpf/gpu_bpf.cuh:30-54). */
typedef struct GpuBpfState GpuBpfState;

/* ---- reference API (gpu_bpf/gpu_bpf_full.cuh) ---------------------------- */

/* gpu_bpf_full.cuh:144 / gpu_bpf_ptx_full.cu:340.  Seeds per-particle PCG32
 * streams and draws x_0 from N(mu, sigma_z^2 / max(1 - rho^2, 1e-6)).
 * nu_obs > 0 selects Student-t observation noise.  nu_state > 0 selects
 * Student-t state noise eps = z / sqrt(chi2 / nu_state), chi2 the sum of
 * round(nu_state) squared normals per particle and tick
 * (gpu_bpf_ptx_full.cu:50-62,392-404; ptx:393-406).  The reference draws
 * those normals with cuRAND; here they come from PCG32 streams seeded with
 * seed * 31337 (stream i * round(nu_state) + k for normal k of particle i,
 * draw t - 1 at tick t), so runs repeat exactly.  round(nu_state) must be
 * 1..64, else NULL.  Default ESS threshold 0.5. */
GpuBpfState* gpu_bpf_create(int n_particles, float rho, float sigma_z, float mu,
                            float nu_state, float nu_obs, int seed);

/* gpu_bpf_full.cuh:148.  NULL-safe. */
void gpu_bpf_destroy(GpuBpfState* state);

/* gpu_bpf_full.cuh:160 / gpu_bpf_ptx_full.cu:764.  One tick, synchronous. */
BpfResult gpu_bpf_step(GpuBpfState* state, float y_t);

/* gpu_bpf_full.cuh:168 / gpu_bpf_ptx_full.cu:550.  One tick, no host sync
 * (the resampling decision is taken on the device). */
void gpu_bpf_step_async(GpuBpfState* state, float y_t);

/* gpu_bpf_full.cuh:174 / gpu_bpf_ptx_full.cu:754.  Waits for the state's
 * stream, then returns the latest tick's result. */
BpfResult gpu_bpf_get_result(GpuBpfState* state);

/* gpu_bpf_full.cuh:297 / gpu_bpf_ptx_full.cu:834.  Creates a filter, steps
 * all ticks, returns the RMSE of h_mean against true_h over ticks >= 100. */
double gpu_bpf_run_rmse(const double* returns, const double* true_h, int n_ticks,
                        int n_particles, float rho, float sigma_z, float mu,
                        float nu_state, float nu_obs, int seed);

/* gpu_bpf_full.cuh:282 / gpu_bpf_ptx_full.cu:858.  Resample when ESS/N is
 * below the threshold; <= 0 resamples every tick. */
void gpu_bpf_set_ess_threshold(GpuBpfState* state, float threshold);

/* gpu_bpf_full.cuh:285 / gpu_bpf_ptx_full.cu:862. */
int gpu_bpf_get_resample_count(GpuBpfState* state);

/* gpu_bpf_full.cuh:207-230 / gpu_bpf_ptx_full.cu:802-831.  set_rho clamps to
 * [0.001, 0.999] as the reference does.  Changes apply from the next tick.
 * With online learning on, set_mu / set_rho also restart the learner's step
 * counter and accumulators, and get_mu / get_rho return the learned values. */
float gpu_bpf_get_mu(GpuBpfState* state);
float gpu_bpf_get_rho(GpuBpfState* state);
void  gpu_bpf_set_mu(GpuBpfState* state, float mu);
void  gpu_bpf_set_rho(GpuBpfState* state, float rho);

/* gpu_bpf_full.cuh:176-204 / gpu_bpf_ptx_full.cu:469-537,777-800 (online
 * mu/rho learning by natural gradient + Robbins-Monro).  Each tick adds the
 * weighted one-step terms of the reference's kernel 14
 * (ptx_assembly/bpf_kernels_full.ptx:1116-1384) to four FP64 accumulators:
 * sum w g, sum w g^2, sum w g d, sum w g^2 d^2 with g = d log p(y_t | x_t)/dx_t
 * and d = x_{t-1} - mu (0 on tick 0, where the reference reads an unset
 * buffer).  They are computed inside the propagate pass (no extra particle
 * traffic).  Every K ticks the tick synchronises once and applies the
 * reference's update to mu (and rho when enabled); the new values apply from
 * the next tick.  mode 0 is off and any other value is on: the reference's
 * code applies the same natural-gradient step for modes 1 and 2.  K <= 0
 * updates every tick.  gpu_bpf_score_forward refuses to run while learning
 * is on. */
void gpu_bpf_enable_mu_learning(GpuBpfState* state, int mode, int K,
                                float c, float t0, float gamma);
void gpu_bpf_disable_mu_learning(GpuBpfState* state);
void gpu_bpf_enable_rho_learning(GpuBpfState* state, int enable);

/* gpu_bpf_full.cuh:232-272 / gpu_bpf_ptx_full.cu:257-334 (sigma_z scale
 * bands).  Particles [0, b0) scale sigma_z by scales[0], [b0, b1) by
 * scales[1], ... with b the rounded cumulative fractions of n_particles
 * (the last band takes the rest); 1 to 4 bands.  As in the reference the
 * setters take no filter: the configuration is process-wide and every filter
 * adopts it at its next tick.  The adaptive variant holds a calm / alert /
 * panic table switched by the surprise |y_t| e^{-h_t/2} of the previous
 * gpu_bpf_step (EMA 0.3, thresholds as given); the regime machine runs per
 * filter (the reference keeps one for the process).  gpu_bpf_run steps tick
 * by tick while adaptive bands are set; gpu_bpf_step_async and
 * gpu_bpf_run_device leave the surprise unchanged, as the reference's
 * step_async does.  Bad arguments record an error and change nothing. */
void gpu_bpf_set_bands(int n_particles, int n_bands, const float* fracs,
                       const float* scales);
void gpu_bpf_set_adaptive_bands(int n_particles,
                                const float* calm_fracs, const float* calm_scales, int calm_nb,
                                const float* alert_fracs, const float* alert_scales, int alert_nb,
                                const float* panic_fracs, const float* panic_scales, int panic_nb,
                                float thresh_alert, float thresh_panic);

/* ---- build-owned extensions ----------------------------------------------- */

/* Thread-local text of the last error ("" if none). */
const char* gpu_bpf_last_error(void);

/* Replaces writing GpuBpfState.silverman_shrink (gpu_bpf_full.cuh:109,
 * tests/test_cond_resample.cu:77).  > 0 turns on the Silverman jitter after
 * each resample (gpu_bpf_ptx_full.cu:708-739): bandwidth shrink * sigma_hat *
 * 1.05922 * N^-0.2, one more draw of every particle's stream, clamp to
 * [mu - 8, mu + 8]; like the reference, a jittered tick waits on the host. */
void gpu_bpf_set_silverman_shrink(GpuBpfState* state, float shrink);

/* The adaptive bands' current regime of this filter: 0 calm, 1 alert,
 * 2 panic (0 without adaptive bands), or -1 with gpu_bpf_last_error(). */
int gpu_bpf_get_band_regime(GpuBpfState* state);

/* The learner's four accumulated sums since its last update (sum over ticks
 * of sum_i w_i g_i, sum w g^2, sum w g d, sum w g^2 d^2; the reference's
 * d_scalars[5..8], gpu_bpf_full.cuh:88-89) and the ticks they cover.
 * Synchronises the filter's stream.  Returns 0, or -1 with gpu_bpf_last_error(). */
int gpu_bpf_get_learn_sums(GpuBpfState* state, double* sums, int* n_ticks);

/* Replaces reading GpuBpfState.sigma_z (gpu_bpf_full.cuh:107). */
float gpu_bpf_get_sigma_z(GpuBpfState* state);
void  gpu_bpf_set_sigma_z(GpuBpfState* state, float sigma_z);

/* Run n_ticks ticks on y[0..n_ticks) with no host synchronisation inside the
 * loop; per-tick results are copied out at the end (either pointer may be
 * NULL).  Returns 0 or a negative error code. */
int gpu_bpf_run(GpuBpfState* state, const float* y, int n_ticks,
                BpfResult* results, BpfTickStats* stats);

/* Same, with the observations already in device memory (y_dev is a device
 * pointer to n_ticks floats).  Enqueues only: call gpu_bpf_synchronize
 * before reading anything; stats of the last 'ring' ticks stay on the device
 * and can be fetched with gpu_bpf_copy_stats. */
int gpu_bpf_run_device(GpuBpfState* state, const float* y_dev, int n_ticks);

/* Device copy of n host floats (an observation series for
 * gpu_bpf_run_device or gpu_bpf_batch_run_device, for callers without a
 * device allocator of their own); NULL on failure.  Free it with
 * gpu_bpf_device_free (NULL-safe). */
float* gpu_bpf_device_upload(const float* host, int n);
void gpu_bpf_device_free(float* dev);
/* Copy n floats of device memory to the host (synchronous); 0 or -1. */
int gpu_bpf_device_download(const float* dev, float* host, int n);

/* Copy the statistics of the last n ticks (n <= ring capacity) to the host. */
int gpu_bpf_copy_stats(GpuBpfState* state, BpfTickStats* stats, int n);

/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc): the allocation-free variant of gpu_bpf_run_device.  It
 * runs a fresh filter -- the one gpu_bpf_create(n_particles, rho, sigma_z,
 * mu, 0, nu_obs, seed) with gpu_bpf_set_ess_threshold(ess_threshold) would
 * be -- over n_ticks observations already in device memory, on `stream`
 * (a hipStream_t; NULL = the default stream), and enqueues only.  Outputs, in
 * caller-owned device memory: the n_particles final particle states in
 * x_out_dev and, unless stats_out_dev is NULL, one BpfTickStats per tick.
 * Results equal gpu_bpf_run_device's on such a filter bit for bit.  Working
 * memory is one library-held device array of 52 MiB per GPU (held until the
 * process ends), so n_particles is at most GPU_BPF_NOALLOC_MAX_N; nothing is
 * allocated or freed per call.  Calls that overlap on one GPU share that
 * array and race: serialise them.  Returns 0, or -1 with
 * gpu_bpf_last_error() (bad sizes are rejected before anything is
 * enqueued). */
#define GPU_BPF_NOALLOC_MAX_N 4194304
int gpu_bpf_run_device_noalloc(int n_particles, float rho, float sigma_z, float mu, float nu_obs,
                               int seed, float ess_threshold, const float* y_dev, int n_ticks,
                               float* x_out_dev, BpfTickStats* stats_out_dev, void* stream);

/* Grow the device ring of per-tick statistics to hold at least `ticks`
 * ticks (default 1024; gpu_bpf_run grows it to its run length).  Records
 * already held are kept.  Allocates; call it before a run of
 * gpu_bpf_run_device whose every tick is to be read back. */
int gpu_bpf_set_stats_capacity(GpuBpfState* state, int ticks);

/* Latest tick's full statistics (synchronises). */
int gpu_bpf_get_stats(GpuBpfState* state, BpfTickStats* stats);

/* Copy the current particle states (N floats) and, when the filter carries
 * weights across ticks, the current log-weights (NULL to skip). */
int gpu_bpf_copy_particles(GpuBpfState* state, float* x, float* log_w);

/* Weight-tile geometry: particles per tile and number of tiles. */
int gpu_bpf_get_layout(GpuBpfState* state, int* tile, int* n_tiles);

/* Particles per lane of the per-tile kernels (1, 2, 4, 8 or 16); only before
 * the first step.  0 restores the automatic choice.  A change of tiling drops
 * a score pass run before it (gpu_bpf_score_backward then asks for a new
 * gpu_bpf_score_forward). */
int gpu_bpf_set_items_per_thread(GpuBpfState* state, int items);

/* Wait for all work enqueued on the state's stream. */
int gpu_bpf_synchronize(GpuBpfState* state);

/* Number of ticks processed so far. */
int gpu_bpf_get_timestep(GpuBpfState* state);

/* Restart at tick 0 without allocating: waits for pending work, redraws x_0
 * from draw 0 of the particle streams with the current mu, rho and sigma_z,
 * and resets the host resampling stream, the resample count, the cumulative
 * log-likelihood, the learner's counters and the band regime.  The ESS
 * threshold, options and buffers stay.  The next ticks then repeat those of a
 * filter created with the same arguments.  Returns 0, or -1 with
 * gpu_bpf_last_error(). */
int gpu_bpf_restart(GpuBpfState* state);

/* Pathwise score of the log-likelihood with the ancestors held fixed
 * (replaces the reference's autograd path, tests/diff_bpf.py:186-233, and its
 * one-step kernel bpf_grad_alpha, ptx_assembly/bpf_kernels_full.ptx:1116).
 * LL = sum_t [logsumexp_i lw_t,i - log N] of the always-resample filter.
 * score_forward runs n_ticks ticks from a fresh filter (it sets the ESS
 * threshold to 0), keeps the particle states every checkpoint_every ticks
 * and returns LL (FP64).  score_backward replays each segment from its
 * checkpoint and sweeps it in reverse (segments replay on three streams
 * ahead of the sweep, which runs on a fourth); score[3] =
 * dLL/d(mu, phi, sigma_z), the same bit for bit whatever checkpoint_every.
 * x0_reparam = 1 also differentiates the stationary initial draw
 * x_0 = mu + sigma_z / sqrt(1 - phi^2) z_0.  Memory: (T / checkpoint_every
 * + 3 * checkpoint_every + 8) * N floats (three segments replay at once); a
 * later score_forward of the same n_ticks and checkpoint_every (after
 * gpu_bpf_restart) keeps these buffers. */
int gpu_bpf_score_forward(GpuBpfState* state, const float* y, int n_ticks,
                          int checkpoint_every, double* loglik);
int gpu_bpf_score_backward(GpuBpfState* state, int x0_reparam, double* score);

/* Estimator of the next gpu_bpf_score_forward: 0 = pathwise (above, the
 * default), 1 = Fisher identity -- each particle carries
 * alpha = grad log p(x_0:t | mu, phi, sigma_z) along its ancestry (initial
 * stationary density included), and the score is sum_i w_i alpha_i of the
 * last tick.  With 1, score_forward computes the score in its single pass
 * (checkpoint_every is not used; memory 6 * N floats) and score_backward
 * returns it (x0_reparam is not used).  No counterpart in the reference. */
int gpu_bpf_set_score_estimator(GpuBpfState* state, int estimator);

/* Per-kernel timing with HIP start/stop events tied to the dispatches of
 * the tick's three kernels (propagate, finalise, resample) on the state's
 * own stream, for every enable-th tick (1 = every tick, 0 = off).  Enable before enqueuing ticks;
 * gpu_bpf_kernel_times synchronises, then returns the summed milliseconds
 * and launch counts of the timed ticks since the last reset (arrays of 3:
 * propagate, finalise, resample). */
int gpu_bpf_enable_kernel_timing(GpuBpfState* state, int enable);
int gpu_bpf_kernel_times(GpuBpfState* state, double* ms, int* launches, int reset);
/* The per-tick milliseconds (3 per tick, NaN where not timed) of the last
 * n_ticks timed ticks; returns how many ticks were copied. */
int gpu_bpf_kernel_tick_times(GpuBpfState* state, float* ms, int n_ticks);

/* Verification aid (no reference counterpart): evaluates the device's
 * per-particle arithmetic element-wise on n host inputs into n host outputs.
 * fn 0 = inverse normal CDF of a uniform (ptx_assembly/bpf_kernels_full.ptx:
 * 166-205), 1 = exp, 2 = exp for arguments <= 0, 3 = log (the FP32 exp/log
 * of the weight and tail paths).  Returns 0, or -1 with gpu_bpf_last_error(). */
int gpu_bpf_eval_device_math(int fn, const float* in, float* out, int n);

/* ---- auxiliary particle filter (gpu_bpf_full.cuh:303-361) ----------------- */

/* gpu_bpf_full.cuh:314-337 (opaque here). */
typedef struct GpuApfState GpuApfState;

/* gpu_bpf_full.cuh:340 / gpu_bpf_ptx_full.cu:1050.  Pitt & Shephard's
 * auxiliary filter with the bootstrap filter's streams: x_0 from draw 0,
 * tick t >= 1 resamples by the first-stage weights g(y_t | mu + phi (x - mu)),
 * propagates with draw t and weights by g(y_t | x') / g(y_t | mp), then
 * resamples again; both resampling offsets come from the host stream of
 * seed * 67890 + 12345 in that order (gpu_bpf_ptx_full.cu:1102-1170).  The
 * reference draws its noise with cuRAND; here it is PCG32 as in
 * gpu_bpf_create (Student-t state noise as there).  NULL on bad arguments. */
GpuApfState* gpu_apf_create(int n_particles, float rho, float sigma_z, float mu,
                            float nu_state, float nu_obs, int seed);
/* gpu_bpf_full.cuh:344.  NULL-safe. */
void gpu_apf_destroy(GpuApfState* state);
/* gpu_bpf_full.cuh:352.  h_mean = second-stage weighted mean; log_lik =
 * max + log(sum w / N) of the second-stage weights, as the reference
 * returns it (gpu_bpf_ptx_full.cu:952-955). */
BpfResult gpu_apf_step(GpuApfState* state, float y_t);
/* gpu_bpf_full.cuh:355-358.  RMSE of h_mean against true_h from tick 100. */
double gpu_apf_run_rmse(const double* returns, const double* true_h, int n_ticks,
                        int n_particles, float rho, float sigma_z, float mu,
                        float nu_state, float nu_obs, int seed);
/* Build-owned: n_ticks ticks without a host wait between them (observations
 * copied once).  stats (nullable) per tick: the second stage's statistics,
 * with ll_incr = log mean first-stage weight + log mean second-stage weight
 * (the auxiliary filter's likelihood increment) and ll_cum its FP64 sum.
 * Returns 0, or -1 with gpu_bpf_last_error(). */
int gpu_apf_run(GpuApfState* state, const float* y, int n_ticks, BpfResult* results,
                BpfTickStats* stats);
/* Build-owned: particle states after the last tick (n floats), and the tile
 * layout (as gpu_bpf_copy_particles / gpu_bpf_get_layout). */
int gpu_apf_copy_particles(GpuApfState* state, float* x);
int gpu_apf_get_layout(GpuApfState* state, int* tile, int* n_tiles);

/* ---- interacting multiple models (gpu_bpf/gpu_bpf.cuh:128-162) ----------- */

#define IMM_MAX_MODELS 64

/* gpu_bpf.cuh:132-136. */
typedef struct {
    float rho, sigma_z, mu, nu_state, nu_obs;
} ImmModelParams;

/* gpu_bpf.cuh:138-142. */
typedef struct {
    float h_mean, vol, log_lik;
    int best_model;
    float best_prob;
} ImmResult;

/* gpu_bpf.cuh:144-149 (opaque here). */
typedef struct GpuImmState GpuImmState;

/* gpu_bpf.cuh:151 / gpu_bpf.cu:961-1005.  K = n_models SV filters of
 * n_particles_per_model particles, filter k seeded seed + 7919 k, ESS
 * threshold 0.5; model probabilities start uniform; transition_matrix
 * (K x K, row j = from model j; NULL: stay 0.95, switch evenly) is taken in
 * log space floored at 1e-30.  The models step as one batch of filters (one
 * launch per tick) when they share nu_obs, have no Student-t state noise and
 * at most GPU_BPF_BATCH_MAX_N particles; else as K filters.  The reference's
 * filters here are its cuRAND-based CUDA-C ones; these are the PCG32
 * filters of gpu_bpf_create.  NULL for more than IMM_MAX_MODELS models. */
GpuImmState* gpu_imm_create(const ImmModelParams* models, int n_models,
                            int n_particles_per_model, const float* transition_matrix, int seed);
/* gpu_bpf.cuh:154 / gpu_bpf.cu:1022-1080.  One tick: Markov interaction of
 * the model probabilities, every model's filter tick, Bayes update with each
 * filter's BpfResult.log_lik, probability-weighted h_mean (FP64 on the host).
 * vol = exp(h_mean / 2); log_lik = log sum_k pi_k^- p_k; best model and its
 * probability. */
ImmResult gpu_imm_step(GpuImmState* state, float y_t);
/* gpu_bpf.cuh:155-156 / gpu_bpf.cu:1086-1094. */
float gpu_imm_get_prob(const GpuImmState* state, int k);
void gpu_imm_get_probs(const GpuImmState* state, float* probs_out);
/* gpu_bpf.cuh:157.  NULL-safe. */
void gpu_imm_destroy(GpuImmState* state);
/* gpu_bpf.cuh:159-162 / gpu_bpf.cu:1100-1122.  n_rho * n_sigma * n_mu models
 * (rho outermost); the caller frees the array with free(). */
ImmModelParams* gpu_imm_build_grid(const float* rhos, int n_rho, const float* sigma_zs, int n_sigma,
                                   const float* mus, int n_mu, float nu_state, float nu_obs,
                                   int* out_n_models);
/* Build-owned: 1 when the models step as one batch of filters. */
int gpu_imm_uses_batch(const GpuImmState* state);

#ifdef __cplusplus
}
#endif
#endif /* GPU_BPF_H */
