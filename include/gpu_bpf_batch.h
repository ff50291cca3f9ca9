/* gpu_bpf_batch.h -- batched inner SV bootstrap filters and the SMC^2 layer
 * on them, for MI355X (gfx950).  Plain C ABI.
 *
 * No reference function maps one to one onto these entry points: the
 * reference steps its SMC^2 inner filters with one fixed particle count per
 * build (smc2/smc2_rbpf_batch.cu:248-398, kernel_rbpf_step_impl<N_INNER>) and
 * its parameter sweeps loop the plain filter (gpu_bpf/gpu_bpf.cu:1022-1080,
 * gpu_imm_step).  A batch runs M filters of uneven particle counts in one
 * launch per tick; inner filter m is the plain filter of gpu_bpf.h created
 * with (n_particles[m], phi[m], sigma_z[m], mu[m], 0, nu_obs, seeds[m]) and
 * the batch's ESS threshold, tick for tick.
 */
#ifndef GPU_BPF_BATCH_H
#define GPU_BPF_BATCH_H

#include "gpu_bpf.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct GpuBpfBatch GpuBpfBatch;

/* Largest particle count of an inner filter. */
#define GPU_BPF_BATCH_MAX_N 16384

/* n_filters inner filters; arrays of n_filters entries.  nu_obs > 0 selects
 * Student-t observations for every filter.  Returns NULL on failure
 * (gpu_bpf_last_error()). */
GpuBpfBatch* gpu_bpf_batch_create(int n_filters, const int* n_particles, const float* mu,
                                  const float* phi, const float* sigma_z, const int* seeds,
                                  float nu_obs, float ess_threshold);
void gpu_bpf_batch_destroy(GpuBpfBatch* batch);

/* Threads per inner filter (64, 128, 256, 512 or 1024); 0 = chosen by the
 * autotuner at the first run.  Returns 0 or -1. */
int gpu_bpf_batch_set_block(GpuBpfBatch* batch, int threads);
int gpu_bpf_batch_get_block(GpuBpfBatch* batch);
/* Particles per tile (256, 512 or 1024; 0 = 256): each wave of an inner
 * filter's workgroup takes whole tiles, 4 tile/256 particles per lane, and
 * each inner filter is the plain filter with the same tile
 * (gpu_bpf_set_items_per_thread(tile / 256)).  Returns 0 or -1. */
int gpu_bpf_batch_set_tile(GpuBpfBatch* batch, int particles);
int gpu_bpf_batch_get_tile(GpuBpfBatch* batch);
/* Per-tick statistics on (1, the default) or off (0).  Off, the ticks form
 * only what the filters need -- the weight sum for the log-likelihood and the
 * CDF, and the sum of squared weights when an ESS threshold is set -- and
 * record nothing in the ring; log-likelihoods, resampling, particles and
 * scores are unchanged bit for bit.  Statistics are available again for the
 * ticks run after they are switched back on.  No counterpart in the
 * reference; SMC^2 runs its inner filters with them off.  Returns 0 or -1. */
int gpu_bpf_batch_set_stats(GpuBpfBatch* batch, int on);
/* Times every (block, tile) pair whose LDS fits one CU on a scratch batch of
 * the same filters, with the same score estimator when this batch's score
 * is enabled (enable it first): two untimed ticks on y[0..1], then
 * ticks - 2 timed ones on the rest of y (ticks >= 3 values); keeps the
 * fastest pair and returns its block (gpu_bpf_batch_get_tile gives its
 * tile), or -1. */
int gpu_bpf_batch_autotune(GpuBpfBatch* batch, const float* y, int ticks);

/* Advance every filter by n_ticks observations (host or device array);
 * returns 0 or -1.  Ticks run as replays of one captured HIP graph. */
int gpu_bpf_batch_run(GpuBpfBatch* batch, const float* y, int n_ticks);
int gpu_bpf_batch_run_device(GpuBpfBatch* batch, const float* y_dev, int n_ticks);
int gpu_bpf_batch_synchronize(GpuBpfBatch* batch);
int gpu_bpf_batch_get_timestep(GpuBpfBatch* batch);

/* Per-tick statistics of the last n_ticks ticks, tick-major
 * (out[t * n_filters + m]); at most the ring capacity (1024 ticks). */
int gpu_bpf_batch_copy_stats(GpuBpfBatch* batch, BpfTickStats* out, int n_ticks);
/* FP64 cumulative log-likelihood of each filter. */
int gpu_bpf_batch_get_loglik(GpuBpfBatch* batch, double* ll);
int gpu_bpf_batch_get_resample_counts(GpuBpfBatch* batch, int* counts);
/* The current particles of filter m (n_particles[m] floats). */
int gpu_bpf_batch_copy_particles(GpuBpfBatch* batch, int filter, float* x);

/* Score of every filter's log-likelihood with respect to (mu, phi,
 * sigma_eta), carried forward as a per-particle 3-vector.  estimator 0:
 * pathwise -- ancestors held fixed and x_0 = mu + sd z differentiated (the
 * quantity of gpu_bpf_score_backward with x0_reparam = 1), accumulated over
 * every tick run since.  estimator 1: Fisher identity -- each particle carries
 * grad log p(x_0:t) along its ancestry and the score is sum_i w_i alpha_i of
 * the latest tick.  Enable before the first tick, with ess_threshold <= 0.
 * get_score writes n_filters x 3 doubles. */
int gpu_bpf_batch_enable_score(GpuBpfBatch* batch, int estimator);
int gpu_bpf_batch_get_score(GpuBpfBatch* batch, double* score);

/* ---- SMC^2 over the batched inner filters ---------------------------------
 * Mirrors the reference's smc2_cuda_* host API (smc2/smc2_rbpf_batch.cuh:
 * 463-486) for the SV bootstrap inner model: theta = (mu, phi, sigma_eta)
 * with a uniform box prior, each theta-particle driving an inner filter whose
 * particle count is drawn log-uniformly in [n_min, n_max] and travels with it
 * on resampling.  Outer resampling at outer ESS < threshold * n_theta
 * (default 0.5), then `moves` correlated pseudo-marginal MH moves (CPMMH:
 * adaptive Gaussian random walk on theta, Crank-Nicolson noise proposal with
 * rho = 0.99; smc2_rbpf_batch.cu:693-981,1246-1294).  The inner noise of the
 * moves is stored on the device (capacity in ticks, default 2048). */
typedef struct GpuSmc2 GpuSmc2;
/* Random numbers of the outer layer (so that runs repeat and can be checked):
 * PCG32 XSH-RR streams seeded (state 0, increment 2 id + 1, one step, + seed,
 * one step) with id 0x5eed5c2 for the prior draws (per theta-particle:
 * log-uniform particle count, mu, phi, sigma_eta, filter seed), the
 * outer-resampling uniform and the fresh filter seeds after it, and id 0x30e5
 * for the moves (per theta-particle: three proposal normals, a proposal seed;
 * then one MH uniform each); uniforms ((r1 >> 5) 2^26 + (r2 >> 6) + 0.5) 2^-53,
 * normals by Box-Muller (cosine branch).  A filter with seed s uses the
 * particle streams and host resampling stream of gpu_bpf_create(..., s);
 * CPMMH proposals draw their normals from the particle streams of the
 * proposal seed and their uniform-walk steps from its stream index 16384. */

GpuSmc2* gpu_smc2_create(int n_theta, int n_min, int n_max, float nu_obs, int seed);
void gpu_smc2_destroy(GpuSmc2* smc2);
/* lo/hi: 3 floats each (mu, phi, sigma_eta); default [-2,0]x[0.90,0.995]x[0.05,0.40]. */
int gpu_smc2_set_prior(GpuSmc2* smc2, const float* lo, const float* hi);
/* kind 0 = CPMMH (stored noise, correlated proposals), 1 = MALA
 * (preconditioned Langevin proposals from the per-filter pathwise scores,
 * fresh noise per proposal); moves per rejuvenation (default 5).  Before
 * gpu_smc2_init_from_prior. */
int gpu_smc2_set_rejuvenation(GpuSmc2* smc2, int kind, int moves);
/* MALA step h: theta' = theta + (h/2) A g + sqrt(h) A^(1/2) xi, A = Cov(theta)
 * of the theta-particles (default 0.5). */
int gpu_smc2_set_mala_step(GpuSmc2* smc2, float h);
int gpu_smc2_set_cpmmh_rho(GpuSmc2* smc2, float rho);
/* Scale of the adaptive theta random walk (default 1) and half-width of the
 * resampling-uniform random walk on the circle (default 0.05). */
int gpu_smc2_set_proposal_scale(GpuSmc2* smc2, float theta_scale, float u_step);
int gpu_smc2_set_outer_threshold(GpuSmc2* smc2, float fraction);
int gpu_smc2_set_capacity(GpuSmc2* smc2, int max_ticks);
int gpu_smc2_init_from_prior(GpuSmc2* smc2);
/* One observation; returns the outer ESS (as smc2_cuda_update), -1 on error. */
float gpu_smc2_update(GpuSmc2* smc2, float y_obs);
int gpu_smc2_run(GpuSmc2* smc2, const float* y, int n_ticks);
float gpu_smc2_get_outer_ess(GpuSmc2* smc2);
int gpu_smc2_get_theta_mean(GpuSmc2* smc2, float* theta_mean);   /* 3 floats */
int gpu_smc2_get_theta_std(GpuSmc2* smc2, float* theta_std);     /* 3 floats */
/* theta [n_theta x 3], outer log-weights, particle counts (any may be NULL). */
int gpu_smc2_copy_theta(GpuSmc2* smc2, float* theta, double* logw, int* n_particles);
/* FP64 cumulative log-likelihood estimate of each theta-particle's filter. */
int gpu_smc2_get_loglik(GpuSmc2* smc2, double* ll);
/* [ticks, outer resamples, moves proposed, moves accepted, replayed ticks,
 *  forward particle-ticks, replayed particle-ticks] */
int gpu_smc2_get_counters(GpuSmc2* smc2, long long* counters);
/* MALA runs: each theta-particle's pathwise score of its filter's
 * log-likelihood with respect to (mu, phi, sigma_eta), n_theta x 3 doubles
 * (the drift of the next move); -1 for CPMMH runs. */
int gpu_smc2_get_score(GpuSmc2* smc2, double* score);
/* Outer-resampling history: for each outer resample so far, the number of
 * ticks done when it happened and the total particle count of the resampled
 * theta-particles' filters (the reference prints its resample ticks,
 * smc2/test_smc2_rbpf.cu).  Fills up to max_entries of each array (either
 * may be NULL) and returns the number of outer resamples, or -1. */
int gpu_smc2_get_history(GpuSmc2* smc2, int* ticks, long long* sum_n, int max_entries);
/* Host only (no GPU): the outer layer's systematic resampling of n
 * log-weights with uniform u in (0, 1) -- slot m takes the first j whose
 * running sum of exp(log_w - max) reaches (u + m) / n of the total
 * (smc2/smc2_rbpf_batch.cu:519-540).  Non-finite log-weights have weight zero
 * and are never chosen; returns -1 when none is finite. */
int gpu_smc2_systematic_ancestors(const double* log_w, int n, double u, int* anc);
int gpu_smc2_set_block(GpuSmc2* smc2, int threads);
/* Particles per tile of the inner filters (as gpu_bpf_batch_set_tile). */
int gpu_smc2_set_tile(GpuSmc2* smc2, int particles);
/* Verification aid (host only, no GPU): the rejuvenation moves' proposal and
 * acceptance code on a closed-form 3-d Gaussian target; kind as above;
 * fault 1 drops the MALA proposal correction.  out: mean[3], cov[9],
 * acceptance rate. */
int gpu_smc2_selftest_mcmc(int kind, int iters, int seed, int fault, double* out);

#ifdef __cplusplus
}
#endif
#endif /* GPU_BPF_BATCH_H */
