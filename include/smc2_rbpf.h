/* smc2_rbpf.h -- C ABI of the SMC^2 learner whose inner filter is a
 * Rao-Blackwellised particle filter (RBPF) over a regime-switching SV model:
 * the regime z~ is carried by particles, the log-volatility h by per-particle
 * Kalman moments, and the observation log(y^2) by the 10-component OCSN
 * mixture (Omori, Chib, Shephard and Nakajima 2007).
 *
 * Drop-in for the host API of smc2/smc2_rbpf_batch.cuh:458-490 (same entry
 * point names, arguments and meaning).  The reference's state struct is
 * public and its tests read and write fields (prior, bounds, curves,
 * K_rejuv, n_resamples, ...; smc2/test_smc2_rbpf.cu:225-244,298-316); here
 * the state is opaque and those pokes become the smc2_cuda_set_model /
 * smc2_cuda_get_model / smc2_cuda_set_rejuv / smc2_cuda_get_counters calls
 * below.
 *
 * Model (smc2/smc2_rbpf_batch.cuh:8-103, smc2/smc2_rbpf_batch.cu:250-398):
 *   z~_t = rho z~_{t-1} + sigma_z eta_t,     z_t = 1.5 (1 + tanh z~_t)
 *   f(z) = base + scale (1 - exp(-rate z))   for theta(z), mu(z), sigma_h(z)
 *   h_t = (1 - theta) h_{t-1} + theta mu(z) + sigma_h(z) eps_t
 *   y_t = log(r_t^2) = h_t + log chi^2_1   (OCSN mixture)
 * Learned: rho, sigma_z, mu_base, sigma_base.  Fixed: the curve shapes.
 *
 * Random numbers are counter-based (a splitmix64 hash of seed, purpose,
 * outer-resample epoch, theta index, particle index and tick), so every
 * draw is a pure function of its coordinates and the CPU oracle reproduces
 * a run bit for bit.  The reference uses cuRAND streams.
 *
 * Errors: entry points returning int give 0 on success and -1 on failure;
 * smc2_cuda_update returns NaN on failure; void setters record the error.
 * The message is read with gpu_bpf_last_error() (include/gpu_bpf.h).
 */
#ifndef SMC2_RBPF_H
#define SMC2_RBPF_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SMC2_N_PARAMS 4   /* rho, sigma_z, mu_base, sigma_base */

/* smc2/smc2_rbpf_batch.cuh:142-147: independent N(mean, std^2) priors. */
typedef struct SVPrior {
    float rho_mean, rho_std;
    float sigma_z_mean, sigma_z_std;
    float mu_base_mean, mu_base_std;
    float sigma_base_mean, sigma_base_std;
} SVPrior;

/* smc2/smc2_rbpf_batch.cuh:153-158: hard support of the learned parameters. */
typedef struct SVBounds {
    float rho_min, rho_max;
    float sigma_z_min, sigma_z_max;
    float mu_base_min, mu_base_max;
    float sigma_base_min, sigma_base_max;
} SVBounds;

/* smc2/smc2_rbpf_batch.cuh:166-171: fixed shapes of mu(z) and sigma_h(z). */
typedef struct SVFixedCurves {
    float mu_scale, mu_rate;
    float sigma_scale, sigma_rate;
} SVFixedCurves;

/* smc2/smc2_rbpf_batch.cuh:176-180: f(z) = base + scale (1 - exp(-rate z)). */
typedef struct SVCurve {
    float base, scale, rate;
} SVCurve;

typedef struct SMC2Counters {
    int n_resamples;       /* outer resampling events */
    int n_rejuv_accepts;   /* accepted CPMMH moves */
    int n_rejuv_total;     /* attempted CPMMH moves (N_theta per move round) */
    int t_current;         /* index of the last observation, -1 before any */
    int t_checkpoint;      /* fixed-lag checkpoint tick, -1 when none */
    int noise_capacity;    /* ticks of stored noise per particle */
} SMC2Counters;

typedef struct SMC2StateCUDA SMC2StateCUDA;

/* smc2/smc2_rbpf_batch.cu:1028-1143.  N_inner in {64, 128, 256, 512, 1024}
 * (the reference: 64..512); N_theta >= 2.  Defaults as the reference: prior,
 * bounds, curves, proposal std, ESS thresholds 0.5, K_rejuv 5, noise
 * capacity 2048, CPMMH rho 0.99, no fixed lag.  NULL on failure. */
SMC2StateCUDA* smc2_cuda_alloc(int N_theta, int N_inner);
/* smc2_rbpf_batch.cu:1145-1186 */
void smc2_cuda_free(SMC2StateCUDA* state);
/* smc2_rbpf_batch.cu:1188-1191.  Seeds the host outer-resampling stream and
 * every device stream (0 selects the fixed default 12345; the reference
 * seeds from the clock). */
void smc2_cuda_set_seed(SMC2StateCUDA* state, uint64_t seed);
/* smc2_rbpf_batch.cu:1193-1219.  Grows the stored noise history (never
 * shrinks); existing histories are kept, row by row. */
void smc2_cuda_set_noise_capacity(SMC2StateCUDA* state, int capacity);
/* smc2_rbpf_batch.cu:1221-1224.  L > 0: checkpoints every L ticks, and CPMMH
 * replays from the checkpoint when it is at most 2L ticks old. */
void smc2_cuda_set_fixed_lag(SMC2StateCUDA* state, int L);
/* smc2_rbpf_batch.cu:1238-1240: correlation of proposed and current noise. */
void smc2_cuda_set_cpmmh_rho(SMC2StateCUDA* state, float rho);
/* smc2_rbpf_batch.cu:1226-1236: random-walk std per parameter, NULL for the
 * defaults {0.01, 0.02, 0.1, 0.02}. */
void smc2_cuda_set_proposal_std(SMC2StateCUDA* state, const float* std);
/* smc2_rbpf_batch.cu:1296-1327: draws theta from the prior (truncated to the
 * bounds by rejection) and the inner particles from the stationary law;
 * resets the counters and the observation history. */
void smc2_cuda_init_from_prior(SMC2StateCUDA* state);
/* smc2_rbpf_batch.cu:1329-1502: one observation y = log(r^2): RBPF tick of
 * every theta-particle, outer reweighting, and when the outer ESS falls below
 * the threshold an outer resample, K_rejuv CPMMH moves and a weight reset.
 * Returns the outer ESS after the update (NaN on failure). */
float smc2_cuda_update(SMC2StateCUDA* state, float y_obs);
/* smc2_rbpf_batch.cu:1504-1523,1525-1551: weighted posterior mean / std of
 * (rho, sigma_z, mu_base, sigma_base). */
void smc2_cuda_get_theta_mean(SMC2StateCUDA* state, float* theta_mean);
void smc2_cuda_get_theta_std(SMC2StateCUDA* state, float* theta_std);
/* smc2_rbpf_batch.cu:1553-1557 */
float smc2_cuda_get_outer_ess(SMC2StateCUDA* state);

/* ---- replacements for the reference's field access ---------------------- */

/* state->prior / bounds / theta_curve / fixed_curves (any may be NULL = keep).
 * Takes effect at the next smc2_cuda_init_from_prior. */
int smc2_cuda_set_model(SMC2StateCUDA* state, const SVPrior* prior, const SVBounds* bounds,
                        const SVCurve* theta_curve, const SVFixedCurves* fixed_curves);
int smc2_cuda_get_model(SMC2StateCUDA* state, SVPrior* prior, SVBounds* bounds,
                        SVCurve* theta_curve, SVFixedCurves* fixed_curves);
/* state->K_rejuv (>= 0) and state->ess_threshold_outer (fraction of N_theta). */
int smc2_cuda_set_rejuv(SMC2StateCUDA* state, int K_rejuv, float ess_threshold_outer);
/* Compile-time SMC2_NOISE_FP16 of smc2/smc2_noise_precision.cuh as a setting:
 * 16 stores the noise histories as IEEE half (half the memory and copy
 * traffic; every noise is used as stored, rounded to nearest even), 32 (the
 * default) as FP32.  Reallocates the histories; call smc2_cuda_init_from_prior
 * afterwards. */
int smc2_cuda_set_noise_precision(SMC2StateCUDA* state, int bits);
/* state->use_adaptive_proposals (default on): 0 keeps the initial diagonal
 * proposal diag(proposal_std) * 2.38 / sqrt(4) instead of the scaled
 * Cholesky factor of the theta-particles' covariance. */
int smc2_cuda_set_adaptive(SMC2StateCUDA* state, int on);
/* state->n_resamples, n_rejuv_accepts, n_rejuv_total, t_current, ... */
int smc2_cuda_get_counters(SMC2StateCUDA* state, SMC2Counters* out);
/* Host copies of the theta-particles (each array N_theta long, any NULL):
 * the learned parameters, the outer log-weights (FP64), the pseudo-marginal
 * log-likelihoods (FP64) and the inner ESS of the last tick. */
int smc2_cuda_get_particles(SMC2StateCUDA* state, float* rho, float* sigma_z, float* mu_base,
                            float* sigma_base, double* log_weight, double* log_likelihood,
                            float* ess_inner);
/* Host copies of the inner particles, N_theta * N_inner each, theta-major
 * (any NULL): regime z~, Kalman mean and variance of h, log-weight. */
int smc2_cuda_get_inner(SMC2StateCUDA* state, float* z_tilde, float* mu_h, float* var_h,
                        float* log_w);

#ifdef __cplusplus
}
#endif

#endif /* SMC2_RBPF_H */
