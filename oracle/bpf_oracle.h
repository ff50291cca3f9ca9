/*
 * bpf_oracle.h -- CPU oracle for the SV bootstrap particle filter.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may use it, and only as the checker / the timed CPU
 * baseline, never as the thing measured on the GPU.
 *
 * It restates, in plain C, the reference filter's per-tick algorithm:
 *   - PCG32 per-particle streams seeded as in ptx_assembly/bpf_kernels_full.ptx:74-85
 *     and stepped as in :140-164 (index-bound, one draw per propagating tick);
 *   - the host resampling stream of gpu_bpf/gpu_bpf_ptx_full.cu:241-251,353;
 *   - the Acklam (2003) inverse normal CDF with the FP32 coefficients of
 *     ptx_assembly/bpf_kernels_full.ptx:166-207;
 *   - the SV transition and Gaussian / Student-t log-weight of :406-450;
 *   - ESS-triggered systematic resampling with weight carry
 *     (gpu_bpf/gpu_bpf_ptx_full.cu:601-604,676-745; target rule of
 *     gpu_bpf/gpu_bpf.cu:297-310);
 *   - the matched-DGP generator of tests/test_bpf_matched_dgp.cu:59-175.
 *
 * Two filter modes:
 *   ORC_MODE_F32 -- the FP32 restatement.  Per-particle arithmetic is written
 *                   with explicit fmaf and correctly rounded + - * / sqrt only,
 *                   so it is bit-reproducible on any IEEE machine; weights are
 *                   formed tile-relative as the GPU forms them; sums, CDF and
 *                   log-likelihood are FP64 sequential (textbook order).
 *   ORC_MODE_F64 -- the FP64 oracle: same normals and same resampling
 *                   uniforms, but states, log-weights, weights, CDF and
 *                   log-likelihood in double with libm exp/log.
 *   ORC_MODE_F64_COUPLED -- the FP64 oracle run beside the FP32 restatement
 *                   and resampled with the FP32 run's decisions and ancestors,
 *                   so both share one genealogy and the difference of their
 *                   outputs is the arithmetic error of the FP32 pipeline.
 *                   Outputs are the FP64 ones.
 * Plus a point-mass (grid) filter giving the exact SV log-likelihood.
 */
#ifndef BPF_ORACLE_H
#define BPF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MODE_F32 0
#define ORC_MODE_F64 1
#define ORC_MODE_F64_COUPLED 2

/* Filter parameters (reference names: rho = phi, sigma_z = sigma_eta). */
typedef struct {
    int   n_particles;
    float rho;
    float sigma_z;
    float mu;
    float nu_obs;        /* 0 = Gaussian observation noise */
    float ess_threshold; /* <= 0: resample every tick */
    int   seed;
    int   tile;          /* particles per weight tile (must equal the GPU's) */
} OrcParams;

/* Per-tick outputs (same fields the GPU library reports). */
typedef struct {
    double mean;      /* sum w x / sum w, pre-resample */
    double var;       /* sum w x^2 / sum w - mean^2 */
    double ess;       /* (sum w)^2 / sum w^2 */
    double ll_incr;   /* log p(y_t | y_{1:t-1}) estimate */
    double ll_cum;    /* FP64 running sum of ll_incr */
    double ll_ref;    /* reference-compatible BpfResult.log_lik */
    float  h_mean;    /* (float)mean */
    float  log_lik;   /* (float)ll_ref */
    int    resampled;
    int    tick;
} OrcTick;

/* ---- PCG32 ---------------------------------------------------------------- */
uint64_t orc_pcg_state(int seed, uint32_t i, uint32_t k);   /* state of particle i before draw k */
uint64_t orc_pcg_state_naive(int seed, uint32_t i, uint32_t k);
uint32_t orc_pcg_output(uint64_t old_state);
uint64_t orc_pcg_host_seed(int seed);
uint32_t orc_pcg_host_next(uint64_t* state);
float    orc_uniform(uint32_t r);

/* ---- scalar math (bit-reproducible FP32) ---------------------------------- */
float  orc_expf(float x);
float  orc_logf(float x);
float  orc_icdf_f32(float u);
double orc_icdf_f64(double u);
float  orc_normal(int seed, uint32_t i, uint32_t k);   /* z of particle i, draw k */
/* element-wise: fn 0 = orc_icdf_f32, 1 and 2 = orc_expf, 3 = orc_logf */
void   orc_math_array(int fn, const float* in, float* out, int n);

/* ---- data generation ------------------------------------------------------ */
void orc_dgp(int n_ticks, double rho, double sigma_z, double mu, double nu_state,
             double nu_obs, uint64_t seed, double* true_h, double* returns);
void orc_dgp_regimes(int n_ticks, double rho, double sigma_z, const double* mu_t,
                     double nu_obs, uint64_t seed, double* true_h, double* returns);

/* ---- filters -------------------------------------------------------------- */
/* Run T ticks.  out[T] receives per-tick outputs; x_out (nullable, N floats)
 * the particle states after the last tick; anc_out (nullable, N ints) the
 * ancestors of the last resample; mu_t (nullable, T floats) overrides mu per
 * tick.  Returns the resample count, or -1 on bad arguments. */
int orc_bpf_run(const OrcParams* p, int mode, const float* y, const float* mu_t,
                int n_ticks, OrcTick* out, float* x_out, double* x64_out,
                int* anc_out, int n_threads);

/* Online mu/rho learner of gpu_bpf/gpu_bpf_ptx_full.cu:469-537,777-800
 * (natural gradient + Robbins-Monro on the one-step terms of the reference's
 * kernel 14, ptx_assembly/bpf_kernels_full.ptx:1116-1384). */
typedef struct {
    int   update_K;      /* ticks between updates (<= 0: every tick) */
    int   learn_rho;     /* also update rho */
    float c, t0, gamma;  /* eta_k = c / (k + t0)^gamma */
    float grad_clip;     /* the reference sets 5 */
} OrcLearn;

/* One learner update in the reference's FP32 arithmetic: sums = the four
 * accumulated sums (FP64, rounded to FP32 first), k = ticks accumulated,
 * *rm_step the step counter (incremented); updates *mu and, if learn_rho,
 * *rho. */
void orc_learn_update(const OrcLearn* lp, const double* sums, int k, int* rm_step,
                      float* mu, float* rho);

/* orc_bpf_run (FP32 restatement only, mode 0) with the learner on from tick
 * 0.  lsum_out (nullable, 4*T doubles) receives each tick's normalised sums
 * sum_i w_i (g, g^2, g d, g^2 d^2)_i with g = d lw/dx of the propagated
 * particle and d = x_prev - mu (0 on tick 0); par_out (nullable, 2*T floats)
 * the (mu, rho) in force after each tick. */
int orc_bpf_run_learn(const OrcParams* p, const OrcLearn* lp, const float* y, int n_ticks,
                      OrcTick* out, double* lsum_out, float* par_out, float* x_out,
                      int n_threads);

/* sigma_z scale bands (gpu_bpf/gpu_bpf_ptx_full.cu:257-334,
 * ptx_assembly/bpf_kernels_full.ptx:297-324).  Config 0 is the static one;
 * with adaptive set, configs 0/1/2 are calm/alert/panic, switched before each
 * propagating tick by the previous tick's surprise |y| e^{-h_mean/2}. */
typedef struct {
    int   adaptive;
    int   n_particles;   /* the count the fractions refer to */
    int   n_bands[3];
    float fracs[3][4];
    float scales[3][4];
    float thr_alert, thr_panic;
} OrcBands;

/* orc_bpf_run (FP32 restatement, mode 0) with bands; regime_out (nullable,
 * T ints) receives the regime each tick propagated with. */
int orc_bpf_run_bands(const OrcParams* p, const OrcBands* bd, const float* y, int n_ticks,
                      OrcTick* out, float* x_out, int* regime_out, int n_threads);

/* orc_bpf_run (FP32 restatement, mode 0) with Student-t state noise
 * (nu_state > 0; chi2 normals from PCG32 streams of seed * 31337, stream
 * i * round(nu_state) + k, draw t - 1 at tick t) and / or the Silverman
 * jitter after every resample (silverman_shrink > 0; one more draw of every
 * particle stream per jitter). */
int orc_bpf_run_ext(const OrcParams* p, float nu_state, float silverman_shrink, const float* y,
                    int n_ticks, OrcTick* out, float* x_out, int n_threads);

/* Auxiliary particle filter (gpu_bpf/gpu_bpf_ptx_full.cu:990-1170) in FP32
 * with this build's PCG32 streams: out[t] holds the second stage's statistics,
 * ll_incr = log mean first-stage weight + log mean second-stage weight, ll_ref
 * the reference's log_lik (second stage only).  Returns n_ticks. */
int orc_apf_run(const OrcParams* p, float nu_state, const float* y, int n_ticks, OrcTick* out,
                float* x_out, int n_threads);

/* Every option at once (each nullable / 0 = off): learner, bands, Student-t
 * state noise, Silverman jitter. */
int orc_bpf_run_all(const OrcParams* p, const OrcLearn* lp, const OrcBands* bd, float nu_state,
                    float silverman_shrink, const float* y, int n_ticks, OrcTick* out,
                    float* par_out, int* regime_out, float* x_out, int n_threads);

/* Systematic resampling given a normalised FP64 CDF: lower_bound of the float
 * targets u + i/n (wrapped at 1) clamped to n-1. */
void orc_resample_ancestors(const double* cdf, int n, float u, int* anc);

/* Score d LL / d(mu, phi, sigma) of the always-resample filter with the
 * ancestors held fixed (FP64 forward-mode tangents).  Ancestors come from the
 * FP32 restatement (as on the GPU) unless anc_in (T*N ints) replays given
 * ones; anc_out (T*N, nullable) records them.  theta64 (nullable) overrides
 * the FP64 parameters (finite differences under common random numbers).
 * x0_reparam: 1 = differentiate through x_0 = mu + sd(phi, sigma) z_0.
 * ll_out: FP64 LL; ll32_out: LL of the FP32 restatement. */
int orc_bpf_score(const OrcParams* p, const float* y, int n_ticks, int x0_reparam,
                  const double* theta64, const int* anc_in, int* anc_out,
                  double* ll_out, double* score_out, double* ll32_out);

/* Plant a mistake in the oracle (tests only): 0 none, 1 drop the -x/2 term
 * of the log-weight, 2 negate phi in the transition, 3 upper_bound instead
 * of lower_bound in resampling, 4 drop the weight carry, 6 drop the factor
 * 1/2 of eta^2 in the learner's d lw/dx, 7 take d = x' - mu (the propagated
 * state) instead of x_prev - mu in the learner, 8 drop the first-stage
 * correction -lw(mp) of the auxiliary filter's second-stage weight.
 * Process-global. */
void orc_set_fault(int code);

/* Threads of orc_bpf_score's loops (1, the default: sequential, fixed
 * summation order; more: OpenMP reductions). */
void orc_set_threads(int n);

/* One tick of the FP32 restatement (plain filter) from a given state: x (in:
 * particles, out: after this tick's resampling if any), l (in: carried
 * log-weights when carry, out: this tick's log-weights), x_pre (out: the
 * propagated states before resampling), anc (out: ancestors, identity
 * without resampling), *hrng the host resampling stream (advanced when
 * resampling), draw t of the particle streams (tick 0 does not propagate).
 * The carried maximum and sum come from l.  out receives the statistics
 * (ll_ref and ll_cum are not set); ll64 (nullable) the FP64 evaluation of the
 * tick's log-likelihood increment from the same state and normals.
 * Returns 1 if the tick resampled, 0 if not, -1 on bad arguments. */
int orc_bpf_tick(const OrcParams* p, float mu, int t, float y, int carry, uint64_t* hrng, float* x,
                 float* l, float* x_pre, int* anc, OrcTick* out, double* ll64, int n_threads);

/* ---- helpers shared with smc2_oracle.c ------------------------------------ */
/* PCG32 jump constants of draw k: state of particle i = a + (2i + 1) g. */
void   orc_pcg_jump(int seed, uint32_t k, uint64_t* a, uint64_t* g);
/* FP32 observation log-weight (Gaussian, or Student-t with nu_obs > 0). */
float  orc_lw(float x, float ysq, float nu_obs, float c_obs, float nph);
/* Tile-relative FP64 weights w (tiles of `tile`, groups of 256 tiles);
 * returns the global max log-weight. */
float  orc_tiled_weights(const float* l, int n, int tile, float* tmax, double* tf, double* w);
/* Normalised CDF of the weights (80-bit running sum, one FP64 rounding). */
void   orc_build_cdf(const double* w, int n, double* cdf);
/* Student-t log-density constant C_obs in FP64. */
double orc_c_obs(double nu);

/* Exact log p(y_{1:T}) of the Gaussian SV model by a point-mass filter. */
double orc_grid_loglik(const float* y, int n_ticks, double rho, double sigma_z,
                       double mu, int n_grid, double* ll_t);

#ifdef __cplusplus
}
#endif
/* ---- SMC^2 with the RBPF / OCSN inner filter (rbpf_oracle.c) ---------- */
typedef struct {
    int n_theta, n_inner, K_rejuv, lag, cap, adaptive, noise_bits;   /* noise_bits 32 or 16 */
    float ess_thr, crho;
    uint64_t seed;              /* nonzero */
    float pstd[4];
    float prior[8];             /* SVPrior field order */
    float bounds[8];            /* SVBounds field order */
    float tcurve[3];            /* theta(z): base, scale, rate */
    float fixed[4];             /* mu_scale, mu_rate, sigma_scale, sigma_rate */
} OrcRbCfg;
typedef struct { int n_resamples, n_acc, n_tot, t_cp; } OrcRbCount;
/* Whole run over T observations (T <= cap).  Outputs (nullable): outer ESS
 * per tick [T], theta [4][n_theta], log-likelihoods and log-weights
 * [n_theta], inner ESS [n_theta], inner state [4][n_theta * n_inner],
 * resample ticks [n_resamples]. */
int orc_rb_smc2_run(const OrcRbCfg* c, const float* y, int T, float* ess_out, float* theta_out,
                    double* ll_out, double* lw_out, float* essi_out, float* inner_out,
                    int* resample_ticks, OrcRbCount* cnt);
int orc_rb_inner_run(const OrcRbCfg* c, const float* theta, const float* y, int T,
                     double* ll_incr, double* ess, float* state_out);
void orc_rb_ocsn(float y, float mp, float vp, float* out3);
void orc_rb_math(int fn, const float* in, float* out, int n);
void orc_rb_ocsn_table(float* p, float* m, float* v);
/* float -> IEEE half (round to nearest even) -> float, element-wise */
void orc_rb_half_roundtrip(const float* in, float* out, int n);

#endif
