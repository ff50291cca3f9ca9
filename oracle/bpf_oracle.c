/*
 * bpf_oracle.c -- CPU oracle for the SV bootstrap particle filter.
 *
 * TEST INFRASTRUCTURE ONLY (see bpf_oracle.h).  The shipped GPU library never
 * links or calls this file.  Each function cites the reference lines it
 * restates.  Compile with -ffp-contract=off: every fused multiply-add below is
 * written as fmaf()/fma() on purpose, so results are bit-reproducible.
 */
#include "bpf_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline float f32_bits(uint32_t b) { float f; memcpy(&f, &b, 4); return f; }

/* Planted-mistake switch for the tests that show each pinning check fails
 * on a plausible error (0 = none).  See orc_set_fault in bpf_oracle.h. */
static int g_fault = 0;
void orc_set_fault(int code) { g_fault = code; }

/* Threads of orc_bpf_score (1 = sequential, the default). */
static int g_threads = 1;
void orc_set_threads(int n) { g_threads = n > 1 ? n : 1; }
static inline uint32_t u32_of(float f) { uint32_t b; memcpy(&b, &f, 4); return b; }

/* ========================================================================= */
/* PCG32                                                                      */
/* ========================================================================= */

#define PCG_MULT 6364136223846793005ULL
#define PCG_HOST_INC 1442695040888963407ULL

/* XSH-RR output of the old state: ptx_assembly/bpf_kernels_full.ptx:146-157. */
uint32_t orc_pcg_output(uint64_t old) {
    uint32_t xsh = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xsh >> rot) | (xsh << ((-rot) & 31u));
}

/* Literal restatement of the reference seeding (ptx:74-85: inc = (2i+1)|1,
 * state = (seed*M + inc)*M + inc, seed widened from int) followed by k steps
 * s <- s*M + inc (ptx:140-144). */
uint64_t orc_pcg_state_naive(int seed, uint32_t i, uint32_t k) {
    uint64_t inc = (uint64_t)(((uint32_t)(2u * i + 1u)) | 1u);
    uint64_t s = (uint64_t)(int64_t)seed * PCG_MULT + inc;
    s = s * PCG_MULT + inc;
    for (uint32_t j = 0; j < k; j++) s = s * PCG_MULT + inc;
    return s;
}

/* Closed form: s_k(i) = seed*M^(k+2) + inc_i * G_(k+2), G_n = sum_{j<n} M^j.
 * (M^n, G_n) come from binary powering of the affine map s -> M s + 1. */
typedef struct { uint64_t a, g; } PcgJump;   /* s_k(i) = a + inc_i * g */

static PcgJump pcg_jump(int seed, uint32_t k) {
    uint64_t n = (uint64_t)k + 2u;
    uint64_t acc_m = 1, acc_g = 0;        /* map for the steps taken so far */
    uint64_t cur_m = PCG_MULT, cur_g = 1; /* map for 2^bit steps */
    while (n) {
        if (n & 1u) { acc_g = acc_g * cur_m + cur_g; acc_m = acc_m * cur_m; }
        cur_g = cur_g * cur_m + cur_g;
        cur_m = cur_m * cur_m;
        n >>= 1;
    }
    PcgJump j = { (uint64_t)(int64_t)seed * acc_m, acc_g };
    return j;
}

static inline uint64_t pcg_inc(uint32_t i) { return (uint64_t)(((uint32_t)(2u * i + 1u)) | 1u); }

uint64_t orc_pcg_state(int seed, uint32_t i, uint32_t k) {
    PcgJump j = pcg_jump(seed, k);
    return j.a + pcg_inc(i) * j.g;
}

/* Host stream for the resampling offset: gpu_bpf/gpu_bpf_ptx_full.cu:241-251
 * (default increment), seeded seed*67890+12345 with no warm-up (:353). */
uint64_t orc_pcg_host_seed(int seed) {
    return (uint64_t)(int64_t)seed * 67890ULL + 12345ULL;
}

uint32_t orc_pcg_host_next(uint64_t* state) {
    uint64_t old = *state;
    *state = old * PCG_MULT + PCG_HOST_INC;
    return orc_pcg_output(old);
}

/* u = (r>>9)*2^-23 clamped to [1e-6, 0x3F7FFBCE]: ptx:159-164. */
float orc_uniform(uint32_t r) {
    float u = (float)(r >> 9) * (1.0f / 8388608.0f);
    u = fmaxf(u, f32_bits(0x358637BDu));
    u = fminf(u, f32_bits(0x3F7FFBCEu));
    return u;
}

/* ========================================================================= */
/* Scalar math                                                                */
/* ========================================================================= */

/* exp(x) after S. L. Moshier's Cephes expf (Cody-Waite reduction with
 * C1 = 0.693359375, C2 = -2.12194440e-4, Cephes' degree-5 polynomial P).
 * z is the integer nearest the exact product x log2 e (one fused
 * multiply-add onto 1.5 * 2^23), and the polynomial closes in Horner form,
 * 1 + r (1 + r P(r)).  Inputs below -86 return 0 (no subnormal results),
 * above 88 return +inf. */
float orc_expf(float x) {
    if (!(x >= -86.0f)) return (x != x) ? x : 0.0f;
    if (x > 88.0f) return INFINITY;
    const float shift = 12582912.0f;
    float z = fmaf(x, 1.44269504088896341f, shift) - shift;
    float r = fmaf(z, -0.693359375f, x);
    r = fmaf(z, 2.12194440e-4f, r);
    float p = 1.9875691500e-4f;
    p = fmaf(p, r, 1.3981999507e-3f);
    p = fmaf(p, r, 8.3334519073e-3f);
    p = fmaf(p, r, 4.1665795894e-2f);
    p = fmaf(p, r, 1.6666665459e-1f);
    p = fmaf(p, r, 5.0000001201e-1f);
    p = fmaf(p, r, 1.0f);
    p = fmaf(p, r, 1.0f);
    int n = (int)z;
    return f32_bits(u32_of(p) + ((uint32_t)n << 23));
}

void orc_math_array(int fn, const float* in, float* out, int n) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        float v = in[i];
        out[i] = fn == 0 ? orc_icdf_f32(v) : (fn == 3 ? orc_logf(v) : orc_expf(v));
    }
}

/* log(x) after Moshier's Cephes logf (frexp to [sqrt(1/2), sqrt(2)), degree-8
 * polynomial, exponent added back in two parts). */
float orc_logf(float x) {
    if (x != x) return x;
    if (x < 0.0f) return NAN;
    if (x == 0.0f) return -INFINITY;
    if (x == INFINITY) return x;
    uint32_t b = u32_of(x);
    int e;
    if ((b >> 23) == 0) {                 /* subnormal: scale by 2^25 */
        b = u32_of(x * 33554432.0f);
        e = (int)(b >> 23) - 126 - 25;
    } else {
        e = (int)(b >> 23) - 126;
    }
    float m = f32_bits((b & 0x007FFFFFu) | 0x3F000000u);   /* [0.5, 1) */
    if (m < 0.707106781186547524f) { e -= 1; m = m + m - 1.0f; }
    else { m = m - 1.0f; }
    float z = m * m;
    float y = 7.0376836292e-2f;
    y = fmaf(y, m, -1.1514610310e-1f);
    y = fmaf(y, m, 1.1676998740e-1f);
    y = fmaf(y, m, -1.2420140846e-1f);
    y = fmaf(y, m, 1.4249322787e-1f);
    y = fmaf(y, m, -1.6668057665e-1f);
    y = fmaf(y, m, 2.0000714765e-1f);
    y = fmaf(y, m, -2.4999993993e-1f);
    y = fmaf(y, m, 3.3333331174e-1f);
    y = (y * m) * z;
    float fe = (float)e;
    y = fmaf(fe, -2.12194440e-4f, y);
    y = fmaf(-0.5f, z, y);
    float r = m + y;
    r = fmaf(fe, 0.693359375f, r);
    return r;
}

/* Acklam (2003) inverse normal CDF, FP32 coefficients of ptx:172-203,
 * central region |u-1/2| < 0.47575 (0x3EF39581, ptx:169), tails
 * q = sqrt(-2 log p) negated for the upper tail (ptx:185-205).  Division and
 * square root are IEEE; log is orc_logf. */
float orc_icdf_f32(float u) {
    float q = u - 0.5f;
    if (fabsf(q) < f32_bits(0x3EF39581u)) {
        float r = q * q;
        float num = fmaf(r, -39.69683028665376f, 220.9460984245205f);
        num = fmaf(r, num, -275.9285104469687f);
        num = fmaf(r, num, 138.3577518672690f);
        num = fmaf(r, num, -30.66479806614716f);
        num = fmaf(r, num, 2.506628277459239f);
        num = q * num;
        float den = fmaf(r, -54.47609879822406f, 161.5858368580409f);
        den = fmaf(r, den, -155.6989798598866f);
        den = fmaf(r, den, 66.80131188771972f);
        den = fmaf(r, den, -13.28068155288572f);
        den = fmaf(r, den, 1.0f);
        return num / den;
    }
    int low = u < 0.5f;
    float p = low ? u : 1.0f - u;
    float t = sqrtf(-2.0f * orc_logf(p));
    float num = fmaf(t, -7.784894002430293e-03f, -3.223964580411365e-01f);
    num = fmaf(t, num, -2.400758277161838e+00f);
    num = fmaf(t, num, -2.549732539343734e+00f);
    num = fmaf(t, num, 4.374664141464968e+00f);
    num = fmaf(t, num, 2.938163982698783e+00f);
    float den = fmaf(t, 7.784695709041462e-03f, 3.224671290700398e-01f);
    den = fmaf(t, den, 2.445134137142996e+00f);
    den = fmaf(t, den, 3.754408661907416e+00f);
    den = fmaf(t, den, 1.0f);
    float z = num / den;
    return low ? z : -z;
}

/* Same algorithm in FP64 with Acklam's full-precision coefficients. */
double orc_icdf_f64(double u) {
    double q = u - 0.5;
    if (fabs(q) < 0.47575) {
        double r = q * q;
        double num = (((((-3.969683028665376e+01 * r + 2.209460984245205e+02) * r
                        - 2.759285104469687e+02) * r + 1.383577518672690e+02) * r
                        - 3.066479806614716e+01) * r + 2.506628277459239e+00) * q;
        double den = ((((-5.447609879822406e+01 * r + 1.615858368580409e+02) * r
                        - 1.556989798598866e+02) * r + 6.680131188771972e+01) * r
                        - 1.328068155288572e+01) * r + 1.0;
        return num / den;
    }
    int low = u < 0.5;
    double p = low ? u : 1.0 - u;
    double t = sqrt(-2.0 * log(p));
    double num = ((((-7.784894002430293e-03 * t - 3.223964580411365e-01) * t
                    - 2.400758277161838e+00) * t - 2.549732539343734e+00) * t
                    + 4.374664141464968e+00) * t + 2.938163982698783e+00;
    double den = (((7.784695709041462e-03 * t + 3.224671290700398e-01) * t
                   + 2.445134137142996e+00) * t + 3.754408661907416e+00) * t + 1.0;
    double z = num / den;
    return low ? z : -z;
}

float orc_normal(int seed, uint32_t i, uint32_t k) {
    return orc_icdf_f32(orc_uniform(orc_pcg_output(orc_pcg_state(seed, i, k))));
}

static inline float normal_j(PcgJump j, uint32_t i) {
    return orc_icdf_f32(orc_uniform(orc_pcg_output(j.a + pcg_inc(i) * j.g)));
}

/* ========================================================================= */
/* DGP: tests/test_bpf_matched_dgp.cu:59-99 (PCG32 + FP64 Box-Muller) and     */
/* :141-175 (SV series, h clamped to [-12, 4]).                                */
/* ========================================================================= */

typedef struct { uint64_t state, inc; } DgpRng;

static uint32_t dgp_next(DgpRng* g) {
    uint64_t old = g->state;
    g->state = old * PCG_MULT + g->inc;
    return orc_pcg_output(old);
}

static void dgp_seed(DgpRng* g, uint64_t seed) {
    g->state = 0;
    g->inc = (seed * 12345ULL) | 1u;
    dgp_next(g);
    g->state += seed * 67890ULL;
    dgp_next(g);
    dgp_next(g);
}

static double dgp_double(DgpRng* g) { return (dgp_next(g) >> 11) * (1.0 / 2097152.0); }

static double dgp_gauss(DgpRng* g) {
    double u1, u2;
    do { u1 = dgp_double(g); } while (u1 < 1e-10);
    u2 = dgp_double(g);
    return sqrt(-2.0 * log(u1)) * cos(2.0 * 3.14159265358979323846 * u2);
}

static double dgp_student_t(DgpRng* g, double nu) {
    if (nu > 100.0) return dgp_gauss(g);
    double z = dgp_gauss(g);
    double chi = 0.0;
    int inu = (int)nu;
    for (int i = 0; i < inu; i++) { double v = dgp_gauss(g); chi += v * v; }
    return z / sqrt(chi / nu);
}

static void dgp_series(int n, double rho, double sigma_z, const double* mu_t, double mu0,
                       double nu_state, double nu_obs, uint64_t seed,
                       double* true_h, double* returns) {
    DgpRng g;
    dgp_seed(&g, seed);
    double m0 = mu_t ? mu_t[0] : mu0;
    double var_stat = (sigma_z * sigma_z) / fmax(1.0 - rho * rho, 1e-6);
    double h = m0 + sqrt(var_stat) * dgp_gauss(&g);
    for (int t = 0; t < n; t++) {
        double mu = mu_t ? mu_t[t] : mu0;
        if (t > 0) {
            double eps = (nu_state > 0.0) ? dgp_student_t(&g, nu_state) : dgp_gauss(&g);
            h = mu + rho * (h - mu) + sigma_z * eps;
        }
        if (h < -12.0) h = -12.0;
        if (h > 4.0) h = 4.0;
        if (true_h) true_h[t] = h;
        double vol = exp(h / 2.0);
        double eta = (nu_obs > 0.0) ? dgp_student_t(&g, nu_obs) : dgp_gauss(&g);
        if (returns) returns[t] = vol * eta;
    }
}

void orc_dgp(int n_ticks, double rho, double sigma_z, double mu, double nu_state,
             double nu_obs, uint64_t seed, double* true_h, double* returns) {
    dgp_series(n_ticks, rho, sigma_z, NULL, mu, nu_state, nu_obs, seed, true_h, returns);
}

void orc_dgp_regimes(int n_ticks, double rho, double sigma_z, const double* mu_t,
                     double nu_obs, uint64_t seed, double* true_h, double* returns) {
    dgp_series(n_ticks, rho, sigma_z, mu_t, 0.0, 0.0, nu_obs, seed, true_h, returns);
}

/* ========================================================================= */
/* Resampling: target_i = u + (float)i/(float)n, wrapped once at 1           */
/* (gpu_bpf/gpu_bpf.cu:297-310, IEEE division); ancestor = lower_bound of the */
/* target in cdf[0..n-1] clamped to n-1 (ptx:816-830).                        */
/* ========================================================================= */

static inline float orc_target(float u, int i, int n) {
    float t = u + (float)i / (float)n;
    if (t >= 1.0f) t -= 1.0f;
    return t;
}

void orc_resample_ancestors(const double* cdf, int n, float u, int* anc) {
    for (int i = 0; i < n; i++) {
        double t = (double)orc_target(u, i, n);
        int lo = 0, hi = n - 1;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            int left = (g_fault == 3) ? (cdf[mid] <= t) : (cdf[mid] < t);
            if (left) lo = mid + 1; else hi = mid;
        }
        anc[i] = lo;
    }
}

/* ========================================================================= */
/* The filter                                                                 */
/* ========================================================================= */

/* Log-weight of ptx:414-450: Gaussian -0.9189385 - y^2 e^{-x}/2 - x/2, or
 * Student-t C_obs - (nu+1)/2 log(1 + y^2 e^{-x}/nu) - x/2. */
static inline float lw_f32(float x, float ysq, float nu_obs, float c_obs, float nph) {
    float eta2 = ysq * orc_expf(-x);
    if (nu_obs > 0.0f) {
        float ta = eta2 / nu_obs + 1.0f;
        float lw = fmaf(-nph, orc_logf(ta), c_obs);
        return fmaf(-0.5f, x, lw);
    }
    float lw = fmaf(-0.5f, eta2, f32_bits(0xBF6B3F8Eu));
    return g_fault == 1 ? lw : fmaf(-0.5f, x, lw);
}

static inline double lw_f64(double x, double y, double nu_obs, double c_obs) {
    double eta2 = y * y * exp(-x);
    if (nu_obs > 0.0)
        return c_obs - 0.5 * (nu_obs + 1.0) * log1p(eta2 / nu_obs) - 0.5 * x;
    return -0.91893853320467274178 - 0.5 * eta2 - (g_fault == 1 ? 0.0 : 0.5 * x);
}

/* Student-t constant C_obs in FP64 (gpu_bpf/gpu_bpf_ptx_full.cu:407-412). */
static double c_obs_f64(double nu) {
    return lgamma((nu + 1.0) / 2.0) - lgamma(nu / 2.0) - 0.5 * log(nu * 3.14159265358979323846);
}

/* Normalised CDF of the weights: an 80-bit running sum divided by its own
 * total, rounded to FP64 once.  This is the exact CDF to within an FP64
 * rounding, so targets resolve as they do on the exact values unless one
 * lies within ~1e-16 of a CDF step (the device's three-level FP64 CDF carries
 * errors of that size). */
static void build_cdf(const double* w, int n, double* cdf) {
    long double tot = 0.0L;
    for (int i = 0; i < n; i++) tot += (long double)w[i];
    long double c = 0.0L;
    for (int i = 0; i < n; i++) {
        c += (long double)w[i];
        cdf[i] = (double)(c / tot);
    }
}

/* Weights as the GPU forms them: w_i = e_i * f_tile * f_group with
 * e_i = exp(l_i - m_tile), f_tile = exp(m_tile - m_group),
 * f_group = exp(m_group - M) (each an FP32 orc_expf; the two scales multiply
 * exactly in FP64), tiles of `tile` particles, groups of 256 tiles.
 * Returns M.  tmax/tf (n_tiles) receive the tile maxima and FP64 scales. */
#define ORC_GROUP 256
static float tiled_weights(const float* l, int n, int tile, float* tmax, double* tf, double* w,
                           int n_threads) {
    const int n_tiles = (n + tile - 1) / tile;
    float gm = -INFINITY;
    for (int b = 0; b < n_tiles; b++) {
        float m = -INFINITY;
        int e = (b + 1) * tile < n ? (b + 1) * tile : n;
        for (int i = b * tile; i < e; i++) m = fmaxf(m, l[i]);
        tmax[b] = m;
        gm = fmaxf(gm, m);
    }
    for (int g0 = 0; g0 < n_tiles; g0 += ORC_GROUP) {
        int g1 = g0 + ORC_GROUP < n_tiles ? g0 + ORC_GROUP : n_tiles;
        float grp = -INFINITY;
        for (int b = g0; b < g1; b++) grp = fmaxf(grp, tmax[b]);
        float fg = (grp == -INFINITY) ? 0.0f : orc_expf(grp - gm);
        for (int b = g0; b < g1; b++) {
            float fb = (tmax[b] == -INFINITY) ? 0.0f : orc_expf(tmax[b] - grp);
            tf[b] = (double)fb * (double)fg;
        }
    }
#pragma omp parallel for num_threads(n_threads) schedule(static)
    for (int i = 0; i < n; i++) {
        int b = i / tile;
        float e = (tmax[b] == -INFINITY) ? 0.0f : orc_expf(l[i] - tmax[b]);
        w[i] = (double)e * tf[b];
    }
    return gm;
}

/* d lw / dx of the observation density at the propagated state, in the form
 * of ptx_assembly/bpf_kernels_full.ptx:1172-1189 (eta^2 = y^2 e^{-x}):
 * Gaussian -1/2 + eta^2/2, Student-t -1/2 + (nu+1)/2 eta^2 / (nu + eta^2). */
static inline float learn_dlw(float x, float ysq, float nu_obs, float nph) {
    float eta2 = ysq * orc_expf(-x);
    if (nu_obs > 0.0f) return fmaf(nph, eta2 / (nu_obs + eta2), -0.5f);
    return fmaf(eta2, g_fault == 6 ? 1.0f : 0.5f, -0.5f);
}

/* gpu_bpf/gpu_bpf_ptx_full.cu:485-537 */
void orc_learn_update(const OrcLearn* lp, const double* sums, int k, int* rm_step,
                      float* mu, float* rho) {
    float sm[4];
    for (int q = 0; q < 4; q++) sm[q] = (float)sums[q];
    (*rm_step)++;
    float eta = lp->c / powf((float)(*rm_step) + lp->t0, lp->gamma);
    float K_f = (float)k;
    float K_inv = 1.0f / K_f;
    {
        float mean_g = sm[0] * K_inv;
        float mean_f = sm[1] * K_inv;
        float dir = mean_g / fmaxf(mean_f, 1e-8f);
        if (lp->grad_clip > 0.0f && fabsf(dir) > lp->grad_clip) dir *= lp->grad_clip / fabsf(dir);
        float snr = fabsf(mean_g) / fmaxf(sqrtf(mean_f / K_f), 1e-10f);
        float gate = fminf(snr / 2.0f, 1.0f);
        float alpha = *mu * (1.0f - *rho);
        alpha += eta * gate * dir;
        float mu_new = alpha / fmaxf(1.0f - *rho, 1e-6f);
        *mu = fmaxf(fminf(mu_new, 2.0f), -5.0f);
    }
    if (lp->learn_rho) {
        float mean_g = sm[2] * K_inv;
        float mean_f = sm[3] * K_inv;
        float dir = mean_g / fmaxf(mean_f, 1e-8f);
        if (lp->grad_clip > 0.0f && fabsf(dir) > lp->grad_clip) dir *= lp->grad_clip / fabsf(dir);
        float snr = fabsf(mean_g) / fmaxf(sqrtf(mean_f / K_f), 1e-10f);
        float gate = fminf(snr / 2.0f, 1.0f);
        float rho_new = *rho + eta * gate * dir;
        *rho = fmaxf(fminf(rho_new, 0.999f), 0.001f);
    }
}

/* MixBandConfig and its builder, gpu_bpf/gpu_bpf_ptx_full.cu:257-275 */
typedef struct { int n; int bnd[4]; float k[4]; } BandCfg;

static BandCfg band_build(int n_particles, int n_bands, const float* fracs, const float* scales) {
    BandCfg c;
    memset(&c, 0, sizeof c);
    c.n = n_bands;
    int cum = 0;
    for (int b = 0; b < n_bands; b++) {
        int count = (b < n_bands - 1) ? (int)(fracs[b] * n_particles + 0.5f) : (n_particles - cum);
        cum += count;
        c.bnd[b] = cum;
        c.k[b] = scales[b];
    }
    for (int b = n_bands; b < 4; b++) {
        c.bnd[b] = cum;
        c.k[b] = 1.0f;
    }
    return c;
}

/* ptx_assembly/bpf_kernels_full.ptx:297-324: one band uses k[0] for all */
static inline float band_scale(const BandCfg* c, int i) {
    if (c->n <= 1) return c->k[0];
    if (i < c->bnd[0]) return c->k[0];
    if (i < c->bnd[1]) return c->k[1];
    if (i < c->bnd[2]) return c->k[2];
    return c->k[3];
}

static int bpf_run_impl(const OrcParams* p, int mode, const float* y, const float* mu_t,
                        int n_ticks, OrcTick* out, float* x_out, double* x64_out,
                        int* anc_out, int n_threads, const OrcLearn* lp, double* lsum_out,
                        float* par_out, const OrcBands* bd, int* regime_out, float nu_state,
                        float shrink);

int orc_bpf_run(const OrcParams* p, int mode, const float* y, const float* mu_t,
                int n_ticks, OrcTick* out, float* x_out, double* x64_out,
                int* anc_out, int n_threads) {
    return bpf_run_impl(p, mode, y, mu_t, n_ticks, out, x_out, x64_out, anc_out, n_threads,
                        NULL, NULL, NULL, NULL, NULL, 0.0f, 0.0f);
}

int orc_bpf_run_all(const OrcParams* p, const OrcLearn* lp, const OrcBands* bd, float nu_state,
                    float silverman_shrink, const float* y, int n_ticks, OrcTick* out,
                    float* par_out, int* regime_out, float* x_out, int n_threads) {
    if (bd)
        for (int c = 0; c < 3; c++)
            if (bd->n_bands[c] < 1 || bd->n_bands[c] > 4) return -1;
    return bpf_run_impl(p, ORC_MODE_F32, y, NULL, n_ticks, out, x_out, NULL, NULL, n_threads,
                        lp, NULL, par_out, bd, regime_out, nu_state, silverman_shrink);
}

int orc_bpf_run_ext(const OrcParams* p, float nu_state, float silverman_shrink, const float* y,
                    int n_ticks, OrcTick* out, float* x_out, int n_threads) {
    return bpf_run_impl(p, ORC_MODE_F32, y, NULL, n_ticks, out, x_out, NULL, NULL, n_threads,
                        NULL, NULL, NULL, NULL, NULL, nu_state, silverman_shrink);
}

int orc_bpf_run_bands(const OrcParams* p, const OrcBands* bd, const float* y, int n_ticks,
                      OrcTick* out, float* x_out, int* regime_out, int n_threads) {
    if (!bd) return -1;
    for (int c = 0; c < 3; c++)
        if (bd->n_bands[c] < 1 || bd->n_bands[c] > 4) return -1;
    return bpf_run_impl(p, ORC_MODE_F32, y, NULL, n_ticks, out, x_out, NULL, NULL, n_threads,
                        NULL, NULL, NULL, bd, regime_out, 0.0f, 0.0f);
}

int orc_bpf_run_learn(const OrcParams* p, const OrcLearn* lp, const float* y, int n_ticks,
                      OrcTick* out, double* lsum_out, float* par_out, float* x_out,
                      int n_threads) {
    if (!lp) return -1;
    return bpf_run_impl(p, ORC_MODE_F32, y, NULL, n_ticks, out, x_out, NULL, NULL, n_threads,
                        lp, lsum_out, par_out, NULL, NULL, 0.0f, 0.0f);
}

static int bpf_run_impl(const OrcParams* p, int mode, const float* y, const float* mu_t,
                        int n_ticks, OrcTick* out, float* x_out, double* x64_out,
                        int* anc_out, int n_threads, const OrcLearn* lp, double* lsum_out,
                        float* par_out, const OrcBands* bd, int* regime_out, float nu_state,
                        float shrink) {
    const int n = p->n_particles;
    if (n <= 0 || n_ticks < 0 || p->tile <= 0) return -1;
    /* Student-t state noise: round(nu_state) normals per particle and tick
     * from PCG32 streams of seed * 31337 (the reference's cuRAND generator
     * seed, gpu_bpf_ptx_full.cu:399-404), stream i * nu_int + k, draw t - 1 */
    const int nu_int = nu_state > 0.0f ? (int)(nu_state + 0.5f) : 0;
    if (nu_state > 0.0f && (nu_int < 1 || nu_int > 64)) return -1;
    if ((nu_int > 0 || shrink > 0.0f) && mode != ORC_MODE_F32) return -1;
    const uint64_t chi_seed = (uint64_t)(int64_t)p->seed * 31337ULL;
    uint32_t dnext = 1;     /* next unused draw of the particle streams (draw 0: x_0) */
    if (lp && (mode != ORC_MODE_F32 || mu_t)) return -1;
    BandCfg bcfg[3];
    int regime = 0;
    float surprise = 0.0f, sur_ema = 0.0f;
    if (bd) {
        if (mode != ORC_MODE_F32) return -1;
        for (int c = 0; c < 3; c++)
            bcfg[c] = band_build(bd->n_particles, bd->n_bands[c], bd->fracs[c], bd->scales[c]);
    }
    const int tile = p->tile;
    const int n_tiles = (n + tile - 1) / tile;
#ifdef _OPENMP
    if (n_threads < 1) n_threads = 1;
#else
    (void)n_threads;
#endif
    const int do32 = (mode != ORC_MODE_F64);     /* FP32 restatement runs */
    const int do64 = (mode != ORC_MODE_F32);     /* FP64 oracle runs */
    const int f64 = do64;                        /* statistics from FP64 */

    float* x = (float*)malloc(sizeof(float) * (size_t)n);
    float* l = (float*)malloc(sizeof(float) * (size_t)n);
    double* xd = (double*)malloc(sizeof(double) * (size_t)n);
    double* ld = (double*)malloc(sizeof(double) * (size_t)n);
    double* w = (double*)malloc(sizeof(double) * (size_t)n);
    double* wd = (double*)malloc(sizeof(double) * (size_t)n);
    double* cdf = (double*)malloc(sizeof(double) * (size_t)n);
    int* anc = (int*)malloc(sizeof(int) * (size_t)n);
    float* xt = (float*)malloc(sizeof(float) * (size_t)n);
    double* xdt = (double*)malloc(sizeof(double) * (size_t)n);
    float* tmax = (float*)malloc(sizeof(float) * (size_t)n_tiles);
    double* tfd = (double*)malloc(sizeof(double) * (size_t)n_tiles);
    float* xp = lp ? (float*)malloc(sizeof(float) * (size_t)n) : NULL;
    if (!x || !l || !xd || !ld || !w || !wd || !cdf || !anc || !xt || !xdt || !tmax || !tfd ||
        (lp && !xp)) {
        free(x); free(l); free(xd); free(ld); free(w); free(wd); free(cdf); free(anc);
        free(xt); free(xdt); free(tmax); free(tfd); free(xp);
        return -1;
    }

    float rho = p->rho;
    const float sig = p->sigma_z;
    float mu_cur = p->mu;              /* moved by the learner */
    double lacc[4] = {0.0, 0.0, 0.0, 0.0};
    int grad_count = 0, rm_step = 0;
    const float nu_obs = p->nu_obs;
    const double c_obs_d = nu_obs > 0.0f ? c_obs_f64((double)nu_obs) : 0.0;
    const float c_obs = (float)c_obs_d;
    const float nph = (nu_obs + 1.0f) * 0.5f;

    /* Stationary start: std computed in float as gpu_bpf_ptx_full.cu:430. */
    {
        float mu0 = mu_t ? mu_t[0] : p->mu;
        float sd = sqrtf((sig * sig) / fmaxf(1.0f - rho * rho, 1e-6f));
        double sdd = sqrt(((double)sig * sig) / fmax(1.0 - (double)rho * rho, 1e-6));
        PcgJump j0 = pcg_jump(p->seed, 0);
#pragma omp parallel for num_threads(n_threads) schedule(static)
        for (int i = 0; i < n; i++) {
            float z = normal_j(j0, (uint32_t)i);
            x[i] = fmaf(sd, z, mu0);
            xd[i] = (double)mu0 + sdd * (double)z;
        }
    }

    uint64_t hrng = orc_pcg_host_seed(p->seed);
    int carry = 0, count = 0;
    float m_prev = 0.0f;
    double md_prev = 0.0, s_prev = 1.0, off = 0.0, ll_cum = 0.0;
    const double ll_floor = (double)f32_bits(0x0D6BF94Eu);

    for (int t = 0; t < n_ticks; t++) {
        const float mu = mu_t ? mu_t[t] : mu_cur;
        const float yt = y[t];
        const float ysq = yt * yt;
        /* tick t >= 1 takes the next draw; a Silverman jitter takes one more */
        const PcgJump jt = pcg_jump(p->seed, t > 0 ? dnext : 0u);
        if (t > 0) dnext++;
        PcgJump jc = {0, 0};
        if (nu_int > 0 && t > 0) {
            jc = pcg_jump(1, (uint32_t)(t - 1));
            jc.a *= chi_seed;
        }
        /* the tick's band configuration; the regime switches on the previous
         * tick's surprise before propagating (gpu_bpf_ptx_full.cu:312-334,585-588) */
        const BandCfg* act = NULL;
        if (bd) {
            if (bd->adaptive) {
                if (t > 0) {
                    sur_ema = 0.3f * surprise + (1.0f - 0.3f) * sur_ema;
                    if (surprise < bd->thr_alert) {
                        regime = 0;
                        sur_ema = surprise;
                    } else if (sur_ema >= bd->thr_panic) {
                        regime = 2;
                    } else if (sur_ema >= bd->thr_alert) {
                        regime = 1;
                    } else {
                        regime = 0;
                    }
                }
                act = &bcfg[regime];
            } else {
                act = &bcfg[0];
            }
            if (regime_out) regime_out[t] = bd->adaptive ? regime : 0;
        }

        double M = 0.0;
        if (do32) {
#pragma omp parallel for num_threads(n_threads) schedule(static)
            for (int i = 0; i < n; i++) {
                float xi = x[i];
                if (xp) xp[i] = xi;
                if (t > 0) {
                    float z = normal_j(jt, (uint32_t)i);
                    if (nu_int > 0) {
                        /* ptx_assembly/bpf_kernels_full.ptx:393-406 with IEEE ops */
                        float chi2 = 0.0f;
                        for (int q = 0; q < nu_int; q++) {
                            const uint64_t j = (uint64_t)i * (uint64_t)nu_int + (uint64_t)q;
                            const float g = orc_icdf_f32(orc_uniform(orc_pcg_output(jc.a + (2u * j + 1u) * jc.g)));
                            chi2 = fmaf(g, g, chi2);
                        }
                        z = z * (1.0f / sqrtf(chi2 / nu_state));
                    }
                    float d = xi - mu;
                    d = (g_fault == 2 ? -rho : rho) * d;
                    xi = mu + d;
                    xi = fmaf(act ? sig * band_scale(act, i) : sig, z, xi);
                    x[i] = xi;
                }
                float lw = lw_f32(xi, ysq, nu_obs, c_obs, nph);
                l[i] = (carry && g_fault != 4) ? lw + (l[i] - m_prev) : lw;
            }
            /* tile maxima (fmaxf: NaN ignored), group and global scales */
            float gm = tiled_weights(l, n, tile, tmax, tfd, w, n_threads);
            M = (double)gm;
            m_prev = gm;
        }
        if (do64) {
            const double mud = mu, rhod = rho, sigd = sig;
            double gm = -INFINITY;
#pragma omp parallel for num_threads(n_threads) if (n_threads > 1) schedule(static) reduction(max:gm)
            for (int i = 0; i < n; i++) {
                if (t > 0) {
                    double z = (double)normal_j(jt, (uint32_t)i);
                    xd[i] = mud + (g_fault == 2 ? -rhod : rhod) * (xd[i] - mud) + sigd * z;
                }
                double lw = lw_f64(xd[i], (double)yt, (double)nu_obs, c_obs_d);
                ld[i] = (carry && g_fault != 4) ? lw + (ld[i] - md_prev) : lw;
                if (ld[i] > gm) gm = ld[i];
            }
            double* wt = do32 ? wd : w;
#pragma omp parallel for num_threads(n_threads) if (n_threads > 1) schedule(static)
            for (int i = 0; i < n; i++) wt[i] = (gm == -INFINITY) ? 0.0 : exp(ld[i] - gm);
            M = gm;
            md_prev = gm;
        }
        /* statistics come from FP64 weights when the FP64 oracle runs */
        const double* ws = (do64 && do32) ? wd : w;

        /* FP64 sums in particle order */
        double S = 0.0, SX = 0.0, SXX = 0.0, Q = 0.0;
        if (n_threads <= 1) {
            for (int i = 0; i < n; i++) {
                double xi = f64 ? xd[i] : (double)x[i];
                double wi = ws[i];
                S += wi;
                SX += wi * xi;
                SXX += wi * xi * xi;
                Q += wi * wi;
            }
        } else {
#pragma omp parallel for num_threads(n_threads) schedule(static) reduction(+:S,SX,SXX,Q)
            for (int i = 0; i < n; i++) {
                double xi = f64 ? xd[i] : (double)x[i];
                double wi = ws[i];
                S += wi;
                SX += wi * xi;
                SXX += wi * xi * xi;
                Q += wi * wi;
            }
        }
        /* the coupled mode takes the resampling decision and the CDF from
         * the FP32 weights, so both filters share one genealogy */
        double S32 = S, Q32 = Q;
        if (do64 && do32) {
            S32 = 0.0; Q32 = 0.0;
#pragma omp parallel for num_threads(n_threads) if (n_threads > 1) schedule(static) reduction(+:S32,Q32)
            for (int i = 0; i < n; i++) { S32 += w[i]; Q32 += w[i] * w[i]; }
        }

        OrcTick r;
        memset(&r, 0, sizeof(r));
        r.tick = t;
        int ok = isfinite(S) && S > 0.0 && isfinite(M);
        if (ok) {
            r.mean = SX / S;
            r.var = SXX / S - r.mean * r.mean;
            r.ess = (S * S) / Q;
            r.ll_ref = off + M + log(fmax(S / (double)n, ll_floor));
            r.ll_incr = carry ? (M + log(S)) - log(s_prev) : M + log(S / (double)n);
        } else {
            r.mean = r.var = r.ess = r.ll_ref = r.ll_incr = NAN;
        }
        ll_cum += r.ll_incr;
        r.ll_cum = ll_cum;
        r.h_mean = (float)r.mean;
        r.log_lik = (float)r.ll_ref;

        const double ess_dec = (S32 * S32) / Q32;
        int resample = ok && isfinite(S32) && S32 > 0.0 &&
                       (p->ess_threshold <= 0.0f || (ess_dec / (double)n) < (double)p->ess_threshold);
        r.resampled = resample;
        if (out) out[t] = r;
        /* surprise of this tick for the next (gpu_bpf_ptx_full.cu:767-769) */
        surprise = fabsf(yt) * expf(-r.h_mean * 0.5f);

        if (lp) {
            /* the learner's one-step sums of this tick on the FP32 particles
             * (w = the tiled weights, normalised by their sum) */
            double L[4] = {0.0, 0.0, 0.0, 0.0};
            for (int i = 0; i < n; i++) {
                const double g = (double)learn_dlw(x[i], ysq, nu_obs, nph);
                const float dpf = (t > 0) ? ((g_fault == 7 ? x[i] : xp[i]) - mu) : 0.0f;
                const double d = (double)dpf;
                const double t0 = w[i] * g;
                const double t1 = t0 * g;
                L[0] += t0;
                L[1] += t1;
                L[2] += t0 * d;
                L[3] += (t1 * d) * d;
            }
            for (int q = 0; q < 4; q++) {
                const double v = (isfinite(S) && S > 0.0) ? L[q] / S : 0.0;
                if (lsum_out) lsum_out[4 * (size_t)t + q] = v;
                lacc[q] += v;
            }
            if (++grad_count >= lp->update_K) {
                orc_learn_update(lp, lacc, grad_count, &rm_step, &mu_cur, &rho);
                lacc[0] = lacc[1] = lacc[2] = lacc[3] = 0.0;
                grad_count = 0;
            }
            if (par_out) {
                par_out[2 * (size_t)t] = mu_cur;
                par_out[2 * (size_t)t + 1] = rho;
            }
        }

        if (resample) {
            uint32_t hr = orc_pcg_host_next(&hrng);
            float u = ((float)(hr >> 9) * (1.0f / 8388608.0f)) / (float)n;
            build_cdf(w, n, cdf);
            if (n_threads <= 1) {
                orc_resample_ancestors(cdf, n, u, anc);
            } else {
#pragma omp parallel for num_threads(n_threads) schedule(static)
                for (int i = 0; i < n; i++) {
                    double tt = (double)orc_target(u, i, n);
                    int lo = 0, hi = n - 1;
                    while (lo < hi) {
                        int mid = (lo + hi) >> 1;
                        if (cdf[mid] < tt) lo = mid + 1; else hi = mid;
                    }
                    anc[i] = lo;
                }
            }
            if (do64) {
#pragma omp parallel for num_threads(n_threads) if (n_threads > 1) schedule(static)
                for (int i = 0; i < n; i++) xdt[i] = xd[anc[i]];
                double* sw = xd; xd = xdt; xdt = sw;
            }
            if (do32) {
#pragma omp parallel for num_threads(n_threads) schedule(static)
                for (int i = 0; i < n; i++) xt[i] = x[anc[i]];
                float* sw = x; x = xt; xt = sw;
            }
            if (shrink > 0.0f) {
                /* Silverman jitter, gpu_bpf_ptx_full.cu:708-739 and
                 * ptx_assembly/bpf_kernels_full.ptx:852-1113 */
                double v = 0.0;
                for (int i = 0; i < n; i++) {
                    const double d = (double)(x[i] - r.h_mean);
                    v = fma(d, d, v);
                }
                const float var_sum = (float)v;
                const float sigma_hat = sqrtf(var_sum / (float)n);
                const float bw = shrink * sigma_hat * 1.05922f * powf((float)n, -0.2f);
                const PcgJump jj = pcg_jump(p->seed, dnext);
                dnext++;
                /* the clamp takes mu after this tick's learner update, which
                 * the reference applies before resampling (gpu_bpf_ptx_full.cu:
                 * 660-664 then :708-739) */
                const float mj = mu_t ? mu : mu_cur;
                const float lo = mj - 8.0f, hi = mj + 8.0f;
#pragma omp parallel for num_threads(n_threads) schedule(static)
                for (int i = 0; i < n; i++)
                    x[i] = fminf(fmaxf(fmaf(bw, normal_j(jj, (uint32_t)i), x[i]), lo), hi);
            }
            if (anc_out) memcpy(anc_out, anc, sizeof(int) * (size_t)n);
            count++;
            carry = 0;
            off = 0.0;
        } else {
            carry = 1;
            off += M;
            s_prev = S;
        }
    }

    if (x_out) memcpy(x_out, x, sizeof(float) * (size_t)n);
    if (x64_out) memcpy(x64_out, xd, sizeof(double) * (size_t)n);
    free(x); free(l); free(xd); free(ld); free(w); free(wd); free(cdf); free(anc);
    free(xt); free(xdt); free(tmax); free(tfd); free(xp);
    return count;
}

/* ========================================================================= */
/* Auxiliary particle filter: gpu_bpf/gpu_bpf_ptx_full.cu:990-1047 (kernels)  */
/* and :1102-1170 (tick), restated with this build's streams and arithmetic.  */
/* ========================================================================= */

/* One resampling of x by the tiled weights w (always): the host stream's
 * next offset, the normalised 80-bit CDF, lower_bound of u + i/n. */
static void apf_resample(float* x, float* xt, const double* w, double* cdf, int* anc, int n,
                         uint64_t* hrng, int n_threads) {
    uint32_t hr = orc_pcg_host_next(hrng);
    float u = ((float)(hr >> 9) * (1.0f / 8388608.0f)) / (float)n;
    build_cdf(w, n, cdf);
#pragma omp parallel for num_threads(n_threads) schedule(static)
    for (int i = 0; i < n; i++) {
        double tt = (double)orc_target(u, i, n);
        int lo = 0, hi = n - 1;
        while (lo < hi) {
            int mid = (lo + hi) >> 1;
            if (cdf[mid] < tt) lo = mid + 1; else hi = mid;
        }
        anc[i] = lo;
    }
#pragma omp parallel for num_threads(n_threads) schedule(static)
    for (int i = 0; i < n; i++) xt[i] = x[anc[i]];
    memcpy(x, xt, sizeof(float) * (size_t)n);
}

int orc_apf_run(const OrcParams* p, float nu_state, const float* y, int n_ticks, OrcTick* out,
                float* x_out, int n_threads) {
    const int n = p->n_particles;
    if (n <= 0 || n_ticks < 0 || p->tile <= 0) return -1;
    const int nu_int = nu_state > 0.0f ? (int)(nu_state + 0.5f) : 0;
    if (nu_state > 0.0f && (nu_int < 1 || nu_int > 64)) return -1;
#ifdef _OPENMP
    if (n_threads < 1) n_threads = 1;
#else
    (void)n_threads;
#endif
    const int tile = p->tile;
    const int n_tiles = (n + tile - 1) / tile;
    float* x = (float*)malloc(sizeof(float) * (size_t)n);
    float* xt = (float*)malloc(sizeof(float) * (size_t)n);
    float* l = (float*)malloc(sizeof(float) * (size_t)n);
    double* w = (double*)malloc(sizeof(double) * (size_t)n);
    double* cdf = (double*)malloc(sizeof(double) * (size_t)n);
    int* anc = (int*)malloc(sizeof(int) * (size_t)n);
    float* tmax = (float*)malloc(sizeof(float) * (size_t)n_tiles);
    double* tfd = (double*)malloc(sizeof(double) * (size_t)n_tiles);
    if (!x || !xt || !l || !w || !cdf || !anc || !tmax || !tfd) {
        free(x); free(xt); free(l); free(w); free(cdf); free(anc); free(tmax); free(tfd);
        return -1;
    }
    const float rho = p->rho, sig = p->sigma_z, mu = p->mu, nu_obs = p->nu_obs;
    const float c_obs = nu_obs > 0.0f ? (float)c_obs_f64((double)nu_obs) : 0.0f;
    const float nph = (nu_obs + 1.0f) * 0.5f;
    const uint64_t chi_seed = (uint64_t)(int64_t)p->seed * 31337ULL;
    {
        float sd = sqrtf((sig * sig) / fmaxf(1.0f - rho * rho, 1e-6f));
        PcgJump j0 = pcg_jump(p->seed, 0);
        for (int i = 0; i < n; i++) x[i] = fmaf(sd, normal_j(j0, (uint32_t)i), mu);
    }
    uint64_t hrng = orc_pcg_host_seed(p->seed);
    const double ll_floor = (double)f32_bits(0x0D6BF94Eu);
    double ll_cum = 0.0;
    for (int t = 0; t < n_ticks; t++) {
        const float ysq = y[t] * y[t];
        double ll1 = 0.0;
        if (t > 0) {
            /* stage 1: weights at the predictive means, then resample */
#pragma omp parallel for num_threads(n_threads) schedule(static)
            for (int i = 0; i < n; i++) {
                float d = x[i] - mu;
                d = rho * d;
                l[i] = lw_f32(mu + d, ysq, nu_obs, c_obs, nph);
            }
            double M1 = (double)tiled_weights(l, n, tile, tmax, tfd, w, n_threads);
            long double S1 = 0.0L;
            for (int i = 0; i < n; i++) S1 += (long double)w[i];
            ll1 = M1 + log((double)S1 / (double)n);
            apf_resample(x, xt, w, cdf, anc, n, &hrng, n_threads);
            /* stage 2: propagate (draw t), weight lw(x') - lw(mp) */
            const PcgJump jt = pcg_jump(p->seed, (uint32_t)t);
            PcgJump jc = {0, 0};
            if (nu_int > 0) {
                jc = pcg_jump(1, (uint32_t)(t - 1));
                jc.a *= chi_seed;
            }
#pragma omp parallel for num_threads(n_threads) schedule(static)
            for (int i = 0; i < n; i++) {
                float d = x[i] - mu;
                d = rho * d;
                const float mp = mu + d;
                float z = normal_j(jt, (uint32_t)i);
                if (nu_int > 0) {
                    float chi2 = 0.0f;
                    for (int q = 0; q < nu_int; q++) {
                        const uint64_t j = (uint64_t)i * (uint64_t)nu_int + (uint64_t)q;
                        const float g = orc_icdf_f32(orc_uniform(orc_pcg_output(jc.a + (2u * j + 1u) * jc.g)));
                        chi2 = fmaf(g, g, chi2);
                    }
                    z = z * (1.0f / sqrtf(chi2 / nu_state));
                }
                const float xn = fmaf(sig, z, mp);
                x[i] = xn;
                l[i] = lw_f32(xn, ysq, nu_obs, c_obs, nph) +
                       (g_fault == 8 ? 0.0f : -lw_f32(mp, ysq, nu_obs, c_obs, nph));
            }
        } else {
            for (int i = 0; i < n; i++) l[i] = lw_f32(x[i], ysq, nu_obs, c_obs, nph);
        }
        const double M = (double)tiled_weights(l, n, tile, tmax, tfd, w, n_threads);
        /* 80-bit running sums: the sequential order stays far below the
         * device's hierarchical FP64 error at millions of particles */
        long double Sl = 0.0L, SXl = 0.0L, SXXl = 0.0L, Ql = 0.0L;
        for (int i = 0; i < n; i++) {
            const long double xi = (long double)x[i], wi = (long double)w[i];
            Sl += wi;
            SXl += wi * xi;
            SXXl += wi * xi * xi;
            Ql += wi * wi;
        }
        const double S = (double)Sl, SX = (double)SXl, SXX = (double)SXXl, Q = (double)Ql;
        OrcTick r;
        memset(&r, 0, sizeof r);
        r.tick = t;
        if (isfinite(S) && S > 0.0 && isfinite(M)) {
            r.mean = SX / S;
            r.var = SXX / S - r.mean * r.mean;
            r.ess = (S * S) / Q;
            r.ll_ref = M + log(fmax(S / (double)n, ll_floor));
            r.ll_incr = M + log(S / (double)n) + ll1;
        } else {
            r.mean = r.var = r.ess = r.ll_ref = r.ll_incr = NAN;
        }
        ll_cum += r.ll_incr;
        r.ll_cum = ll_cum;
        r.h_mean = (float)r.mean;
        r.log_lik = (float)r.ll_ref;
        r.resampled = 1;
        if (out) out[t] = r;
        apf_resample(x, xt, w, cdf, anc, n, &hrng, n_threads);
    }
    if (x_out) memcpy(x_out, x, sizeof(float) * (size_t)n);
    free(x); free(xt); free(l); free(w); free(cdf); free(anc); free(tmax); free(tfd);
    return n_ticks;
}

/* ========================================================================= */
/* Point-mass filter: exact log p(y_{1:T}) of the Gaussian SV model with the  */
/* reference's tick convention (tick 0 weights the stationary law, ticks >= 1 */
/* propagate then weight; gpu_bpf/gpu_bpf_ptx_full.cu:591-599).               */
/* ========================================================================= */

double orc_grid_loglik(const float* y, int n_ticks, double rho, double sigma_z,
                       double mu, int n_grid, double* ll_t) {
    if (n_grid < 16) return NAN;
    double sd = sqrt(sigma_z * sigma_z / fmax(1.0 - rho * rho, 1e-6));
    double lo = mu - 10.0 * sd - 4.0 * sigma_z, hi = mu + 10.0 * sd + 4.0 * sigma_z;
    double dx = (hi - lo) / (n_grid - 1);
    double* g = (double*)malloc(sizeof(double) * (size_t)n_grid);
    double* pi = (double*)malloc(sizeof(double) * (size_t)n_grid);
    double* pr = (double*)malloc(sizeof(double) * (size_t)n_grid);
    if (!g || !pi || !pr) { free(g); free(pi); free(pr); return NAN; }
    for (int i = 0; i < n_grid; i++) g[i] = lo + dx * i;
    /* stationary prior, normalised on the grid */
    double z = 0.0;
    for (int i = 0; i < n_grid; i++) {
        double d = (g[i] - mu) / sd;
        pi[i] = exp(-0.5 * d * d);
        z += pi[i];
    }
    for (int i = 0; i < n_grid; i++) pi[i] /= z;
    int band = (int)ceil(9.0 * sigma_z / dx) + 1;
    double total = 0.0;
    for (int t = 0; t < n_ticks; t++) {
        if (t > 0) {
            /* predict: pr[i] = sum_j K(g_i | g_j) pi[j], kernel normalised per j */
            for (int i = 0; i < n_grid; i++) pr[i] = 0.0;
            for (int j = 0; j < n_grid; j++) {
                if (pi[j] == 0.0) continue;
                double m = mu + rho * (g[j] - mu);
                int c = (int)floor((m - lo) / dx);
                int a = c - band < 0 ? 0 : c - band;
                int b = c + band >= n_grid ? n_grid - 1 : c + band;
                double kz = 0.0;
                for (int i = a; i <= b; i++) {
                    double d = (g[i] - m) / sigma_z;
                    kz += exp(-0.5 * d * d);
                }
                if (kz <= 0.0) continue;
                double sc = pi[j] / kz;
                for (int i = a; i <= b; i++) {
                    double d = (g[i] - m) / sigma_z;
                    pr[i] += sc * exp(-0.5 * d * d);
                }
            }
            for (int i = 0; i < n_grid; i++) pi[i] = pr[i];
        }
        double yy = (double)y[t];
        double s = 0.0;
        for (int i = 0; i < n_grid; i++) {
            double lw = -0.91893853320467274178 - 0.5 * yy * yy * exp(-g[i]) - 0.5 * g[i];
            pi[i] *= exp(lw);
            s += pi[i];
        }
        double l = log(s);
        if (ll_t) ll_t[t] = l;
        total += l;
        for (int i = 0; i < n_grid; i++) pi[i] /= s;
    }
    free(g); free(pi); free(pr);
    return total;
}

/* ========================================================================= */
/* Score of the log-likelihood (pathwise, ancestors held fixed).              */
/*                                                                            */
/* LL = sum_t [ logsumexp_i lw(x_t,i, y_t) - log N ] of the always-resample   */
/* filter (the quantity tests/diff_bpf.py:218-222 differentiates).  With the  */
/* ancestors a_t frozen, x_t,i = mu + phi (x_{t-1,a_{t-1,i}} - mu) +          */
/* sigma z_t,i is smooth in theta = (mu, phi, sigma); x_0 = mu + sd z_0 with  */
/* sd = sigma / sqrt(max(1 - phi^2, 1e-6)) (reparameterised when x0_reparam). */
/* The score is carried forward as per-particle tangents D = dx/dtheta:       */
/*   D_t = (1 - phi + phi D~_mu, (x~ - mu) + phi D~_phi, z + phi D~_sigma),   */
/*   dLL += sum_i W_t,i lw'(x_t,i) D_t,i,                                      */
/* D~ = D gathered by the ancestors.  FP64 throughout.                        */
/* ========================================================================= */

static inline double dlw_dx_f64(double x, double y, double nu_obs) {
    double eta2 = y * y * exp(-x);
    if (nu_obs > 0.0) return 0.5 * (nu_obs + 1.0) * eta2 / (nu_obs + eta2) - 0.5;
    return 0.5 * eta2 - 0.5;
}

int orc_bpf_score(const OrcParams* p, const float* y, int n_ticks, int x0_reparam,
                  const double* theta64, const int* anc_in, int* anc_out,
                  double* ll_out, double* score_out, double* ll32_out) {
    const int n = p->n_particles;
    if (n <= 0 || n_ticks <= 0 || p->tile <= 0) return -1;
    const int tile = p->tile;
    const int n_tiles = (n + tile - 1) / tile;
    const int replay = anc_in != NULL;
    const int nt = g_threads;
    /* FP64 parameters: the FP32 ones, or theta64 (finite differences) */
    const double mud = theta64 ? theta64[0] : (double)p->mu;
    const double rhod = theta64 ? theta64[1] : (double)p->rho;
    const double sigd = theta64 ? theta64[2] : (double)p->sigma_z;
    const float rho = p->rho, sig = p->sigma_z, mu = p->mu;
    const float nu_obs = p->nu_obs;
    const double c_obs_d = nu_obs > 0.0f ? c_obs_f64((double)nu_obs) : 0.0;
    const float c_obs = (float)c_obs_d;
    const float nph = (nu_obs + 1.0f) * 0.5f;

    float* x = (float*)malloc(sizeof(float) * (size_t)n);
    float* xt = (float*)malloc(sizeof(float) * (size_t)n);
    float* l = (float*)malloc(sizeof(float) * (size_t)n);
    double* w = (double*)malloc(sizeof(double) * (size_t)n);
    double* cdf = (double*)malloc(sizeof(double) * (size_t)n);
    int* anc = (int*)malloc(sizeof(int) * (size_t)n);
    double* xd = (double*)malloc(sizeof(double) * (size_t)n);
    double* xdt = (double*)malloc(sizeof(double) * (size_t)n);
    double* D = (double*)malloc(sizeof(double) * 3 * (size_t)n);
    double* Dt = (double*)malloc(sizeof(double) * 3 * (size_t)n);
    double* lw = (double*)malloc(sizeof(double) * (size_t)n);
    float* tmax = (float*)malloc(sizeof(float) * (size_t)n_tiles);
    double* tfd = (double*)malloc(sizeof(double) * (size_t)n_tiles);
    if (!x || !xt || !l || !w || !cdf || !anc || !xd || !xdt || !D || !Dt || !lw || !tmax || !tfd) {
        free(x); free(xt); free(l); free(w); free(cdf); free(anc); free(xd); free(xdt);
        free(D); free(Dt); free(lw); free(tmax); free(tfd);
        return -1;
    }

    /* initial state */
    {
        float sd = sqrtf((sig * sig) / fmaxf(1.0f - rho * rho, 1e-6f));
        double om = 1.0 - rhod * rhod;
        double sdd, dsd_dphi, dsd_dsig;
        if (om > 1e-6) {
            sdd = sigd / sqrt(om);
            dsd_dsig = 1.0 / sqrt(om);
            dsd_dphi = sigd * rhod / (om * sqrt(om));
        } else {
            sdd = sigd / sqrt(1e-6);
            dsd_dsig = 1.0 / sqrt(1e-6);
            dsd_dphi = 0.0;
        }
        PcgJump j0 = pcg_jump(p->seed, 0);
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static)
        for (int i = 0; i < n; i++) {
            float z = normal_j(j0, (uint32_t)i);
            x[i] = fmaf(sd, z, mu);
            xd[i] = mud + sdd * (double)z;
            D[3 * i + 0] = x0_reparam ? 1.0 : 0.0;
            D[3 * i + 1] = x0_reparam ? (double)z * dsd_dphi : 0.0;
            D[3 * i + 2] = x0_reparam ? (double)z * dsd_dsig : 0.0;
        }
    }

    uint64_t hrng = orc_pcg_host_seed(p->seed);
    double ll = 0.0, ll32 = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0;
    for (int t = 0; t < n_ticks; t++) {
        const float yt = y[t];
        const float ysq = yt * yt;
        const PcgJump jt = pcg_jump(p->seed, (uint32_t)t);
        /* propagate (FP32 restatement for the genealogy, FP64 + tangents) */
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static)
        for (int i = 0; i < n; i++) {
            if (t > 0) {
                float z = normal_j(jt, (uint32_t)i);
                if (!replay) {
                    float d = x[i] - mu;
                    d = rho * d;
                    float xi = mu + d;
                    x[i] = fmaf(sig, z, xi);
                }
                double pre = xd[i];
                double* Di = &D[3 * i];
                xd[i] = mud + rhod * (pre - mud) + sigd * (double)z;
                Di[0] = (1.0 - rhod) + rhod * Di[0];
                Di[1] = (pre - mud) + rhod * Di[1];
                Di[2] = (double)z + rhod * Di[2];
            }
        }
        /* FP64 weights, log-likelihood and score increment */
        double M = -INFINITY;
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static) reduction(max:M)
        for (int i = 0; i < n; i++) {
            lw[i] = lw_f64(xd[i], (double)yt, (double)nu_obs, c_obs_d);
            if (lw[i] > M) M = lw[i];
        }
        double S = 0.0;
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static) reduction(+:S)
        for (int i = 0; i < n; i++) S += exp(lw[i] - M);
        ll += M + log(S / (double)n);
        double a0 = 0.0, a1 = 0.0, a2 = 0.0;
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static) reduction(+:a0,a1,a2)
        for (int i = 0; i < n; i++) {
            double Wg = exp(lw[i] - M) / S * dlw_dx_f64(xd[i], (double)yt, (double)nu_obs);
            a0 += Wg * D[3 * i + 0];
            a1 += Wg * D[3 * i + 1];
            a2 += Wg * D[3 * i + 2];
        }
        g0 += a0; g1 += a1; g2 += a2;
        /* ancestors: replayed, or from the FP32 restatement (as the GPU) */
        const int* at;
        if (replay) {
            at = anc_in + (size_t)t * (size_t)n;
        } else {
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static)
            for (int i = 0; i < n; i++) l[i] = lw_f32(x[i], ysq, nu_obs, c_obs, nph);
            float gm = tiled_weights(l, n, tile, tmax, tfd, w, nt);
            double S32 = 0.0;
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static) reduction(+:S32)
            for (int i = 0; i < n; i++) S32 += w[i];
            ll32 += (double)gm + log(S32 / (double)n);
            uint32_t hr = orc_pcg_host_next(&hrng);
            float u = ((float)(hr >> 9) * (1.0f / 8388608.0f)) / (float)n;
            build_cdf(w, n, cdf);
            if (nt > 1) {
#pragma omp parallel for num_threads(nt) schedule(static)
                for (int i = 0; i < n; i++) {
                    const double tt = (double)orc_target(u, i, n);
                    int lo = 0, hi = n - 1;
                    while (lo < hi) {
                        const int mid = (lo + hi) >> 1;
                        if (g_fault == 3 ? (cdf[mid] <= tt) : (cdf[mid] < tt)) lo = mid + 1; else hi = mid;
                    }
                    anc[i] = lo;
                }
            } else {
                orc_resample_ancestors(cdf, n, u, anc);
            }
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static)
            for (int i = 0; i < n; i++) xt[i] = x[anc[i]];
            float* sw = x; x = xt; xt = sw;
            at = anc;
        }
        if (anc_out) memcpy(anc_out + (size_t)t * (size_t)n, at, sizeof(int) * (size_t)n);
#pragma omp parallel for num_threads(nt) if (nt > 1) schedule(static)
        for (int i = 0; i < n; i++) {
            int a = at[i];
            xdt[i] = xd[a];
            Dt[3 * i + 0] = D[3 * a + 0];
            Dt[3 * i + 1] = D[3 * a + 1];
            Dt[3 * i + 2] = D[3 * a + 2];
        }
        double* s1 = xd; xd = xdt; xdt = s1;
        double* s2 = D; D = Dt; Dt = s2;
    }
    if (ll_out) *ll_out = ll;
    if (ll32_out) *ll32_out = ll32;
    if (score_out) { score_out[0] = g0; score_out[1] = g1; score_out[2] = g2; }
    free(x); free(xt); free(l); free(w); free(cdf); free(anc); free(xd); free(xdt);
    free(D); free(Dt); free(lw); free(tmax); free(tfd);
    return 0;
}

/* ========================================================================= */
/* Helpers shared with smc2_oracle.c (the batched inner filters' restatement) */
/* ========================================================================= */

void orc_pcg_jump(int seed, uint32_t k, uint64_t* a, uint64_t* g) {
    PcgJump j = pcg_jump(seed, k);
    *a = j.a;
    *g = j.g;
}

float orc_lw(float x, float ysq, float nu_obs, float c_obs, float nph) {
    return lw_f32(x, ysq, nu_obs, c_obs, nph);
}

float orc_tiled_weights(const float* l, int n, int tile, float* tmax, double* tf, double* w) {
    return tiled_weights(l, n, tile, tmax, tf, w, 1);
}

void orc_build_cdf(const double* w, int n, double* cdf) { build_cdf(w, n, cdf); }

double orc_c_obs(double nu) { return c_obs_f64(nu); }

/* ========================================================================= */
/* One tick from a given state (plain filter), for tick-by-tick checks of    */
/* the GPU at sizes where whole-run parity would be broken by a near-tie     */
/* ancestor (gpu_bpf/gpu_bpf_ptx_full.cu:591-748 for one tick).               */
/* ========================================================================= */

int orc_bpf_tick(const OrcParams* p, float mu, int t, float y, int carry, uint64_t* hrng, float* x,
                 float* l, float* x_pre, int* anc, OrcTick* out, double* ll64, int n_threads) {
    const int n = p->n_particles, tile = p->tile;
    if (n <= 0 || tile <= 0 || t < 0 || !hrng || !x || !l || !x_pre || !anc || !out) return -1;
#ifdef _OPENMP
    if (n_threads < 1) n_threads = 1;
#else
    (void)n_threads;
#endif
    const int n_tiles = (n + tile - 1) / tile;
    float* tmax = (float*)malloc(sizeof(float) * (size_t)n_tiles);
    double* tf = (double*)malloc(sizeof(double) * (size_t)n_tiles);
    double* w = (double*)malloc(sizeof(double) * (size_t)n);
    double* cdf = (double*)malloc(sizeof(double) * (size_t)n);
    if (!tmax || !tf || !w || !cdf) {
        free(tmax); free(tf); free(w); free(cdf);
        return -1;
    }
    const float rho = p->rho, sig = p->sigma_z, nu_obs = p->nu_obs;
    const double c_obs_d = nu_obs > 0.0f ? c_obs_f64((double)nu_obs) : 0.0;
    const float c_obs = (float)c_obs_d, nph = (nu_obs + 1.0f) * 0.5f;
    const float ysq = y * y;
    /* carried weights: m_prev = their max, s_prev = their tiled FP64 sum */
    float m_prev = 0.0f;
    double s_prev = 1.0, s_prev64 = 1.0;
    if (carry) {
        m_prev = tiled_weights(l, n, tile, tmax, tf, w, n_threads);
        double a = 0.0, b = 0.0;
#pragma omp parallel for num_threads(n_threads) schedule(static) reduction(+:a,b)
        for (int i = 0; i < n; i++) {
            a += w[i];
            b += exp((double)l[i] - (double)m_prev);
        }
        s_prev = a;
        s_prev64 = b;
    }
    const PcgJump jt = pcg_jump(p->seed, (uint32_t)t);
    double gm64 = -INFINITY;
#pragma omp parallel for num_threads(n_threads) schedule(static) reduction(max:gm64)
    for (int i = 0; i < n; i++) {
        float xi = x[i];
        double xd = (double)xi;
        if (t > 0) {
            const float z = normal_j(jt, (uint32_t)i);
            float d = xi - mu;
            d = rho * d;
            xi = mu + d;
            xi = fmaf(sig, z, xi);
            xd = (double)mu + (double)rho * (xd - (double)mu) + (double)sig * (double)z;
        }
        x_pre[i] = xi;
        const float lw = lw_f32(xi, ysq, nu_obs, c_obs, nph);
        const float lprev = l[i];
        l[i] = carry ? lw + (lprev - m_prev) : lw;
        /* FP64 log-weight of the same state and noise, carried on the FP32 weights */
        double l64 = lw_f64(xd, (double)y, (double)nu_obs, c_obs_d);
        if (carry) l64 += (double)lprev - (double)m_prev;
        w[i] = l64;           /* staged: FP64 log-weights */
        if (l64 > gm64) gm64 = l64;
    }
    double S64 = 0.0;
#pragma omp parallel for num_threads(n_threads) schedule(static) reduction(+:S64)
    for (int i = 0; i < n; i++) S64 += exp(w[i] - gm64);
    if (ll64) *ll64 = carry ? (gm64 + log(S64)) - log(s_prev64) : gm64 + log(S64 / (double)n);
    const float gm = tiled_weights(l, n, tile, tmax, tf, w, n_threads);
    const double M = (double)gm;
    double S = 0.0, SX = 0.0, SXX = 0.0, Q = 0.0;
#pragma omp parallel for num_threads(n_threads) schedule(static) reduction(+:S,SX,SXX,Q)
    for (int i = 0; i < n; i++) {
        const double xi = (double)x_pre[i], wi = w[i];
        S += wi;
        SX += wi * xi;
        SXX += wi * xi * xi;
        Q += wi * wi;
    }
    OrcTick r;
    memset(&r, 0, sizeof r);
    r.tick = t;
    const int ok = isfinite(S) && S > 0.0 && isfinite(M);
    if (ok) {
        r.mean = SX / S;
        r.var = SXX / S - r.mean * r.mean;
        r.ess = (S * S) / Q;
        r.ll_incr = carry ? (M + log(S)) - log(s_prev) : M + log(S / (double)n);
    } else {
        r.mean = r.var = r.ess = r.ll_incr = NAN;
    }
    r.h_mean = (float)r.mean;
    const int resample = ok && (p->ess_threshold <= 0.0f || (r.ess / (double)n) < (double)p->ess_threshold);
    r.resampled = resample;
    if (resample) {
        const uint32_t hr = orc_pcg_host_next(hrng);
        const float u = ((float)(hr >> 9) * (1.0f / 8388608.0f)) / (float)n;
        build_cdf(w, n, cdf);
#pragma omp parallel for num_threads(n_threads) schedule(static)
        for (int i = 0; i < n; i++) {
            const double tt = (double)orc_target(u, i, n);
            int lo = 0, hi = n - 1;
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (cdf[mid] < tt) lo = mid + 1; else hi = mid;
            }
            anc[i] = lo;
            x[i] = x_pre[lo];
        }
    } else {
#pragma omp parallel for num_threads(n_threads) schedule(static)
        for (int i = 0; i < n; i++) {
            anc[i] = i;
            x[i] = x_pre[i];
        }
    }
    *out = r;
    free(tmax); free(tf); free(w); free(cdf);
    return resample;
}
