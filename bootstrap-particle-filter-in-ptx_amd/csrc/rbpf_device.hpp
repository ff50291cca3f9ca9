// rbpf_device.hpp -- per-particle arithmetic of the Rao-Blackwellised inner
// filter (regime particle z~, Kalman moments of h, OCSN observation mixture)
// behind include/smc2_rbpf.h.  As in bpf_device.hpp every operation is an
// IEEE +, -, *, /, sqrt, an explicit fmaf, or integer work, compiled with
// -ffp-contract=off, so the CPU oracle repeats each result bit for bit.
#pragma once

#include "bpf_device.hpp"

namespace bpf {
namespace rb {

constexpr int kOcsnK = 10;
constexpr int kSortEvery = 4;        // SORT_EVERY_K, smc2/smc2_rbpf_batch.cuh:116
constexpr int kParams = 4;

// OCSN mixture of log chi^2_1 (Omori, Chib, Shephard, Nakajima 2007, Table 1;
// the values of smc2/smc2_rbpf_batch.cu:31-42): weights, means, variances.
// The log-weights the kernels use include -log(2 pi)/2, so the mixture
// log-likelihood is a proper log-density of y = log r^2.
constexpr float kOcsnP[kOcsnK] = {0.00609f, 0.04775f, 0.13057f, 0.20674f, 0.22715f,
                                  0.18842f, 0.12047f, 0.05591f, 0.01575f, 0.00115f};
constexpr float kOcsnM[kOcsnK] = {1.92677f,  1.34744f,  0.73504f,  0.02266f,  -0.85173f,
                                  -1.97278f, -3.46788f, -5.55246f, -8.68384f, -14.65000f};
constexpr float kOcsnV[kOcsnK] = {0.11265f, 0.17788f, 0.26768f, 0.40611f, 0.62699f,
                                  0.98583f, 1.57469f, 2.54498f, 4.16591f, 7.33342f};
// (float)(log((double)kOcsnP[k]) - log(2 pi) / 2), exactly (the host checks
// its libm against these at set-up; compile-time constants keep the replay
// loop's scalar registers free)
constexpr float kOcsnLogW[kOcsnK] = {-0x1.81486ep+2f, -0x1.faf8b4p+1f, -0x1.7a3660p+1f, -0x1.3f63c2p+1f,
                                     -0x1.3356b2p+1f, -0x1.4b4440p+1f, -0x1.84847cp+1f, -0x1.e6c716p+1f,
                                     -0x1.44787ap+2f, -0x1.ebf6b0p+2f};

// Fixed part of the model, passed by value to every kernel.
struct Model {
    float tb, ts, tr;          // theta(z) curve
    float ms, mr, ss, sr;      // shapes of mu(z) and sigma_h(z)
    float lw0;                 // -log(N_inner), the log-weight after a resample
};

// ---- counter-based random numbers --------------------------------------
// splitmix64 finaliser (Steele, Lea, Flood 2014).
__device__ __host__ inline uint64_t sm64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
// Key of a stream: key(parent, coordinate).
__device__ __host__ inline uint64_t key(uint64_t a, uint64_t b) { return sm64(a ^ sm64(b)); }
__device__ inline uint32_t bits(uint64_t k, uint32_t ctr) {
    return (uint32_t)(sm64(k + (uint64_t)ctr * 0xD1B54A32D192ED03ull) >> 32);
}
__device__ inline float uniform(uint64_t k, uint32_t ctr) { return pcg_uniform(bits(k, ctr)); }
__device__ inline float normal(uint64_t k, uint32_t ctr) { return icdf(uniform(k, ctr)); }

// Stream purposes (first coordinate under the seed).
enum Purpose : uint64_t { kPrior = 1, kZ = 2, kU0 = 3, kMove = 4, kFreshZ = 5, kFreshU = 6 };

// ---- model functions ------------------------------------------------------
// tanh from one exp: sign(x) (1 - 2 / (e^{2|x|} + 1)); absolute error below
// 2^-23 (the subtraction near 0 costs relative, not absolute, accuracy).
__device__ inline float tanh_rb(float x) {
    float ax = __builtin_fabsf(x);
    float e = expf_cw(ax + ax);
    float r = 1.0f - 2.0f / (e + 1.0f);
    return __builtin_copysignf(r, x);
}

// z = 1.5 (1 + tanh z~) (smc2/smc2_rbpf_batch.cuh:392-394).
__device__ inline float z_of(float zt) { return 1.5f * (1.0f + tanh_rb(zt)); }

// f(z) = base + scale (1 - exp(-rate z)) (smc2/smc2_rbpf_batch.cuh:388-390).
__device__ inline float curve(float base, float scale, float rate, float z) {
    return base + scale * (1.0f - expf_cw(-(rate * z)));
}

// erfc after Numerical Recipes' erfcc (Press et al., 2nd ed., section 6.2):
// Chebyshev-fitted exponent, fractional error below 1.2e-7.
__device__ inline float erfc_nr(float x) {
    float z = __builtin_fabsf(x);
    float t = 1.0f / (1.0f + 0.5f * z);
    float p = 0.17087277f;
    p = __builtin_fmaf(t, p, -0.82215223f);
    p = __builtin_fmaf(t, p, 1.48851587f);
    p = __builtin_fmaf(t, p, -1.13520398f);
    p = __builtin_fmaf(t, p, 0.27886807f);
    p = __builtin_fmaf(t, p, -0.18628806f);
    p = __builtin_fmaf(t, p, 0.09678418f);
    p = __builtin_fmaf(t, p, 0.37409196f);
    p = __builtin_fmaf(t, p, 1.00002368f);
    p = __builtin_fmaf(t, p, -1.26551223f);
    float ans = t * expf_cw(p - z * z);
    return x >= 0.0f ? ans : 2.0f - ans;
}

// Systematic-resampling offset from its stored normal: Phi(n) clamped to
// [1e-7, 1 - 1e-7] (u0_from_noise, smc2/smc2_rbpf_batch.cuh:402-405).
__device__ inline float u0_of(float n) {
    float u = 0.5f * erfc_nr(-(n * 0.70710678f));
    return fminf(fmaxf(u, 1e-7f), 1.0f - 1e-7f);
}

// Stationary initial state of one inner particle
// (smc2/smc2_rbpf_batch.cu:208-229): z~ ~ N(0, sigma_z^2 / (1 - rho^2)),
// h ~ N(mu(z), sigma_h(z)^2 / (1 - phi^2)).
__device__ inline void init_particle(float rho, float sz, float mub, float sgb, float n0,
                                     const Model& md, float& zt, float& mu, float& var) {
    float omr = fmaxf(1.0f - rho * rho, 1e-6f);
    float sstd = sz / __builtin_sqrtf(omr);
    zt = sstd * n0;
    float z = z_of(zt);
    float th = curve(md.tb, md.ts, md.tr, z);
    float muz = curve(mub, md.ms, md.mr, z);
    float sh = curve(sgb, md.ss, md.sr, z);
    float phi = 1.0f - th;
    mu = muz;
    var = (sh * sh) / fmaxf(1.0f - phi * phi, 1e-6f);
}

// Regime transition and Kalman prediction (smc2/smc2_rbpf_batch.cu:358-370).
__device__ inline void predict(float rho, float sz, float mub, float sgb, float zn, const Model& md,
                               float& zt, float mu, float var, float& mp, float& vp) {
    float a = rho * zt;
    float b = sz * zn;
    zt = a + b;
    float z = z_of(zt);
    float th = curve(md.tb, md.ts, md.tr, z);
    float muz = curve(mub, md.ms, md.mr, z);
    float sh = curve(sgb, md.ss, md.sr, z);
    float phi = 1.0f - th;
    mp = phi * mu + th * muz;
    vp = fmaxf(phi * phi * var + sh * sh, 1e-8f);
}

// OCSN update (smc2/smc2_rbpf_batch.cu:64-122): component log-likelihoods
// log p_k N(y; mp + m_k, vp + v_k), their log-sum, and the moment-matched
// posterior mean and variance of h.
__device__ inline void ocsn_update(float y, float mp, float vp, const Model& md, float& mu_post,
                                   float& var_post, float& lnorm) {
    const float v = fmaxf(vp, 1e-6f);
    // per component: log-likelihood, posterior mean m_k and second moment
    // var_k + m_k^2 (the same expressions as the reference's second loop,
    // formed here so only three values per component stay live)
    float val[kOcsnK], mk[kOcsnK], s2[kOcsnK];
    float mx = -__builtin_inff();
    const float d = y - mp;
#pragma unroll
    for (int k = 0; k < kOcsnK; ++k) {
        const float S = v + kOcsnV[k];
        const float inv = 1.0f / S;
        const float inn = d - kOcsnM[k];
        const float q = inn * inn * inv;
        val[k] = kOcsnLogW[k] - 0.5f * (logf_cw(S) + q);
        mx = fmaxf(mx, val[k]);
        const float K = v * inv;
        mk[k] = mp + K * inn;
        const float vk = (1.0f - K) * v;
        s2[k] = vk + mk[k] * mk[k];
    }
    // component weights e_k / sum e, e_k = exp(val_k - max) (the reference
    // recomputes exp(val_k - lnorm); the quotient is the same weight)
    float se = 0.0f;
#pragma unroll
    for (int k = 0; k < kOcsnK; ++k) {
        val[k] = expf_cw(val[k] - mx);
        se += val[k];
    }
    lnorm = mx + logf_cw(se);
    const float rse = 1.0f / se;
    float mo = 0.0f, e2 = 0.0f;
#pragma unroll
    for (int k = 0; k < kOcsnK; ++k) {
        const float w = val[k] * rse;
        mo += w * mk[k];
        e2 += w * s2[k];
    }
    mu_post = mo;
    var_post = fmaxf(e2 - mo * mo, 1e-6f);
}

// Order-preserving key of a float (total order, NaN above +inf).
__device__ inline uint32_t sort_key(float f) {
    uint32_t b = f32_bits(f);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

}  // namespace rb
}  // namespace bpf
