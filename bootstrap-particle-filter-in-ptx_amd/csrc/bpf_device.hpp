// bpf_device.hpp -- per-particle arithmetic of the SV bootstrap filter for
// gfx950.  Every operation is an IEEE correctly rounded +, -, *, /, sqrt, an
// explicit fmaf, or integer/bit work, and the file is compiled with
// -ffp-contract=off, so each result is reproducible bit for bit by any IEEE
// CPU that performs the same sequence.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bpf {

constexpr int kBlock = 256;          // threads per block of the per-tile kernels
constexpr int kWave = 64;            // CDNA wavefront width
constexpr uint64_t kPcgMult = 6364136223846793005ULL;
constexpr uint64_t kPcgHostInc = 1442695040888963407ULL;

__device__ __host__ inline float bits_f32(uint32_t b) { return __builtin_bit_cast(float, b); }
__device__ __host__ inline uint32_t f32_bits(float f) { return __builtin_bit_cast(uint32_t, f); }

// PCG32 XSH-RR output of the pre-step state
// (ptx_assembly/bpf_kernels_full.ptx:146-157).
__device__ __host__ inline uint32_t pcg_output(uint64_t old) {
    uint32_t xsh = (uint32_t)(((old >> 18u) ^ old) >> 27u);
    uint32_t rot = (uint32_t)(old >> 59u);
    return (xsh >> rot) | (xsh << ((0u - rot) & 31u));
}

// Uniform of ptx:159-164: (r >> 9) * 2^-23 clamped to [1e-6, 0x3F7FFBCE].
// u is never NaN, so one median-of-three equals the max-then-min clamp.
__device__ inline float pcg_uniform(uint32_t r) {
    float u = (float)(r >> 9) * (1.0f / 8388608.0f);
    return __builtin_amdgcn_fmed3f(u, bits_f32(0x358637BDu), bits_f32(0x3F7FFBCEu));
}

// The same draw as q = u - 1/2 before the clamp, in two instructions: the
// 23 bits r >> 9 under the exponent of 1.0 form the float 1 + (r >> 9) 2^-23
// exactly (one funnel shift), and subtracting 1.5 from it is exact.  The
// clamp of pcg_uniform moves only uniforms with |u - 1/2| > 0.4999, all in
// the inverse CDF's tails, so a central draw's q equals pcg_uniform(r) - 1/2
// bit for bit, and pcg_uniform(r) = uniform_from_half(q) for every r.
__device__ inline float pcg_half(uint32_t r) {
    return bits_f32(__builtin_amdgcn_alignbit(0x7Fu, r, 9u)) - 1.5f;
}
__device__ inline float uniform_from_half(float q) {
    return __builtin_amdgcn_fmed3f(q + 0.5f, bits_f32(0x358637BDu), bits_f32(0x3F7FFBCEu));
}

// exp(x) after S. L. Moshier's Cephes expf: Cody-Waite reduction by ln 2
// (C1 = 0.693359375, C2 = -2.12194440e-4), Cephes' degree-5 polynomial P,
// exponent added in the bits.  x < -86 gives 0 (no subnormal results), x > 88 gives
// +inf, NaN propagates.  The two selects at the end decide every x outside
// [-86, 88], so the reduction runs unclamped: within the range the integer
// part z = rint(x log2 e) is found by adding and subtracting 1.5 * 2^23
// (round to nearest even, exact for |z| < 2^22) and read back from the bits
// of the sum, and 2^z is applied by ldexp, exact for these normal results;
// outside the range the discarded lanes compute finite or NaN values
// without any undefined conversion, and a NaN x stays NaN through ldexp.
// The reduction's product and shift are one fused multiply-add (z is the
// integer nearest the exact x log2 e), and the polynomial closes in Horner
// form, 1 + r (1 + r P(r)).
__device__ inline float expf_cw_core(float x, float& zs) {
    constexpr float kShift = 12582912.0f;        // 1.5 * 2^23
    zs = __builtin_fmaf(x, 1.44269504088896341f, kShift);
    float z = zs - kShift;
    float r = __builtin_fmaf(z, -0.693359375f, x);
    r = __builtin_fmaf(z, 2.12194440e-4f, r);
    float p = 1.9875691500e-4f;
    p = __builtin_fmaf(p, r, 1.3981999507e-3f);
    p = __builtin_fmaf(p, r, 8.3334519073e-3f);
    p = __builtin_fmaf(p, r, 4.1665795894e-2f);
    p = __builtin_fmaf(p, r, 1.6666665459e-1f);
    p = __builtin_fmaf(p, r, 5.0000001201e-1f);
    p = __builtin_fmaf(p, r, 1.0f);
    p = __builtin_fmaf(p, r, 1.0f);
    return __builtin_amdgcn_ldexpf(p, (int)(f32_bits(zs) - 0x4B400000u));
}

__device__ inline float expf_cw(float x) {
    float zs;
    float res = expf_cw_core(x, zs);
    res = (x > 88.0f) ? __builtin_inff() : res;
    res = (x < -86.0f) ? 0.0f : res;       // NaN fails both tests and stays NaN
    return res;
}

// expf_cw for arguments that cannot exceed 0 (a log-weight minus a maximum
// taken over the same values, or -inf, or NaN): the overflow select of
// expf_cw can never fire, so the results are those of expf_cw.
__device__ inline float expf_cw_nonpos(float x) {
    float zs;
    const float res = expf_cw_core(x, zs);
    return (x < -86.0f) ? 0.0f : res;
}

// log(x) after Moshier's Cephes logf: x = m 2^e with m in [sqrt(1/2),
// sqrt(2)), degree-8 polynomial in m-1, e*ln2 added in two parts.
__device__ inline float logf_cw(float x) {
    uint32_t b = f32_bits(x);
    int sub = (b >> 23) == 0;                         // zero or subnormal
    uint32_t bs = sub ? f32_bits(x * 33554432.0f) : b;
    int e = (int)(bs >> 23) - 126 - (sub ? 25 : 0);
    float m = bits_f32((bs & 0x007FFFFFu) | 0x3F000000u);
    int lo = m < 0.707106781186547524f;
    e -= lo;
    m = lo ? (m + m - 1.0f) : (m - 1.0f);
    float z = m * m;
    float y = 7.0376836292e-2f;
    y = __builtin_fmaf(y, m, -1.1514610310e-1f);
    y = __builtin_fmaf(y, m, 1.1676998740e-1f);
    y = __builtin_fmaf(y, m, -1.2420140846e-1f);
    y = __builtin_fmaf(y, m, 1.4249322787e-1f);
    y = __builtin_fmaf(y, m, -1.6668057665e-1f);
    y = __builtin_fmaf(y, m, 2.0000714765e-1f);
    y = __builtin_fmaf(y, m, -2.4999993993e-1f);
    y = __builtin_fmaf(y, m, 3.3333331174e-1f);
    y = (y * m) * z;
    float fe = (float)e;
    y = __builtin_fmaf(fe, -2.12194440e-4f, y);
    y = __builtin_fmaf(-0.5f, z, y);
    float r = m + y;
    r = __builtin_fmaf(fe, 0.693359375f, r);
    // special values
    r = (x == __builtin_inff()) ? x : r;
    r = (x == 0.0f) ? -__builtin_inff() : r;
    r = (x < 0.0f || x != x) ? __builtin_nanf("") : r;
    return r;
}

// Correctly rounded n / d for operands whose quotient needs no scaling:
// d in [2^-60, 2^60], |n| in {0} U [2^-60, 2^60], both finite.  This is the
// reciprocal refinement and two Markstein corrections of the IEEE division
// sequence the compiler emits for '/', without its operand scaling and its
// special-value fix-up, which are the identity for such operands.  The
// central inverse-CDF rational (|n| <= 0.3, d in [0.0026, 1]) qualifies; the
// device inverse CDF is checked against IEEE division for every uniform the
// generator can produce.
__device__ inline float div_rn_normal(float n, float d) {
    float y = __builtin_amdgcn_rcpf(d);
    float e = __builtin_fmaf(-d, y, 1.0f);
    y = __builtin_fmaf(e, y, y);
    float q = n * y;
    float rem = __builtin_fmaf(-d, q, n);
    q = __builtin_fmaf(rem, y, q);
    rem = __builtin_fmaf(-d, q, n);
    return __builtin_fmaf(rem, y, q);
}

// n / d for the central inverse-CDF rational: the reciprocal refined once
// and one Markstein correction.  It is the correctly rounded quotient for
// every central draw the generator can produce (all 2^23 uniforms are
// checked against IEEE division on the device); div_rn_normal adds the
// second correction that makes it so for any operands in its range.
__device__ inline float div_central(float n, float d) {
    float y = __builtin_amdgcn_rcpf(d);
    float e = __builtin_fmaf(-d, y, 1.0f);
    y = __builtin_fmaf(e, y, y);
    float q = n * y;
    float rem = __builtin_fmaf(-d, q, n);
    return __builtin_fmaf(rem, y, q);
}

// Acklam (2003) inverse normal CDF with the FP32 coefficients of ptx:172-203.
// Central region |u - 1/2| < 0.47575 (0x3EF39581, ptx:169): rational in
// r = q^2.  Valid for every q; callers select it only for central lanes.
__device__ inline float icdf_central(float q) {
    float r = q * q;
    float num = __builtin_fmaf(r, -39.69683028665376f, 220.9460984245205f);
    num = __builtin_fmaf(r, num, -275.9285104469687f);
    num = __builtin_fmaf(r, num, 138.3577518672690f);
    num = __builtin_fmaf(r, num, -30.66479806614716f);
    num = __builtin_fmaf(r, num, 2.506628277459239f);
    num = q * num;
    float den = __builtin_fmaf(r, -54.47609879822406f, 161.5858368580409f);
    den = __builtin_fmaf(r, den, -155.6989798598866f);
    den = __builtin_fmaf(r, den, 66.80131188771972f);
    den = __builtin_fmaf(r, den, -13.28068155288572f);
    den = __builtin_fmaf(r, den, 1.0f);
    return div_central(num, den);
}

__device__ inline bool icdf_is_central(float u) {
    return __builtin_fabsf(u - 0.5f) < bits_f32(0x3EF39581u);
}
__device__ inline bool icdf_is_central_half(float q) {
    return __builtin_fabsf(q) < bits_f32(0x3EF39581u);
}

// Tails (ptx:185-205): t = sqrt(-2 log p), p = min(u, 1-u), negated above 1/2.
__device__ inline float icdf_tail(float u) {
    int low = u < 0.5f;
    float p = low ? u : 1.0f - u;
    float t = __builtin_sqrtf(-2.0f * logf_cw(p));
    float num = __builtin_fmaf(t, -7.784894002430293e-03f, -3.223964580411365e-01f);
    num = __builtin_fmaf(t, num, -2.400758277161838e+00f);
    num = __builtin_fmaf(t, num, -2.549732539343734e+00f);
    num = __builtin_fmaf(t, num, 4.374664141464968e+00f);
    num = __builtin_fmaf(t, num, 2.938163982698783e+00f);
    float den = __builtin_fmaf(t, 7.784695709041462e-03f, 3.224671290700398e-01f);
    den = __builtin_fmaf(t, den, 2.445134137142996e+00f);
    den = __builtin_fmaf(t, den, 3.754408661907416e+00f);
    den = __builtin_fmaf(t, den, 1.0f);
    // t in [2.72, 5.26] for the clamped tail uniforms: den in [36, 140] and
    // |num| = |z| den in [71, 662], inside div_rn_normal's operand range
    float z = div_rn_normal(num, den);
    return low ? z : -z;
}

__device__ inline float icdf(float u) {
    return icdf_is_central(u) ? icdf_central(u - 0.5f) : icdf_tail(u);
}

// SV transition of ptx:406-412: x' = mu + phi (x - mu) + sigma eps.
__device__ inline float transition(float x, float mu, float phi, float sig, float z) {
    float d = x - mu;
    d = phi * d;
    float xn = mu + d;
    return __builtin_fmaf(sig, z, xn);
}

// Observation log-weight of ptx:414-450.  Gaussian:
//   -0.9189385 - y^2 e^{-x} / 2 - x / 2
// Student-t (nu_obs > 0): C_obs - (nu+1)/2 log(1 + y^2 e^{-x} / nu) - x / 2.
struct ObsModel {
    float ysq;      // y_t^2
    float nu;       // 0 = Gaussian
    float c_obs;    // Student-t normalising constant (host FP64 lgamma)
    float nph;      // (nu + 1) / 2
};

// Log-weights given eta^2 = y^2 e^{-x}.
__device__ inline float log_weight_gauss_eta(float x, float eta2) {
    float lw = __builtin_fmaf(-0.5f, eta2, bits_f32(0xBF6B3F8Eu));
    return __builtin_fmaf(-0.5f, x, lw);
}

__device__ inline float log_weight_t_eta(float x, float eta2, const ObsModel& o) {
    float ta = eta2 / o.nu + 1.0f;
    float lw = __builtin_fmaf(-o.nph, logf_cw(ta), o.c_obs);
    return __builtin_fmaf(-0.5f, x, lw);
}

__device__ inline float log_weight_gauss(float x, float ysq) {
    return log_weight_gauss_eta(x, ysq * expf_cw(-x));
}

__device__ inline float log_weight_t(float x, const ObsModel& o) {
    return log_weight_t_eta(x, o.ysq * expf_cw(-x), o);
}

// d lw / dx of the observation log-density at the propagated state, in the
// reference's form (ptx_assembly/bpf_kernels_full.ptx:1172-1189): with
// eta^2 = y^2 e^{-x}, Gaussian -1/2 + eta^2/2, Student-t
// -1/2 + (nu+1)/2 * eta^2 / (nu + eta^2) (correctly rounded division).
template <bool STUDENT>
__device__ inline float learn_dlw_eta(float eta2, const ObsModel& o) {
    if (STUDENT) return __builtin_fmaf(o.nph, eta2 / (o.nu + eta2), -0.5f);
    return __builtin_fmaf(eta2, 0.5f, -0.5f);
}

__device__ inline float log_weight(float x, const ObsModel& o) {
    return o.nu > 0.0f ? log_weight_t(x, o) : log_weight_gauss(x, o.ysq);
}

// Systematic-resampling target of gpu_bpf/gpu_bpf.cu:300-302, u + i/N in
// FP32 with the IEEE-rounded quotient, before the wrap at 1.  The quotient is
// formed without a float division, as float(double(float(i)) * R) with
// R = 1 / double(float(N)) rounded to double.  For a = float(i) and
// b = float(N) (24-bit significands) the exact a/b is never a midpoint
// between adjacent floats and lies more than 2^-50 a/b from every midpoint
// (a - b m is a nonzero multiple of 2^(e_a - 25)), while the product is
// within 2^-52 a/b of a/b; both therefore round to the same float, the IEEE
// quotient.
__device__ __host__ inline double target_recip(int n) { return 1.0 / (double)(float)n; }
__device__ __host__ inline float target_at(float u, int i, double rn) {
    return u + (float)((double)(float)i * rn);
}
__device__ __host__ inline float target_unwrapped(float u, int i, int n) {
    return target_at(u, i, target_recip(n));
}

}  // namespace bpf
