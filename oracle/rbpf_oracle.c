/*
 * rbpf_oracle.c -- CPU oracle for the SMC^2 learner with the
 * Rao-Blackwellised (RBPF / OCSN) inner filter.
 *
 * TEST INFRASTRUCTURE ONLY: nothing in the shipped library links, loads or
 * calls this code; only tests/ and __graft_entry__.smoke() use it, as the
 * checker.
 *
 * It restates the algorithm of smc2/smc2_rbpf_batch.cu with the arithmetic
 * the HIP library uses, so a run repeats bit for bit:
 *   - per-particle FP32 model functions (rbpf_device.hpp twins): regime map
 *     z = 1.5 (1 + tanh z~) (smc2_rbpf_batch.cuh:392-394), curves
 *     (:388-390), stationary start (smc2_rbpf_batch.cu:208-229), transition
 *     and Kalman prediction (:358-370), OCSN moment-matched update (:64-122),
 *     resampling offset Phi(n) (smc2_rbpf_batch.cuh:402-405);
 *   - the inner tick (:305-393): systematic resample every tick on an FP64
 *     CDF, sort by mu_h every 4th tick (smc2/smc2_sorting.cuh:47-106),
 *     FP64 normalisation with the kernel's association order (wave64
 *     butterfly sums and Hillis-Steele scans, wave totals in wave order);
 *   - the outer layer (:1329-1502): FP64 weights and ESS, systematic
 *     resample from the xorshift64* host stream (:1009-1018), adaptive
 *     Cholesky proposal (:1248-1294), CPMMH moves with correlated noise and
 *     fixed-lag checkpoints (:651-1003).
 * Random numbers are the library's counter-based streams (splitmix64 keys,
 * Acklam inverse CDF).  FP64 log / exp come from libm here and from the
 * device library on the GPU, so FP64 log-likelihoods agree to a few ulp.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "bpf_oracle.h"

#define RB_K 10
#define RB_SORT_EVERY 4
#define RB_NP 4

static const float OC_P[RB_K] = {0.00609f, 0.04775f, 0.13057f, 0.20674f, 0.22715f,
                                 0.18842f, 0.12047f, 0.05591f, 0.01575f, 0.00115f};
static const float OC_M[RB_K] = {1.92677f,  1.34744f,  0.73504f,  0.02266f,  -0.85173f,
                                 -1.97278f, -3.46788f, -5.55246f, -8.68384f, -14.65000f};
static const float OC_V[RB_K] = {0.11265f, 0.17788f, 0.26768f, 0.40611f, 0.62699f,
                                 0.98583f, 1.57469f, 2.54498f, 4.16591f, 7.33342f};

typedef struct {
    float tb, ts, tr, ms, mr, ss, sr;
    float om[RB_K], ov[RB_K], olw[RB_K];
    float lw0;
} RbModel;

static void rb_model(const OrcRbCfg* c, RbModel* m) {
    m->tb = c->tcurve[0];
    m->ts = c->tcurve[1];
    m->tr = c->tcurve[2];
    m->ms = c->fixed[0];
    m->mr = c->fixed[1];
    m->ss = c->fixed[2];
    m->sr = c->fixed[3];
    const double h = 0.5 * log(2.0 * 3.14159265358979323846);
    for (int k = 0; k < RB_K; ++k) {
        m->om[k] = OC_M[k];
        m->ov[k] = OC_V[k];
        m->olw[k] = (float)(log((double)OC_P[k]) - h);
    }
    m->lw0 = -(float)log((double)c->n_inner);
}

/* ---- counter-based streams -------------------------------------------- */
static inline uint64_t sm64(uint64_t z) {
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}
static inline uint64_t rkey(uint64_t a, uint64_t b) { return sm64(a ^ sm64(b)); }
static inline float runif(uint64_t k, uint32_t ctr) {
    return orc_uniform((uint32_t)(sm64(k + (uint64_t)ctr * 0xD1B54A32D192ED03ull) >> 32));
}
static inline float rnorm(uint64_t k, uint32_t ctr) { return orc_icdf_f32(runif(k, ctr)); }
enum { P_PRIOR = 1, P_Z = 2, P_U0 = 3, P_MOVE = 4, P_FZ = 5, P_FU = 6 };

/* ---- noise storage ------------------------------------------------------ */
/* float -> IEEE binary16 with round-to-nearest-even (overflow to inf, NaN
 * kept, subnormal halves) -> float: the value a half-precision noise history
 * holds (SMC2_NOISE_FP16 of smc2/smc2_noise_precision.cuh). */
static float half_rt(float x) {
    uint32_t b;
    memcpy(&b, &x, 4);
    const uint32_t sign = (b >> 16) & 0x8000u;
    const uint32_t ab = b & 0x7FFFFFFFu;
    uint32_t h;
    if (ab > 0x7F800000u) {
        h = sign | 0x7E00u;
    } else if (ab >= 0x47800000u) {               /* >= 65536: infinity */
        h = sign | 0x7C00u;
    } else if (ab >= 0x38800000u) {               /* normal half range, >= 2^-14 */
        const uint32_t e = (ab >> 23) - 127u + 15u;
        const uint32_t mant = ab & 0x7FFFFFu;
        uint32_t hv = (e << 10) | (mant >> 13);
        const uint32_t rem = mant & 0x1FFFu;
        if (rem > 0x1000u || (rem == 0x1000u && (hv & 1u))) hv += 1u;   /* may carry to inf */
        h = sign | hv;
    } else {                                      /* subnormal half or zero: units of 2^-24 */
        float f;
        memcpy(&f, &ab, 4);
        h = sign | (uint32_t)rintf(f * 16777216.0f);
    }
    const uint32_t e = (h >> 10) & 31u, m = h & 0x3FFu;
    uint32_t ob;
    if (e == 0u) {
        float v = (float)m * (1.0f / 16777216.0f);
        memcpy(&ob, &v, 4);
        ob |= sign << 16;
    } else if (e == 31u) {
        ob = (sign << 16) | 0x7F800000u | (m << 13);
    } else {
        ob = (sign << 16) | ((e - 15u + 127u) << 23) | (m << 13);
    }
    float o;
    memcpy(&o, &ob, 4);
    return o;
}

void orc_rb_half_roundtrip(const float* in, float* out, int n) {
    for (int i = 0; i < n; ++i) out[i] = half_rt(in[i]);
}

/* the value a noise history stores and the filter uses */
static int g_noise16 = 0;
static inline float nstore(float v) { return g_noise16 ? half_rt(v) : v; }

/* ---- model functions (FP32) ------------------------------------------- */
static inline float rb_tanh(float x) {
    float ax = fabsf(x);
    float e = orc_expf(ax + ax);
    float r = 1.0f - 2.0f / (e + 1.0f);
    return copysignf(r, x);
}
static inline float rb_z(float zt) { return 1.5f * (1.0f + rb_tanh(zt)); }
static inline float rb_curve(float base, float scale, float rate, float z) {
    return base + scale * (1.0f - orc_expf(-(rate * z)));
}
static inline float rb_erfc(float x) {
    float z = fabsf(x);
    float t = 1.0f / (1.0f + 0.5f * z);
    float p = 0.17087277f;
    p = fmaf(t, p, -0.82215223f);
    p = fmaf(t, p, 1.48851587f);
    p = fmaf(t, p, -1.13520398f);
    p = fmaf(t, p, 0.27886807f);
    p = fmaf(t, p, -0.18628806f);
    p = fmaf(t, p, 0.09678418f);
    p = fmaf(t, p, 0.37409196f);
    p = fmaf(t, p, 1.00002368f);
    p = fmaf(t, p, -1.26551223f);
    float ans = t * orc_expf(p - z * z);
    return x >= 0.0f ? ans : 2.0f - ans;
}
static inline float rb_u0(float n) {
    float u = 0.5f * rb_erfc(-(n * 0.70710678f));
    return fminf(fmaxf(u, 1e-7f), 1.0f - 1e-7f);
}
static inline void rb_init_particle(float rho, float sz, float mub, float sgb, float n0,
                                    const RbModel* md, float* zt, float* mu, float* var) {
    float omr = fmaxf(1.0f - rho * rho, 1e-6f);
    float sstd = sz / sqrtf(omr);
    *zt = sstd * n0;
    float z = rb_z(*zt);
    float th = rb_curve(md->tb, md->ts, md->tr, z);
    float muz = rb_curve(mub, md->ms, md->mr, z);
    float sh = rb_curve(sgb, md->ss, md->sr, z);
    float phi = 1.0f - th;
    *mu = muz;
    *var = (sh * sh) / fmaxf(1.0f - phi * phi, 1e-6f);
}
static inline void rb_predict(float rho, float sz, float mub, float sgb, float zn,
                              const RbModel* md, float* zt, float mu, float var, float* mp,
                              float* vp) {
    float a = rho * *zt;
    float b = sz * zn;
    *zt = a + b;
    float z = rb_z(*zt);
    float th = rb_curve(md->tb, md->ts, md->tr, z);
    float muz = rb_curve(mub, md->ms, md->mr, z);
    float sh = rb_curve(sgb, md->ss, md->sr, z);
    float phi = 1.0f - th;
    *mp = phi * mu + th * muz;
    *vp = fmaxf(phi * phi * var + sh * sh, 1e-8f);
}
static inline void rb_ocsn(float y, float mp, float vp, const RbModel* md, float* mu_post,
                           float* var_post, float* lnorm) {
    const float v = fmaxf(vp, 1e-6f);
    float val[RB_K], inn[RB_K], inv[RB_K];
    float mx = -INFINITY;
    const float d = y - mp;
    for (int k = 0; k < RB_K; ++k) {
        const float S = v + md->ov[k];
        inv[k] = 1.0f / S;
        inn[k] = d - md->om[k];
        const float q = inn[k] * inn[k] * inv[k];
        val[k] = md->olw[k] - 0.5f * (orc_logf(S) + q);
        mx = fmaxf(mx, val[k]);
    }
    float e[RB_K];
    float se = 0.0f;
    for (int k = 0; k < RB_K; ++k) {
        e[k] = orc_expf(val[k] - mx);
        se += e[k];
    }
    *lnorm = mx + orc_logf(se);
    const float rse = 1.0f / se;
    float mo = 0.0f, e2 = 0.0f;
    for (int k = 0; k < RB_K; ++k) {
        const float w = e[k] * rse;
        const float K = v * inv[k];
        const float mk = mp + K * inn[k];
        const float vk = (1.0f - K) * v;
        mo += w * mk;
        e2 += w * (vk + mk * mk);
    }
    *mu_post = mo;
    *var_post = fmaxf(e2 - mo * mo, 1e-6f);
}
static inline uint32_t rb_sort_key(float f) {
    uint32_t b;
    memcpy(&b, &f, 4);
    return (b & 0x80000000u) ? ~b : (b | 0x80000000u);
}

/* ---- block reductions in the kernel's association order --------------- */
static double wave_butterfly(const double* x) {
    double v[64], nv[64];
    memcpy(v, x, sizeof(v));
    for (int off = 32; off >= 1; off >>= 1) {
        for (int i = 0; i < 64; ++i) nv[i] = v[i] + v[i ^ off];
        memcpy(v, nv, sizeof(v));
    }
    return v[0];
}
static double block_sum(const double* x, int n) {
    double tot = 0.0;
    for (int w = 0; w < n / 64; ++w) {
        double t = wave_butterfly(x + 64 * w);
        tot = (w == 0) ? t : tot + t;
    }
    return tot;
}
/* inclusive scan: per-wave Hillis-Steele, plus the running sum of the wave
 * totals in wave order; *total is the sum of all wave totals */
static void block_scan(const double* x, int n, double* c, double* total) {
    double T = 0.0;
    for (int w = 0; w < n / 64; ++w) {
        double v[64], nv[64];
        memcpy(v, x + 64 * w, sizeof(v));
        for (int off = 1; off < 64; off <<= 1) {
            for (int i = 0; i < 64; ++i) nv[i] = v[i] + ((i >= off) ? v[i - off] : 0.0);
            memcpy(v, nv, sizeof(v));
        }
        const double P = T;
        for (int i = 0; i < 64; ++i) c[64 * w + i] = v[i] + P;
        T += v[63];
    }
    *total = T;
}
static int cmp_u64(const void* a, const void* b) {
    uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
    return (x > y) - (x < y);
}

/* Scratch of one theta-particle's tick. */
typedef struct {
    float *z0, *mu0, *var0;
    double *w, *c;
    uint64_t* k;
} RbScratch;

static void scratch_alloc(RbScratch* s, int ni) {
    s->z0 = malloc(sizeof(float) * ni);
    s->mu0 = malloc(sizeof(float) * ni);
    s->var0 = malloc(sizeof(float) * ni);
    s->w = malloc(sizeof(double) * ni);
    s->c = malloc(sizeof(double) * ni);
    s->k = malloc(sizeof(uint64_t) * ni);
}
static void scratch_free(RbScratch* s) {
    free(s->z0);
    free(s->mu0);
    free(s->var0);
    free(s->w);
    free(s->c);
    free(s->k);
}

/* One inner tick of one theta-particle (k_rb_step / k_rb_move's tick). */
static void rb_tick(int ni, float* zt, float* mu, float* var, float* lw, const float* zn, float u0,
                    float y, int t, const float* th, const RbModel* md, RbScratch* s,
                    double* ll_incr, double* ess) {
    float m0 = -INFINITY;
    for (int i = 0; i < ni; ++i) m0 = fmaxf(m0, lw[i]);
    for (int i = 0; i < ni; ++i) s->w[i] = (double)orc_expf(lw[i] - m0);
    double S0;
    block_scan(s->w, ni, s->c, &S0);
    for (int i = 0; i < ni; ++i) s->c[i] = (i == ni - 1) ? 1.0 : s->c[i] / S0;
    memcpy(s->z0, zt, sizeof(float) * ni);
    memcpy(s->mu0, mu, sizeof(float) * ni);
    memcpy(s->var0, var, sizeof(float) * ni);
    for (int i = 0; i < ni; ++i) {
        const float u = (u0 + (float)i) / (float)ni;
        const double uu = (double)u;
        int lo = 0, hi = ni - 1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (s->c[mid] < uu) lo = mid + 1; else hi = mid;
        }
        zt[i] = s->z0[lo];
        mu[i] = s->mu0[lo];
        var[i] = s->var0[lo];
        lw[i] = md->lw0;
    }
    if (t % RB_SORT_EVERY == 0) {
        memcpy(s->z0, zt, sizeof(float) * ni);
        memcpy(s->mu0, mu, sizeof(float) * ni);
        memcpy(s->var0, var, sizeof(float) * ni);
        for (int i = 0; i < ni; ++i) s->k[i] = ((uint64_t)rb_sort_key(mu[i]) << 32) | (uint32_t)i;
        qsort(s->k, (size_t)ni, sizeof(uint64_t), cmp_u64);
        for (int i = 0; i < ni; ++i) {
            const int src = (int)(uint32_t)s->k[i];
            zt[i] = s->z0[src];
            mu[i] = s->mu0[src];
            var[i] = s->var0[src];
        }
    }
    for (int i = 0; i < ni; ++i) {
        float mp, vp, ln;
        rb_predict(th[0], th[1], th[2], th[3], zn[i], md, &zt[i], mu[i], var[i], &mp, &vp);
        rb_ocsn(y, mp, vp, md, &mu[i], &var[i], &ln);
        lw[i] = lw[i] + ln;
    }
    float m1 = -INFINITY;
    for (int i = 0; i < ni; ++i) m1 = fmaxf(m1, lw[i]);
    for (int i = 0; i < ni; ++i) {
        const float w1 = orc_expf(lw[i] - m1);
        s->w[i] = (double)w1;
        s->c[i] = (double)w1 * (double)w1;
    }
    const double S = block_sum(s->w, ni), Q = block_sum(s->c, ni);
    *ess = (S * S) / Q;
    *ll_incr = (double)m1 + log(S);
}

static double rb_log_prior(const float* th, const OrcRbCfg* c) {
    const float* b = c->bounds;
    for (int q = 0; q < RB_NP; ++q)
        if (th[q] < b[2 * q] || th[q] > b[2 * q + 1]) return -HUGE_VAL;
    const float* p = c->prior;
    double s = 0.0;
    double d[RB_NP];
    for (int q = 0; q < RB_NP; ++q) d[q] = ((double)th[q] - (double)p[2 * q]) / (double)p[2 * q + 1];
    s = d[0] * d[0] + d[1] * d[1] + d[2] * d[2] + d[3] * d[3];
    return -0.5 * s;
}

static uint64_t xs64(uint64_t* s) {
    uint64_t x = *s;
    x ^= x >> 12;
    x ^= x << 25;
    x ^= x >> 27;
    *s = x;
    return x * 0x2545F4914F6CDD1DULL;
}
static float xs64_uniform(uint64_t* s) {
    return (float)((xs64(s) >> 11) + 1) * (1.0f / 9007199254740994.0f);
}

/* ---- whole SMC^2 run ------------------------------------------------- */
typedef struct {
    int nt, ni, cap;
    float* th[2][RB_NP];
    double* ll[2];
    float* essi[2];
    float* in[2][4];
    float* cp[2][4];
    double* cpll[2];
    float *zn[2], *un[2];
    int tb, ib, nb, cb;
} RbState;

static void st_alloc(RbState* S, int nt, int ni, int cap) {
    memset(S, 0, sizeof(*S));
    S->nt = nt;
    S->ni = ni;
    S->cap = cap;
    const size_t tot = (size_t)nt * ni;
    for (int b = 0; b < 2; ++b) {
        for (int q = 0; q < RB_NP; ++q) S->th[b][q] = calloc(nt, sizeof(float));
        S->ll[b] = calloc(nt, sizeof(double));
        S->essi[b] = calloc(nt, sizeof(float));
        for (int q = 0; q < 4; ++q) {
            S->in[b][q] = calloc(tot, sizeof(float));
            S->cp[b][q] = calloc(tot, sizeof(float));
        }
        S->cpll[b] = calloc(nt, sizeof(double));
        S->zn[b] = calloc(tot * (size_t)(cap + 1), sizeof(float));
        S->un[b] = calloc((size_t)nt * (cap + 1), sizeof(float));
    }
}
static void st_free(RbState* S) {
    for (int b = 0; b < 2; ++b) {
        for (int q = 0; q < RB_NP; ++q) free(S->th[b][q]);
        free(S->ll[b]);
        free(S->essi[b]);
        for (int q = 0; q < 4; ++q) {
            free(S->in[b][q]);
            free(S->cp[b][q]);
        }
        free(S->cpll[b]);
        free(S->zn[b]);
        free(S->un[b]);
    }
}

static void outer_weights(const double* lw, double* w, int n, float* ess) {
    double mx = -HUGE_VAL;
    for (int m = 0; m < n; ++m) mx = fmax(mx, lw[m]);
    double S = 0.0;
    for (int m = 0; m < n; ++m) {
        w[m] = exp(lw[m] - mx);
        S += w[m];
    }
    double Q = 0.0;
    for (int m = 0; m < n; ++m) {
        w[m] /= S;
        Q += w[m] * w[m];
    }
    *ess = (float)(1.0 / fmax(Q, 1e-300));
}

static void adapt_chol(const RbState* S, float* chol, int adaptive, const float* pstd) {
    const int n = S->nt;
    const double sc = sqrt((double)(2.38f * 2.38f / (float)RB_NP));
    double mean[RB_NP] = {0}, cov[RB_NP][RB_NP] = {{0}}, L[RB_NP][RB_NP] = {{0}};
    if (!adaptive) {
        for (int i = 0; i < RB_NP; ++i)
            for (int j = 0; j < RB_NP; ++j) chol[i * RB_NP + j] = (i == j) ? (float)((double)pstd[i] * sc) : 0.0f;
        return;
    }
    float* const* th = S->th[S->tb];
    for (int q = 0; q < RB_NP; ++q) {
        double a = 0.0;
        for (int m = 0; m < n; ++m) a += (double)th[q][m];
        mean[q] = a / n;
    }
    for (int i = 0; i < RB_NP; ++i)
        for (int j = 0; j <= i; ++j) {
            double a = 0.0;
            for (int m = 0; m < n; ++m) a += ((double)th[i][m] - mean[i]) * ((double)th[j][m] - mean[j]);
            cov[i][j] = cov[j][i] = a / (n - 1);
        }
    for (int i = 0; i < RB_NP; ++i) cov[i][i] += 1e-6;
    int ok = 1;
    for (int i = 0; i < RB_NP; ++i)
        for (int j = 0; j <= i; ++j) {
            double a = 0.0;
            for (int k = 0; k < j; ++k) a += L[i][k] * L[j][k];
            if (i == j) {
                double v = cov[i][i] - a;
                if (v <= 0.0) {
                    ok = 0;
                    v = 1e-8;
                }
                L[i][j] = sqrt(v);
            } else {
                L[i][j] = (L[j][j] > 1e-10) ? (cov[i][j] - a) / L[j][j] : 0.0;
            }
        }
    if (!ok)
        for (int i = 0; i < RB_NP; ++i)
            for (int j = 0; j < RB_NP; ++j) L[i][j] = (i == j) ? sqrt(fmax(cov[i][i], 1e-8)) : 0.0;
    for (int i = 0; i < RB_NP; ++i)
        for (int j = 0; j < RB_NP; ++j) chol[i * RB_NP + j] = (float)(L[i][j] * sc);
}

static void rb_init(RbState* S, const OrcRbCfg* c, const RbModel* md) {
    const int ni = S->ni;
    const uint64_t kp = rkey(c->seed, P_PRIOR);
    const uint64_t kz = rkey(rkey(c->seed, P_Z), 0), ku = rkey(rkey(c->seed, P_U0), 0);
    const long long zrow = (long long)(S->cap + 1) * ni;
    for (int m = 0; m < S->nt; ++m) {
        const uint64_t km = rkey(kp, (uint64_t)m);
        float th[RB_NP];
        for (int at = 0; at < 1000; ++at) {
            int ok = 1;
            for (int q = 0; q < RB_NP; ++q) {
                th[q] = c->prior[2 * q] + c->prior[2 * q + 1] * rnorm(km, (uint32_t)(4 * at + q));
                ok = ok && th[q] >= c->bounds[2 * q] && th[q] <= c->bounds[2 * q + 1];
            }
            if (ok) break;
        }
        for (int q = 0; q < RB_NP; ++q) S->th[0][q][m] = th[q];
        S->ll[0][m] = 0.0;
        S->essi[0][m] = (float)ni;
        S->un[0][(long long)m * (S->cap + 1)] = nstore(rnorm(rkey(ku, (uint64_t)m), 0));
        for (int i = 0; i < ni; ++i) {
            const long long g = (long long)m * ni + i;
            const float n0 = nstore(rnorm(rkey(kz, (uint64_t)g), 0));
            S->zn[0][m * zrow + i] = n0;
            rb_init_particle(th[0], th[1], th[2], th[3], n0, md, &S->in[0][0][g], &S->in[0][1][g],
                             &S->in[0][2][g]);
            S->in[0][3][g] = md->lw0;
        }
    }
}

/* forward tick t of every theta-particle; lli[m] receives the increments */
static void rb_step(RbState* S, const OrcRbCfg* c, const RbModel* md, int t, float y, int epoch,
                    double* lli) {
    const int ni = S->ni;
    const uint64_t kz = rkey(rkey(c->seed, P_Z), (uint64_t)epoch);
    const uint64_t ku = rkey(rkey(c->seed, P_U0), (uint64_t)epoch);
    const long long zrow = (long long)(S->cap + 1) * ni;
#pragma omp parallel for schedule(dynamic)
    for (int m = 0; m < S->nt; ++m) {
        RbScratch sc;
        scratch_alloc(&sc, ni);
        float* zn = S->zn[S->nb] + m * zrow + (long long)(t + 1) * ni;
        for (int i = 0; i < ni; ++i)
            zn[i] = nstore(rnorm(rkey(kz, (uint64_t)((long long)m * ni + i)), (uint32_t)(t + 1)));
        const float n = nstore(rnorm(rkey(ku, (uint64_t)m), (uint32_t)(t + 1)));
        S->un[S->nb][(long long)m * (S->cap + 1) + t + 1] = n;
        const float u0 = rb_u0(n);
        float th[RB_NP];
        for (int q = 0; q < RB_NP; ++q) th[q] = S->th[S->tb][q][m];
        float** in = S->in[S->ib];
        const long long g = (long long)m * ni;
        double l, e;
        rb_tick(ni, in[0] + g, in[1] + g, in[2] + g, in[3] + g, zn, u0, y, t, th, md, &sc, &l, &e);
        lli[m] = l;
        S->ll[S->tb][m] += l;
        S->essi[S->tb][m] = (float)e;
        scratch_free(&sc);
    }
}

/* one CPMMH move of every theta-particle; acc[m] = accepted */
static void rb_move(RbState* S, const OrcRbCfg* c, const RbModel* md, const float* y, int t_cur,
                    int r, int k, int t_cp, const float* chol, int* acc) {
    const int ni = S->ni;
    const uint64_t kmove = rkey(rkey(rkey(c->seed, P_MOVE), (uint64_t)r), (uint64_t)k);
    const uint64_t kfz = rkey(rkey(rkey(c->seed, P_FZ), (uint64_t)r), (uint64_t)k);
    const uint64_t kfu = rkey(rkey(rkey(c->seed, P_FU), (uint64_t)r), (uint64_t)k);
    const long long zrow = (long long)(S->cap + 1) * ni;
    const float crho = c->crho, cscale = sqrtf(1.0f - c->crho * c->crho);
    const float* zc = S->zn[S->nb];
    float* zo = S->zn[1 - S->nb];
    const float* uc = S->un[S->nb];
    float* uo = S->un[1 - S->nb];
#pragma omp parallel for schedule(dynamic)
    for (int m = 0; m < S->nt; ++m) {
        const uint64_t km = rkey(kmove, (uint64_t)m);
        float th[RB_NP], z[RB_NP], pert[RB_NP], tp[RB_NP];
        for (int q = 0; q < RB_NP; ++q) th[q] = S->th[S->tb][q][m];
        const double lpc = rb_log_prior(th, c);
        for (int q = 0; q < RB_NP; ++q) z[q] = rnorm(km, (uint32_t)q);
        const float umix = runif(km, 4);
        if (umix > 0.05f) {
            for (int i = 0; i < RB_NP; ++i) {
                float s = 0.0f;
                for (int j = 0; j <= i; ++j) s += chol[i * RB_NP + j] * z[j];
                pert[i] = s;
            }
        } else {
            for (int i = 0; i < RB_NP; ++i) pert[i] = c->pstd[i] * z[i];
        }
        for (int q = 0; q < RB_NP; ++q) tp[q] = th[q] + pert[q];
        const double lpp = rb_log_prior(tp, c);
        if (!isfinite(lpp)) {
            acc[m] = 0;
            continue;
        }
        RbScratch sc;
        scratch_alloc(&sc, ni);
        float* zt = malloc(sizeof(float) * ni);
        float* mu = malloc(sizeof(float) * ni);
        float* var = malloc(sizeof(float) * ni);
        float* lw = malloc(sizeof(float) * ni);
        float* zp = malloc(sizeof(float) * ni);
        const long long g0 = (long long)m * ni;
        const long long zb = (long long)m * zrow, ub = (long long)m * (S->cap + 1);
        int t0;
        if (t_cp >= 0) {
            for (int i = 0; i < ni; ++i) {
                zt[i] = S->cp[S->cb][0][g0 + i];
                mu[i] = S->cp[S->cb][1][g0 + i];
                var[i] = S->cp[S->cb][2][g0 + i];
                lw[i] = S->cp[S->cb][3][g0 + i];
            }
            t0 = t_cp + 1;
        } else {
            for (int i = 0; i < ni; ++i) {
                const float v = nstore(crho * zc[zb + i] + cscale * rnorm(rkey(kfz, (uint64_t)(g0 + i)), 0));
                zo[zb + i] = v;
                rb_init_particle(tp[0], tp[1], tp[2], tp[3], v, md, &zt[i], &mu[i], &var[i]);
                lw[i] = md->lw0;
            }
            t0 = 0;
        }
        double llacc = 0.0, ess = (double)ni;
        const uint64_t ku = rkey(kfu, (uint64_t)m);
        for (int t = t0; t <= t_cur; ++t) {
            const long long zi = zb + (long long)(t + 1) * ni;
            for (int i = 0; i < ni; ++i) {
                zp[i] = nstore(crho * zc[zi + i] + cscale * rnorm(rkey(kfz, (uint64_t)(g0 + i)), (uint32_t)(t + 1)));
                zo[zi + i] = zp[i];
            }
            const float up = nstore(crho * uc[ub + t + 1] + cscale * rnorm(ku, (uint32_t)(t + 1)));
            uo[ub + t + 1] = up;
            double l;
            rb_tick(ni, zt, mu, var, lw, zp, rb_u0(up), y[t], t, tp, md, &sc, &l, &ess);
            llacc += l;
        }
        const double llb = (t_cp >= 0) ? S->cpll[S->cb][m] : 0.0;
        const double llc = S->ll[S->tb][m] - llb;
        const double la = (llacc + lpp) - (llc + lpc);
        const float u = runif(km, 5);
        const int ok = log((double)u) < la;
        acc[m] = ok;
        if (ok) {
            for (int q = 0; q < RB_NP; ++q) S->th[S->tb][q][m] = tp[q];
            S->ll[S->tb][m] = llb + llacc;
            S->essi[S->tb][m] = (float)ess;
            for (int i = 0; i < ni; ++i) {
                S->in[S->ib][0][g0 + i] = zt[i];
                S->in[S->ib][1][g0 + i] = mu[i];
                S->in[S->ib][2][g0 + i] = var[i];
                S->in[S->ib][3][g0 + i] = lw[i];
            }
        }
        free(zt);
        free(mu);
        free(var);
        free(lw);
        free(zp);
        scratch_free(&sc);
    }
}

static void gather_rows(const float* src, float* dst, const int* anc, int rows, long long stride,
                        long long len) {
    for (int m = 0; m < rows; ++m)
        memcpy(dst + (long long)m * stride, src + (long long)anc[m] * stride, sizeof(float) * len);
}

int orc_rb_smc2_run(const OrcRbCfg* c, const float* y, int T, float* ess_out, float* theta_out,
                    double* ll_out, double* lw_out, float* essi_out, float* inner_out,
                    int* resample_ticks, OrcRbCount* cnt) {
    const int nt = c->n_theta, ni = c->n_inner;
    if (nt < 2 || ni < 64 || ni % 64 || T < 0 || c->cap < T) return -1;
    if (c->noise_bits != 16 && c->noise_bits != 32) return -1;
    g_noise16 = c->noise_bits == 16;
    RbModel md;
    rb_model(c, &md);
    RbState S;
    st_alloc(&S, nt, ni, c->cap);
    rb_init(&S, c, &md);
    double* lw = calloc(nt, sizeof(double));
    double* w = calloc(nt, sizeof(double));
    double* lli = calloc(nt, sizeof(double));
    int* anc = calloc(nt, sizeof(int));
    int* acc = calloc(nt, sizeof(int));
    double* cdf = calloc(nt, sizeof(double));
    float chol[RB_NP * RB_NP];
    adapt_chol(&S, chol, 0, c->pstd);
    uint64_t host = 0x853C49E6748FEA9BULL ^ (c->seed ? c->seed : 12345ull);
    int n_res = 0, n_acc = 0, n_tot = 0, t_cp = -1;
    float ess = 0.0f;
    outer_weights(lw, w, nt, &ess);
    const long long zrow = (long long)(c->cap + 1) * ni;
    for (int t = 0; t < T; ++t) {
        rb_step(&S, c, &md, t, y[t], n_res, lli);
        for (int m = 0; m < nt; ++m) lw[m] += lli[m];
        outer_weights(lw, w, nt, &ess);
        if (ess < c->ess_thr * (float)nt) {
            if (resample_ticks) resample_ticks[n_res] = t;
            n_res += 1;
            const float u = xs64_uniform(&host);
            double a = 0.0;
            for (int m = 0; m < nt; ++m) {
                a += w[m];
                cdf[m] = a;
            }
            cdf[nt - 1] = 1.0;
            for (int i = 0; i < nt; ++i) {
                const double ui = (double)((u + (float)i) / (float)nt);
                int lo = 0, hi = nt - 1;
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    if (cdf[mid] < ui) lo = mid + 1; else hi = mid;
                }
                anc[i] = lo;
            }
            for (int q = 0; q < RB_NP; ++q) gather_rows(S.th[S.tb][q], S.th[1 - S.tb][q], anc, nt, 1, 1);
            for (int m = 0; m < nt; ++m) {
                S.ll[1 - S.tb][m] = S.ll[S.tb][anc[m]];
                S.essi[1 - S.tb][m] = S.essi[S.tb][anc[m]];
            }
            for (int q = 0; q < 4; ++q) gather_rows(S.in[S.ib][q], S.in[1 - S.ib][q], anc, nt, ni, ni);
            gather_rows(S.zn[S.nb], S.zn[1 - S.nb], anc, nt, zrow, (long long)(t + 2) * ni);
            gather_rows(S.un[S.nb], S.un[1 - S.nb], anc, nt, c->cap + 1, t + 2);
            S.tb = 1 - S.tb;
            S.ib = 1 - S.ib;
            S.nb = 1 - S.nb;
            if (c->lag > 0 && t_cp >= 0) {
                for (int q = 0; q < 4; ++q) gather_rows(S.cp[S.cb][q], S.cp[1 - S.cb][q], anc, nt, ni, ni);
                for (int m = 0; m < nt; ++m) S.cpll[1 - S.cb][m] = S.cpll[S.cb][anc[m]];
                S.cb = 1 - S.cb;
            }
            for (int m = 0; m < nt; ++m) lw[m] = 0.0;
            int use_cp = 0;
            if (c->lag > 0 && t_cp >= 0) {
                const int steps = t - t_cp;
                use_cp = steps > 0 && steps <= 2 * c->lag;
            }
            adapt_chol(&S, chol, c->adaptive, c->pstd);
            const int t_start = use_cp ? t_cp + 1 : 0;
            for (int k = 0; k < c->K_rejuv; ++k) {
                rb_move(&S, c, &md, y, t, n_res, k, use_cp ? t_cp : -1, chol, acc);
                for (int m = 0; m < nt; ++m) {
                    if (!acc[m]) continue;
                    memcpy(S.zn[S.nb] + m * zrow + (long long)t_start * ni,
                           S.zn[1 - S.nb] + m * zrow + (long long)t_start * ni,
                           sizeof(float) * (size_t)(t + 2 - t_start) * ni);
                    memcpy(S.un[S.nb] + (long long)m * (c->cap + 1) + t_start,
                           S.un[1 - S.nb] + (long long)m * (c->cap + 1) + t_start,
                           sizeof(float) * (size_t)(t + 2 - t_start));
                    n_acc += 1;
                }
                n_tot += nt;
            }
            for (int m = 0; m < nt; ++m) lw[m] = 0.0;
            outer_weights(lw, w, nt, &ess);
        }
        if (c->lag > 0 && t > 0 && t % c->lag == 0 && t > t_cp) {
            for (int q = 0; q < 4; ++q) memcpy(S.cp[S.cb][q], S.in[S.ib][q], sizeof(float) * (size_t)nt * ni);
            memcpy(S.cpll[S.cb], S.ll[S.tb], sizeof(double) * nt);
            t_cp = t;
        }
        if (ess_out) ess_out[t] = ess;
    }
    if (theta_out)
        for (int q = 0; q < RB_NP; ++q) memcpy(theta_out + (size_t)q * nt, S.th[S.tb][q], sizeof(float) * nt);
    if (ll_out) memcpy(ll_out, S.ll[S.tb], sizeof(double) * nt);
    if (lw_out) memcpy(lw_out, lw, sizeof(double) * nt);
    if (essi_out) memcpy(essi_out, S.essi[S.tb], sizeof(float) * nt);
    if (inner_out)
        for (int q = 0; q < 4; ++q)
            memcpy(inner_out + (size_t)q * nt * ni, S.in[S.ib][q], sizeof(float) * (size_t)nt * ni);
    if (cnt) {
        cnt->n_resamples = n_res;
        cnt->n_acc = n_acc;
        cnt->n_tot = n_tot;
        cnt->t_cp = t_cp;
    }
    free(lw);
    free(w);
    free(lli);
    free(anc);
    free(acc);
    free(cdf);
    st_free(&S);
    return 0;
}

/* One theta-particle (index 0, first epoch) filtered over y with theta held
 * fixed: the inner filter alone, for the likelihood pins. */
int orc_rb_inner_run(const OrcRbCfg* c, const float* theta, const float* y, int T,
                     double* ll_incr, double* ess, float* state_out) {
    const int ni = c->n_inner;
    if (ni < 64 || ni % 64 || T < 0) return -1;
    if (c->noise_bits != 16 && c->noise_bits != 32) return -1;
    g_noise16 = c->noise_bits == 16;
    RbModel md;
    rb_model(c, &md);
    const uint64_t kz = rkey(rkey(c->seed, P_Z), 0), ku = rkey(rkey(c->seed, P_U0), 0);
    float* zt = malloc(sizeof(float) * ni);
    float* mu = malloc(sizeof(float) * ni);
    float* var = malloc(sizeof(float) * ni);
    float* lw = malloc(sizeof(float) * ni);
    float* zn = malloc(sizeof(float) * ni);
    RbScratch sc;
    scratch_alloc(&sc, ni);
    for (int i = 0; i < ni; ++i) {
        rb_init_particle(theta[0], theta[1], theta[2], theta[3], nstore(rnorm(rkey(kz, (uint64_t)i), 0)), &md,
                         &zt[i], &mu[i], &var[i]);
        lw[i] = md.lw0;
    }
    for (int t = 0; t < T; ++t) {
        for (int i = 0; i < ni; ++i) zn[i] = nstore(rnorm(rkey(kz, (uint64_t)i), (uint32_t)(t + 1)));
        const float u0 = rb_u0(nstore(rnorm(rkey(ku, 0), (uint32_t)(t + 1))));
        double l, e;
        rb_tick(ni, zt, mu, var, lw, zn, u0, y[t], t, theta, &md, &sc, &l, &e);
        if (ll_incr) ll_incr[t] = l;
        if (ess) ess[t] = e;
    }
    if (state_out) {
        memcpy(state_out, zt, sizeof(float) * ni);
        memcpy(state_out + ni, mu, sizeof(float) * ni);
        memcpy(state_out + 2 * ni, var, sizeof(float) * ni);
        memcpy(state_out + 3 * ni, lw, sizeof(float) * ni);
    }
    free(zt);
    free(mu);
    free(var);
    free(lw);
    free(zn);
    scratch_free(&sc);
    return 0;
}

void orc_rb_ocsn(float y, float mp, float vp, float* out3) {
    OrcRbCfg c;
    memset(&c, 0, sizeof(c));
    c.n_inner = 64;
    RbModel md;
    rb_model(&c, &md);
    rb_ocsn(y, mp, vp, &md, &out3[0], &out3[1], &out3[2]);
}

/* fn 0 tanh, 1 erfc, 2 resampling offset Phi(n) clamped, 3 z = 1.5 (1 + tanh) */
void orc_rb_math(int fn, const float* in, float* out, int n) {
    for (int i = 0; i < n; ++i) {
        const float x = in[i];
        out[i] = fn == 0 ? rb_tanh(x) : fn == 1 ? rb_erfc(x) : fn == 2 ? rb_u0(x) : rb_z(x);
    }
}

void orc_rb_ocsn_table(float* p, float* m, float* v) {
    memcpy(p, OC_P, sizeof(OC_P));
    memcpy(m, OC_M, sizeof(OC_M));
    memcpy(v, OC_V, sizeof(OC_V));
}
