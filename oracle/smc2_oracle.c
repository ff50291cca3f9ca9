/*
 * smc2_oracle.c -- CPU oracle of SMC^2 over batched SV bootstrap filters.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the shipped library links, loads or
 * calls this code; tests/ use it as the checker of the GPU library's gpu_smc2_*
 * path.
 *
 * The inner filters are the single filter's FP32 tick of bpf_oracle.c with
 * 256-particle tiles in one reduction group, resampled every tick, driven by
 * one of three noise sources:
 *   - the PCG32 particle streams of the filter's seed (draw t at tick t,
 *     draw 0 for x_0) and the host resampling stream seed * 67890 + 12345
 *     (gpu_bpf/gpu_bpf_ptx_full.cu:241-251,353), optionally recorded;
 *   - a Crank-Nicolson proposal around the recorded normals,
 *     z' = fmaf(sqrt(1 - rho^2), xi, rho z) with xi from the PCG32 streams of a
 *     proposal seed, and a circular random walk of the recorded resampling
 *     uniforms, u' = frac(u + step (2 v - 1)) with v the uniform of stream
 *     index 16384 of the proposal seed (the correlated pseudo-marginal noise
 *     move of smc2/smc2_rbpf_batch.cu:693-882);
 * and, for MALA, FP32 forward-mode tangents d = dx/d(mu, phi, sigma)
 * carried through propagation (d' = phi d + (1 - phi, x - mu, z)) and
 * resampling, from d_0 = (1, z dsd/dphi, z dsd/dsigma) of
 * x_0 = mu + sd(phi, sigma) z, with the pathwise score
 * sum_t sum_i w_ti g_ti d_ti / sum_i w_ti, g = d lw / dx (the gradient path
 * of tests/diff_bpf.py:186-233 with ancestors held fixed).
 *
 * The outer layer restates smc2/smc2_rbpf_batch.cu:
 *   - outer log-weights: the sums of each inner filter's log-likelihood
 *     increments since the last outer resample (:385-397);
 *   - after every tick the outer ESS (sum w)^2 / sum w^2, w = exp(lw - max),
 *     and below threshold * M (or NaN) an outer resample (:486-513);
 *   - systematic resampling of whole inner filters: parameters, particle
 *     count, particles, tangents, score, filter state and recorded noise
 *     (:519-580); non-finite outer log-weights weigh nothing; afterwards every
 *     slot continues on fresh streams;
 *   - K moves: CPMMH with the adaptive random walk (2.38^2 / 3) s^2 Cov(theta)
 *     (:1246-1294) and the noise proposal above (:693-926), or preconditioned
 *     MALA theta' = theta + (h/2) A g + sqrt(h) A^(1/2) xi, A = Cov(theta), with
 *     fresh noise and the reverse proposal density in the acceptance ratio;
 *     acceptance inside the uniform prior box only (:906-926, :959-981).
 * Its random numbers follow the streams the library documents for the outer
 * layer (include/gpu_bpf_batch.h): PCG32 XSH-RR with stream id 0x5eed5c2
 * (prior draws, outer-resampling uniform, fresh seeds) and 0x30e5 (proposal
 * normals, proposal seeds, MH uniforms), 53-bit uniforms, Box-Muller normals.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "bpf_oracle.h"

#define S2_TILE 256
#define S2_MAXN 16384
#define S2_MULT 6364136223846793005ULL

static int g_sfault = 0;
static int g_tile = S2_TILE;   /* tile of the run in progress (orc_sv_smc2_run) */

/* Plant a mistake (tests only): 0 none, 1 drop the MALA proposal-density
 * correction, 2 Crank-Nicolson with 1 - rho^2 in place of sqrt(1 - rho^2),
 * 3 outer resampling with the offset u / M instead of (u + m) / M. */
void orc_smc2_set_fault(int code) { g_sfault = code; }

/* ---- outer-layer random numbers -------------------------------------------- */

typedef struct { uint64_t s, inc; } Hrng;

static uint32_t hr_next(Hrng* r) {
    uint64_t old = r->s;
    r->s = old * S2_MULT + r->inc;
    return orc_pcg_output(old);
}

static void hr_seed(Hrng* r, uint64_t seed, uint64_t stream) {
    r->s = 0;
    r->inc = (stream << 1u) | 1u;
    hr_next(r);
    r->s += seed;
    hr_next(r);
}

static double hr_uniform(Hrng* r) {
    uint64_t a = hr_next(r) >> 5;
    uint64_t b = hr_next(r) >> 6;
    return ((double)(a * 67108864ull + b) + 0.5) / 9007199254740992.0;
}

static double hr_normal(Hrng* r) {
    double u1 = hr_uniform(r);
    double u2 = hr_uniform(r);
    return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

/* ---- one inner filter ------------------------------------------------------ */

typedef struct {
    int n, seed;
    float mu, phi, sig;
    uint64_t hrng;      /* host resampling stream */
    int tick;
    double ll_cum;
    float* x;           /* cap */
    float* d[3];        /* tangents (cap each) or NULL */
    double score[3];
} Slot;

typedef struct {
    float* l;
    double* w;
    double* cdf;
    int* anc;
    float* xt;
    float* dt[3];
    float tmax[S2_MAXN / S2_TILE];
    double tf[S2_MAXN / S2_TILE];
} Work;

/* Noise of one slot's tick.  mode 0: PCG32 of the slot's seed; 1: the same,
 * recorded to z_out / u_out; 2: proposal around z_in / u_in with the streams of
 * prop_seed, recorded to z_out / u_out. */
typedef struct {
    int mode;
    const float* z_in;
    float* z_out;
    float u_in;
    float* u_out;
    int prop_seed;
    float cn_rho, cn_s, u_step;
} Noise;

typedef struct {
    float nu, c_obs, nph;
} Obs;

static inline float transition(float x, float mu, float phi, float sig, float z) {
    float d = x - mu;
    d = phi * d;
    float xn = mu + d;
    return fmaf(sig, z, xn);
}

static inline float dlw_dx(float x, float ysq, const Obs* o) {
    float eta2 = ysq * orc_expf(-x);
    if (o->nu > 0.0f) return o->nph * eta2 / (o->nu + eta2) - 0.5f;
    return fmaf(0.5f, eta2, -0.5f);
}

static inline float pcg_normal_j(uint64_t a, uint64_t g, uint32_t i) {
    return orc_icdf_f32(orc_uniform(orc_pcg_output(a + (uint64_t)(2u * i + 1u) * g)));
}

static inline float cn_step(const Noise* nz, float xi, float z) {
    float s = (g_sfault == 2) ? nz->cn_s * nz->cn_s : nz->cn_s;
    return fmaf(s, xi, nz->cn_rho * z);
}

/* x_0 = mu + sd z (draw 0), tangents of it; the slot's tick-0 normals. */
static void slot_init(Slot* s, const Noise* nz) {
    const float sd = sqrtf((s->sig * s->sig) / fmaxf(1.0f - s->phi * s->phi, 1e-6f));
    const double om = 1.0 - (double)s->phi * (double)s->phi;
    const float dsd_dsig = (float)(om > 1e-6 ? 1.0 / sqrt(om) : 1.0 / sqrt(1e-6));
    const float dsd_dphi = (float)(om > 1e-6 ? (double)s->sig * (double)s->phi / (om * sqrt(om)) : 0.0);
    uint64_t a, g, pa = 0, pg = 0;
    orc_pcg_jump(s->seed, 0u, &a, &g);
    if (nz->mode == 2) orc_pcg_jump(nz->prop_seed, 0u, &pa, &pg);
    for (int i = 0; i < s->n; i++) {
        float z;
        if (nz->mode == 2) {
            z = cn_step(nz, pcg_normal_j(pa, pg, (uint32_t)i), nz->z_in[i]);
            nz->z_out[i] = z;
        } else {
            z = pcg_normal_j(a, g, (uint32_t)i);
            if (nz->mode == 1) nz->z_out[i] = z;
        }
        s->x[i] = fmaf(sd, z, s->mu);
        if (s->d[0]) {
            s->d[0][i] = 1.0f;
            s->d[1][i] = z * dsd_dphi;
            s->d[2][i] = z * dsd_dsig;
        }
    }
    s->tick = 0;
    s->ll_cum = 0.0;
    s->score[0] = s->score[1] = s->score[2] = 0.0;
}

/* One tick (always resampled); returns the log-likelihood increment. */
static double slot_tick(Slot* s, float y, const Obs* o, const Noise* nz, Work* wk) {
    const int n = s->n, t = s->tick;
    const float ysq = y * y;
    if (t > 0) {
        uint64_t a, g;
        orc_pcg_jump(nz->mode == 2 ? nz->prop_seed : s->seed, (uint32_t)t, &a, &g);
        for (int i = 0; i < n; i++) {
            float z;
            if (nz->mode == 2) {
                z = cn_step(nz, pcg_normal_j(a, g, (uint32_t)i), nz->z_in[i]);
                nz->z_out[i] = z;
            } else {
                z = pcg_normal_j(a, g, (uint32_t)i);
                if (nz->mode == 1) nz->z_out[i] = z;
            }
            const float xi = s->x[i];
            if (s->d[0]) {
                s->d[0][i] = fmaf(s->phi, s->d[0][i], 1.0f - s->phi);
                s->d[1][i] = fmaf(s->phi, s->d[1][i], xi - s->mu);
                s->d[2][i] = fmaf(s->phi, s->d[2][i], z);
            }
            s->x[i] = transition(xi, s->mu, s->phi, s->sig, z);
        }
    }
    for (int i = 0; i < n; i++) wk->l[i] = orc_lw(s->x[i], ysq, o->nu, o->c_obs, o->nph);
    const float M = orc_tiled_weights(wk->l, n, g_tile, wk->tmax, wk->tf, wk->w);
    double S = 0.0, Q = 0.0;
    for (int i = 0; i < n; i++) {
        S += wk->w[i];
        Q += wk->w[i] * wk->w[i];
    }
    (void)Q;
    const int ok = isfinite(S) && S > 0.0 && isfinite((double)M);
    const double incr = ok ? (double)M + log(S / (double)n) : NAN;
    s->ll_cum += incr;
    if (ok && s->d[0]) {
        double sc[3] = {0.0, 0.0, 0.0};
        for (int i = 0; i < n; i++) {
            const double eg = wk->w[i] * (double)dlw_dx(s->x[i], ysq, o);
            for (int k = 0; k < 3; k++) sc[k] += eg * (double)s->d[k][i];
        }
        for (int k = 0; k < 3; k++) s->score[k] += sc[k] / S;
    }
    if (ok) {
        float ur;
        if (nz->mode == 2) {
            uint64_t a, g;
            orc_pcg_jump(nz->prop_seed, (uint32_t)t, &a, &g);
            const float v = orc_uniform(orc_pcg_output(a + (uint64_t)(2u * S2_MAXN + 1u) * g));
            float w = nz->u_in + nz->u_step * (2.0f * v - 1.0f);
            w = w - floorf(w);
            ur = (w < 1.0f) ? w : 0.0f;
        } else {
            ur = (float)(orc_pcg_host_next(&s->hrng) >> 9) * (1.0f / 8388608.0f);
        }
        if (nz->u_out) *nz->u_out = ur;
        const float u = ur / (float)n;
        orc_build_cdf(wk->w, n, wk->cdf);
        orc_resample_ancestors(wk->cdf, n, u, wk->anc);
        for (int i = 0; i < n; i++) wk->xt[i] = s->x[wk->anc[i]];
        memcpy(s->x, wk->xt, sizeof(float) * (size_t)n);
        if (s->d[0]) {
            for (int k = 0; k < 3; k++) {
                for (int i = 0; i < n; i++) wk->dt[k][i] = s->d[k][wk->anc[i]];
                memcpy(s->d[k], wk->dt[k], sizeof(float) * (size_t)n);
            }
        }
    }
    s->tick = t + 1;
    return incr;
}

static void slot_copy(Slot* dst, const Slot* src) {
    float* x = dst->x;
    float* d[3] = {dst->d[0], dst->d[1], dst->d[2]};
    *dst = *src;
    dst->x = x;
    memcpy(dst->x, src->x, sizeof(float) * (size_t)src->n);
    for (int k = 0; k < 3; k++) {
        dst->d[k] = d[k];
        if (d[k] && src->d[k]) memcpy(d[k], src->d[k], sizeof(float) * (size_t)src->n);
    }
}

/* ---- outer-layer proposals --------------------------------------------------- */

typedef struct { double A[9], L[9]; } Precond;

/* Cholesky factor of scale2 * (C + ridge): ridge 1e-12 + 1e-9 C_ii on the
 * diagonal; a diagonal factor if C is not positive definite. */
static void make_precond(const double cov[9], double scale2, Precond* P) {
    double C[9], L[9];
    for (int k = 0; k < 9; k++) C[k] = cov[k];
    for (int i = 0; i < 3; i++) C[4 * i] += 1e-12 + 1e-9 * C[4 * i];
    for (int k = 0; k < 9; k++) L[k] = 0.0;
    int ok = 1;
    for (int i = 0; i < 3 && ok; i++)
        for (int j = 0; j <= i; j++) {
            double sum = C[3 * i + j];
            for (int k = 0; k < j; k++) sum -= L[3 * i + k] * L[3 * j + k];
            if (i == j) {
                if (!(sum > 0.0)) {
                    ok = 0;
                    break;
                }
                L[3 * i + i] = sqrt(sum);
            } else {
                L[3 * i + j] = sum / L[3 * j + j];
            }
        }
    if (!ok) {
        for (int k = 0; k < 9; k++) {
            L[k] = 0.0;
            if (k % 4 != 0) C[k] = 0.0;
        }
        for (int i = 0; i < 3; i++) L[4 * i] = sqrt(fmax(C[4 * i], 1e-12));
    }
    const double sc = sqrt(scale2);
    for (int k = 0; k < 9; k++) {
        P->L[k] = sc * L[k];
        P->A[k] = scale2 * C[k];
    }
}

static void mala_mean(const Precond* P, double h, const double th[3], const double g[3], double out[3]) {
    for (int i = 0; i < 3; i++) {
        double ag = 0.0;
        for (int k = 0; k < 3; k++) ag += P->A[3 * i + k] * g[k];
        out[i] = th[i] + 0.5 * h * ag;
    }
}

static double mala_logq(const Precond* P, double h, const double from[3], const double gf[3],
                        const double to[3]) {
    double mean[3], d[3], r[3];
    mala_mean(P, h, from, gf, mean);
    for (int i = 0; i < 3; i++) d[i] = to[i] - mean[i];
    for (int i = 0; i < 3; i++) {
        double v = d[i];
        for (int k = 0; k < i; k++) v -= P->L[3 * i + k] * r[k];
        r[i] = v / P->L[3 * i + i];
    }
    return -(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) / (2.0 * h);
}

static void propose(int mala, const Precond* P, double h, const double th[3], const double g[3],
                    const double xi[3], double out[3]) {
    double base[3];
    if (mala) mala_mean(P, h, th, g, base);
    else for (int i = 0; i < 3; i++) base[i] = th[i];
    const double sc = mala ? sqrt(h) : 1.0;
    for (int i = 0; i < 3; i++) {
        double d = 0.0;
        for (int k = 0; k <= i; k++) d += P->L[3 * i + k] * xi[k];
        out[i] = base[i] + sc * d;
    }
}

/* Systematic resampling of outer log-weights; -1 when none is finite. */
static int outer_ancestors(const double* lw, int M, double u, int* anc, double* cdf) {
    double mx = -INFINITY;
    for (int m = 0; m < M; m++)
        if (isfinite(lw[m]) && lw[m] > mx) mx = lw[m];
    if (!isfinite(mx)) return -1;
    double acc = 0.0;
    int last = 0;
    for (int m = 0; m < M; m++) {
        const double w = isfinite(lw[m]) ? exp(lw[m] - mx) : 0.0;
        acc += w;
        cdf[m] = acc;
        if (w > 0.0) last = m;
    }
    int j = 0;
    for (int m = 0; m < M; m++) {
        const double t = (g_sfault == 3 ? u : u + m) / M * acc;
        while (j < last && cdf[j] < t) ++j;
        anc[m] = j;
    }
    return 0;
}

/* ---- the run ----------------------------------------------------------------- */

typedef struct {
    int n_theta, n_min, n_max, seed;
    int kind;            /* 0 CPMMH, 1 MALA */
    int moves;
    float nu_obs, cn_rho, u_step, prop_scale, mala_h, outer_thr;
    float lo[3], hi[3];
    int tile;            /* particles per tile of the inner filters (0: 256) */
} OrcSmc2Cfg;

typedef struct {
    long long ticks, outer_resamples, proposed, accepted, replay_ticks, pt_forward, pt_replay;
} OrcSmc2Count;

static void slots_free(Slot* s, int M) {
    if (!s) return;
    for (int m = 0; m < M; m++) {
        free(s[m].x);
        for (int k = 0; k < 3; k++) free(s[m].d[k]);
    }
    free(s);
}

static Slot* slots_alloc(int M, int cap, int tangents) {
    Slot* s = (Slot*)calloc((size_t)M, sizeof(Slot));
    if (!s) return NULL;
    for (int m = 0; m < M; m++) {
        s[m].x = (float*)calloc((size_t)cap, sizeof(float));
        if (!s[m].x) { slots_free(s, M); return NULL; }
        for (int k = 0; k < 3 && tangents; k++) {
            s[m].d[k] = (float*)calloc((size_t)cap, sizeof(float));
            if (!s[m].d[k]) { slots_free(s, M); return NULL; }
        }
    }
    return s;
}

/* Whole run over T observations.  Outputs (each nullable): theta [3 M],
 * particle counts [M], outer log-weights [M], inner FP64 log-likelihoods [M],
 * scores [3 M] (MALA), outer ESS after each tick [T] (before any resample),
 * the outer-resample history (tick, total particles) up to hist_cap entries.
 * Returns the number of outer resamples, or -1. */
int orc_sv_smc2_run(const OrcSmc2Cfg* c, const float* y, int T, float* theta_out, int* n_out,
                    double* logw_out, double* ll_out, double* score_out, double* ess_out,
                    int* hist_tick, long long* hist_sum_n, int hist_cap, OrcSmc2Count* cnt) {
    const int M = c->n_theta;
    if (M <= 0 || c->n_min <= 0 || c->n_max < c->n_min || c->n_max > S2_MAXN || T < 0) return -1;
    if (c->tile != 0 && c->tile != 256 && c->tile != 512 && c->tile != 1024) return -1;
    g_tile = c->tile ? c->tile : S2_TILE;
    const int mala = c->kind == 1;
    const int cap = (c->n_max + S2_TILE - 1) / S2_TILE * S2_TILE;
    const int store = !mala;
    Obs o;
    o.nu = c->nu_obs > 0.0f ? c->nu_obs : 0.0f;
    o.c_obs = o.nu > 0.0f ? (float)orc_c_obs((double)o.nu) : 0.0f;
    o.nph = (o.nu + 1.0f) * 0.5f;
    const float cn_s = sqrtf(fmaxf(1.0f - c->cn_rho * c->cn_rho, 0.0f));

    Slot* main = slots_alloc(M, cap, mala);
    Slot* rep = slots_alloc(M, cap, mala);
    Slot* tmp = slots_alloc(M, cap, mala);
    double* logw = (double*)calloc((size_t)M, sizeof(double));
    double* ocdf = (double*)calloc((size_t)M, sizeof(double));
    int* anc = (int*)calloc((size_t)M, sizeof(int));
    int* zcur = (int*)calloc((size_t)M, sizeof(int));
    int* pseed = (int*)calloc((size_t)M, sizeof(int));
    int* acc = (int*)calloc((size_t)M, sizeof(int));
    const size_t zrows = store ? (size_t)(T > 0 ? T : 1) : 1;
    float* zpool[2] = {NULL, NULL};
    float* upool[2] = {NULL, NULL};
    for (int p = 0; p < 2 && store; p++) {
        zpool[p] = (float*)calloc(zrows * (size_t)M * cap, sizeof(float));
        upool[p] = (float*)calloc(zrows * (size_t)M, sizeof(float));
    }
    Work wk;
    memset(&wk, 0, sizeof wk);
    wk.l = (float*)malloc(sizeof(float) * S2_MAXN);
    wk.w = (double*)malloc(sizeof(double) * S2_MAXN);
    wk.cdf = (double*)malloc(sizeof(double) * S2_MAXN);
    wk.anc = (int*)malloc(sizeof(int) * S2_MAXN);
    wk.xt = (float*)malloc(sizeof(float) * S2_MAXN);
    for (int k = 0; k < 3; k++) wk.dt[k] = (float*)malloc(sizeof(float) * S2_MAXN);
    int rc = 0;
    if (!main || !rep || !tmp || !logw || !ocdf || !anc || !zcur || !pseed || !acc ||
        (store && (!zpool[0] || !zpool[1] || !upool[0] || !upool[1])) || !wk.l || !wk.w || !wk.cdf ||
        !wk.anc || !wk.xt || !wk.dt[0] || !wk.dt[1] || !wk.dt[2]) {
        rc = -1;
        goto done;
    }
#define ZROW(p, t, m) (zpool[p] + ((size_t)(t) * M + (size_t)(m)) * cap)
#define UCELL(p, t, m) (upool[p] + (size_t)(t) * M + (size_t)(m))

    OrcSmc2Count k0;
    memset(&k0, 0, sizeof k0);
    Hrng rng, mrng;
    hr_seed(&rng, (uint64_t)(int64_t)c->seed, 0x5eed5c2ull);
    hr_seed(&mrng, (uint64_t)(int64_t)c->seed, 0x30e5ull);

    /* prior draws */
    for (int m = 0; m < M; m++) {
        Slot* s = &main[m];
        const double lu = log((double)c->n_min) + hr_uniform(&rng) * (log((double)c->n_max) - log((double)c->n_min));
        long r = lround(exp(lu));
        int nn = (int)r;
        if (nn < c->n_min) nn = c->n_min;
        if (nn > c->n_max) nn = c->n_max;
        s->n = nn;
        s->mu = (float)(c->lo[0] + hr_uniform(&rng) * (c->hi[0] - c->lo[0]));
        s->phi = (float)(c->lo[1] + hr_uniform(&rng) * (c->hi[1] - c->lo[1]));
        s->sig = (float)(c->lo[2] + hr_uniform(&rng) * (c->hi[2] - c->lo[2]));
        s->seed = (int)(hr_next(&rng) & 0x7fffffffu);
        s->hrng = (uint64_t)(int64_t)s->seed * 67890ULL + 12345ULL;
        Noise nz;
        memset(&nz, 0, sizeof nz);
        nz.mode = store ? 1 : 0;
        nz.z_out = store ? ZROW(0, 0, m) : NULL;
        slot_init(s, &nz);
    }

    int nres = 0;
    for (int t = 0; t < T;) {
        long long sum_n = 0;
        for (int m = 0; m < M; m++) {
            Noise nz;
            memset(&nz, 0, sizeof nz);
            nz.mode = store ? 1 : 0;
            if (store) {
                nz.z_out = ZROW(zcur[m], t, m);
                nz.u_out = UCELL(zcur[m], t, m);
            }
            logw[m] += slot_tick(&main[m], y[t], &o, &nz, &wk);
            sum_n += main[m].n;
        }
        k0.pt_forward += sum_n;
        k0.ticks += 1;
        t += 1;
        /* outer ESS */
        double mx = -INFINITY, S = 0.0, Q = 0.0;
        for (int m = 0; m < M; m++) mx = fmax(mx, logw[m]);
        for (int m = 0; m < M; m++) {
            const double w = exp(logw[m] - mx);
            S += w;
            Q += w * w;
        }
        const double ess = (S * S) / Q;
        if (ess_out) ess_out[t - 1] = ess;
        if (ess >= (double)c->outer_thr * (double)M) continue;

        /* ---- outer resample after `rows` ticks ---- */
        const int rows = t;
        const double u = hr_uniform(&rng);
        if (outer_ancestors(logw, M, u, anc, ocdf)) {
            rc = -1;
            goto done;
        }
        for (int m = 0; m < M; m++) {
            const int src = anc[m];
            slot_copy(&tmp[m], &main[src]);
            if (store) {
                for (int r = 0; r < rows; r++) {
                    memcpy(ZROW(1 - zcur[m], r, m), ZROW(zcur[src], r, src), sizeof(float) * (size_t)main[src].n);
                    *UCELL(1 - zcur[m], r, m) = *UCELL(zcur[src], r, src);
                }
            }
        }
        {
            Slot* sw = main;
            main = tmp;
            tmp = sw;
        }
        for (int m = 0; m < M; m++) {
            main[m].seed = (int)(hr_next(&rng) & 0x7fffffffu);
            if (store) zcur[m] = 1 - zcur[m];
        }
        long long sn = 0;
        for (int m = 0; m < M; m++) {
            main[m].hrng = (uint64_t)(int64_t)main[m].seed * 67890ULL + 12345ULL;
            logw[m] = 0.0;
            sn += main[m].n;
        }
        if (hist_tick && nres < hist_cap) hist_tick[nres] = rows;
        if (hist_sum_n && nres < hist_cap) hist_sum_n[nres] = sn;
        nres++;
        k0.outer_resamples += 1;

        /* ---- K moves ---- */
        for (int mv = 0; mv < c->moves; mv++) {
            double mean[3] = {0.0, 0.0, 0.0}, cov[9];
            for (int m = 0; m < M; m++) {
                mean[0] += main[m].mu;
                mean[1] += main[m].phi;
                mean[2] += main[m].sig;
            }
            for (int k = 0; k < 3; k++) mean[k] /= M;
            for (int k = 0; k < 9; k++) cov[k] = 0.0;
            for (int m = 0; m < M; m++) {
                const double d[3] = {main[m].mu - mean[0], main[m].phi - mean[1], main[m].sig - mean[2]};
                for (int i = 0; i < 3; i++)
                    for (int j = 0; j < 3; j++) cov[3 * i + j] += d[i] * d[j];
            }
            for (int k = 0; k < 9; k++) cov[k] /= (M > 1 ? M - 1 : 1);
            Precond P;
            make_precond(cov, mala ? 1.0 : (double)c->prop_scale * c->prop_scale * 2.38 * 2.38 / 3.0, &P);
            const double h = c->mala_h;
            for (int m = 0; m < M; m++) {
                double xi[3];
                xi[0] = hr_normal(&mrng);
                xi[1] = hr_normal(&mrng);
                xi[2] = hr_normal(&mrng);
                const double th[3] = {main[m].mu, main[m].phi, main[m].sig};
                const double g0[3] = {0.0, 0.0, 0.0};
                double tp[3];
                propose(mala, &P, h, th, mala ? main[m].score : g0, xi, tp);
                Slot* r = &rep[m];
                r->mu = (float)tp[0];
                r->phi = (float)tp[1];
                r->sig = (float)tp[2];
                r->n = main[m].n;
                pseed[m] = (int)(hr_next(&mrng) & 0x7fffffffu);
                r->seed = mala ? pseed[m] : main[m].seed;
                r->hrng = (uint64_t)(int64_t)r->seed * 67890ULL + 12345ULL;
                Noise nz;
                memset(&nz, 0, sizeof nz);
                if (store) {
                    nz.mode = 2;
                    nz.z_in = ZROW(zcur[m], 0, m);
                    nz.z_out = ZROW(1 - zcur[m], 0, m);
                    nz.prop_seed = pseed[m];
                    nz.cn_rho = c->cn_rho;
                    nz.cn_s = cn_s;
                }
                slot_init(r, &nz);
            }
            for (int t2 = 0; t2 < rows; t2++) {
                for (int m = 0; m < M; m++) {
                    Noise nz;
                    memset(&nz, 0, sizeof nz);
                    if (store) {
                        nz.mode = 2;
                        nz.z_in = ZROW(zcur[m], t2, m);
                        nz.z_out = ZROW(1 - zcur[m], t2, m);
                        nz.u_in = *UCELL(zcur[m], t2, m);
                        nz.u_out = UCELL(1 - zcur[m], t2, m);
                        nz.prop_seed = pseed[m];
                        nz.cn_rho = c->cn_rho;
                        nz.cn_s = cn_s;
                        nz.u_step = c->u_step;
                    }
                    (void)slot_tick(&rep[m], y[t2], &o, &nz, &wk);
                }
            }
            int n_acc = 0;
            long long np = 0;
            for (int m = 0; m < M; m++) {
                const Slot* r = &rep[m];
                const double th[3] = {main[m].mu, main[m].phi, main[m].sig};
                const double tp[3] = {r->mu, r->phi, r->sig};
                const double llc = main[m].ll_cum, llp = r->ll_cum;
                double la = llp - llc;
                if (mala && g_sfault != 1)
                    la += mala_logq(&P, h, tp, r->score, th) - mala_logq(&P, h, th, main[m].score, tp);
                const double lu = log(hr_uniform(&mrng));
                const int inside = r->mu >= c->lo[0] && r->mu <= c->hi[0] && r->phi >= c->lo[1] &&
                                   r->phi <= c->hi[1] && r->sig >= c->lo[2] && r->sig <= c->hi[2];
                acc[m] = inside && isfinite(llp) && (lu < la || !isfinite(llc));
            }
            for (int m = 0; m < M; m++) {
                if (acc[m]) {
                    const uint64_t keep = main[m].hrng;
                    slot_copy(&main[m], &rep[m]);
                    if (!mala) {
                        main[m].hrng = keep;      /* the forward stream continues */
                        zcur[m] = 1 - zcur[m];
                    }
                    ++n_acc;
                }
                np += main[m].n;
            }
            k0.proposed += M;
            k0.accepted += n_acc;
            k0.replay_ticks += rows;
            k0.pt_replay += np * rows;
        }
    }
    rc = nres;
    for (int m = 0; m < M; m++) {
        if (theta_out) {
            theta_out[3 * m] = main[m].mu;
            theta_out[3 * m + 1] = main[m].phi;
            theta_out[3 * m + 2] = main[m].sig;
        }
        if (n_out) n_out[m] = main[m].n;
        if (logw_out) logw_out[m] = logw[m];
        if (ll_out) ll_out[m] = main[m].ll_cum;
        if (score_out)
            for (int k = 0; k < 3; k++) score_out[3 * m + k] = main[m].score[k];
    }
    if (cnt) *cnt = k0;
done:
    slots_free(main, M);
    slots_free(rep, M);
    slots_free(tmp, M);
    free(logw); free(ocdf); free(anc); free(zcur); free(pseed); free(acc);
    for (int p = 0; p < 2; p++) { free(zpool[p]); free(upool[p]); }
    free(wk.l); free(wk.w); free(wk.cdf); free(wk.anc); free(wk.xt);
    for (int k = 0; k < 3; k++) free(wk.dt[k]);
    return rc;
#undef ZROW
#undef UCELL
}

/* One inner filter with the tangents and score, on PCG32 noise (mode 0):
 * the pieces above on their own, for pinning against the single-filter
 * oracle and against finite differences.  Returns the log-likelihood;
 * score (nullable) receives the pathwise score, x_out (nullable) the
 * particles. */
double orc_sv_inner_run(int n, float mu, float phi, float sig, float nu_obs, int seed, const float* y,
                        int T, double* score, float* x_out) {
    if (n <= 0 || n > S2_MAXN) return NAN;
    Slot* s = slots_alloc(1, S2_MAXN, 1);
    Work wk;
    memset(&wk, 0, sizeof wk);
    wk.l = (float*)malloc(sizeof(float) * S2_MAXN);
    wk.w = (double*)malloc(sizeof(double) * S2_MAXN);
    wk.cdf = (double*)malloc(sizeof(double) * S2_MAXN);
    wk.anc = (int*)malloc(sizeof(int) * S2_MAXN);
    wk.xt = (float*)malloc(sizeof(float) * S2_MAXN);
    for (int k = 0; k < 3; k++) wk.dt[k] = (float*)malloc(sizeof(float) * S2_MAXN);
    double ll = NAN;
    if (s && wk.l && wk.w && wk.cdf && wk.anc && wk.xt && wk.dt[0] && wk.dt[1] && wk.dt[2]) {
        Obs o;
        o.nu = nu_obs > 0.0f ? nu_obs : 0.0f;
        o.c_obs = o.nu > 0.0f ? (float)orc_c_obs((double)o.nu) : 0.0f;
        o.nph = (o.nu + 1.0f) * 0.5f;
        s->n = n;
        s->mu = mu;
        s->phi = phi;
        s->sig = sig;
        s->seed = seed;
        s->hrng = (uint64_t)(int64_t)seed * 67890ULL + 12345ULL;
        Noise nz;
        memset(&nz, 0, sizeof nz);
        slot_init(s, &nz);
        for (int t = 0; t < T; t++) (void)slot_tick(s, y[t], &o, &nz, &wk);
        ll = s->ll_cum;
        if (score)
            for (int k = 0; k < 3; k++) score[k] = s->score[k];
        if (x_out) memcpy(x_out, s->x, sizeof(float) * (size_t)n);
    }
    slots_free(s, 1);
    free(wk.l); free(wk.w); free(wk.cdf); free(wk.anc); free(wk.xt);
    for (int k = 0; k < 3; k++) free(wk.dt[k]);
    return ll;
}

/* The outer systematic resampling on its own (for pinning by brute force). */
int orc_smc2_ancestors(const double* lw, int M, double u, int* anc) {
    double* cdf = (double*)malloc(sizeof(double) * (size_t)(M > 0 ? M : 1));
    if (!cdf || M <= 0) {
        free(cdf);
        return -1;
    }
    const int rc = outer_ancestors(lw, M, u, anc, cdf);
    free(cdf);
    return rc;
}

/* The moves' code on closed-form targets.  kind 0 (random walk) / 1 (MALA):
 * a chain on N(m, C) in three dimensions, the exact log-density as the
 * log-likelihood and its gradient as the score, A = C; out = mean[3],
 * cov[9], acceptance rate.  kind 2: the Crank-Nicolson noise move
 * z <- cn(z, xi) with rho = 0.6 iterated from z = 0, which must keep N(0, 1)
 * invariant; out = mean, variance, lag-1 correlation of z (expected
 * 0, 1, rho). */
int orc_smc2_selftest(int kind, int iters, int seed, double* out) {
    if (iters <= 0 || !out || kind < 0 || kind > 2) return -1;
    Hrng r;
    hr_seed(&r, (uint64_t)(int64_t)seed, 0x7e57ull);
    if (kind == 2) {
        Noise nz;
        memset(&nz, 0, sizeof nz);
        nz.cn_rho = 0.6f;
        nz.cn_s = sqrtf(fmaxf(1.0f - nz.cn_rho * nz.cn_rho, 0.0f));
        float z = 0.0f;
        double s1 = 0.0, s2 = 0.0, s12 = 0.0;
        for (int it = 0; it < iters; it++) {
            const float zn = cn_step(&nz, (float)hr_normal(&r), z);
            s1 += zn;
            s2 += (double)zn * zn;
            s12 += (double)zn * z;
            z = zn;
        }
        const double m = s1 / iters;
        out[0] = m;
        out[1] = s2 / iters - m * m;
        out[2] = (s12 / iters - m * m) / out[1];
        return 0;
    }
    const int mala = kind == 1;
    const double mu0[3] = {-1.0, 0.95, 0.2};
    const double C[9] = {0.04, 0.002, -0.003, 0.002, 0.0004, 0.0001, -0.003, 0.0001, 0.0025};
    /* C^-1 by Gauss-Jordan elimination */
    double Ci[9], A[18];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 6; j++) A[6 * i + j] = j < 3 ? C[3 * i + j] : (j - 3 == i ? 1.0 : 0.0);
    for (int p = 0; p < 3; p++) {
        const double piv = A[6 * p + p];
        for (int j = 0; j < 6; j++) A[6 * p + j] /= piv;
        for (int i = 0; i < 3; i++) {
            if (i == p) continue;
            const double f = A[6 * i + p];
            for (int j = 0; j < 6; j++) A[6 * i + j] -= f * A[6 * p + j];
        }
    }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) Ci[3 * i + j] = A[6 * i + 3 + j];
    Precond P;
    make_precond(C, mala ? 1.0 : 2.38 * 2.38 / 3.0, &P);
    const double h = 0.8;
    double th[3] = {mu0[0], mu0[1], mu0[2]}, g[3], lp = 0.0;
    {   /* log-density and gradient at the start */
        double d[3] = {th[0] - mu0[0], th[1] - mu0[1], th[2] - mu0[2]}, q = 0.0;
        for (int i = 0; i < 3; i++) {
            double v = 0.0;
            for (int k = 0; k < 3; k++) v += Ci[3 * i + k] * d[k];
            g[i] = -v;
            q += d[i] * v;
        }
        lp = -0.5 * q;
    }
    double s1[3] = {0, 0, 0}, s2[9] = {0};
    long long acc = 0;
    for (int it = 0; it < iters; it++) {
        double xi[3];
        xi[0] = hr_normal(&r);
        xi[1] = hr_normal(&r);
        xi[2] = hr_normal(&r);
        double tp[3], gp[3], q = 0.0;
        propose(mala, &P, h, th, g, xi, tp);
        double d[3] = {tp[0] - mu0[0], tp[1] - mu0[1], tp[2] - mu0[2]};
        for (int i = 0; i < 3; i++) {
            double v = 0.0;
            for (int k = 0; k < 3; k++) v += Ci[3 * i + k] * d[k];
            gp[i] = -v;
            q += d[i] * v;
        }
        const double lpp = -0.5 * q;
        double la = lpp - lp;
        if (mala && g_sfault != 1) la += mala_logq(&P, h, tp, gp, th) - mala_logq(&P, h, th, g, tp);
        if (log(hr_uniform(&r)) < la) {
            for (int k = 0; k < 3; k++) {
                th[k] = tp[k];
                g[k] = gp[k];
            }
            lp = lpp;
            ++acc;
        }
        for (int i = 0; i < 3; i++) {
            s1[i] += th[i];
            for (int j = 0; j < 3; j++) s2[3 * i + j] += th[i] * th[j];
        }
    }
    for (int i = 0; i < 3; i++) out[i] = s1[i] / iters;
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) out[3 + 3 * i + j] = s2[3 * i + j] / iters - out[i] * out[j];
    out[12] = (double)acc / iters;
    return 0;
}
