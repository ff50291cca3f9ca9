// gpu_smc2_api.hip -- SMC^2 over the batched inner filters (the gpu_smc2_*
// entry points of include/gpu_bpf_batch.h).
//
// theta-particles (mu, phi, sigma_eta) each drive one inner SV filter with
// its own particle count, all advanced by one launch per tick.  After every
// tick the outer log-weights (sums of the inner log-likelihood increments,
// smc2/smc2_rbpf_batch.cu:385-397) are checked on the device; when the outer
// ESS falls below the threshold the check raises a halt flag, the remaining
// launches of the captured run return at once, and the host resamples the
// theta-particles (systematic, smc2_rbpf_batch.cu:486-540) and rejuvenates
// them with K correlated pseudo-marginal MH moves (smc2_rbpf_batch.cu:
// 693-981): Gaussian random-walk theta proposals with the adaptive
// covariance 2.38^2/d * Cov(theta) (smc2_rbpf_batch.cu:1246-1294) and a
// Crank-Nicolson proposal of the stored inner noise (rho = 0.99), each move a
// batched replay of ticks 0..t.
#include <hip/hip_runtime.h>

#include <math.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "api_util.hpp"
#include "batch_host.hpp"

using namespace bpf;

namespace {

constexpr int kRunTicks = 32;         // [tick, outer check] pairs per captured graph

// Host PCG32 (XSH-RR, increment from the stream id) for the outer layer's
// uniforms, proposals and seeds.
struct HostRng {
    uint64_t s = 0, inc = 1;
    void seed(uint64_t seed, uint64_t stream) {
        s = 0;
        inc = (stream << 1u) | 1u;
        next();
        s += seed;
        next();
    }
    uint32_t next() {
        const uint64_t old = s;
        s = old * kPcgMult + inc;
        const uint32_t xs = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (xs >> rot) | (xs << ((0u - rot) & 31u));
    }
    double uniform() {      // in (0, 1), 53 bits
        const uint64_t a = next() >> 5, b = next() >> 6;
        return ((double)(a * 67108864ull + b) + 0.5) / 9007199254740992.0;
    }
    double normal() {
        const double u1 = uniform(), u2 = uniform();
        return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
    }
};

struct TickGraph {
    hipGraph_t g = nullptr;
    hipGraphExec_t e = nullptr;
    const float* y = nullptr;         // history pointer the capture used
    int block = 0, tile = 0;
    void reset() {
        if (e) (void)hipGraphExecDestroy(e);
        if (g) (void)hipGraphDestroy(g);
        e = nullptr;
        g = nullptr;
    }
};

}  // namespace

struct GpuSmc2 {
    int M = 0, nmin = 0, nmax = 0, seed = 0;
    float nu_obs = 0.0f;
    int kind = 0;                      // 0 = CPMMH, 1 = MALA
    int moves = 5;
    float cn_rho = 0.99f;
    float u_step = 0.05f;
    float prop_scale = 1.0f;
    float mala_h = 0.5f;
    float outer_thr = 0.5f;
    float lo[3] = {-2.0f, 0.90f, 0.05f};
    float hi[3] = {0.0f, 0.995f, 0.40f};
    int t_cap = 0;                     // ticks of noise the pools hold
    GpuBpfBatch* main = nullptr;       // the theta-particles' filters
    GpuBpfBatch* rep = nullptr;        // replays, and scratch for outer resampling
    double* d_logw = nullptr;
    int* d_halt = nullptr;             // [flag, ticks done when raised]
    double* d_ess = nullptr;
    float* d_zpool[2] = {nullptr, nullptr};
    float* d_upool[2] = {nullptr, nullptr};
    int* d_zcur = nullptr;
    int* d_prop_seed = nullptr;
    int* d_anc = nullptr;
    int* d_acc = nullptr;
    std::vector<int> zcur;
    HostRng rng;                       // prior draws, outer resampling, fresh seeds
    HostRng mrng;                      // moves: proposals, MH uniforms
    TickGraph run_graph, rep_graph;
    double last_ess = 0.0;
    long long ticks = 0, outer_resamples = 0, proposed = 0, accepted = 0;
    long long replay_ticks = 0, pt_forward = 0, pt_replay = 0;
    std::vector<int> hist_tick;        // per outer resample: ticks done when it happened
    std::vector<long long> hist_sum_n; // and the particle count of the resampled population
    bool ready = false;
};

static void smc2_free(GpuSmc2* s) {
    if (!s) return;
    if (s->main && s->main->stream) (void)hipStreamSynchronize(s->main->stream);
    s->run_graph.reset();
    s->rep_graph.reset();
    void* ptrs[] = {s->d_logw, s->d_halt, s->d_ess, s->d_zpool[0], s->d_zpool[1], s->d_upool[0],
                    s->d_upool[1], s->d_zcur, s->d_prop_seed, s->d_anc, s->d_acc};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    destroy_batch(s->main);
    destroy_batch(s->rep);
    delete s;
}

static bool in_prior(const GpuSmc2* s, const float* th) {
    for (int k = 0; k < 3; ++k)
        if (!(th[k] >= s->lo[k] && th[k] <= s->hi[k])) return false;
    return true;
}

static bool stores_noise(const GpuSmc2* s) { return s->kind == 0; }

// Launch arguments of the theta-particles' ticks (with the outer check) and
// of the replays.
static BatchArgs main_args(const GpuSmc2* s) {
    BatchArgs a = tick_args(s->main);
    a.noise = stores_noise(s) ? kNoisePcgStore : kNoisePcg;
    a.zpool[0] = s->d_zpool[0];
    a.zpool[1] = s->d_zpool[1];
    a.upool[0] = s->d_upool[0];
    a.upool[1] = s->d_upool[1];
    a.zcur = s->d_zcur;
    a.outer_logw = s->d_logw;
    a.halt = s->d_halt;
    return a;
}

static BatchArgs rep_args(const GpuSmc2* s) {
    BatchArgs a = tick_args(s->rep);
    a.y = s->main->d_y;
    a.noise = stores_noise(s) ? kNoisePropose : kNoisePcg;
    a.zpool[0] = s->d_zpool[0];
    a.zpool[1] = s->d_zpool[1];
    a.upool[0] = s->d_upool[0];
    a.upool[1] = s->d_upool[1];
    a.zcur = s->d_zcur;
    a.cn_rho = s->cn_rho;
    a.cn_s = sqrtf(fmaxf(1.0f - s->cn_rho * s->cn_rho, 0.0f));
    a.u_step = s->u_step;
    a.prop_seed = s->d_prop_seed;
    return a;
}

static int enqueue_run(GpuSmc2* s, TickGraph& tg, const BatchArgs& a, bool check, int n_ticks,
                       hipStream_t st) {
    GpuBpfBatch* b = check ? s->main : s->rep;
    if (b->block == 0) b->block = 256;
    if (b->tile == 0) b->tile = kBatchTile;
    if (!batch_tick_fits(b->block, b->tile, a.score != nullptr)) {
        set_err("gpu_smc2: %d threads with tiles of %d particles need more LDS than one CU holds", b->block,
                b->tile);
        return -1;
    }
    if (n_ticks >= kRunTicks && (!tg.e || tg.y != a.y || tg.block != b->block || tg.tile != b->tile)) {
        tg.reset();
        HIP_TRY(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        hipError_t e = hipSuccess;
        for (int k = 0; k < kRunTicks && e == hipSuccess; ++k) {
            e = launch_batch_tick(a, s->M, b->block, b->tile, st);
            if (e == hipSuccess && check)
                e = launch_outer_check(s->d_logw, s->main->d_ctl, s->M, s->outer_thr, s->d_halt, s->d_ess, st);
        }
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(st, &g);
        if (e != hipSuccess || e2 != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            set_err("smc2 graph capture failed: %s", hipGetErrorString(e != hipSuccess ? e : e2));
            return -1;
        }
        tg.g = g;
        HIP_TRY(hipGraphInstantiate(&tg.e, g, nullptr, nullptr, 0));
        tg.y = a.y;
        tg.block = b->block;
        tg.tile = b->tile;
    }
    int done = 0;
    while (n_ticks - done >= kRunTicks) {
        HIP_TRY(hipGraphLaunch(tg.e, st));
        done += kRunTicks;
    }
    for (; done < n_ticks; ++done) {
        HIP_TRY(launch_batch_tick(a, s->M, b->block, b->tile, st));
        if (check) HIP_TRY(launch_outer_check(s->d_logw, s->main->d_ctl, s->M, s->outer_thr, s->d_halt, s->d_ess, st));
    }
    return 0;
}

static int upload_slots(GpuSmc2* s, GpuBpfBatch* b) {
    const int M = s->M;
    hipStream_t st = b->stream;
    HIP_TRY(hipMemcpyAsync(b->d_n, b->n.data(), M * sizeof(int), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(b->d_mu, b->mu.data(), M * sizeof(float), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(b->d_phi, b->phi.data(), M * sizeof(float), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(b->d_sig, b->sig.data(), M * sizeof(float), hipMemcpyHostToDevice, st));
    HIP_TRY(hipMemcpyAsync(b->d_seed, b->seed.data(), M * sizeof(int), hipMemcpyHostToDevice, st));
    b->order.resize(M);
    std::iota(b->order.begin(), b->order.end(), 0);
    std::stable_sort(b->order.begin(), b->order.end(), [&](int i, int j) { return b->n[i] > b->n[j]; });
    HIP_TRY(hipMemcpyAsync(b->d_order, b->order.data(), M * sizeof(int), hipMemcpyHostToDevice, st));
    return 0;
}

static int fresh_ctl(GpuSmc2* s, GpuBpfBatch* b) {
    std::vector<Ctl> c(s->M);
    memset((void*)c.data(), 0, s->M * sizeof(Ctl));
    for (int m = 0; m < s->M; ++m) {
        c[m].host_rng = (unsigned long long)(long long)b->seed[m] * 67890ULL + 12345ULL;
        c[m].s_prev = 1.0;
    }
    HIP_TRY(hipMemcpyAsync(b->d_ctl, c.data(), s->M * sizeof(Ctl), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return 0;
}

static int read_ll(GpuBpfBatch* b, int M, std::vector<double>& ll) {
    std::vector<Ctl> c(M);
    HIP_TRY(hipStreamSynchronize(b->stream));
    HIP_TRY(hipMemcpy(c.data(), b->d_ctl, M * sizeof(Ctl), hipMemcpyDeviceToHost));
    ll.resize(M);
    for (int m = 0; m < M; ++m) ll[m] = c[m].ll_cum;
    return 0;
}

static void theta_moments(const GpuSmc2* s, double mean[3], double cov[9]) {
    const GpuBpfBatch* b = s->main;
    const int M = s->M;
    for (int k = 0; k < 3; ++k) mean[k] = 0.0;
    for (int m = 0; m < M; ++m) {
        mean[0] += b->mu[m];
        mean[1] += b->phi[m];
        mean[2] += b->sig[m];
    }
    for (int k = 0; k < 3; ++k) mean[k] /= M;
    for (int k = 0; k < 9; ++k) cov[k] = 0.0;
    for (int m = 0; m < M; ++m) {
        const double d[3] = {b->mu[m] - mean[0], b->phi[m] - mean[1], b->sig[m] - mean[2]};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) cov[3 * i + j] += d[i] * d[j];
    }
    for (int k = 0; k < 9; ++k) cov[k] /= (M > 1 ? M - 1 : 1);
}

// ---------------------------------------------------------------------------
// Proposals.  A = scale2 * Cov(theta) (+ a tiny ridge) with Cholesky factor
// L; random walk: theta' = theta + L xi; preconditioned MALA:
// theta' = theta + (h/2) A g + sqrt(h) L xi, log q = -|L^-1 (theta' - mean)|^2 / 2h.
// ---------------------------------------------------------------------------
struct Precond {
    double A[9], L[9];
};

static void make_precond(const double cov[9], double scale2, Precond& P) {
    // factor Cov + ridge, then scale (scale2 = 0 gives exactly theta' = theta)
    double C[9], L[9];
    for (int k = 0; k < 9; ++k) C[k] = cov[k];
    for (int i = 0; i < 3; ++i) C[4 * i] += 1e-12 + 1e-9 * C[4 * i];
    for (int k = 0; k < 9; ++k) L[k] = 0.0;
    bool ok = true;
    for (int i = 0; i < 3 && ok; ++i)
        for (int j = 0; j <= i; ++j) {
            double sum = C[3 * i + j];
            for (int k = 0; k < j; ++k) sum -= L[3 * i + k] * L[3 * j + k];
            if (i == j) {
                if (!(sum > 0.0)) {
                    ok = false;
                    break;
                }
                L[3 * i + i] = sqrt(sum);
            } else {
                L[3 * i + j] = sum / L[3 * j + j];
            }
        }
    if (!ok) {      // diagonal fallback
        for (int k = 0; k < 9; ++k) {
            L[k] = 0.0;
            if (k % 4 != 0) C[k] = 0.0;
        }
        for (int i = 0; i < 3; ++i) L[4 * i] = sqrt(std::max(C[4 * i], 1e-12));
    }
    const double sc = sqrt(scale2);
    for (int k = 0; k < 9; ++k) {
        P.L[k] = sc * L[k];
        P.A[k] = scale2 * C[k];
    }
}

static void mala_mean(const Precond& P, double h, const double th[3], const double g[3], double out[3]) {
    for (int i = 0; i < 3; ++i) {
        double ag = 0.0;
        for (int k = 0; k < 3; ++k) ag += P.A[3 * i + k] * g[k];
        out[i] = th[i] + 0.5 * h * ag;
    }
}

// log q(to | from) up to a constant (the same constant both ways).
static double mala_logq(const Precond& P, double h, const double from[3], const double g_from[3],
                        const double to[3]) {
    double mean[3], d[3], r[3];
    mala_mean(P, h, from, g_from, mean);
    for (int i = 0; i < 3; ++i) d[i] = to[i] - mean[i];
    for (int i = 0; i < 3; ++i) {        // forward substitution L r = d
        double v = d[i];
        for (int k = 0; k < i; ++k) v -= P.L[3 * i + k] * r[k];
        r[i] = v / P.L[3 * i + i];
    }
    return -(r[0] * r[0] + r[1] * r[1] + r[2] * r[2]) / (2.0 * h);
}

static void propose(int kind, const Precond& P, double h, const double th[3], const double g[3],
                    const double xi[3], double out[3]) {
    double base[3];
    if (kind == 1) {
        mala_mean(P, h, th, g, base);
    } else {
        for (int i = 0; i < 3; ++i) base[i] = th[i];
    }
    const double sc = (kind == 1) ? sqrt(h) : 1.0;
    for (int i = 0; i < 3; ++i) {
        double d = 0.0;
        for (int k = 0; k <= i; ++k) d += P.L[3 * i + k] * xi[k];
        out[i] = base[i] + sc * d;
    }
}

// log acceptance ratio of a move from (theta, ll, g) to (theta', ll', g'):
// the pseudo-marginal ratio, with the MALA proposal correction (omitted
// when `drop_q` -- a planted fault for the self test).
static double log_alpha(int kind, const Precond& P, double h, const double th[3], double ll,
                        const double g[3], const double tp[3], double llp, const double gp[3],
                        bool drop_q = false) {
    double la = llp - ll;
    if (kind == 1 && !drop_q) la += mala_logq(P, h, tp, gp, th) - mala_logq(P, h, th, g, tp);
    return la;
}

// Systematic resampling of n log-weights with one uniform u in (0, 1):
// slot m takes the first j whose running weight sum reaches (u + m) / n of the
// total.  Weights are exp(lw - max) over the finite entries; a non-finite
// log-weight (NaN from a NaN observation, -inf from a zero likelihood) has
// weight zero and is never an ancestor.  Returns -1 when no entry is finite.
static int systematic_ancestors(const double* lw, int n, double u, int* anc) {
    double mx = -INFINITY;
    for (int m = 0; m < n; ++m)
        if (std::isfinite(lw[m]) && lw[m] > mx) mx = lw[m];
    if (!std::isfinite(mx)) return -1;
    std::vector<double> cdf(n);
    double acc = 0.0;
    int last = 0;               // last entry of positive weight: the walk never passes it
    for (int m = 0; m < n; ++m) {
        const double w = std::isfinite(lw[m]) ? exp(lw[m] - mx) : 0.0;
        acc += w;
        cdf[m] = acc;
        if (w > 0.0) last = m;
    }
    int j = 0;
    for (int m = 0; m < n; ++m) {
        const double t = (u + m) / n * acc;
        while (j < last && cdf[j] < t) ++j;     // a zero weight never reaches t > 0 first
        anc[m] = j;
    }
    return 0;
}

// Outer systematic resampling after `rows` ticks: slot m takes slot anc[m]'s
// whole filter; future ticks of every slot draw from fresh streams.
static int outer_resample(GpuSmc2* s, int rows) {
    const int M = s->M;
    GpuBpfBatch* mb = s->main;
    GpuBpfBatch* rb = s->rep;
    hipStream_t st = mb->stream;
    std::vector<double> lw(M);
    HIP_TRY(hipMemcpy(lw.data(), s->d_logw, M * sizeof(double), hipMemcpyDeviceToHost));
    std::vector<int> anc(M);
    const double u = s->rng.uniform();
    if (systematic_ancestors(lw.data(), M, u, anc.data())) {
        set_err("gpu_smc2: every outer log-weight is non-finite at tick %d (NaN or impossible "
                "observations); the theta-particles cannot be resampled", rows);
        return -1;
    }
    HIP_TRY(hipMemcpyAsync(s->d_anc, anc.data(), M * sizeof(int), hipMemcpyHostToDevice, st));
    GatherArgs g;
    memset((void*)&g, 0, sizeof g);
    g.anc = s->d_anc;
    g.n_src = mb->d_n; g.mu_src = mb->d_mu; g.phi_src = mb->d_phi; g.sig_src = mb->d_sig;
    g.n_dst = rb->d_n; g.mu_dst = rb->d_mu; g.phi_dst = rb->d_phi; g.sig_dst = rb->d_sig;
    g.ctl_src = mb->d_ctl; g.ctl_dst = rb->d_ctl;
    g.x_src[0] = mb->d_x[0]; g.x_src[1] = mb->d_x[1]; g.x_dst = rb->d_x[0];
    g.l_src = mb->d_l; g.l_dst = rb->d_l;
    g.zpool[0] = s->d_zpool[0]; g.zpool[1] = s->d_zpool[1];
    g.upool[0] = s->d_upool[0]; g.upool[1] = s->d_upool[1];
    g.zcur = s->d_zcur;
    g.rows = stores_noise(s) ? rows : 0;
    g.cap = mb->cap;
    g.n_slots = M;
    if (mb->d_score) {
        for (int h = 0; h < 2; ++h)
            for (int k = 0; k < 3; ++k) g.dx_src[h][k] = mb->d_dx[h][k];
        for (int k = 0; k < 3; ++k) g.dx_dst[k] = rb->d_dx[0][k];
        g.score_src = mb->d_score;
        g.score_dst = rb->d_score;
    }
    HIP_TRY(launch_smc2_gather(g, st));
    if (mb->d_score) {
        for (int k = 0; k < 3; ++k)
            HIP_TRY(hipMemcpyAsync(mb->d_dx[0][k], rb->d_dx[0][k], (size_t)M * mb->cap * sizeof(float),
                                   hipMemcpyDeviceToDevice, st));
        HIP_TRY(hipMemcpyAsync(mb->d_score, rb->d_score, (size_t)M * 3 * sizeof(double),
                               hipMemcpyDeviceToDevice, st));
    }
    const size_t slots = (size_t)M * mb->cap;
    HIP_TRY(hipMemcpyAsync(mb->d_n, rb->d_n, M * sizeof(int), hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(mb->d_mu, rb->d_mu, M * sizeof(float), hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(mb->d_phi, rb->d_phi, M * sizeof(float), hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(mb->d_sig, rb->d_sig, M * sizeof(float), hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(mb->d_ctl, rb->d_ctl, M * sizeof(Ctl), hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(mb->d_x[0], rb->d_x[0], slots * sizeof(float), hipMemcpyDeviceToDevice, st));
    HIP_TRY(hipMemcpyAsync(mb->d_l, rb->d_l, slots * sizeof(float), hipMemcpyDeviceToDevice, st));
    // host mirrors, fresh streams for the ticks ahead, noise pools swapped
    std::vector<int> n(M);
    std::vector<float> mu(M), phi(M), sig(M);
    for (int m = 0; m < M; ++m) {
        n[m] = mb->n[anc[m]];
        mu[m] = mb->mu[anc[m]];
        phi[m] = mb->phi[anc[m]];
        sig[m] = mb->sig[anc[m]];
        mb->seed[m] = (int)(s->rng.next() & 0x7fffffffu);
        if (stores_noise(s)) s->zcur[m] = 1 - s->zcur[m];
    }
    mb->n = n;
    mb->mu = mu;
    mb->phi = phi;
    mb->sig = sig;
    if (upload_slots(s, mb)) return -1;
    HIP_TRY(hipMemcpyAsync(s->d_zcur, s->zcur.data(), M * sizeof(int), hipMemcpyHostToDevice, st));
    std::vector<Ctl> c(M);
    HIP_TRY(hipStreamSynchronize(st));
    HIP_TRY(hipMemcpy(c.data(), mb->d_ctl, M * sizeof(Ctl), hipMemcpyDeviceToHost));
    for (int m = 0; m < M; ++m) c[m].host_rng = (unsigned long long)(long long)mb->seed[m] * 67890ULL + 12345ULL;
    HIP_TRY(hipMemcpy(mb->d_ctl, c.data(), M * sizeof(Ctl), hipMemcpyHostToDevice));
    HIP_TRY(hipMemsetAsync(s->d_logw, 0, M * sizeof(double), st));
    s->outer_resamples += 1;
    long long sum_n = 0;
    for (int m = 0; m < M; ++m) sum_n += mb->n[m];
    s->hist_tick.push_back(rows);
    s->hist_sum_n.push_back(sum_n);
    return 0;
}

// One correlated pseudo-marginal MH move of every slot, replaying ticks
// 0..rows-1 with proposed parameters and noise.
static int read_score(GpuBpfBatch* b, int M, std::vector<double>& g) {
    g.assign((size_t)M * 3, 0.0);
    if (!b->d_score) return 0;
    HIP_TRY(hipStreamSynchronize(b->stream));
    HIP_TRY(hipMemcpy(g.data(), b->d_score, (size_t)M * 3 * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

// One pseudo-marginal MH move of every slot, replaying ticks 0..rows-1 with
// proposed parameters: CPMMH (random walk, Crank-Nicolson noise around the
// stored noise) or MALA (score-driven drift, fresh noise, replay with
// tangents for the reverse proposal density).
static int move(GpuSmc2* s, int rows) {
    const int M = s->M;
    const int kind = s->kind;
    GpuBpfBatch* mb = s->main;
    GpuBpfBatch* rb = s->rep;
    double mean[3], cov[9];
    theta_moments(s, mean, cov);
    Precond P;
    make_precond(cov, kind == 1 ? 1.0 : (double)s->prop_scale * s->prop_scale * 2.38 * 2.38 / 3.0, P);
    const double h = s->mala_h;
    std::vector<double> g;
    if (read_score(mb, M, g)) return -1;
    std::vector<int> pseed(M);
    for (int m = 0; m < M; ++m) {
        const double xi[3] = {s->mrng.normal(), s->mrng.normal(), s->mrng.normal()};
        const double th[3] = {mb->mu[m], mb->phi[m], mb->sig[m]};
        double tp[3];
        propose(kind, P, h, th, &g[3 * m], xi, tp);
        rb->mu[m] = (float)tp[0];
        rb->phi[m] = (float)tp[1];
        rb->sig[m] = (float)tp[2];
        rb->n[m] = mb->n[m];
        pseed[m] = (int)(s->mrng.next() & 0x7fffffffu);
        rb->seed[m] = (kind == 1) ? pseed[m] : mb->seed[m];
    }
    if (upload_slots(s, rb)) return -1;
    HIP_TRY(hipMemcpyAsync(s->d_prop_seed, pseed.data(), M * sizeof(int), hipMemcpyHostToDevice, rb->stream));
    if (fresh_ctl(s, rb)) return -1;
    if (rb->d_score) HIP_TRY(hipMemsetAsync(rb->d_score, 0, (size_t)M * 3 * sizeof(double), rb->stream));
    BatchInitArgs ia;
    memset((void*)&ia, 0, sizeof ia);
    ia.n = rb->d_n; ia.seed = rb->d_seed; ia.mu = rb->d_mu; ia.phi = rb->d_phi; ia.sig = rb->d_sig;
    ia.x0 = rb->d_x[0];
    ia.cap = rb->cap;
    ia.n_slots = M;
    ia.jm = kPcgMult * kPcgMult;
    ia.jg = 1u + kPcgMult;
    ia.noise = (kind == 1) ? kNoisePcg : kNoisePropose;
    ia.zpool[0] = s->d_zpool[0]; ia.zpool[1] = s->d_zpool[1];
    ia.zcur = s->d_zcur;
    ia.cn_rho = s->cn_rho;
    ia.cn_s = sqrtf(fmaxf(1.0f - s->cn_rho * s->cn_rho, 0.0f));
    ia.prop_seed = s->d_prop_seed;
    if (rb->d_score)
        for (int k = 0; k < 3; ++k) ia.dx0[k] = rb->d_dx[0][k];
    HIP_TRY(launch_batch_init(ia, rb->stream));
    if (enqueue_run(s, s->rep_graph, rep_args(s), false, rows, rb->stream)) return -1;
    std::vector<double> llp, llc, gp;
    if (read_ll(rb, M, llp) || read_ll(mb, M, llc) || read_score(rb, M, gp)) return -1;
    std::vector<Ctl> cm(M);
    HIP_TRY(hipMemcpy(cm.data(), mb->d_ctl, M * sizeof(Ctl), hipMemcpyDeviceToHost));
    std::vector<int> acc(M, 0);
    int n_acc = 0;
    for (int m = 0; m < M; ++m) {
        const double th[3] = {mb->mu[m], mb->phi[m], mb->sig[m]};
        const double tp[3] = {rb->mu[m], rb->phi[m], rb->sig[m]};
        const float tpf[3] = {rb->mu[m], rb->phi[m], rb->sig[m]};
        const double la = log_alpha(kind, P, h, th, llc[m], &g[3 * m], tp, llp[m], &gp[3 * m]);
        const double lu = log(s->mrng.uniform());
        if (in_prior(s, tpf) && std::isfinite(llp[m]) && (lu < la || !std::isfinite(llc[m]))) {
            acc[m] = 1;
            ++n_acc;
            mb->mu[m] = tpf[0];
            mb->phi[m] = tpf[1];
            mb->sig[m] = tpf[2];
            if (kind == 1) mb->seed[m] = rb->seed[m];
            else s->zcur[m] = 1 - s->zcur[m];
        }
    }
    hipStream_t st = mb->stream;
    HIP_TRY(hipMemcpyAsync(s->d_acc, acc.data(), M * sizeof(int), hipMemcpyHostToDevice, st));
    AcceptArgs a;
    memset((void*)&a, 0, sizeof a);
    a.acc = s->d_acc;
    a.mu_src = rb->d_mu; a.phi_src = rb->d_phi; a.sig_src = rb->d_sig;
    a.mu_dst = mb->d_mu; a.phi_dst = mb->d_phi; a.sig_dst = mb->d_sig;
    a.ctl_src = rb->d_ctl; a.ctl_dst = mb->d_ctl;
    a.x_src[0] = rb->d_x[0]; a.x_src[1] = rb->d_x[1];
    a.x_dst[0] = mb->d_x[0]; a.x_dst[1] = mb->d_x[1];
    a.l_src = rb->d_l; a.l_dst = mb->d_l;
    a.cap = mb->cap;
    if (mb->d_score) {
        for (int h2 = 0; h2 < 2; ++h2)
            for (int k = 0; k < 3; ++k) {
                a.dx_src[h2][k] = rb->d_dx[h2][k];
                a.dx_dst[h2][k] = mb->d_dx[h2][k];
            }
        a.score_src = rb->d_score;
        a.score_dst = mb->d_score;
    }
    if (kind == 1) {
        a.seed_src = rb->d_seed;
        a.seed_dst = mb->d_seed;
    }
    HIP_TRY(launch_smc2_accept(a, M, st));
    HIP_TRY(hipMemcpyAsync(s->d_zcur, s->zcur.data(), M * sizeof(int), hipMemcpyHostToDevice, st));
    HIP_TRY(hipStreamSynchronize(st));
    if (kind == 0) {
        // accepted slots continue their own host streams for future ticks
        std::vector<Ctl> c(M);
        HIP_TRY(hipMemcpy(c.data(), mb->d_ctl, M * sizeof(Ctl), hipMemcpyDeviceToHost));
        for (int m = 0; m < M; ++m)
            if (acc[m]) c[m].host_rng = cm[m].host_rng;
        HIP_TRY(hipMemcpy(mb->d_ctl, c.data(), M * sizeof(Ctl), hipMemcpyHostToDevice));
    }
    s->proposed += M;
    s->accepted += n_acc;
    s->replay_ticks += rows;
    long long np = 0;
    for (int m = 0; m < M; ++m) np += mb->n[m];
    s->pt_replay += np * rows;
    return 0;
}

static int rejuvenate(GpuSmc2* s, int rows) {
    if (outer_resample(s, rows)) return -1;
    for (int k = 0; k < s->moves; ++k)
        if (move(s, rows)) return -1;
    return 0;
}

// n_ticks more observations (host array).
static int smc2_run(GpuSmc2* s, const float* y, int n_ticks) {
    GpuBpfBatch* mb = s->main;
    if (stores_noise(s) && mb->timestep + n_ticks > s->t_cap) {
        set_err("gpu_smc2: %d ticks exceed the noise store of %d ticks (gpu_smc2_set_capacity)",
                mb->timestep + n_ticks, s->t_cap);
        return -1;
    }
    if (append_y(mb, y, n_ticks, hipMemcpyHostToDevice)) return -1;
    const int t_end = mb->timestep + n_ticks;
    int t = mb->timestep;
    while (t < t_end) {
        const int chunk = std::min(kRunTicks, t_end - t);
        if (enqueue_run(s, s->run_graph, main_args(s), true, chunk, mb->stream)) return -1;
        int halt[2];
        HIP_TRY(hipStreamSynchronize(mb->stream));
        HIP_TRY(hipMemcpy(halt, s->d_halt, sizeof halt, hipMemcpyDeviceToHost));
        HIP_TRY(hipMemcpy(&s->last_ess, s->d_ess, sizeof(double), hipMemcpyDeviceToHost));
        const int done = halt[0] ? halt[1] : t + chunk;
        long long np = 0;
        for (int m = 0; m < s->M; ++m) np += mb->n[m];
        s->pt_forward += np * (done - t);
        s->ticks += done - t;
        t = done;
        mb->timestep = t;
        if (halt[0]) {
            if (rejuvenate(s, t)) return -1;
            HIP_TRY(hipMemsetAsync(s->d_halt, 0, 2 * sizeof(int), mb->stream));
            s->last_ess = (double)s->M;
        }
    }
    HIP_TRY(hipStreamSynchronize(mb->stream));
    return 0;
}

extern "C" {

GpuSmc2* gpu_smc2_create(int n_theta, int n_min, int n_max, float nu_obs, int seed) {
    clear_err();
    if (n_theta <= 0 || n_theta > 65536 || n_min <= 0 || n_max < n_min || n_max > kBatchMaxN) {
        set_err("gpu_smc2_create: need 0 < n_theta <= 65536 and 0 < n_min <= n_max <= %d", kBatchMaxN);
        return nullptr;
    }
    GpuSmc2* s = new GpuSmc2();
    s->M = n_theta;
    s->nmin = n_min;
    s->nmax = n_max;
    s->seed = seed;
    s->nu_obs = nu_obs > 0.0f ? nu_obs : 0.0f;
    s->rng.seed((uint64_t)(int64_t)seed, 0x5eed5c2ull);
    s->mrng.seed((uint64_t)(int64_t)seed, 0x30e5ull);
    return s;
}

void gpu_smc2_destroy(GpuSmc2* s) { smc2_free(s); }

int gpu_smc2_set_prior(GpuSmc2* s, const float* lo, const float* hi) {
    if (!s || !lo || !hi || s->ready) {
        set_err("gpu_smc2_set_prior: bad arguments or already initialised");
        return -1;
    }
    for (int k = 0; k < 3; ++k) {
        if (!(lo[k] < hi[k])) {
            set_err("gpu_smc2_set_prior: empty interval for parameter %d", k);
            return -1;
        }
        s->lo[k] = lo[k];
        s->hi[k] = hi[k];
    }
    return 0;
}

int gpu_smc2_set_rejuvenation(GpuSmc2* s, int kind, int moves) {
    if (!s || s->ready || (kind != 0 && kind != 1) || moves < 0) {
        set_err("gpu_smc2_set_rejuvenation: kind 0 (CPMMH) or 1 (MALA), moves >= 0, before "
                "gpu_smc2_init_from_prior");
        return -1;
    }
    s->kind = kind;
    s->moves = moves;
    return 0;
}

int gpu_smc2_set_cpmmh_rho(GpuSmc2* s, float rho) {
    if (!s || !(rho >= 0.0f && rho <= 1.0f)) {
        set_err("gpu_smc2_set_cpmmh_rho: rho in [0, 1]");
        return -1;
    }
    s->cn_rho = rho;
    return 0;
}

int gpu_smc2_set_proposal_scale(GpuSmc2* s, float theta_scale, float u_step) {
    if (!s || !(theta_scale >= 0.0f) || !(u_step >= 0.0f && u_step <= 0.5f)) {
        set_err("gpu_smc2_set_proposal_scale: theta_scale >= 0, u_step in [0, 0.5]");
        return -1;
    }
    s->prop_scale = theta_scale;
    s->u_step = u_step;
    return 0;
}

int gpu_smc2_set_mala_step(GpuSmc2* s, float h) {
    if (!s || !(h > 0.0f)) {
        set_err("gpu_smc2_set_mala_step: h > 0");
        return -1;
    }
    s->mala_h = h;
    return 0;
}

int gpu_smc2_set_outer_threshold(GpuSmc2* s, float frac) {
    if (!s || !(frac >= 0.0f && frac <= 1.0f)) {
        set_err("gpu_smc2_set_outer_threshold: fraction in [0, 1]");
        return -1;
    }
    s->outer_thr = frac;
    return 0;
}

int gpu_smc2_set_capacity(GpuSmc2* s, int max_ticks) {
    if (!s || max_ticks <= 0 || s->ready) {
        set_err("gpu_smc2_set_capacity: positive, before gpu_smc2_init_from_prior");
        return -1;
    }
    s->t_cap = max_ticks;
    return 0;
}

int gpu_smc2_init_from_prior(GpuSmc2* s) {
    if (!s || s->ready) {
        set_err("gpu_smc2_init_from_prior: null or already initialised");
        return -1;
    }
    const int M = s->M;
    std::vector<int> n(M), seeds(M);
    std::vector<float> mu(M), phi(M), sig(M);
    for (int m = 0; m < M; ++m) {
        const double lu = log((double)s->nmin) + s->rng.uniform() * (log((double)s->nmax) - log((double)s->nmin));
        n[m] = std::min(s->nmax, std::max(s->nmin, (int)lround(exp(lu))));
        mu[m] = (float)(s->lo[0] + s->rng.uniform() * (s->hi[0] - s->lo[0]));
        phi[m] = (float)(s->lo[1] + s->rng.uniform() * (s->hi[1] - s->lo[1]));
        sig[m] = (float)(s->lo[2] + s->rng.uniform() * (s->hi[2] - s->lo[2]));
        seeds[m] = (int)(s->rng.next() & 0x7fffffffu);
    }
    // slots hold any particle count of the range (particle counts travel
    // with the theta-particles when they are resampled)
    std::vector<int> ncap(M, s->nmax);
    s->main = new_batch(M, ncap.data(), mu.data(), phi.data(), sig.data(), seeds.data(), s->nu_obs, 0.0f);
    s->main->n = n;
    s->rep = new_batch(M, ncap.data(), mu.data(), phi.data(), sig.data(), seeds.data(), s->nu_obs, 0.0f);
    s->rep->n = n;
    // the outer layer reads the inner filters' log-likelihoods only: no
    // per-tick mean, variance or ESS, and no statistics ring
    s->main->stats = 0;
    s->rep->stats = 0;
    if (create_impl(s->main) || create_impl(s->rep)) return -1;
    if (stores_noise(s)) {
        if (s->t_cap <= 0) s->t_cap = 2048;
        const size_t zn = (size_t)s->t_cap * M * s->main->cap;
        HIP_TRY(hipMalloc(&s->d_zpool[0], zn * sizeof(float)));
        HIP_TRY(hipMalloc(&s->d_zpool[1], zn * sizeof(float)));
        HIP_TRY(hipMalloc(&s->d_upool[0], (size_t)s->t_cap * M * sizeof(float)));
        HIP_TRY(hipMalloc(&s->d_upool[1], (size_t)s->t_cap * M * sizeof(float)));
    }
    HIP_TRY(hipMalloc(&s->d_logw, M * sizeof(double)));
    HIP_TRY(hipMalloc(&s->d_halt, 2 * sizeof(int)));
    HIP_TRY(hipMalloc(&s->d_ess, sizeof(double)));
    HIP_TRY(hipMalloc(&s->d_zcur, M * sizeof(int)));
    HIP_TRY(hipMalloc(&s->d_prop_seed, M * sizeof(int)));
    HIP_TRY(hipMalloc(&s->d_anc, M * sizeof(int)));
    HIP_TRY(hipMalloc(&s->d_acc, M * sizeof(int)));
    s->zcur.assign(M, 0);
    hipStream_t st = s->main->stream;
    HIP_TRY(hipMemsetAsync(s->d_logw, 0, M * sizeof(double), st));
    HIP_TRY(hipMemsetAsync(s->d_halt, 0, 2 * sizeof(int), st));
    HIP_TRY(hipMemsetAsync(s->d_zcur, 0, M * sizeof(int), st));
    if (upload_slots(s, s->main) || fresh_ctl(s, s->main)) return -1;
    if (!stores_noise(s) && (enable_score(s->main, 0) || enable_score(s->rep, 0))) return -1;
    // x_0 of every slot (stored as tick 0 of the noise when moves replay it)
    BatchInitArgs ia;
    memset((void*)&ia, 0, sizeof ia);
    ia.n = s->main->d_n; ia.seed = s->main->d_seed;
    ia.mu = s->main->d_mu; ia.phi = s->main->d_phi; ia.sig = s->main->d_sig;
    ia.x0 = s->main->d_x[0];
    ia.cap = s->main->cap;
    ia.n_slots = M;
    ia.jm = kPcgMult * kPcgMult;
    ia.jg = 1u + kPcgMult;
    ia.noise = stores_noise(s) ? kNoisePcgStore : kNoisePcg;
    ia.zpool[0] = s->d_zpool[0]; ia.zpool[1] = s->d_zpool[1];
    ia.zcur = s->d_zcur;
    if (s->main->d_score) {
        HIP_TRY(hipMemsetAsync(s->main->d_score, 0, (size_t)M * 3 * sizeof(double), st));
        for (int k = 0; k < 3; ++k) ia.dx0[k] = s->main->d_dx[0][k];
    }
    HIP_TRY(launch_batch_init(ia, st));
    HIP_TRY(hipStreamSynchronize(st));
    s->last_ess = (double)M;
    s->ready = true;
    return 0;
}

int gpu_smc2_run(GpuSmc2* s, const float* y, int n_ticks) {
    if (!s || !s->ready || (!y && n_ticks > 0) || n_ticks < 0) {
        set_err("gpu_smc2_run: not initialised or bad arguments");
        return -1;
    }
    return n_ticks ? smc2_run(s, y, n_ticks) : 0;
}

float gpu_smc2_update(GpuSmc2* s, float y) {
    if (gpu_smc2_run(s, &y, 1)) return -1.0f;
    return (float)s->last_ess;
}

float gpu_smc2_get_outer_ess(GpuSmc2* s) { return s ? (float)s->last_ess : -1.0f; }

int gpu_smc2_copy_theta(GpuSmc2* s, float* theta, double* logw, int* n_particles) {
    if (!s || !s->ready) {
        set_err("gpu_smc2_copy_theta: not initialised");
        return -1;
    }
    HIP_TRY(hipStreamSynchronize(s->main->stream));
    for (int m = 0; m < s->M; ++m) {
        if (theta) {
            theta[3 * m] = s->main->mu[m];
            theta[3 * m + 1] = s->main->phi[m];
            theta[3 * m + 2] = s->main->sig[m];
        }
        if (n_particles) n_particles[m] = s->main->n[m];
    }
    if (logw) HIP_TRY(hipMemcpy(logw, s->d_logw, s->M * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int gpu_smc2_get_loglik(GpuSmc2* s, double* ll) {
    if (!s || !s->ready || !ll) {
        set_err("gpu_smc2_get_loglik: not initialised or null");
        return -1;
    }
    std::vector<double> v;
    if (read_ll(s->main, s->M, v)) return -1;
    std::copy(v.begin(), v.end(), ll);
    return 0;
}

static int theta_stats(GpuSmc2* s, float* mean3, float* std3) {
    if (!s || !s->ready) {
        set_err("gpu_smc2: not initialised");
        return -1;
    }
    std::vector<double> lw(s->M);
    HIP_TRY(hipStreamSynchronize(s->main->stream));
    HIP_TRY(hipMemcpy(lw.data(), s->d_logw, s->M * sizeof(double), hipMemcpyDeviceToHost));
    // non-finite log-weights count as zero weight, as in outer resampling;
    // with none finite every theta-particle counts equally
    double mx = -INFINITY;
    for (double v : lw)
        if (std::isfinite(v) && v > mx) mx = v;
    double W = 0.0, m1[3] = {0, 0, 0}, m2[3] = {0, 0, 0};
    for (int m = 0; m < s->M; ++m) {
        const double w = !std::isfinite(mx) ? 1.0 : (std::isfinite(lw[m]) ? exp(lw[m] - mx) : 0.0);
        const double th[3] = {s->main->mu[m], s->main->phi[m], s->main->sig[m]};
        W += w;
        for (int k = 0; k < 3; ++k) {
            m1[k] += w * th[k];
            m2[k] += w * th[k] * th[k];
        }
    }
    for (int k = 0; k < 3; ++k) {
        const double mean = m1[k] / W;
        if (mean3) mean3[k] = (float)mean;
        if (std3) std3[k] = (float)sqrt(std::max(m2[k] / W - mean * mean, 0.0));
    }
    return 0;
}

int gpu_smc2_get_theta_mean(GpuSmc2* s, float* mean3) { return theta_stats(s, mean3, nullptr); }
int gpu_smc2_get_theta_std(GpuSmc2* s, float* std3) { return theta_stats(s, nullptr, std3); }

int gpu_smc2_get_counters(GpuSmc2* s, long long* out) {
    if (!s || !out) {
        set_err("gpu_smc2_get_counters: null");
        return -1;
    }
    const long long v[7] = {s->ticks, s->outer_resamples, s->proposed, s->accepted,
                            s->replay_ticks, s->pt_forward, s->pt_replay};
    memcpy(out, v, sizeof v);
    return 0;
}

int gpu_smc2_get_score(GpuSmc2* s, double* score) {
    if (!s || !s->ready || !score) {
        set_err("gpu_smc2_get_score: not initialised or null");
        return -1;
    }
    if (!s->main->d_score) {
        set_err("gpu_smc2_get_score: the theta-particles carry scores only with MALA rejuvenation");
        return -1;
    }
    std::vector<double> g;
    if (read_score(s->main, s->M, g)) return -1;
    std::copy(g.begin(), g.end(), score);
    return 0;
}

int gpu_smc2_get_history(GpuSmc2* s, int* ticks, long long* sum_n, int max_entries) {
    if (!s || max_entries < 0) {
        set_err("gpu_smc2_get_history: null state or negative size");
        return -1;
    }
    const int k = (int)s->hist_tick.size();
    for (int i = 0; i < k && i < max_entries; ++i) {
        if (ticks) ticks[i] = s->hist_tick[i];
        if (sum_n) sum_n[i] = s->hist_sum_n[i];
    }
    return k;
}

int gpu_smc2_systematic_ancestors(const double* log_w, int n, double u, int* anc) {
    if (!log_w || !anc || n <= 0 || !(u > 0.0 && u < 1.0)) {
        set_err("gpu_smc2_systematic_ancestors: need n > 0 entries and u in (0, 1)");
        return -1;
    }
    if (systematic_ancestors(log_w, n, u, anc)) {
        set_err("gpu_smc2_systematic_ancestors: no finite log-weight");
        return -1;
    }
    return 0;
}

// The move code on a closed-form target: N(m, C) in three dimensions with
// the exact log-density as "log-likelihood" and its gradient as "score",
// preconditioner A = C.  Writes [mean(3), cov(9), acceptance rate].  fault 1
// drops the MALA proposal correction (a planted mistake the check must see).
int gpu_smc2_selftest_mcmc(int kind, int iters, int seed, int fault, double* out) {
    if ((kind != 0 && kind != 1) || iters <= 0 || !out) {
        set_err("gpu_smc2_selftest_mcmc: bad arguments");
        return -1;
    }
    const double mu0[3] = {-1.0, 0.95, 0.2};
    const double C[9] = {0.04, 0.002, -0.003, 0.002, 0.0004, 0.0001, -0.003, 0.0001, 0.0025};
    // C^-1 by cofactors
    const double det = C[0] * (C[4] * C[8] - C[5] * C[7]) - C[1] * (C[3] * C[8] - C[5] * C[6]) +
                       C[2] * (C[3] * C[7] - C[4] * C[6]);
    double Ci[9];
    Ci[0] = (C[4] * C[8] - C[5] * C[7]) / det;
    Ci[1] = (C[2] * C[7] - C[1] * C[8]) / det;
    Ci[2] = (C[1] * C[5] - C[2] * C[4]) / det;
    Ci[3] = (C[5] * C[6] - C[3] * C[8]) / det;
    Ci[4] = (C[0] * C[8] - C[2] * C[6]) / det;
    Ci[5] = (C[2] * C[3] - C[0] * C[5]) / det;
    Ci[6] = (C[3] * C[7] - C[4] * C[6]) / det;
    Ci[7] = (C[1] * C[6] - C[0] * C[7]) / det;
    Ci[8] = (C[0] * C[4] - C[1] * C[3]) / det;
    auto logp = [&](const double* t, double* g) {
        double d[3] = {t[0] - mu0[0], t[1] - mu0[1], t[2] - mu0[2]};
        double q = 0.0;
        for (int i = 0; i < 3; ++i) {
            double v = 0.0;
            for (int k = 0; k < 3; ++k) v += Ci[3 * i + k] * d[k];
            g[i] = -v;
            q += d[i] * v;
        }
        return -0.5 * q;
    };
    Precond P;
    make_precond(C, kind == 1 ? 1.0 : 2.38 * 2.38 / 3.0, P);
    const double h = 0.8;
    HostRng r;
    r.seed((uint64_t)(int64_t)seed, 0x7e57ull);
    double th[3] = {mu0[0], mu0[1], mu0[2]}, g[3];
    double lp = logp(th, g);
    double s1[3] = {0, 0, 0}, s2[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    long long acc = 0;
    for (int it = 0; it < iters; ++it) {
        const double xi[3] = {r.normal(), r.normal(), r.normal()};
        double tp[3], gp[3];
        propose(kind, P, h, th, g, xi, tp);
        const double lpp = logp(tp, gp);
        const double la = log_alpha(kind, P, h, th, lp, g, tp, lpp, gp, fault == 1);
        if (log(r.uniform()) < la) {
            for (int k = 0; k < 3; ++k) {
                th[k] = tp[k];
                g[k] = gp[k];
            }
            lp = lpp;
            ++acc;
        }
        for (int i = 0; i < 3; ++i) {
            s1[i] += th[i];
            for (int j = 0; j < 3; ++j) s2[3 * i + j] += th[i] * th[j];
        }
    }
    for (int i = 0; i < 3; ++i) out[i] = s1[i] / iters;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) out[3 + 3 * i + j] = s2[3 * i + j] / iters - out[i] * out[j];
    out[12] = (double)acc / iters;
    return 0;
}

int gpu_smc2_set_block(GpuSmc2* s, int threads) {
    if (!s || !s->ready) {
        set_err("gpu_smc2_set_block: not initialised");
        return -1;
    }
    if (gpu_bpf_batch_set_block(s->main, threads) || gpu_bpf_batch_set_block(s->rep, threads)) return -1;
    return 0;
}

int gpu_smc2_set_tile(GpuSmc2* s, int particles) {
    if (!s || !s->ready) {
        set_err("gpu_smc2_set_tile: not initialised");
        return -1;
    }
    if (gpu_bpf_batch_set_tile(s->main, particles) || gpu_bpf_batch_set_tile(s->rep, particles)) return -1;
    return 0;
}

}  // extern "C"
