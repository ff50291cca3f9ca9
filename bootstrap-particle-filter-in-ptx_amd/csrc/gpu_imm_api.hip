// gpu_imm_api.hip -- interacting multiple models (IMM) over K SV bootstrap
// filters, behind the reference's gpu_imm_* entry points
// (gpu_bpf/gpu_bpf.cuh:128-162; behaviour of gpu_bpf/gpu_bpf.cu:943-1122).
//
// The model layer is a discrete hidden Markov chain over the K parameter sets
// (Blom & Bar-Shalom's IMM with the filters as the mode-conditioned
// estimators).  With P the row-stochastic transition matrix (row j = from
// model j) and p_{t-1} the model posterior, one tick is
//
//   prediction   q_k   = sum_j P_jk p_{t-1,j}
//   evidence     L_k   = p(y_t | y_{0:t-1}, model k)   (filter k's BpfResult.log_lik)
//   update       p_t,k = q_k L_k / Z,  Z = sum_k q_k L_k
//   outputs      log Z, the posterior-weighted state mean sum_k p_t,k h_k,
//                and the most probable model.
//
// Prediction is a matrix-vector product in linear FP64.  The entries of P are
// floored at 1e-30 (the reference's floor), so q_k >= 1e-30 max_j p_j >=
// 1e-30 / K and no predicted probability underflows.  The update runs in the
// log domain with the largest log q_k + log L_k subtracted before
// exponentiating, because filter log-likelihoods can differ by hundreds.
//
// The K model filters step together each tick.  When they fit the batched
// filters (no Student-t state noise, one observation model, at most
// GPU_BPF_BATCH_MAX_N particles) they are one GpuBpfBatch: one launch per
// tick for all models instead of K streams; otherwise K GpuBpfState filters,
// all enqueued before any is waited on.  Filter k is the plain filter
// gpu_bpf_create(n, rho_k, sigma_z_k, mu_k, nu_state_k, nu_obs_k,
// seed + 7919 k) with ESS threshold 0.5 either way (the batch's inner filters
// equal it tick for tick).
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "api_util.hpp"
#include "../../include/gpu_bpf_batch.h"

using namespace bpf;

namespace {

// Seed offset between consecutive model filters and the default chain
// (gpu_bpf/gpu_bpf.cu:983,997-1001): stay with probability 0.95, move to each
// other model with an equal share of the rest.
constexpr int kSeedStride = 7919;
constexpr double kDefaultStay = 0.95;
constexpr double kTransFloor = 1e-30;

// Posterior over the K models of a discrete Markov chain, updated by one
// likelihood per model and tick.
class ModelChain {
public:
    void reset(int K, const float* trans) {
        K_ = K;
        P_.resize((size_t)K * K);
        for (int from = 0; from < K; from++)
            for (int to = 0; to < K; to++) {
                double v;
                if (trans) {
                    v = (double)trans[(size_t)from * K + to];
                } else {
                    const double move = K > 1 ? (1.0 - kDefaultStay) / (double)(K - 1) : 0.0;
                    v = from == to ? kDefaultStay : move;
                }
                // fmax also maps a NaN entry to the floor
                P_[(size_t)from * K + to] = fmax(v, kTransFloor);
            }
        post_.assign(K, 1.0 / (double)K);
        prior_.assign(K, 0.0);
        logpost_.assign(K, -log((double)K));
    }

    // q = P^T p: the model probabilities before this tick's observations
    void predict() {
        for (int to = 0; to < K_; to++) prior_[to] = 0.0;
        for (int from = 0; from < K_; from++) {
            const double pf = post_[from];
            if (pf == 0.0) continue;
            const double* row = &P_[(size_t)from * K_];
            for (int to = 0; to < K_; to++) prior_[to] += row[to] * pf;
        }
    }

    // Bayes update with per-model log evidence; returns log Z
    double absorb(const double* log_ev) {
        double top = -INFINITY;
        for (int k = 0; k < K_; k++) {
            logpost_[k] = log(prior_[k]) + log_ev[k];
            if (logpost_[k] > top) top = logpost_[k];
        }
        if (!(top > -INFINITY)) {
            // every model rules the observation out (or it is NaN): keep the
            // prediction, report the NaN / -inf evidence
            for (int k = 0; k < K_; k++) {
                post_[k] = prior_[k];
                logpost_[k] = log(prior_[k]);
            }
            return top == -INFINITY ? -INFINITY : NAN;
        }
        double z = 0.0;
        for (int k = 0; k < K_; k++) z += exp(logpost_[k] - top);
        const double log_z = top + log(z);
        for (int k = 0; k < K_; k++) {
            logpost_[k] -= log_z;
            post_[k] = exp(logpost_[k]);
        }
        return log_z;
    }

    // posterior mean of per-model values, and the most probable model
    double mean_of(const float* v) const {
        double acc = 0.0;
        for (int k = 0; k < K_; k++) acc += post_[k] * (double)v[k];
        return acc;
    }
    int mode() const {
        int arg = 0;
        for (int k = 1; k < K_; k++)
            if (logpost_[k] > logpost_[arg]) arg = k;
        return arg;
    }
    double prob(int k) const { return post_[k]; }

private:
    int K_ = 0;
    std::vector<double> P_;        // floored transition matrix, row = from
    std::vector<double> post_;     // p_t
    std::vector<double> prior_;    // q_t
    std::vector<double> logpost_;  // log p_t (decides the mode without rounding ties)
};

}  // namespace

struct GpuImmState {
    int K = 0;
    int n = 0;
    int timestep = 0;
    ModelChain chain;
    GpuBpfBatch* batch = nullptr;            // all models in one launch per tick, or
    std::vector<GpuBpfState*> filters;       // one filter per model
    std::vector<double> log_ev;              // per tick: each model's log-likelihood
    std::vector<float> h;                    // per tick: each model's state mean
    std::vector<BpfTickStats> stats;         // batch read-back
};

static void imm_free(GpuImmState* s) {
    if (!s) return;
    if (s->batch) gpu_bpf_batch_destroy(s->batch);
    for (GpuBpfState* f : s->filters) gpu_bpf_destroy(f);
    delete s;
}

// The batched filters take every model when they share one observation model,
// have Gaussian state noise and fit a batch slot.
static bool models_fit_batch(const ImmModelParams* m, int K, int n) {
    if (n > GPU_BPF_BATCH_MAX_N) return false;
    const float nu = m[0].nu_obs > 0.0f ? m[0].nu_obs : 0.0f;
    for (int k = 0; k < K; k++) {
        if (m[k].nu_state > 0.0f) return false;
        if ((m[k].nu_obs > 0.0f ? m[k].nu_obs : 0.0f) != nu) return false;
    }
    return true;
}

extern "C" {

GpuImmState* gpu_imm_create(const ImmModelParams* models, int n_models, int n_particles_per_model,
                            const float* transition_matrix, int seed) {
    clear_err();
    if (!models || n_models < 1 || n_models > IMM_MAX_MODELS || n_particles_per_model <= 0) {
        set_err("gpu_imm_create: need 1..%d models and a positive particle count", IMM_MAX_MODELS);
        return nullptr;
    }
    GpuImmState* s = new GpuImmState();
    s->K = n_models;
    s->n = n_particles_per_model;
    s->chain.reset(n_models, transition_matrix);
    s->log_ev.resize(n_models);
    s->h.resize(n_models);
    if (models_fit_batch(models, n_models, n_particles_per_model)) {
        std::vector<int> counts(n_models, n_particles_per_model), seeds(n_models);
        std::vector<float> mu(n_models), phi(n_models), sig(n_models);
        for (int k = 0; k < n_models; k++) {
            seeds[k] = seed + k * kSeedStride;
            mu[k] = models[k].mu;
            phi[k] = models[k].rho;
            sig[k] = models[k].sigma_z;
        }
        const float nu = models[0].nu_obs > 0.0f ? models[0].nu_obs : 0.0f;
        s->batch = gpu_bpf_batch_create(n_models, counts.data(), mu.data(), phi.data(), sig.data(),
                                        seeds.data(), nu, 0.5f);
        if (!s->batch) {
            imm_free(s);
            return nullptr;
        }
        s->stats.resize(n_models);
        return s;
    }
    s->filters.reserve(n_models);
    for (int k = 0; k < n_models; k++) {
        const ImmModelParams& m = models[k];
        GpuBpfState* f = gpu_bpf_create(n_particles_per_model, m.rho, m.sigma_z, m.mu, m.nu_state,
                                        m.nu_obs, seed + k * kSeedStride);
        if (!f) {
            imm_free(s);
            return nullptr;
        }
        s->filters.push_back(f);
    }
    return s;
}

void gpu_imm_destroy(GpuImmState* s) { imm_free(s); }

// Steps every model filter on y_t and fills s->log_ev and s->h; 0 or -1.
static int step_models(GpuImmState* s, float y_t) {
    if (s->batch) {
        if (gpu_bpf_batch_run(s->batch, &y_t, 1) || gpu_bpf_batch_copy_stats(s->batch, s->stats.data(), 1))
            return -1;
        for (int k = 0; k < s->K; k++) {
            s->log_ev[k] = (double)s->stats[k].log_lik;
            s->h[k] = s->stats[k].h_mean;
        }
        return 0;
    }
    for (GpuBpfState* f : s->filters) gpu_bpf_step_async(f, y_t);
    for (int k = 0; k < s->K; k++) {
        const BpfResult r = gpu_bpf_get_result(s->filters[k]);
        s->log_ev[k] = (double)r.log_lik;
        s->h[k] = r.h_mean;
    }
    return 0;
}

ImmResult gpu_imm_step(GpuImmState* s, float y_t) {
    ImmResult out;
    out.h_mean = out.vol = out.log_lik = out.best_prob = __builtin_nanf("");
    out.best_model = -1;
    if (!s) return out;
    s->chain.predict();
    if (step_models(s, y_t)) return out;
    const double log_z = s->chain.absorb(s->log_ev.data());
    const double hm = s->chain.mean_of(s->h.data());
    const int top = s->chain.mode();
    s->timestep++;
    out.h_mean = (float)hm;
    out.vol = expf(0.5f * out.h_mean);
    out.log_lik = (float)log_z;
    out.best_model = top;
    out.best_prob = (float)s->chain.prob(top);
    return out;
}

float gpu_imm_get_prob(const GpuImmState* s, int k) {
    if (!s || k < 0 || k >= s->K) return 0.0f;
    return (float)s->chain.prob(k);
}

void gpu_imm_get_probs(const GpuImmState* s, float* probs_out) {
    if (!s || !probs_out) return;
    for (int k = 0; k < s->K; k++) probs_out[k] = (float)s->chain.prob(k);
}

int gpu_imm_uses_batch(const GpuImmState* s) { return s && s->batch ? 1 : 0; }

// Cartesian product of the three parameter lists; model index
// ((r * n_sigma) + q) * n_mu + m holds (rhos[r], sigma_zs[q], mus[m]), i.e.
// rho varies slowest (the reference's order).  The caller frees the array
// with free().
ImmModelParams* gpu_imm_build_grid(const float* rhos, int n_rho, const float* sigma_zs, int n_sigma,
                                   const float* mus, int n_mu, float nu_state, float nu_obs,
                                   int* out_n_models) {
    clear_err();
    if (out_n_models) *out_n_models = 0;
    if (!rhos || !sigma_zs || !mus || n_rho <= 0 || n_sigma <= 0 || n_mu <= 0 || !out_n_models) {
        set_err("gpu_imm_build_grid: bad arguments");
        return nullptr;
    }
    const long long total = (long long)n_rho * n_sigma * n_mu;
    if (total > 0x7fffffffLL) {
        set_err("gpu_imm_build_grid: %lld models do not fit an int", total);
        return nullptr;
    }
    ImmModelParams* grid = (ImmModelParams*)calloc((size_t)total, sizeof(ImmModelParams));
    if (!grid) {
        set_err("gpu_imm_build_grid: out of memory");
        return nullptr;
    }
    for (long long i = 0; i < total; i++) {
        const long long m = i % n_mu, q = (i / n_mu) % n_sigma, r = i / ((long long)n_mu * n_sigma);
        grid[i] = ImmModelParams{rhos[r], sigma_zs[q], mus[m], nu_state, nu_obs};
    }
    *out_n_models = (int)total;
    return grid;
}

}  // extern "C"
