// gpu_imm_api.hip -- interacting multiple models (IMM) over K SV bootstrap
// filters: gpu_bpf/gpu_bpf.cu:943-1122 (gpu_imm_*).
//
// The K model filters step together each tick.  When they fit the batched
// filters (no Student-t state noise, one observation model, at most
// GPU_BPF_BATCH_MAX_N particles) they are one GpuBpfBatch: one launch per
// tick for all models instead of K streams; otherwise K GpuBpfState filters,
// all enqueued before any is waited on.  Filter k is the plain filter
// gpu_bpf_create(n, rho_k, sigma_z_k, mu_k, nu_state_k, nu_obs_k,
// seed + 7919 k) with ESS threshold 0.5 either way (the batch's inner filters
// equal it tick for tick).  The model probabilities are mixed on the host in
// FP64 as the reference does.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "api_util.hpp"
#include "../../include/gpu_bpf_batch.h"

using namespace bpf;

struct GpuImmState {
    int K = 0;
    int n = 0;
    int timestep = 0;
    std::vector<double> log_pi, log_pi_pred, log_T;   // log_T[j * K + k] = log T(j -> k)
    GpuBpfBatch* batch = nullptr;
    std::vector<GpuBpfState*> filters;
};

// gpu_bpf/gpu_bpf.cu:943-952
static double log_sum_exp(const double* x, int n) {
    double mx = -1e30;
    for (int i = 0; i < n; i++)
        if (x[i] > mx) mx = x[i];
    double s = 0.0;
    for (int i = 0; i < n; i++) s += exp(x[i] - mx);
    return mx + log(s);
}

static void imm_free(GpuImmState* s) {
    if (!s) return;
    if (s->batch) gpu_bpf_batch_destroy(s->batch);
    for (GpuBpfState* f : s->filters) gpu_bpf_destroy(f);
    delete s;
}

extern "C" {

GpuImmState* gpu_imm_create(const ImmModelParams* models, int n_models, int n_particles_per_model,
                            const float* transition_matrix, int seed) {
    clear_err();
    if (!models || n_models < 1 || n_models > IMM_MAX_MODELS || n_particles_per_model <= 0) {
        set_err("gpu_imm_create: need 1..%d models and a positive particle count", IMM_MAX_MODELS);
        return nullptr;
    }
    const int K = n_models, n = n_particles_per_model;
    GpuImmState* s = new GpuImmState();
    s->K = K;
    s->n = n;
    s->log_pi.assign(K, -log((double)K));
    s->log_pi_pred.assign(K, 0.0);
    s->log_T.assign((size_t)K * K, 0.0);
    if (transition_matrix) {
        for (int i = 0; i < K * K; i++) s->log_T[i] = log(fmax((double)transition_matrix[i], 1e-30));
    } else {
        const double p_stay = 0.95;
        const double p_switch = (K > 1) ? (1.0 - p_stay) / (K - 1) : 0.0;
        for (int i = 0; i < K; i++)
            for (int j = 0; j < K; j++) s->log_T[i * K + j] = log(fmax((i == j) ? p_stay : p_switch, 1e-30));
    }
    bool batch_ok = n <= GPU_BPF_BATCH_MAX_N;
    const float nu_obs0 = models[0].nu_obs > 0.0f ? models[0].nu_obs : 0.0f;
    for (int k = 0; k < K; k++) {
        const float nu_o = models[k].nu_obs > 0.0f ? models[k].nu_obs : 0.0f;
        if (models[k].nu_state > 0.0f || nu_o != nu_obs0) batch_ok = false;
    }
    if (batch_ok) {
        std::vector<int> np(K, n), seeds(K);
        std::vector<float> mu(K), phi(K), sig(K);
        for (int k = 0; k < K; k++) {
            seeds[k] = seed + k * 7919;
            mu[k] = models[k].mu;
            phi[k] = models[k].rho;
            sig[k] = models[k].sigma_z;
        }
        s->batch = gpu_bpf_batch_create(K, np.data(), mu.data(), phi.data(), sig.data(), seeds.data(),
                                        nu_obs0, 0.5f);
        if (!s->batch) {
            imm_free(s);
            return nullptr;
        }
    } else {
        for (int k = 0; k < K; k++) {
            GpuBpfState* f = gpu_bpf_create(n, models[k].rho, models[k].sigma_z, models[k].mu,
                                            models[k].nu_state, models[k].nu_obs, seed + k * 7919);
            if (!f) {
                imm_free(s);
                return nullptr;
            }
            s->filters.push_back(f);
        }
    }
    return s;
}

void gpu_imm_destroy(GpuImmState* s) { imm_free(s); }

// gpu_bpf/gpu_bpf.cu:1022-1080
ImmResult gpu_imm_step(GpuImmState* s, float y_t) {
    ImmResult r;
    r.h_mean = r.vol = r.log_lik = __builtin_nanf("");
    r.best_model = -1;
    r.best_prob = __builtin_nanf("");
    if (!s) return r;
    const int K = s->K;
    // interaction: pi_k^- = sum_j T(j -> k) pi_j
    std::vector<double> terms(K);
    for (int k = 0; k < K; k++) {
        for (int j = 0; j < K; j++) terms[j] = s->log_T[(size_t)j * K + k] + s->log_pi[j];
        s->log_pi_pred[k] = log_sum_exp(terms.data(), K);
    }
    // every model's tick, then their results
    std::vector<double> log_liks(K);
    std::vector<float> h_ests(K);
    if (s->batch) {
        std::vector<BpfTickStats> st(K);
        if (gpu_bpf_batch_run(s->batch, &y_t, 1) || gpu_bpf_batch_copy_stats(s->batch, st.data(), 1))
            return r;
        for (int k = 0; k < K; k++) {
            h_ests[k] = st[k].h_mean;
            log_liks[k] = (double)st[k].log_lik;
        }
    } else {
        for (int k = 0; k < K; k++) gpu_bpf_step_async(s->filters[k], y_t);
        for (int k = 0; k < K; k++) {
            const BpfResult b = gpu_bpf_get_result(s->filters[k]);
            h_ests[k] = b.h_mean;
            log_liks[k] = (double)b.log_lik;
        }
    }
    // update: log pi_k = log pi_k^- + log p(y | model k) - log Z
    std::vector<double> log_joint(K);
    for (int k = 0; k < K; k++) log_joint[k] = s->log_pi_pred[k] + log_liks[k];
    const double log_Z = log_sum_exp(log_joint.data(), K);
    for (int k = 0; k < K; k++) s->log_pi[k] = log_joint[k] - log_Z;
    // mixed output
    double h_mixed = 0.0;
    int best_k = 0;
    double best_log_pi = -1e30;
    for (int k = 0; k < K; k++) {
        const double pi_k = exp(s->log_pi[k]);
        h_mixed += pi_k * (double)h_ests[k];
        if (s->log_pi[k] > best_log_pi) {
            best_log_pi = s->log_pi[k];
            best_k = k;
        }
    }
    s->timestep++;
    r.h_mean = (float)h_mixed;
    r.vol = expf((float)h_mixed * 0.5f);
    r.log_lik = (float)log_Z;
    r.best_model = best_k;
    r.best_prob = (float)exp(best_log_pi);
    return r;
}

float gpu_imm_get_prob(const GpuImmState* s, int k) {
    if (!s || k < 0 || k >= s->K) return 0.0f;
    return (float)exp(s->log_pi[k]);
}

void gpu_imm_get_probs(const GpuImmState* s, float* probs_out) {
    if (!s || !probs_out) return;
    for (int k = 0; k < s->K; k++) probs_out[k] = (float)exp(s->log_pi[k]);
}

int gpu_imm_uses_batch(const GpuImmState* s) { return s && s->batch ? 1 : 0; }

// gpu_bpf/gpu_bpf.cu:1100-1122 (the caller frees the array with free())
ImmModelParams* gpu_imm_build_grid(const float* rhos, int n_rho, const float* sigma_zs, int n_sigma,
                                   const float* mus, int n_mu, float nu_state, float nu_obs,
                                   int* out_n_models) {
    if (!rhos || !sigma_zs || !mus || n_rho <= 0 || n_sigma <= 0 || n_mu <= 0 || !out_n_models) {
        set_err("gpu_imm_build_grid: bad arguments");
        if (out_n_models) *out_n_models = 0;
        return nullptr;
    }
    const int total = n_rho * n_sigma * n_mu;
    ImmModelParams* grid = (ImmModelParams*)malloc((size_t)total * sizeof(ImmModelParams));
    if (!grid) {
        set_err("gpu_imm_build_grid: out of memory");
        *out_n_models = 0;
        return nullptr;
    }
    int idx = 0;
    for (int r = 0; r < n_rho; r++)
        for (int q = 0; q < n_sigma; q++)
            for (int m = 0; m < n_mu; m++) {
                grid[idx].rho = rhos[r];
                grid[idx].sigma_z = sigma_zs[q];
                grid[idx].mu = mus[m];
                grid[idx].nu_state = nu_state;
                grid[idx].nu_obs = nu_obs;
                idx++;
            }
    *out_n_models = total;
    return grid;
}

}  // extern "C"
