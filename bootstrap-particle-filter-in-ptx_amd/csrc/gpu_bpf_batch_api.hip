// gpu_bpf_batch_api.hip -- host runtime behind include/gpu_bpf_batch.h.
//
// A batch holds M inner filters in slots of `cap` particles (cap = the
// largest particle count rounded up to a tile), one control block per slot,
// and the observation history in device memory.  A tick is one launch of
// k_batch_tick whose arguments do not depend on the tick (each workgroup
// reads its slot's tick from the control block), so runs of ticks replay one
// instantiated HIP graph of kGraphTicks launches.
#include <hip/hip_runtime.h>

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <vector>

#include "api_util.hpp"
#include "batch_host.hpp"

using namespace bpf;

static_assert(GPU_BPF_BATCH_MAX_N == kBatchMaxN, "header and kernels disagree on the largest filter");

namespace {

constexpr int kGraphTicks = 32;        // launches per captured graph
constexpr int kBlocks[5] = {64, 128, 256, 512, 1024};
constexpr int kTiles[3] = {kBatchTile, 2 * kBatchTile, 4 * kBatchTile};

double c_obs_student(double nu) {
    return lgamma((nu + 1.0) / 2.0) - lgamma(nu / 2.0) - 0.5 * log(nu * 3.14159265358979323846);
}

}  // namespace

namespace bpf {

void free_graph(GpuBpfBatch* b) {
    if (b->graph_exec) (void)hipGraphExecDestroy(b->graph_exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->graph_exec = nullptr;
    b->graph = nullptr;
}

void destroy_batch(GpuBpfBatch* b) {
    if (!b) return;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    free_graph(b);
    void* ptrs[] = {b->d_order, b->d_n, b->d_seed, b->d_mu, b->d_phi, b->d_sig, b->d_ctl,
                    b->d_x[0], b->d_x[1], b->d_l, b->d_ring, b->d_y, b->d_score};
    for (void* p : ptrs)
        if (p) (void)hipFree(p);
    for (int h = 0; h < 2; ++h)
        for (int k = 0; k < 3; ++k)
            if (b->d_dx[h][k]) (void)hipFree(b->d_dx[h][k]);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

BatchArgs tick_args(const GpuBpfBatch* b) {
    BatchArgs a;
    memset((void*)&a, 0, sizeof a);
    a.order = b->d_order;
    a.n = b->d_n;
    a.seed = b->d_seed;
    a.mu = b->d_mu;
    a.phi = b->d_phi;
    a.sig = b->d_sig;
    a.ctl = b->d_ctl;
    a.x0 = b->d_x[0];
    a.x1 = b->d_x[1];
    a.lw = b->d_l;
    a.ring = b->stats ? b->d_ring : nullptr;   // no ring: no mean, variance or ESS
    a.ring_cap = b->ring_cap;
    a.cap = b->cap;
    a.n_slots = b->M;
    a.y = b->d_y;
    a.obs = b->obs;
    a.threshold = b->threshold;
    a.noise = kNoisePcg;
    if (b->d_score) {
        for (int k = 0; k < 3; ++k) {
            a.dx0[k] = b->d_dx[0][k];
            a.dx1[k] = b->d_dx[1][k];
        }
        a.score = b->d_score;
        a.fisher = b->fisher;
    }
    return a;
}

// Tangents and score accumulators; x_0 is drawn again with its tangents.
int enable_score(GpuBpfBatch* b, int fisher) {
    b->fisher = fisher;
    if (b->d_score) return 0;
    const size_t slots = (size_t)b->M * b->cap;
    for (int h = 0; h < 2; ++h)
        for (int k = 0; k < 3; ++k) {
            HIP_TRY(hipMalloc(&b->d_dx[h][k], slots * sizeof(float)));
            HIP_TRY(hipMemsetAsync(b->d_dx[h][k], 0, slots * sizeof(float), b->stream));
        }
    HIP_TRY(hipMalloc(&b->d_score, (size_t)b->M * 3 * sizeof(double)));
    HIP_TRY(hipMemsetAsync(b->d_score, 0, (size_t)b->M * 3 * sizeof(double), b->stream));
    BatchInitArgs ia;
    memset((void*)&ia, 0, sizeof ia);
    ia.n = b->d_n;
    ia.seed = b->d_seed;
    ia.mu = b->d_mu;
    ia.phi = b->d_phi;
    ia.sig = b->d_sig;
    ia.x0 = b->d_x[0];
    ia.cap = b->cap;
    ia.n_slots = b->M;
    ia.jm = kPcgMult * kPcgMult;
    ia.jg = 1u + kPcgMult;
    ia.noise = kNoisePcg;
    for (int k = 0; k < 3; ++k) ia.dx0[k] = b->d_dx[0][k];
    ia.fisher = fisher;
    HIP_TRY(launch_batch_init(ia, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    free_graph(b);
    return 0;
}

int create_impl(GpuBpfBatch* b) {
    HIP_TRY(hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking));
    const int M = b->M;
    const size_t slots = (size_t)M * b->cap;
    HIP_TRY(hipMalloc(&b->d_order, M * sizeof(int)));
    HIP_TRY(hipMalloc(&b->d_n, M * sizeof(int)));
    HIP_TRY(hipMalloc(&b->d_seed, M * sizeof(int)));
    HIP_TRY(hipMalloc(&b->d_mu, M * sizeof(float)));
    HIP_TRY(hipMalloc(&b->d_phi, M * sizeof(float)));
    HIP_TRY(hipMalloc(&b->d_sig, M * sizeof(float)));
    HIP_TRY(hipMalloc(&b->d_ctl, M * sizeof(Ctl)));
    HIP_TRY(hipMalloc(&b->d_x[0], slots * sizeof(float)));
    HIP_TRY(hipMalloc(&b->d_x[1], slots * sizeof(float)));
    HIP_TRY(hipMalloc(&b->d_l, slots * sizeof(float)));
    if (b->stats) HIP_TRY(hipMalloc(&b->d_ring, (size_t)b->ring_cap * M * sizeof(BpfTickStats)));
    HIP_TRY(hipMemsetAsync(b->d_x[1], 0, slots * sizeof(float), b->stream));
    HIP_TRY(hipMemsetAsync(b->d_l, 0, slots * sizeof(float), b->stream));
    HIP_TRY(hipMemcpyAsync(b->d_order, b->order.data(), M * sizeof(int), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(b->d_n, b->n.data(), M * sizeof(int), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(b->d_seed, b->seed.data(), M * sizeof(int), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(b->d_mu, b->mu.data(), M * sizeof(float), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(b->d_phi, b->phi.data(), M * sizeof(float), hipMemcpyHostToDevice, b->stream));
    HIP_TRY(hipMemcpyAsync(b->d_sig, b->sig.data(), M * sizeof(float), hipMemcpyHostToDevice, b->stream));
    std::vector<Ctl> c(M);
    memset((void*)c.data(), 0, M * sizeof(Ctl));
    for (int m = 0; m < M; ++m) {
        c[m].host_rng = (unsigned long long)(long long)b->seed[m] * 67890ULL + 12345ULL;
        c[m].s_prev = 1.0;
    }
    HIP_TRY(hipMemcpyAsync(b->d_ctl, c.data(), M * sizeof(Ctl), hipMemcpyHostToDevice, b->stream));
    BatchInitArgs ia;
    memset((void*)&ia, 0, sizeof ia);
    ia.n = b->d_n;
    ia.seed = b->d_seed;
    ia.mu = b->d_mu;
    ia.phi = b->d_phi;
    ia.sig = b->d_sig;
    ia.x0 = b->d_x[0];
    ia.cap = b->cap;
    ia.n_slots = M;
    ia.jm = kPcgMult * kPcgMult;       // draw 0
    ia.jg = 1u + kPcgMult;
    ia.noise = kNoisePcg;
    HIP_TRY(launch_batch_init(ia, b->stream));
    HIP_TRY(hipStreamSynchronize(b->stream));
    return 0;
}

// Appends n observations to the device history (grows by doubling).
int append_y(GpuBpfBatch* b, const float* y, int n_ticks, hipMemcpyKind kind) {
    const int need = b->timestep + n_ticks;
    if (need > b->y_cap) {
        int cap = std::max(need, std::max(1024, 2 * b->y_cap));
        float* ny = nullptr;
        HIP_TRY(hipMalloc(&ny, (size_t)cap * sizeof(float)));
        if (b->d_y) {
            HIP_TRY(hipMemcpyAsync(ny, b->d_y, (size_t)b->timestep * sizeof(float),
                                   hipMemcpyDeviceToDevice, b->stream));
            HIP_TRY(hipStreamSynchronize(b->stream));
            (void)hipFree(b->d_y);
        }
        b->d_y = ny;
        b->y_cap = cap;
        free_graph(b);                  // the captured launches hold the old pointer
    }
    HIP_TRY(hipMemcpyAsync(b->d_y + b->timestep, y, (size_t)n_ticks * sizeof(float), kind, b->stream));
    return 0;
}

static int capture_graph(GpuBpfBatch* b) {
    free_graph(b);
    const BatchArgs a = tick_args(b);
    HIP_TRY(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int k = 0; k < kGraphTicks && e == hipSuccess; ++k)
        e = launch_batch_tick(a, b->M, b->block, b->tile, b->stream);
    hipGraph_t g = nullptr;
    hipError_t e2 = hipStreamEndCapture(b->stream, &g);
    if (e != hipSuccess || e2 != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        set_err("batch graph capture failed: %s", hipGetErrorString(e != hipSuccess ? e : e2));
        return -1;
    }
    b->graph = g;
    HIP_TRY(hipGraphInstantiate(&b->graph_exec, g, nullptr, nullptr, 0));
    b->graph_block = b->block;
    b->graph_tile = b->tile;
    b->graph_stats = b->stats;
    return 0;
}

int run_ticks(GpuBpfBatch* b, int n_ticks) {
    if (b->block == 0) b->block = 256;
    if (b->tile == 0) b->tile = kBatchTile;
    if (!batch_tick_fits(b->block, b->tile, b->d_score != nullptr)) {
        set_err("batch: %d threads with tiles of %d particles need more LDS than one CU holds", b->block,
                b->tile);
        return -1;
    }
    if (n_ticks >= kGraphTicks && (!b->graph_exec || b->graph_block != b->block || b->graph_tile != b->tile ||
                                   b->graph_stats != b->stats)) {
        if (capture_graph(b)) return -1;
    }
    int done = 0;
    while (n_ticks - done >= kGraphTicks) {
        HIP_TRY(hipGraphLaunch(b->graph_exec, b->stream));
        done += kGraphTicks;
    }
    const BatchArgs a = tick_args(b);
    for (; done < n_ticks; ++done) HIP_TRY(launch_batch_tick(a, b->M, b->block, b->tile, b->stream));
    b->timestep += n_ticks;
    return 0;
}

GpuBpfBatch* new_batch(int M, const int* n, const float* mu, const float* phi, const float* sig,
                              const int* seeds, float nu_obs, float thr) {
    GpuBpfBatch* b = new GpuBpfBatch();
    b->M = M;
    b->n.assign(n, n + M);
    b->mu.assign(mu, mu + M);
    b->phi.assign(phi, phi + M);
    b->sig.assign(sig, sig + M);
    b->seed.assign(seeds, seeds + M);
    b->threshold = thr;
    const int nmax = *std::max_element(b->n.begin(), b->n.end());
    b->cap = (nmax + kBatchTileMax - 1) / kBatchTileMax * kBatchTileMax;
    b->order.resize(M);
    std::iota(b->order.begin(), b->order.end(), 0);
    std::stable_sort(b->order.begin(), b->order.end(), [&](int i, int j) { return b->n[i] > b->n[j]; });
    const float nu = nu_obs > 0.0f ? nu_obs : 0.0f;
    b->obs.nu = nu;
    b->obs.c_obs = nu > 0.0f ? (float)c_obs_student((double)nu) : 0.0f;
    b->obs.nph = (nu + 1.0f) * 0.5f;
    return b;
}

}  // namespace bpf

extern "C" {

GpuBpfBatch* gpu_bpf_batch_create(int n_filters, const int* n_particles, const float* mu,
                                  const float* phi, const float* sigma_z, const int* seeds,
                                  float nu_obs, float ess_threshold) {
    clear_err();
    if (n_filters <= 0 || !n_particles || !mu || !phi || !sigma_z || !seeds) {
        set_err("gpu_bpf_batch_create: n_filters must be positive and every array given");
        return nullptr;
    }
    for (int m = 0; m < n_filters; ++m) {
        if (n_particles[m] <= 0 || n_particles[m] > kBatchMaxN) {
            set_err("gpu_bpf_batch_create: filter %d has %d particles (1..%d)", m, n_particles[m],
                    kBatchMaxN);
            return nullptr;
        }
    }
    GpuBpfBatch* b = new_batch(n_filters, n_particles, mu, phi, sigma_z, seeds, nu_obs, ess_threshold);
    if (create_impl(b)) {
        destroy_batch(b);
        return nullptr;
    }
    return b;
}

void gpu_bpf_batch_destroy(GpuBpfBatch* b) { destroy_batch(b); }

int gpu_bpf_batch_set_block(GpuBpfBatch* b, int threads) {
    if (!b) {
        set_err("gpu_bpf_batch_set_block: null batch");
        return -1;
    }
    if (threads != 0 && threads != 64 && threads != 128 && threads != 256 && threads != 512 &&
        threads != 1024) {
        set_err("gpu_bpf_batch_set_block: %d is not 0, 64, 128, 256, 512 or 1024", threads);
        return -1;
    }
    b->block = threads;
    return 0;
}

int gpu_bpf_batch_get_block(GpuBpfBatch* b) { return b ? b->block : -1; }

int gpu_bpf_batch_set_tile(GpuBpfBatch* b, int particles) {
    if (!b) {
        set_err("gpu_bpf_batch_set_tile: null batch");
        return -1;
    }
    if (particles != 0 && particles != kTiles[0] && particles != kTiles[1] && particles != kTiles[2]) {
        set_err("gpu_bpf_batch_set_tile: %d is not 0, 256, 512 or 1024", particles);
        return -1;
    }
    b->tile = particles;
    return 0;
}

int gpu_bpf_batch_get_tile(GpuBpfBatch* b) { return b ? (b->tile ? b->tile : kBatchTile) : -1; }

int gpu_bpf_batch_set_stats(GpuBpfBatch* b, int on) {
    if (!b || (on != 0 && on != 1)) {
        set_err("gpu_bpf_batch_set_stats: a batch and 0 or 1");
        return -1;
    }
    if (on && !b->d_ring) HIP_TRY(hipMalloc(&b->d_ring, (size_t)b->ring_cap * b->M * sizeof(BpfTickStats)));
    if (on && !b->stats) b->stats_from = b->timestep;
    b->stats = on;
    return 0;
}

int gpu_bpf_batch_autotune(GpuBpfBatch* b, const float* y, int ticks) {
    // two untimed ticks, then ticks - 2 timed ones: y holds `ticks` values
    if (!b || !y || ticks < 3) {
        set_err("gpu_bpf_batch_autotune: need a batch and at least 3 observations");
        return -1;
    }
    float best_ms = 0.0f;
    int best = 0, best_tile = 0;
    for (int tile : kTiles) {
        for (int blk : kBlocks) {
            // the scratch batch runs the kernel this batch will run: with the
            // score's tangent or Fisher vectors when the score is enabled;
            // (block, tile) pairs whose LDS does not fit one CU are skipped
            if (!batch_tick_fits(blk, tile, b->d_score != nullptr)) continue;
            GpuBpfBatch* s = new_batch(b->M, b->n.data(), b->mu.data(), b->phi.data(), b->sig.data(),
                                       b->seed.data(), b->obs.nu, b->threshold);
            if (create_impl(s) || (b->d_score && enable_score(s, b->fisher))) {
                destroy_batch(s);
                return -1;
            }
            s->block = blk;
            s->tile = tile;
            s->stats = b->stats;
            hipEvent_t e0 = nullptr, e1 = nullptr;
            float ms = 0.0f;
            int rc = -1;
            if (hipEventCreate(&e0) == hipSuccess && hipEventCreate(&e1) == hipSuccess &&
                append_y(s, y, ticks, hipMemcpyHostToDevice) == 0 && run_ticks(s, 2) == 0 &&
                hipEventRecord(e0, s->stream) == hipSuccess && run_ticks(s, ticks - 2) == 0 &&
                hipEventRecord(e1, s->stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
                hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
                rc = 0;
            if (e0) (void)hipEventDestroy(e0);
            if (e1) (void)hipEventDestroy(e1);
            destroy_batch(s);
            if (rc) {
                set_err("gpu_bpf_batch_autotune: timing run failed");
                return -1;
            }
            if (best == 0 || ms < best_ms) {
                best_ms = ms;
                best = blk;
                best_tile = tile;
            }
        }
    }
    b->block = best;
    b->tile = best_tile;
    return best;
}

int gpu_bpf_batch_run(GpuBpfBatch* b, const float* y, int n_ticks) {
    if (!b || (!y && n_ticks > 0) || n_ticks < 0) {
        set_err("gpu_bpf_batch_run: bad arguments");
        return -1;
    }
    if (n_ticks == 0) return 0;
    if (append_y(b, y, n_ticks, hipMemcpyHostToDevice)) return -1;
    return run_ticks(b, n_ticks);
}

int gpu_bpf_batch_run_device(GpuBpfBatch* b, const float* y_dev, int n_ticks) {
    if (!b || (!y_dev && n_ticks > 0) || n_ticks < 0) {
        set_err("gpu_bpf_batch_run_device: bad arguments");
        return -1;
    }
    if (n_ticks == 0) return 0;
    if (append_y(b, y_dev, n_ticks, hipMemcpyDeviceToDevice)) return -1;
    return run_ticks(b, n_ticks);
}

int gpu_bpf_batch_synchronize(GpuBpfBatch* b) {
    if (!b) {
        set_err("gpu_bpf_batch_synchronize: null batch");
        return -1;
    }
    HIP_TRY(hipStreamSynchronize(b->stream));
    return 0;
}

int gpu_bpf_batch_get_timestep(GpuBpfBatch* b) { return b ? b->timestep : -1; }

int gpu_bpf_batch_copy_stats(GpuBpfBatch* b, BpfTickStats* out, int n_ticks) {
    if (!b || !out || n_ticks < 0) {
        set_err("gpu_bpf_batch_copy_stats: bad arguments");
        return -1;
    }
    const int avail = b->stats ? std::min(b->timestep - b->stats_from, b->ring_cap) : 0;
    if (n_ticks > avail) {
        set_err("gpu_bpf_batch_copy_stats: %d ticks requested, %d available%s", n_ticks, avail,
                b->stats ? "" : " (statistics are off)");
        return -1;
    }
    HIP_TRY(hipStreamSynchronize(b->stream));
    const size_t row = (size_t)b->M * sizeof(BpfTickStats);
    for (int k = 0; k < n_ticks; ++k) {
        const int t = b->timestep - n_ticks + k;
        HIP_TRY(hipMemcpy((char*)out + (size_t)k * row, (const char*)b->d_ring + (size_t)(t % b->ring_cap) * row,
                          row, hipMemcpyDeviceToHost));
    }
    return 0;
}

static int copy_ctl(GpuBpfBatch* b, std::vector<Ctl>& c) {
    c.resize(b->M);
    HIP_TRY(hipStreamSynchronize(b->stream));
    HIP_TRY(hipMemcpy(c.data(), b->d_ctl, (size_t)b->M * sizeof(Ctl), hipMemcpyDeviceToHost));
    return 0;
}

int gpu_bpf_batch_get_loglik(GpuBpfBatch* b, double* ll) {
    if (!b || !ll) {
        set_err("gpu_bpf_batch_get_loglik: bad arguments");
        return -1;
    }
    std::vector<Ctl> c;
    if (copy_ctl(b, c)) return -1;
    for (int m = 0; m < b->M; ++m) ll[m] = c[m].ll_cum;
    return 0;
}

int gpu_bpf_batch_get_resample_counts(GpuBpfBatch* b, int* counts) {
    if (!b || !counts) {
        set_err("gpu_bpf_batch_get_resample_counts: bad arguments");
        return -1;
    }
    std::vector<Ctl> c;
    if (copy_ctl(b, c)) return -1;
    for (int m = 0; m < b->M; ++m) counts[m] = c[m].resample_count;
    return 0;
}

int gpu_bpf_batch_enable_score(GpuBpfBatch* b, int estimator) {
    if (!b || b->timestep != 0 || b->threshold > 0.0f || (estimator != 0 && estimator != 1) ||
        (b->d_score && estimator != b->fisher)) {
        set_err("gpu_bpf_batch_enable_score: estimator 0 or 1, once, before the first tick, with "
                "resampling every tick (ess_threshold <= 0)");
        return -1;
    }
    return enable_score(b, estimator);
}

int gpu_bpf_batch_get_score(GpuBpfBatch* b, double* score) {
    if (!b || !score || !b->d_score) {
        set_err("gpu_bpf_batch_get_score: score not enabled");
        return -1;
    }
    HIP_TRY(hipStreamSynchronize(b->stream));
    HIP_TRY(hipMemcpy(score, b->d_score, (size_t)b->M * 3 * sizeof(double), hipMemcpyDeviceToHost));
    return 0;
}

int gpu_bpf_batch_copy_particles(GpuBpfBatch* b, int filter, float* x) {
    if (!b || !x || filter < 0 || filter >= b->M) {
        set_err("gpu_bpf_batch_copy_particles: bad arguments");
        return -1;
    }
    std::vector<Ctl> c;
    if (copy_ctl(b, c)) return -1;
    const float* src = b->d_x[c[filter].cur] + (size_t)filter * b->cap;
    HIP_TRY(hipMemcpy(x, src, (size_t)b->n[filter] * sizeof(float), hipMemcpyDeviceToHost));
    return 0;
}

}  // extern "C"
