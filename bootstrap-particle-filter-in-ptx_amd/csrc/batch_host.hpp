// batch_host.hpp -- host-side state of a batch of inner filters, shared by
// the batch runtime and the SMC^2 driver built on it.
#pragma once

#include <hip/hip_runtime.h>

#include <vector>

#include "../../include/gpu_bpf_batch.h"
#include "bpf_types.hpp"

namespace bpf {
hipError_t launch_batch_init(const BatchInitArgs& a, hipStream_t st);
hipError_t launch_batch_tick(const BatchArgs& a, int n_blocks, int block, int tile, hipStream_t st);
bool batch_tick_fits(int block, int tile, bool tangents);
hipError_t launch_outer_check(const double* logw, const Ctl* ctl, int M, float threshold, int* halt,
                              double* ess_out, hipStream_t st);
hipError_t launch_smc2_gather(const GatherArgs& a, hipStream_t st);
hipError_t launch_smc2_accept(const AcceptArgs& a, int n_slots, hipStream_t st);
}  // namespace bpf

struct GpuBpfBatch {
    int M = 0, cap = 0, block = 0, tile = 0;   // tile: particles per wave tile (0 = 256)
    float threshold = 0.0f;
    bpf::ObsModel obs{};
    hipStream_t stream = nullptr;
    std::vector<int> n, seed, order;
    std::vector<float> mu, phi, sig;
    int* d_order = nullptr;
    int* d_n = nullptr;
    int* d_seed = nullptr;
    float* d_mu = nullptr;
    float* d_phi = nullptr;
    float* d_sig = nullptr;
    bpf::Ctl* d_ctl = nullptr;
    float* d_x[2] = {nullptr, nullptr};
    float* d_l = nullptr;
    BpfTickStats* d_ring = nullptr;
    int ring_cap = 1024;
    int stats = 1;                     // per-tick statistics recorded in the ring
    int stats_from = 0;                // first tick of the current recording span
    float* d_y = nullptr;              // observation history, indexed by tick
    int y_cap = 0;
    int timestep = 0;
    float* d_dx[2][3] = {{nullptr, nullptr, nullptr}, {nullptr, nullptr, nullptr}};   // tangents
    double* d_score = nullptr;         // 3 per slot when the score is enabled
    int fisher = 0;                    // score estimator: 0 pathwise, 1 Fisher identity
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int graph_block = 0, graph_tile = 0, graph_stats = 1;
};


namespace bpf {
void free_graph(GpuBpfBatch* b);
void destroy_batch(GpuBpfBatch* b);
BatchArgs tick_args(const GpuBpfBatch* b);
int create_impl(GpuBpfBatch* b);
int enable_score(GpuBpfBatch* b, int fisher);
int append_y(GpuBpfBatch* b, const float* y, int n_ticks, hipMemcpyKind kind);
int run_ticks(GpuBpfBatch* b, int n_ticks);
GpuBpfBatch* new_batch(int M, const int* n, const float* mu, const float* phi, const float* sig,
                       const int* seeds, float nu_obs, float thr);
}  // namespace bpf
