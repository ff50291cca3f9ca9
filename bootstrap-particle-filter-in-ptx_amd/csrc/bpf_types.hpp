// bpf_types.hpp -- data shared by the host runtime and the gfx950 kernels.
#pragma once

#include <stdint.h>

#include "../../include/gpu_bpf.h"
#include "bpf_device.hpp"

namespace bpf {

// Tiles per reduction group: one finalise block (one thread per tile) folds
// each group of tile partials into a group partial, and the block that
// finishes last reduces G / kGroup group records instead of G tile records.
constexpr int kGroup = 256;
// Largest group count the finalise kernel's last block holds (256 threads x
// 4 groups): 2^30 particles at 16 per lane.
constexpr int kMaxGroups = 1024;
// The resample pass hands out tiles from kClaimRanges contiguous ranges, each
// with its own work counter (counters kClaimStride ints apart, in separate
// cache lines): atomics on one address serialise, so one counter for all
// tiles would bound the pass.
constexpr int kClaimRanges = 8;
constexpr int kClaimStride = 64;

// Device-resident control block.  Written only by the finalise kernel's last
// block each tick and read by the next kernels on the same stream, so
// the tick needs no host round trip (the reference syncs twice per tick,
// gpu_bpf/gpu_bpf_ptx_full.cu:677-685,766).
struct Ctl {
    int cur;             // x buffer holding the current particles
    int carry;           // 1: the propagate pass adds (l_prev - m_prev)
    int resample;        // decision of the latest tick
    int src, dst;        // resample reads x[src], writes x[dst]
    int wrap_start;      // first i whose unwrapped target u + i/N is >= 1
    int resample_count;
    int tick;            // ticks finalised so far
    float m_prev;        // max of the log-weights carried into the next tick
    float u;             // resampling offset u / N of the latest resample
    float M;             // global max log-weight of the latest tick
    float u_raw;         // uniform behind u (u = u_raw / N)
    double S;            // sum exp(l - M) of the latest tick
    double s_prev;       // sum exp(l - m_prev) of the carried weights
    double log_s_prev;   // log(s_prev) (0 = log 1 at the start)
    double off;          // raw log-weight offset (reference log_lik semantics)
    double ll_cum;       // FP64 cumulative log-likelihood
    unsigned long long host_rng;   // host-side PCG32 stream of the reference
    // observation source of ticks whose launch arguments come from this
    // block (captured tick graphs): tick t reads y_base[t - y_tick0]
    const float* y_base;
    int y_tick0;
    BpfTickStats last;
};

// Per-tile partial results of the propagate pass (structure of arrays), and
// the per-group folds of them.
struct Partials {
    float* m;      // max log-weight of the tile
    double* s;     // sum e,        e = exp(l - m)
    double* sx;    // sum e x
    double* sxx;   // sum e x^2
    double* sq;    // sum e^2
    float* gm;     // group max of the tile maxima
    double* gs;    // sum over the group's tiles of f s, f = exp(m - gm)
    double* gsx;
    double* gsxx;
    double* gsq;   // sum f^2 sq
    double* tf;    // per tile: f = exp(m - gm) (FP32 value held in FP64)
    double* tp;    // per tile: inclusive in-group prefix of f s, non-decreasing
    int* gcnt;     // arrival counter (0 between launches)
    int n_tiles;
    int n_groups;
};

// sigma_z scale bands of the reference's MixBandConfig (gpu_bpf_full.cuh:
// 51-55): particle i < bnd[0] scales sigma_z by k[0], i < bnd[1] by k[1],
// i < bnd[2] by k[2], else by k[3]; with n <= 1 every particle takes k[0]
// (ptx_assembly/bpf_kernels_full.ptx:297-324).
struct Bands {
    int n;
    int bnd[3];
    float k[4];
};

__device__ __host__ inline int band_of(const Bands& b, long long i) {
    return i < b.bnd[0] ? 0 : (i < b.bnd[1] ? 1 : (i < b.bnd[2] ? 2 : 3));
}

// k of particle i's band (selects, no indexed access)
__device__ __host__ inline float band_k(const Bands& b, long long i) {
    return i < b.bnd[0] ? b.k[0] : (i < b.bnd[1] ? b.k[1] : (i < b.bnd[2] ? b.k[2] : b.k[3]));
}

struct PropArgs {
    float* x0;
    float* x1;
    float* lw;
    Partials part;
    const Ctl* ctl;
    const float* y_ptr;   // observation in device memory, or null
    float y;              // observation by value when y_ptr is null
    // 1: the tick's observation, draw constants and propagation flag come
    // from the control block (ctl->tick, ctl->y_base), not from y_ptr / y /
    // pcg_a / pcg_g / do_prop, so one captured launch serves every tick
    int tick_from_ctl;
    int seed;             // PCG32 seed of the particle streams (tick_from_ctl)
    uint64_t pcg_a;       // seed * M^(k+2)      (state = pcg_a + inc_i * pcg_g)
    uint64_t pcg_g;       // sum_{j<k+2} M^j
    float mu, phi, sig;
    ObsModel obs;
    int n;
    int do_prop;          // 0 on tick 0 (ptx_assembly/bpf_kernels_full.ptx:294)
    Bands bands;          // sigma_z scales of the tick's transition
    // Student-t state noise (k_propagate's TSTATE variant): eps = z / sqrt(chi2
    // / nu_state), chi2 the sum of nu_int squared normals of particle i, from
    // PCG32 streams j = i * nu_int + k of seed * 31337 (state = chi_a + inc_j *
    // chi_g, inc_j = 2j + 1) in place of the reference's cuRAND normals
    float nu_state;
    int nu_int;
    uint64_t chi_a, chi_g;
    int apf_stage;        // 0 bootstrap; 1 / 2: auxiliary-filter stage (k_propagate APF1 / APF2)
    int store_l;          // keep log-weights for a possible carry
    float* x_save;        // optional copy of the pre-propagation states
    // Fisher-identity score (null: off): per particle a 3-vector
    // alpha = grad log p(x_0:t) (component c of buffer q at alpha[q][c]),
    // double-buffered like x, and per tile FP64 sums of e * alpha_c
    // (component c at fsum + c * n_tiles)
    float* alpha[2][3];
    double* fsum;
    // online mu/rho learning (null: off): per tile FP64 sums of e * v_q with
    // the reference's kernel-14 terms v = (g, g^2, g d, g^2 d^2), g = d lw/dx
    // of the propagated particle and d = x_prev - mu (component q at
    // lsum + q * n_tiles)
    double* lsum;
};

// Silverman jitter after a resample (gpu_bpf/gpu_bpf_ptx_full.cu:708-739):
// k_jitter_var writes per-block FP64 sums of (x - mean)^2; k_jitter adds
// bw * z (z from one more draw of each particle's stream) and clamps to
// [mu - 8, mu + 8].
struct JitterArgs {
    float* x;
    double* part;         // k_jitter_var: one sum per block
    float mean;
    float bw;
    float lo, hi;
    uint64_t pcg_a, pcg_g;
    int n;
};

struct FinArgs {
    Partials part;
    Ctl* ctl;
    BpfTickStats* ring;
    double* gptil;        // normalised exclusive group prefix, n_groups + 1 entries
    int ring_cap;
    int n;
    float threshold;      // ESS/N threshold; <= 0 resamples every tick
    float* u_log;         // optional: resampling offset of this tick
    int* w_log;           // optional: wrap start of this tick
    int* tile_ctr;        // resample work counters, cleared on resampling ticks
    // online learning (null: off): the propagate pass's tile sums (4 x
    // n_tiles), per-group folds of them (4 x n_groups), and the accumulators
    // to which the tick's normalised sums are added (4)
    const double* lsum;
    double* glsum;
    double* lacc;
};

struct ResArgs {
    float* x0;
    float* x1;
    const float* lw;      // stored log-weights, or null: recompute from x
    Partials part;
    const double* gptil;
    Ctl* ctl;
    int* tile_ctr;        // work counters of the tile ranges (kClaimRanges, kClaimStride apart)
    int one_tile;         // 1: block b does tile b only (grid = n_tiles), no counters
    const float* y_ptr;
    float y;
    int tick_from_ctl;    // 1: the observation of tick ctl->tick - 1 from ctl->y_base
    ObsModel obs;
    int n;
    float* m_log;         // optional per-tile copies for the score's replay
    double* ft_log;
    double* b_log;        // n_tiles + 1
    float* alpha[2][3];   // Fisher-identity vectors, gathered with x (null: off)
};

// One reverse-mode tick of the pathwise score.
struct AdjArgs {
    const float* pre;        // pre-propagation states of the tick (x_0 at tick 0)
    const float* xbar_next;  // adjoints of tick t+1, or null at the last tick
    float* xbar_out;         // adjoints of tick t
    const float* m;          // tile maxima of the tick
    const double* ftil;      // tile weight scales of the tick
    const double* btil;      // tile CDF prefix of the tick (n_tiles + 1)
    double* theta_acc;       // per-tile accumulators d/d(mu, phi, sigma)
    uint64_t pcg_a, pcg_g;   // draw t (draw 0 at tick 0)
    float mu, phi, sig;
    float y;
    const float* u_ptr;      // resampling offset of the tick (device)
    const int* w_ptr;        // wrap start of the tick (device)
    ObsModel obs;
    double dsd_dphi, dsd_dsig;   // stationary sd derivatives (x_0 reparameterisation)
    int x0_reparam;
    int tick;
    int n;
    int n_tiles;
};

struct InitArgs {
    float* x;             // x_0 (null: not written)
    uint64_t pcg_a, pcg_g;
    float mu, sd;
    int n, n_pad;
    float* alpha[3];      // optional grad log N(x_0; mu, sd^2), zero on padding
    float dsd_dphi, dsd_dsig;
};

// Fisher-identity score of the latest tick: sum_i w_i alpha_i from the tile
// sums and the tile / group scales of the finalise pass.
struct FisherArgs {
    Partials part;
    const double* fsum;
    const Ctl* ctl;
    double* out;          // 3 doubles
};

// ---------------------------------------------------------------------------
// Batched inner filters: one workgroup runs one filter's whole tick.  A tile
// holds 256 KB particles (KB = 1, 2 or 4: the single filter's tile at KB
// particles per lane), filters hold at most kBatchMaxN particles, i.e. at
// most 64 tiles of the smallest size, so each inner filter is one reduction
// group and reproduces the single filter with the same tile.  Slots hold a
// multiple of the largest tile, so the tile can change between runs.
// ---------------------------------------------------------------------------
constexpr int kBatchTile = 256;
constexpr int kBatchTileMax = 4 * kBatchTile;
constexpr int kBatchMaxTiles = 64;
constexpr int kBatchMaxN = kBatchTile * kBatchMaxTiles;

// Noise source of a batched tick: the PCG32 streams, the streams with the
// normals and the resampling uniform stored for later replays, the stored
// noise, or a Crank-Nicolson proposal around the stored noise (written to
// the slot's other noise pool).
enum BatchNoise { kNoisePcg = 0, kNoisePcgStore = 1, kNoiseStored = 2, kNoisePropose = 3 };

struct BatchArgs {
    const int* order;        // slot run by each workgroup (largest filters first)
    const int* n;            // particles of each slot's filter
    const int* seed;         // PCG32 seed of each slot (device streams, host stream)
    const float* mu;
    const float* phi;
    const float* sig;
    Ctl* ctl;                // one control block per slot
    float* x0;               // particles of slot m at m * cap, ping-pong by ctl[m].cur
    float* x1;
    float* lw;               // log-weights of slot m at m * cap
    BpfTickStats* ring;      // ring[(tick % ring_cap) * n_slots + m], or null
    int ring_cap;
    int cap;
    int n_slots;
    const float* y;          // observation series in device memory, indexed by tick
    ObsModel obs;
    float threshold;
    // The tick t is the slot's ctl[m].tick: draw t of the streams (jump
    // constants computed on the device), y[t], noise-pool row t.  Nothing in
    // the arguments changes from tick to tick, so a captured launch replays.
    // noise pools (kNoisePcgStore / kNoiseStored / kNoisePropose)
    int noise;
    float* zpool[2];         // normals of slot m, tick t: (t * n_slots + m) * cap + i
    float* upool[2];         // resampling uniform of slot m, tick t: t * n_slots + m
    const int* zcur;         // per slot: pool holding its current noise
    float cn_rho;            // proposal: z' = rho z + sqrt(1 - rho^2) xi
    float cn_s;              // sqrt(1 - rho^2)
    float u_step;            // proposal: u' = frac(u + u_step (2 v - 1))
    const int* prop_seed;    // per slot: seed of the proposal streams xi, v
    // SMC^2 outer layer
    double* outer_logw;      // per slot: += the tick's log-likelihood increment, or null
    const int* halt;         // non-zero: every workgroup returns at once, or null
    // forward-mode pathwise score (resampling every tick): tangents
    // dx/d(mu, phi, sigma) of every particle, ping-pong with x, and per slot
    // the FP64 accumulated d(log-likelihood)/d(mu, phi, sigma)
    float* dx0[3];
    float* dx1[3];
    double* score;           // 3 per slot, or null: no tangents
    int fisher;              // the 3-vectors are Fisher-identity path statistics instead
};

struct BatchInitArgs {
    const int* n;
    const int* seed;
    const float* mu;
    const float* phi;
    const float* sig;
    float* x0;
    int cap;
    int n_slots;
    uint64_t jm, jg;         // draw 0
    int noise;
    float* zpool[2];
    const int* zcur;
    float cn_rho, cn_s;
    const int* prop_seed;
    float* dx0[3];           // tangents of x_0 (x_0 = mu + sd(phi, sigma) z), or null
    int fisher;              // dx0 receives grad log p(x_0) instead
};

// Slot m receives slot anc[m]'s whole inner filter: parameters, particle
// count, control block, particles (into buffer 0), log-weights, and rows
// 0..rows-1 of its stored noise (into the pool that slot m does not use).
struct GatherArgs {
    const int* anc;
    const int* n_src; const float* mu_src; const float* phi_src; const float* sig_src;
    int* n_dst; float* mu_dst; float* phi_dst; float* sig_dst;
    const Ctl* ctl_src; Ctl* ctl_dst;
    const float* x_src[2]; float* x_dst;
    const float* l_src; float* l_dst;
    float* zpool[2]; float* upool[2]; const int* zcur;
    int rows;              // ticks of stored noise to move (0: none)
    int cap, n_slots;
    const float* dx_src[2][3]; float* dx_dst[3];   // tangents (null: none)
    const double* score_src; double* score_dst;
};

// Accepted moves: slot m takes the replayed filter (parameters, control
// block, particles, log-weights) where acc[m] is set.
struct AcceptArgs {
    const int* acc;
    const float* mu_src; const float* phi_src; const float* sig_src;
    float* mu_dst; float* phi_dst; float* sig_dst;
    const Ctl* ctl_src; Ctl* ctl_dst;
    const float* x_src[2]; float* x_dst[2];
    const float* l_src; float* l_dst;
    int cap;
    const float* dx_src[2][3]; float* dx_dst[2][3];   // tangents (null: none)
    const double* score_src; double* score_dst;
    const int* seed_src; int* seed_dst;               // null: seeds kept
};

}  // namespace bpf
