// bpf_kernels.hip -- gfx950 kernels of one bootstrap-filter tick.
//
// A tick is at most three launches on the filter's stream:
//   k_propagate  (every tick)  read x [and l], write x' [and l'], per-tile
//                              partials {max, sum e, sum e x, sum e x^2, sum e^2}
//   k_finalize   (1 block)     global max / sums in FP64, statistics, FP64
//                              log-likelihood, ESS decision, host-stream
//                              uniform, FP64 tile prefix for the CDF
//   k_resample   (every tick)  exits at once unless the finalise pass chose to
//                              resample; else per tile: FP64 scan of the
//                              weights, claim of the systematic targets that
//                              fall in the tile's CDF interval, coalesced
//                              write of the offspring into the other buffer
// The reference issues 12-16 launches and two host syncs per tick
// (gpu_bpf/gpu_bpf_ptx_full.cu:591-745).
#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include "bpf_common.hpp"

namespace bpf {

// ---------------------------------------------------------------------------
// Initialisation: x_0 = mu + sd * z, z from draw 0 of each particle's stream
// (ptx_assembly/bpf_kernels_full.ptx:103-218, sd of gpu_bpf_ptx_full.cu:430).
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_init(InitArgs a) {
    long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n_pad) return;
    float x = a.mu;
    float al[3] = {0.0f, 0.0f, 0.0f};
    if (i < a.n) {
        uint64_t s = a.pcg_a + (uint64_t)(2 * i + 1) * a.pcg_g;
        float z = icdf(pcg_uniform(pcg_output(s)));
        x = __builtin_fmaf(a.sd, z, a.mu);
        // grad log N(x_0; mu, sd^2) with sd = sd(phi, sigma)
        const float z2m1 = __builtin_fmaf(z, z, -1.0f);
        al[0] = z / a.sd;
        al[1] = z2m1 * (a.dsd_dphi / a.sd);
        al[2] = z2m1 * (a.dsd_dsig / a.sd);
    }
    if (a.x) a.x[i] = x;
    if (a.alpha[0]) {
#pragma unroll
        for (int c = 0; c < 3; ++c) a.alpha[c][i] = al[c];
    }
}

// ---------------------------------------------------------------------------
// The K standard normals of this lane for one draw index: PCG32 state by
// jump-ahead (state = pcg_a + inc_i * pcg_g, inc_i = 2i + 1; the next particle
// of the lane adds 2 pcg_g, the lane's next VEC group 2 BLOCK VEC pcg_g),
// uniform, Acklam ICDF.  Every item is evaluated as central; the tail items
// (4.85% of draws) are compacted through the wave's own LDS queue, in item
// order, by ballot rank, evaluated densely by the wave and read back over
// the central value, so no wave runs the long tail branch for a single lane.
// Item (v, c) of lane tid is particle base + (v*BLOCK+tid)*VEC+c.  s_q holds
// BLOCK*K floats (kWave*K per wave).  No block barrier; the whole wave must
// call it.
// ---------------------------------------------------------------------------

template <int BLOCK, int K>
__device__ inline void lane_normals(uint64_t pcg_a, uint64_t pcg_g, long long base, int tid,
                                    float* s_q, float (&z)[K]) {
    constexpr int VEC = K < 4 ? K : 4;
    constexpr int NV = K / VEC;
    const int lane = tid & (kWave - 1);
    // the wave's queue, as a wave-uniform offset so that each queue address
    // is one shift-add of the lane's rank onto a scalar
    float* q_w = s_q + __builtin_amdgcn_readfirstlane((tid / kWave) * (kWave * K));
    const uint64_t g2 = pcg_g * 2u;
    const uint64_t gv = pcg_g * (uint64_t)(2 * BLOCK * VEC);
    uint64_t sv = pcg_a + (uint64_t)(2 * (base + (long long)tid * VEC) + 1) * pcg_g;
    bool tail[K];
    uint64_t bal[K];                   // the wave's tail lanes of item k
    int tot = 0;                       // tail items queued so far by the wave
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        uint64_t s = sv;
        sv += gv;
#pragma unroll
        for (int c = 0; c < VEC; ++c) {
            const int k = v * VEC + c;
            const float q = pcg_half(pcg_output(s));
            s += g2;
            tail[k] = !icdf_is_central_half(q);
            z[k] = icdf_central(q);     // replaced below for tail items
            bal[k] = __ballot(tail[k]);
            if (tail[k]) q_w[tot + lanes_below(bal[k], 0)] = q;
            tot += __popcll(bal[k]);
        }
    }
    if (tot == 0) return;
    wave_lds_sync();
    for (int e = lane; e < tot; e += kWave) q_w[e] = icdf_tail(uniform_from_half(q_w[e]));
    wave_lds_sync();
    int pos = 0;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        if (tail[k]) z[k] = q_w[pos + lanes_below(bal[k], 0)];
        pos += __popcll(bal[k]);
    }
}

// Agent-scope (every XCD) relaxed loads and stores: a finalise block hands
// its group record to the last block within the same launch.  The stores
// write through the XCD's L2 and the loads bypass it, so no release or
// acquire fence (a write-back or invalidation of the whole L2) is needed.
template <class T>
__device__ inline T ld_agent(const T* p) {
    return __hip_atomic_load(const_cast<T*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <class T>
__device__ inline void st_agent(T* p, T v) {
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------------------
// Group fold (one block of kGroup threads, thread t <-> tile g*kGroup + t):
// the group's tile partials, written by the previous kernel, rescaled to the
// group maximum, with each tile's scale f and in-group inclusive prefix of
// f*s (made non-decreasing within each wave; across waves the running offset
// ends every wave exactly where the next begins) for the resample pass, and
// the group's sums.
// ---------------------------------------------------------------------------

// Returns the calling thread's tile scale f (0 past the last tile); thread 0
// also receives the group record it stores (rec_m, rec_s).
__device__ inline double group_fold(const Partials& p, int g, int tid, float* s_f, double* s_a,
                                    double (*s_d)[kGroup / kWave], float& rec_m, double (&rec_s)[4]) {
    constexpr int NW = kGroup / kWave;
    const int lane = tid & (kWave - 1), wid = tid / kWave;
    const int t = g * kGroup + tid;
    const bool in = t < p.n_tiles;
    constexpr float NEG_INF = -__builtin_inff();
    const float m = in ? p.m[t] : NEG_INF;
    const double s = in ? p.s[t] : 0.0;
    const double sx = in ? p.sx[t] : 0.0;
    const double sxx = in ? p.sxx[t] : 0.0;
    const double sq = in ? p.sq[t] : 0.0;
    const float wm = wave_max(m);
    if (lane == 0) s_f[wid] = wm;
    __syncthreads();
    float gmax = s_f[0];
#pragma unroll
    for (int w = 1; w < NW; ++w) gmax = fmaxf(gmax, s_f[w]);
    const double f = (m == NEG_INF) ? 0.0 : (double)expf_cw_nonpos(m - gmax);
    double vs = f * s, vx = f * sx, vxx = f * sxx, vq = (f * f) * sq;
    const double pin = wave_incl_max(wave_incl_scan(vs, lane), lane);
    if (lane == kWave - 1) s_a[wid] = pin;
    vs = wave_sum(vs);
    vx = wave_sum(vx);
    vxx = wave_sum(vxx);
    vq = wave_sum(vq);
    if (lane == 0) {
        s_d[0][wid] = vs;
        s_d[1][wid] = vx;
        s_d[2][wid] = vxx;
        s_d[3][wid] = vq;
    }
    __syncthreads();
    double off = 0.0;
    for (int w = 0; w < wid; ++w) off += s_a[w];
    if (in) {
        p.tf[t] = f;
        p.tp[t] = off + pin;
    }
    if (tid == 0) {
        double r[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            double acc = s_d[q][0];
#pragma unroll
            for (int w = 1; w < NW; ++w) acc += s_d[q][w];
            r[q] = acc;
        }
        st_agent(p.gm + g, gmax);
        st_agent(p.gs + g, r[0]);
        st_agent(p.gsx + g, r[1]);
        st_agent(p.gsxx + g, r[2]);
        st_agent(p.gsq + g, r[3]);
        rec_m = gmax;
#pragma unroll
        for (int q = 0; q < 4; ++q) rec_s[q] = r[q];
    }
    return f;
}

// ---------------------------------------------------------------------------
// Propagate + weight + tile partials.
// Tile b holds particles [b*TILE, (b+1)*TILE); lane `tid` owns VEC-wide
// groups at (v*kBlock + tid)*VEC, so every load and store instruction of a
// wave covers one contiguous 64*VEC*4-byte run.
// ICDF tails (|u-1/2| >= 0.47575, 4.85% of draws) would make nearly every
// wave execute the ~60-instruction tail branch; instead each wave compacts its
// tail uniforms into its own LDS queue and evaluates them densely.
// ---------------------------------------------------------------------------

// Log-weights of the lane's K particles (Gaussian or Student-t chosen once
// per block, so the unrolled items interleave), carry added, padding set to
// -inf when MASK (the tile's last particles are padding); returns the lane
// maximum.  Item (v, c) is tile particle (v*kBlock + tid)*VEC + c; rem = live
// particles of the tile.
template <int K, bool STUDENT, bool MASK>
__device__ inline float lane_log_weights(const float (&x)[K], float (&l)[K], const ObsModel& obs,
                                         float moff, int tid, int rem) {
    constexpr int VEC = K < 4 ? K : 4;
    constexpr int NV = K / VEC;
    float mloc = -__builtin_inff();
#pragma unroll
    for (int v = 0; v < NV; ++v) {
#pragma unroll
        for (int c = 0; c < VEC; ++c) {
            const int k = v * VEC + c;
            const int li = (v * kBlock + tid) * VEC + c;
            const float lw = STUDENT ? log_weight_t(x[k], obs) : log_weight_gauss(x[k], obs.ysq);
            const float lk = lw + (l[k] - moff);
            l[k] = (!MASK || li < rem) ? lk : -__builtin_inff();
            mloc = fmaxf(mloc, l[k]);
        }
    }
    return mloc;
}

// FP64 sums e, e x, e x^2, e^2 of the lane's particles, e = exp(l - mb);
// padding has l = -inf, so e = 0, and its x is replaced by 0 when MASK.
template <int K, bool MASK>
__device__ inline void lane_weight_sums(const float (&x)[K], const float (&l)[K], float mb, int tid,
                                        int rem, double& S, double& SX, double& SXX, double& Q) {
    constexpr int VEC = K < 4 ? K : 4;
    constexpr int NV = K / VEC;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
#pragma unroll
        for (int c = 0; c < VEC; ++c) {
            const int k = v * VEC + c;
            const int li = (v * kBlock + tid) * VEC + c;
            const float e = expf_cw_nonpos(l[k] - mb);
            const double de = (double)e;
            const double dx = (!MASK || li < rem) ? (double)x[k] : 0.0;
            const double t = de * dx;
            S += de;
            SX += t;
            SXX = fma(t, dx, SXX);
            Q = fma(de, de, Q);
        }
    }
}

// FP64 sums e * alpha_c of the lane's particles (padding contributes zero).
template <int K>
__device__ inline void lane_fisher_sums(const float (&l)[K], const float (&al)[3][K], float mb, int tid,
                                        int rem, double (&F)[3]) {
    constexpr int VEC = K < 4 ? K : 4;
    constexpr int NV = K / VEC;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
#pragma unroll
        for (int c = 0; c < VEC; ++c) {
            const int k = v * VEC + c;
            const int li = (v * kBlock + tid) * VEC + c;
            const double de = (li < rem) ? (double)expf_cw_nonpos(l[k] - mb) : 0.0;
#pragma unroll
            for (int q = 0; q < 3; ++q) F[q] = fma(de, (double)al[q][k], F[q]);
        }
    }
}

// Learning variant of lane_log_weights: also returns g = d lw/dx at each
// particle from the log-weight's own eta^2 (learn_dlw_eta; the log-weight
// bits are those of lane_log_weights).
template <int K, bool STUDENT, bool MASK>
__device__ inline float lane_log_weights_grad(const float (&x)[K], float (&l)[K], float (&g)[K],
                                              const ObsModel& obs, float moff, int tid, int rem) {
    constexpr int VEC = K < 4 ? K : 4;
    constexpr int NV = K / VEC;
    float mloc = -__builtin_inff();
#pragma unroll
    for (int v = 0; v < NV; ++v) {
#pragma unroll
        for (int c = 0; c < VEC; ++c) {
            const int k = v * VEC + c;
            const int li = (v * kBlock + tid) * VEC + c;
            const float eta2 = obs.ysq * expf_cw(-x[k]);
            const float lw = STUDENT ? log_weight_t_eta(x[k], eta2, obs) : log_weight_gauss_eta(x[k], eta2);
            g[k] = learn_dlw_eta<STUDENT>(eta2, obs);
            const float lk = lw + (l[k] - moff);
            l[k] = (!MASK || li < rem) ? lk : -__builtin_inff();
            mloc = fmaxf(mloc, l[k]);
        }
    }
    return mloc;
}

// lane_weight_sums plus the FP64 learning sums e * v_q, v = (g, g^2, g d,
// g^2 d^2) with d = x_prev - mu (ptx_assembly/bpf_kernels_full.ptx:1191-1208
// with w = e); padding contributes zero.
template <int K, bool MASK>
__device__ inline void lane_weight_learn_sums(const float (&x)[K], const float (&l)[K],
                                              const float (&g)[K], const float (&dp)[K], float mb,
                                              int tid, int rem, double& S, double& SX, double& SXX,
                                              double& Q, double (&L)[4]) {
    constexpr int VEC = K < 4 ? K : 4;
    constexpr int NV = K / VEC;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
#pragma unroll
        for (int c = 0; c < VEC; ++c) {
            const int k = v * VEC + c;
            const int li = (v * kBlock + tid) * VEC + c;
            const bool live = !MASK || li < rem;
            const double de = live ? (double)expf_cw_nonpos(l[k] - mb) : 0.0;
            const double dx = live ? (double)x[k] : 0.0;
            const double t = de * dx;
            S += de;
            SX += t;
            SXX = fma(t, dx, SXX);
            Q = fma(de, de, Q);
            const double gg = live ? (double)g[k] : 0.0;
            const double d = live ? (double)dp[k] : 0.0;
            const double t0 = de * gg;
            const double t1 = t0 * gg;
            L[0] += t0;
            L[1] += t1;
            L[2] += t0 * d;
            L[3] += (t1 * d) * d;
        }
    }
}

// FISH: also carry the Fisher-identity vector alpha = grad log p(x_0:t) of
// each particle through the transition (alpha += grad log N(x'; mu + phi
// (x - mu), sigma^2), with z = (x' - mu - phi (x - mu)) / sigma the drawn
// normal) and write the per-tile sums of e * alpha.
// LEARN: also write the per-tile sums of online learning (lane_weight_learn_sums);
// d = x_prev - mu is 0 on tick 0, which propagates nothing.
// TSTATE: Student-t state noise, eps = z / sqrt(chi2 / nu) with chi2 from
// nu_int further normals per particle (ptx_assembly/bpf_kernels_full.ptx:
// 393-406, chi2 as gpu_bpf/gpu_bpf_ptx_full.cu:50-62).
// APF1 / APF2: the two stages of an auxiliary-filter tick
// (gpu_bpf/gpu_bpf_ptx_full.cu:990-1047,1102-1170).  APF1 propagates nothing
// and weights each particle at its predictive mean mp = mu + phi (x - mu);
// APF2 propagates and weights by lw(x') - lw(mp), mp from the (resampled)
// pre-propagation state.
enum PropMode { kFish = 1, kLearn = 2, kTState = 4, kApf1 = 8, kApf2 = 16 };

template <int K, int MODE>
__global__ __launch_bounds__(kBlock) void k_propagate(PropArgs a) {
    constexpr bool FISH = MODE & kFish, LEARN = MODE & kLearn, TSTATE = MODE & kTState;
    constexpr bool APF1 = MODE & kApf1, APF2 = MODE & kApf2;
    static_assert(!(FISH && LEARN) && !(FISH && TSTATE), "one extra statistic per launch");
    static_assert(!((APF1 || APF2) && (FISH || LEARN)) && !(APF1 && (APF2 || TSTATE)),
                  "auxiliary-filter stages take no extra statistic");
    constexpr int VEC = K < 4 ? K : 4;
    constexpr int NV = K / VEC;
    constexpr int TILE = kBlock * K;
    constexpr int NW = kBlock / kWave;
    // x and l stream through with non-temporal loads and stores: measured on
    // MI355X at 51.2M particles, 149 us per launch against 156 us with plain
    // ones (non-temporal stores alone: 156 us, loads alone: 152 us)
    constexpr bool NTL = true, NTS = true;
    __shared__ float s_q[TILE];
    __shared__ float s_m[NW];
    __shared__ double s_d[FISH ? 7 : (LEARN ? 8 : 4)][NW];

    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1);
    const int wid = tid / kWave;
    const int b = blockIdx.x;
    const long long base = (long long)b * TILE;

    const int cur = a.ctl->cur;
    const int carry = a.ctl->carry;
    const float mprev = a.ctl->m_prev;
    float* __restrict__ xb = cur ? a.x1 : a.x0;

    float x[K], l[K];
#pragma unroll
    for (int k = 0; k < K; ++k) l[k] = 0.0f;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        long long off = base + (long long)(v * kBlock + tid) * VEC;
        if constexpr (NTL) load_vec_nt<VEC>(xb + off, &x[v * VEC]);
        else load_vec<VEC>(xb + off, &x[v * VEC]);
    }
    if (carry) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            long long off = base + (long long)(v * kBlock + tid) * VEC;
            if constexpr (NTL) load_vec_nt<VEC>(a.lw + off, &l[v * VEC]);
            else load_vec<VEC>(a.lw + off, &l[v * VEC]);
        }
    }

    if (a.x_save) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            long long off = base + (long long)(v * kBlock + tid) * VEC;
            store_vec<VEC>(a.x_save + off, &x[v * VEC]);
        }
    }

    ObsModel obs = a.obs;
    uint64_t pcg_a = a.pcg_a, pcg_g = a.pcg_g;
    int do_prop = a.do_prop;
    if (a.tick_from_ctl) {
        // a captured tick: this tick's index, observation and draw constants
        // from the control block
        const int tick = a.ctl->tick;
        const float yv = a.ctl->y_base[tick - a.ctl->y_tick0];
        obs.ysq = yv * yv;
        uint64_t A, G;
        pcg_draw_constants((uint32_t)tick, A, G);
        pcg_a = (uint64_t)(int64_t)a.seed * A;
        pcg_g = G;
        do_prop = tick > 0;
    } else {
        const float yv = a.y_ptr ? *a.y_ptr : a.y;
        obs.ysq = yv * yv;
    }

    float al[FISH ? 3 : 1][K];
    if constexpr (FISH) {
#pragma unroll
        for (int q = 0; q < 3; ++q)
#pragma unroll
            for (int v = 0; v < NV; ++v) {
                long long off = base + (long long)(v * kBlock + tid) * VEC;
                load_vec<VEC>(a.alpha[cur][q] + off, &al[q][v * VEC]);
            }
    }

    float dp[LEARN ? K : 1];
    if constexpr (LEARN) {
#pragma unroll
        for (int k = 0; k < K; ++k) dp[k] = do_prop ? x[k] - a.mu : 0.0f;
    }
    // auxiliary filter: predictive means mp = mu + phi (x - mu), the
    // transition without its noise term
    float mp[(APF1 || APF2) ? K : 1];
    if constexpr (APF1 || APF2) {
#pragma unroll
        for (int k = 0; k < K; ++k) {
            float d = x[k] - a.mu;
            d = a.phi * d;
            mp[k] = a.mu + d;
        }
    }
    if constexpr (APF2) {
        // second-stage weight lw(x') - lw(mp): start l at -lw(mp)
#pragma unroll
        for (int k = 0; k < K; ++k) l[k] = -log_weight(mp[k], obs);
    }

    if (!APF1 && do_prop) {
        float z[K];
        lane_normals<kBlock, K>(pcg_a, pcg_g, base, tid, s_q, z);
        if constexpr (FISH) {
#pragma unroll
            for (int k = 0; k < K; ++k) {
                const float zs = z[k] / a.sig;
                al[0][k] = __builtin_fmaf(zs, 1.0f - a.phi, al[0][k]);
                al[1][k] = __builtin_fmaf(zs, x[k] - a.mu, al[1][k]);
                al[2][k] += __builtin_fmaf(z[k], z[k], -1.0f) / a.sig;
            }
#pragma unroll
            for (int q = 0; q < 3; ++q)
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    long long off = base + (long long)(v * kBlock + tid) * VEC;
                    store_vec<VEC>(a.alpha[cur][q] + off, &al[q][v * VEC]);
                }
        }
        if constexpr (TSTATE) {
            // chi2 = sum of nu_int squared normals, in order (FP32)
#pragma unroll
            for (int v = 0; v < NV; ++v)
#pragma unroll
                for (int c = 0; c < VEC; ++c) {
                    const int k = v * VEC + c;
                    const long long i = base + (long long)(v * kBlock + tid) * VEC + c;
                    float chi2 = 0.0f;
                    for (int q = 0; q < a.nu_int; ++q) {
                        const uint64_t j = (uint64_t)i * (uint64_t)a.nu_int + (uint64_t)q;
                        const float g = icdf(pcg_uniform(pcg_output(a.chi_a + (2u * j + 1u) * a.chi_g)));
                        chi2 = __builtin_fmaf(g, g, chi2);
                    }
                    z[k] = z[k] * (1.0f / __builtin_sqrtf(chi2 / a.nu_state));
                }
        }
        // sigma_z * k_band: one scale for the tile unless a band boundary
        // falls inside it (default: one band, k = 1)
        float sig_t = a.sig * a.bands.k[0];
        bool split = false;
        if (a.bands.n > 1) {
            const long long last = base + TILE - 1;
            sig_t = a.sig * band_k(a.bands, base);
            split = band_of(a.bands, base) != band_of(a.bands, last);
        }
        if (split) {
#pragma unroll
            for (int v = 0; v < NV; ++v)
#pragma unroll
                for (int c = 0; c < VEC; ++c) {
                    const int k = v * VEC + c;
                    const long long i = base + (long long)(v * kBlock + tid) * VEC + c;
                    x[k] = transition(x[k], a.mu, a.phi, a.sig * band_k(a.bands, i), z[k]);
                }
        } else {
#pragma unroll
            for (int k = 0; k < K; ++k) x[k] = transition(x[k], a.mu, a.phi, sig_t, z[k]);
        }
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            long long off = base + (long long)(v * kBlock + tid) * VEC;
            if constexpr (NTS) store_vec_nt<VEC>(xb + off, &x[v * VEC]);
            else store_vec<VEC>(xb + off, &x[v * VEC]);
        }
    }

    // log-weights, with the carry of gpu_bpf_ptx_full.cu:601-604 restated
    // relative to the previous maximum; without a carry l holds zeros and the
    // offset is zero, so the same expression adds exact zeros
    const float moff = carry ? mprev : 0.0f;
    const int rem = (a.n - base >= TILE) ? TILE : (int)(a.n - base);   // live particles
    const bool full = rem == TILE;
    float mloc_l;
    float g[LEARN ? K : 1];
    if constexpr (APF1) {
        // first stage: log-weights at the predictive means
        if (obs.nu > 0.0f) mloc_l = lane_log_weights<K, true, true>(mp, l, obs, moff, tid, rem);
        else mloc_l = lane_log_weights<K, false, true>(mp, l, obs, moff, tid, rem);
    } else if constexpr (LEARN) {
        if (obs.nu > 0.0f) mloc_l = lane_log_weights_grad<K, true, true>(x, l, g, obs, moff, tid, rem);
        else mloc_l = lane_log_weights_grad<K, false, true>(x, l, g, obs, moff, tid, rem);
    } else {
        if (obs.nu > 0.0f) mloc_l = lane_log_weights<K, true, true>(x, l, obs, moff, tid, rem);
        else if (full) mloc_l = lane_log_weights<K, false, false>(x, l, obs, moff, tid, rem);
        else mloc_l = lane_log_weights<K, false, true>(x, l, obs, moff, tid, rem);
    }
    if (a.store_l) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            long long off = base + (long long)(v * kBlock + tid) * VEC;
            if constexpr (NTS) store_vec_nt<VEC>(a.lw + off, &l[v * VEC]);
            else store_vec<VEC>(a.lw + off, &l[v * VEC]);
        }
    }

    const float mloc = wave_max(mloc_l);
    if (lane == 0) s_m[wid] = mloc;
    __syncthreads();
    float mb = s_m[0];
#pragma unroll
    for (int w = 1; w < NW; ++w) mb = fmaxf(mb, s_m[w]);

    double S = 0.0, SX = 0.0, SXX = 0.0, Q = 0.0;
    double L[LEARN ? 4 : 1] = {};
    if (mb != -__builtin_inff()) {      // a tile without live weight keeps zero sums
        if constexpr (LEARN) {
            lane_weight_learn_sums<K, true>(x, l, g, dp, mb, tid, rem, S, SX, SXX, Q, L);
        } else {
            if (full) lane_weight_sums<K, false>(x, l, mb, tid, rem, S, SX, SXX, Q);
            else lane_weight_sums<K, true>(x, l, mb, tid, rem, S, SX, SXX, Q);
        }
    }
    S = wave_sum(S);
    SX = wave_sum(SX);
    SXX = wave_sum(SXX);
    Q = wave_sum(Q);
    if (lane == 0) {
        s_d[0][wid] = S;
        s_d[1][wid] = SX;
        s_d[2][wid] = SXX;
        s_d[3][wid] = Q;
    }
    if constexpr (FISH) {
        double F[3] = {0.0, 0.0, 0.0};
        if (mb != -__builtin_inff()) lane_fisher_sums<K>(l, al, mb, tid, rem, F);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            F[q] = wave_sum(F[q]);
            if (lane == 0) s_d[4 + q][wid] = F[q];
        }
    }
    if constexpr (LEARN) {
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            L[q] = wave_sum(L[q]);
            if (lane == 0) s_d[4 + q][wid] = L[q];
        }
    }
    __syncthreads();
    if (tid == 0) {
        double r[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            double acc = s_d[q][0];
#pragma unroll
            for (int w = 1; w < NW; ++w) acc += s_d[q][w];
            r[q] = acc;
        }
        a.part.m[b] = mb;
        a.part.s[b] = r[0];
        a.part.sx[b] = r[1];
        a.part.sxx[b] = r[2];
        a.part.sq[b] = r[3];
        if constexpr (FISH) {
#pragma unroll
            for (int q = 0; q < 3; ++q) {
                double acc = s_d[4 + q][0];
#pragma unroll
                for (int w = 1; w < NW; ++w) acc += s_d[4 + q][w];
                a.fsum[(size_t)q * a.part.n_tiles + b] = acc;
            }
        }
        if constexpr (LEARN) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                double acc = s_d[4 + q][0];
#pragma unroll
                for (int w = 1; w < NW; ++w) acc += s_d[4 + q][w];
                a.lsum[(size_t)q * a.part.n_tiles + b] = acc;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Finalise: one block per reduction group folds the group's tiles; the block
// that arrives last (arrival counter; each block's agent-scope record stores
// are acknowledged before it arrives) reduces the group records, takes the
// tick's decisions and, on resampling ticks, builds the FP64 group prefix.
// ---------------------------------------------------------------------------
constexpr int kFinBlock = kGroup;
constexpr int kFinPer = kMaxGroups / kFinBlock;   // groups per thread of the last block

__global__ __launch_bounds__(kFinBlock) void k_finalize(FinArgs a) {
    constexpr int NW = kFinBlock / kWave;
    __shared__ float s_f[NW];
    __shared__ double s_a[NW];
    __shared__ double s_d[4][NW];
    __shared__ int s_flag;
    __shared__ double s_S;
    __shared__ int s_last;
    __shared__ double s_l[4][NW];
    __shared__ double s_glq[4];        // this block's group learning sums

    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const int Gg = a.part.n_groups;
    constexpr float NEG_INF = -__builtin_inff();

    // A filter of one reduction group (up to 256 tiles) is folded by this one
    // block: it is its own last block, and thread 0 holds the group record,
    // so the arrival counter and the record's round trip through memory are
    // skipped (the record is still stored for the resample pass).
    const bool single = gridDim.x == 1;
    float rec_m = NEG_INF;
    double rec_s[4] = {0.0, 0.0, 0.0, 0.0};
    const double ft = group_fold(a.part, blockIdx.x, tid, s_f, s_a, s_d, rec_m, rec_s);
    if (a.lsum) {
        // the group's online-learning sums: tile sums scaled by the tile scales
        const int t = blockIdx.x * kGroup + tid;
        const int G = a.part.n_tiles;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            double v = (t < G) ? ft * a.lsum[(size_t)q * G + t] : 0.0;
            v = wave_sum(v);
            if (lane == 0) s_l[q][wid] = v;
        }
        __syncthreads();
        if (tid < 4) {
            double acc = s_l[tid][0];
#pragma unroll
            for (int w = 1; w < NW; ++w) acc += s_l[tid][w];
            st_agent(a.glsum + (size_t)tid * Gg + blockIdx.x, acc);
            s_glq[tid] = acc;
        }
        __syncthreads();
    }
    if (single && !a.lsum) {
        // One group: the global fold below reduces to the group record itself
        // (its scale exp(gm - M) is exactly 1, every other term +0, the group
        // prefix is {0, 1}), so thread 0 applies it directly; adding +0 keeps
        // the fold's one effect on the record, -0 becoming +0.
        if (tid == 0) {
            Ctl c = *a.ctl;
            finalize_tick(c, a.n, a.threshold, rec_m, rec_s[0] + 0.0, rec_s[1] + 0.0, rec_s[2] + 0.0,
                          rec_s[3] + 0.0);
            *a.ctl = c;
            a.ring[c.last.tick % a.ring_cap] = c.last;
            if (c.resample) {
                if (a.u_log) *a.u_log = c.u;
                if (a.w_log) *a.w_log = c.wrap_start;
                a.gptil[0] = 0.0;
                a.gptil[1] = 1.0;
#pragma unroll
                for (int r = 0; r < kClaimRanges; ++r) a.tile_ctr[r * kClaimStride] = 0;
            }
        }
        return;
    }
    if (!single) {
        if (tid == 0) {
            // the group record (and its learning sums, stored by lanes 1-3 of
            // this wave) acknowledged before the group counts itself in; the
            // signal fences keep the compiler from moving stores or the
            // counter update across the wait
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            __builtin_amdgcn_s_waitcnt(0);
            __atomic_signal_fence(__ATOMIC_SEQ_CST);
            const int old = atomicAdd(a.part.gcnt, 1);
            s_last = (old == (int)gridDim.x - 1);
        }
        __syncthreads();
        if (!s_last) return;
        // the records are read below at agent scope: no acquire fence
        if (tid == 0) *a.part.gcnt = 0;
    }

    // thread tid owns the contiguous groups [tid*C, tid*C + C), C <= kFinPer;
    // their records are read once
    const int C = (Gg + kFinBlock - 1) / kFinBlock;
    const int g0 = tid * C;
    float gm[kFinPer];
    double gs[kFinPer], gsx[kFinPer], gsxx[kFinPer], gsq[kFinPer];
    float m = NEG_INF;
#pragma unroll
    for (int j = 0; j < kFinPer; ++j) {
        const int g = g0 + j;
        const bool in = j < C && g < Gg;
        if (single) {             // group 0 is thread 0's (C = 1)
            gm[j] = in ? rec_m : NEG_INF;
            gs[j] = in ? rec_s[0] : 0.0;
            gsx[j] = in ? rec_s[1] : 0.0;
            gsxx[j] = in ? rec_s[2] : 0.0;
            gsq[j] = in ? rec_s[3] : 0.0;
        } else {
            gm[j] = in ? __hip_atomic_load(&a.part.gm[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : NEG_INF;
            gs[j] = in ? __hip_atomic_load(&a.part.gs[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0.0;
            gsx[j] = in ? __hip_atomic_load(&a.part.gsx[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0.0;
            gsxx[j] = in ? __hip_atomic_load(&a.part.gsxx[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0.0;
            gsq[j] = in ? __hip_atomic_load(&a.part.gsq[g], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : 0.0;
        }
        m = fmaxf(m, gm[j]);
    }
    Ctl c;
    if (tid == 0) c = *a.ctl;          // read once, written back once

    m = wave_max(m);
    if (lane == 0) s_f[wid] = m;
    __syncthreads();
    float M = s_f[0];
#pragma unroll
    for (int w = 1; w < NW; ++w) M = fmaxf(M, s_f[w]);

    // group weights v_g = f_g gs_g, f_g = exp(gm_g - M)
    double v[kFinPer];
    double S = 0.0, SX = 0.0, SXX = 0.0, Q = 0.0;
#pragma unroll
    for (int j = 0; j < kFinPer; ++j) {
        const double f = (gm[j] == NEG_INF) ? 0.0 : (double)expf_cw_nonpos(gm[j] - M);
        v[j] = f * gs[j];
        S += v[j];
        SX += f * gsx[j];
        SXX += f * gsxx[j];
        Q += (f * f) * gsq[j];
    }
    S = wave_sum(S);
    SX = wave_sum(SX);
    SXX = wave_sum(SXX);
    Q = wave_sum(Q);
    if (lane == 0) {
        s_d[0][wid] = S;
        s_d[1][wid] = SX;
        s_d[2][wid] = SXX;
        s_d[3][wid] = Q;
    }
    if (a.lsum) {
        // online learning: sum over groups of f_g * (group sums), as the weights
        double Lq[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int j = 0; j < kFinPer; ++j) {
            const int g = g0 + j;
            if (j < C && g < Gg) {
                const double f = (gm[j] == NEG_INF) ? 0.0 : (double)expf_cw_nonpos(gm[j] - M);
#pragma unroll
                for (int q = 0; q < 4; ++q)
                    Lq[q] = fma(f, single ? s_glq[q]
                                          : __hip_atomic_load(&a.glsum[(size_t)q * Gg + g], __ATOMIC_RELAXED,
                                                              __HIP_MEMORY_SCOPE_AGENT), Lq[q]);
            }
        }
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            Lq[q] = wave_sum(Lq[q]);
            if (lane == 0) s_l[q][wid] = Lq[q];
        }
    }
    __syncthreads();

    if (tid == 0 && a.lsum) {
        // the reference's kernel-14 accumulators d_scalars[5..8]
        // (ptx_assembly/bpf_kernels_full.ptx:1373-1377), normalised by the
        // tick's weight sum; a tick without finite positive weight adds nothing
        double St = s_d[0][0];
#pragma unroll
        for (int w = 1; w < NW; ++w) St += s_d[0][w];
        if (St > 0.0 && isfinite(St)) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                double r = s_l[q][0];
#pragma unroll
                for (int w = 1; w < NW; ++w) r += s_l[q][w];
                a.lacc[q] += r / St;
            }
        }
    }
    if (tid == 0) {
        double t4[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            double acc = s_d[q][0];
#pragma unroll
            for (int w = 1; w < NW; ++w) acc += s_d[q][w];
            t4[q] = acc;
        }
        finalize_tick(c, a.n, a.threshold, M, t4[0], t4[1], t4[2], t4[3]);
        *a.ctl = c;
        a.ring[c.last.tick % a.ring_cap] = c.last;
        if (c.resample) {
            if (a.u_log) *a.u_log = c.u;
            if (a.w_log) *a.w_log = c.wrap_start;
        }
        s_flag = c.resample;
        s_S = c.S;
    }
    __syncthreads();
    if (!s_flag) return;
    if (tid < kClaimRanges) a.tile_ctr[tid * kClaimStride] = 0;

    // FP64 exclusive prefix of the group weights, normalised by S, clamped to
    // [0, 1] and made non-decreasing by a prefix maximum (the block scan's
    // association differs between groups, so rounding alone could break the
    // order)
    S = s_S;
    double loc = 0.0;
#pragma unroll
    for (int j = 0; j < kFinPer; ++j) loc += v[j];
    const double incl = wave_incl_scan(loc, lane);
    if (lane == kWave - 1) s_a[wid] = incl;
    __syncthreads();
    double B = wave_prev(incl, 0.0);
    for (int w = 0; w < wid; ++w) B += s_a[w];
    double bt[kFinPer];
    double hi_v = 0.0;
#pragma unroll
    for (int j = 0; j < kFinPer; ++j) {
        bt[j] = fmin(fmax(B / S, 0.0), 1.0);
        if (j < C) hi_v = fmax(hi_v, bt[j]);
        B += v[j];
    }
    __syncthreads();
    const double pm = wave_incl_max(hi_v, lane);
    double pre = wave_prev(pm, 0.0);
    if (lane == kWave - 1) s_a[wid] = pm;
    __syncthreads();
    for (int w = 0; w < wid; ++w) pre = fmax(pre, s_a[w]);
#pragma unroll
    for (int j = 0; j < kFinPer; ++j) {
        const int g = g0 + j;
        pre = fmax(pre, bt[j]);
        if (j < C && g < Gg) a.gptil[g] = (g == 0) ? 0.0 : pre;
    }
    if (tid == 0) a.gptil[Gg] = 1.0;
}

// ---------------------------------------------------------------------------
// Resample.  Tile b owns the CDF interval (Bt[b], Bt[b+1]] (tile 0 also owns
// 0); its in-tile CDF is c_j = min(Bt[b] + ft[b] * L_j, Bt[b+1]) with L_j the
// FP64 inclusive scan of e_j = exp(l_j - m_b), made non-decreasing, and its
// last valid element pinned to Bt[b+1].  Output i (target t_i = u + i/N,
// wrapped once at 1) is written by the tile whose interval holds t_i, with
// ancestor = lower_bound(c, t_i): the reference's rule
// (gpu_bpf/gpu_bpf.cu:297-310, clamp to N-1 as ptx:816-830) on this CDF.
// ---------------------------------------------------------------------------

struct NoPre {
    __device__ void operator()() const {}
};

// Builds the tile's CDF in s_c (TILE doubles) from the lane's weights e[K]
// (coalesced layout) and returns the tile's output ranges in s_rng:
// [0,1) unwrapped targets, [2,3) wrapped targets.  Returns the local index of
// the tile's last valid particle.  s_c doubles as the staging area of e.
// next_out (nullable): thread 0 stores `next` there before the first barrier;
// pre() runs on every thread right after that barrier (the caller's next
// tile loads, issued while this tile's CDF is built).
// Contains block barriers; every thread of the block must call it.
template <int BLOCK, int K, class Pre = NoPre>
__device__ inline int tile_cdf(const float (&e)[K], int b, int n, int n_tiles, double ft,
                               double Bb, double Bb1, float u, int W, double* s_c,
                               double* s_w, int* s_rng, int* next_out = nullptr, int next = 0,
                               Pre pre = Pre()) {
    constexpr int VEC = K < 4 ? K : 4;
    constexpr int NV = K / VEC;
    constexpr int TILE = BLOCK * K;
    constexpr int PAD = TILE + TILE / 16 + 1;
    static_assert(PAD * sizeof(float) <= (TILE + TILE / 16) * sizeof(double), "e staging must fit in s_c");
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const long long base = (long long)b * TILE;
    float* s_e = reinterpret_cast<float*>(s_c);

#pragma unroll
    for (int v = 0; v < NV; ++v) {
#pragma unroll
        for (int c = 0; c < VEC; ++c) {
            const int loc = (v * BLOCK + tid) * VEC + c;
            s_e[loc + (loc >> 4)] = e[v * VEC + c];
        }
    }
    if (tid == 0 && next_out) *next_out = next;
    __syncthreads();
    pre();

    // lane-contiguous FP64 scan: lane tid owns tile positions [tid*K, tid*K+K)
    float ev[K];
    double tot = 0.0;
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const int j = tid * K + q;
        ev[q] = s_e[j + (j >> 4)];
        tot += (double)ev[q];
    }
    double incl = wave_incl_scan(tot, lane);
    if (lane == kWave - 1) s_w[wid] = incl;
    __syncthreads();
    double excl = incl - tot;
    for (int w = 0; w < wid; ++w) excl += s_w[w];
    excl = (tid == 0) ? 0.0 : excl;

    const long long lastll = (b == n_tiles - 1) ? ((long long)n - 1 - base) : (long long)(TILE - 1);
    const int last = (int)lastll;
    // c_j = min(Bb + ft L_j, Bb1) is non-decreasing within the lane; the
    // lane's last value first (the same sequential sum, recomputed below
    // rather than held in registers)
    double L = excl;
#pragma unroll
    for (int q = 0; q < K; ++q) L += (double)ev[q];
    // make the in-tile CDF non-decreasing across lanes (prefix max)
    double pm = wave_incl_max(fmin(fma(ft, L, Bb), Bb1), lane);
    double prev = wave_prev(pm, Bb);
    __syncthreads();   // s_w reads and s_e (aliasing s_c) reads are done
    if (lane == kWave - 1) s_w[wid] = pm;
    __syncthreads();
    for (int w = 0; w < wid; ++w) prev = fmax(prev, s_w[w]);
    L = excl;
#pragma unroll
    for (int q = 0; q < K; ++q) {
        const int j = tid * K + q;
        L += (double)ev[q];
        double c = fmax(fmin(fma(ft, L, Bb), Bb1), prev);
        c = (j == last) ? Bb1 : c;
        c = (j > last) ? __builtin_inf() : c;
        s_c[pad16(j)] = c;
    }
    if (tid < 4) {
        // lanes 0, 1: unwrapped targets [0, W) at or below Bb, Bb1; lanes 2,
        // 3: wrapped targets [W, n) likewise (the first and last tiles own
        // their range ends)
        const bool up = tid & 1, wr = tid >= 2;
        const int lo = wr ? W : 0, hi = wr ? n : W;
        const bool end = up ? (b == n_tiles - 1) : (b == 0);
        s_rng[tid] = end ? (up ? hi : lo)
                         : count_targets_le(up ? Bb1 : Bb, u, wr ? 1.0f : 0.0f, lo, hi, n, target_recip(n));
    }
    __syncthreads();
    return last;
}

// Weight scale and CDF interval of tile b, from the finalise pass's group
// prefix and the tile's stored scale and in-group prefix: tile b owns the
// interval (b0, b1] of the normalised CDF and its b1 is, bit for bit, tile
// b+1's b0 (same operations on the same stored values).
struct TileScale {
    double ft;   // normalised weight of particle j = e_j * ft
    double b0, b1;
};

__device__ inline TileScale tile_scale(const Partials& p, const double* gptil, float M, double S,
                                       int b) {
    constexpr float NEG_INF = -__builtin_inff();
    const int g = b / kGroup;
    const int l = b - g * kGroup;
    const float gm = p.gm[g];
    const double fg = (gm == NEG_INF) ? 0.0 : (double)expf_cw_nonpos(gm - M);
    const double rS = 1.0 / S;
    const double g0 = gptil[g], g1 = gptil[g + 1];
    const double pex = (l == 0) ? 0.0 : p.tp[b - 1];
    const double pin = p.tp[b];
    TileScale ts;
    ts.ft = (p.tf[b] * fg) * rS;
    ts.b0 = fmin(fma(fg * pex, rS, g0), g1);
    ts.b1 = (l == kGroup - 1 || b == p.n_tiles - 1) ? g1 : fmin(fma(fg * pin, rS, g0), g1);
    return ts;
}

// Next tile for a block that works in range r (tiles [r R, (r+1) R)): from
// its own range while that lasts, then from the others in turn; n_tiles when
// all are exhausted.  r and tried (ranges found exhausted) are the block's
// state (thread 0).
__device__ inline int claim_tile(int* ctr, int& r, int& tried, int n_tiles) {
    const int R = (n_tiles + kClaimRanges - 1) / kClaimRanges;
    while (tried < kClaimRanges) {
        const int lo = r * R;
        const int len = min(R, n_tiles - lo);
        if (len > 0) {
            const int t = atomicAdd(&ctr[r * kClaimStride], 1);
            if (t < len) return lo + t;
        }
        r = (r + 1) % kClaimRanges;
        ++tried;
    }
    return n_tiles;
}

// FISH: the Fisher-identity vectors are gathered with x (staged in LDS).
template <int BLOCK, int K, bool FISH>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(FISH ? 1 : 4))) void k_resample(ResArgs a) {
    Ctl* ctl = a.ctl;
    if (!ctl->resample) return;
    constexpr int VEC = K < 4 ? K : 4;
    constexpr int NV = K / VEC;
    constexpr int TILE = BLOCK * K;
    constexpr int TP = TILE + TILE / 16;   // padded length
    constexpr int NW = BLOCK / kWave;
    __shared__ float s_x[TP];
    __shared__ float s_o[FISH ? 1 : TP];   // staged offspring of one chunk
    __shared__ double s_c[TP];
    __shared__ double s_w[NW];
    __shared__ int s_rng[4];
    __shared__ int s_first, s_tile;      // first tile; next tile (written each tile)
    __shared__ float s_al[FISH ? 3 : 1][FISH ? TP : 1];

    const int tid = threadIdx.x;
    const int n = a.n;
    const double rn = target_recip(n);   // target quotients i/n as i * rn
    const float* __restrict__ xs = ctl->src ? a.x1 : a.x0;
    float* __restrict__ xd = ctl->dst ? a.x1 : a.x0;
    const float* als[3];
    float* ald[3];
#pragma unroll
    for (int q = 0; q < 3; ++q) {
        als[q] = FISH ? a.alpha[ctl->src][q] : nullptr;
        ald[q] = FISH ? a.alpha[ctl->dst][q] : nullptr;
    }
    const float u = ctl->u;
    const int W = ctl->wrap_start;
    const float Mg = ctl->M;
    const double Sg = ctl->S;
    ObsModel obs = a.obs;
    {
        // a captured tick runs after the finalise pass advanced ctl->tick
        // the device observation and the by-value one are selected as values:
        // a select of the two addresses would copy a.y to scratch
        const float y_dev = a.y_ptr ? *a.y_ptr : 0.0f;
        const float yv = a.tick_from_ctl ? ctl->y_base[ctl->tick - 1 - ctl->y_tick0]
                                         : (a.y_ptr ? y_dev : a.y);
        obs.ysq = yv * yv;
    }

    // A grid of resident blocks takes tiles from a work counter (reset by
    // the finalise pass): one launch wave, so a tick without resampling costs
    // one short wave of early exits, and heavy tiles do not pile up on a few
    // blocks.  Each block claims its next tile while it works on the current
    // one, publishes the claim with the tile's CDF (before any store, so
    // waiting on the claim waits on no store) and loads the next tile's
    // states and log-weights while it writes the current tile's offspring;
    // every block leaves once the counter has passed the last tile.
    // (block-uniform values read from LDS go through readfirstlane, so that
    // tile bases are scalar and per-lane addresses 32-bit offsets)
    int cr = blockIdx.x % kClaimRanges, ctried = 0;   // blocks alternate between XCDs
    if (tid == 0) s_first = a.one_tile ? (int)blockIdx.x : claim_tile(a.tile_ctr, cr, ctried, a.part.n_tiles);
    __syncthreads();
    int b = __builtin_amdgcn_readfirstlane(s_first);
    float xv[K], lv[K];
    if (b < a.part.n_tiles) {
        const float* xb = xs + (long long)b * TILE;
        const float* lb = a.lw + (long long)b * TILE;
#pragma unroll
        for (int v = 0; v < NV; ++v) {
            const int loc0 = (v * BLOCK + tid) * VEC;
            load_vec<VEC>(xb + loc0, &xv[v * VEC]);
            if (a.lw) load_vec<VEC>(lb + loc0, &lv[v * VEC]);
        }
    }
    while (b < a.part.n_tiles) {
        int nb = 0;
        if (tid == 0) nb = a.one_tile ? a.part.n_tiles : claim_tile(a.tile_ctr, cr, ctried, a.part.n_tiles);
        const long long base = (long long)b * TILE;
        if constexpr (FISH) {
#pragma unroll
            for (int q = 0; q < 3; ++q)
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    const int loc0 = (v * BLOCK + tid) * VEC;
                    float t[VEC];
                    load_vec<VEC>(als[q] + base + loc0, t);
#pragma unroll
                    for (int c = 0; c < VEC; ++c) s_al[q][pad16(loc0 + c)] = t[c];
                }
        }
        const float mb = a.part.m[b];
        const bool dead = (mb == -__builtin_inff());
        const TileScale ts = tile_scale(a.part, a.gptil, Mg, Sg, b);
        if (a.m_log && tid == 0) {
            a.m_log[b] = mb;
            a.ft_log[b] = ts.ft;
            a.b_log[b] = ts.b0;
            if (b == a.part.n_tiles - 1) a.b_log[b + 1] = ts.b1;
        }

        // log-weights: stored (ESS mode) or recomputed from x (a block-uniform
        // branch, so only one of the two is evaluated)
        if (!a.lw) {
#pragma unroll
            for (int k = 0; k < K; ++k) lv[k] = log_weight(xv[k], obs);
        }
        float e[K];
#pragma unroll
        for (int v = 0; v < NV; ++v) {
#pragma unroll
            for (int c = 0; c < VEC; ++c) {
                const int k = v * VEC + c;
                const int loc = (v * BLOCK + tid) * VEC + c;
                bool valid = base + loc < n;
                e[k] = (dead || !valid) ? 0.0f : expf_cw_nonpos(lv[k] - mb);
                s_x[pad16(loc)] = xv[k];
            }
        }
        // the next tile's loads are issued once its index is published (the
        // first barrier of the CDF build) and overlap the CDF, the range
        // counts and this tile's stores
        auto prefetch = [&]() {
            const int bq = __builtin_amdgcn_readfirstlane(s_tile);
            if (bq < a.part.n_tiles) {
                const float* xb = xs + (long long)bq * TILE;
                const float* lb = a.lw + (long long)bq * TILE;
#pragma unroll
                for (int v = 0; v < NV; ++v) {
                    const int loc0 = (v * BLOCK + tid) * VEC;
                    load_vec<VEC>(xb + loc0, &xv[v * VEC]);
                    if (a.lw) load_vec<VEC>(lb + loc0, &lv[v * VEC]);
                }
            }
        };
        const int last = tile_cdf<BLOCK, K>(e, b, n, a.part.n_tiles, ts.ft, ts.b0, ts.b1, u, W, s_c, s_w,
                                            s_rng, &s_tile, nb, prefetch);
        const int ilo = __builtin_amdgcn_readfirstlane(s_rng[0]);
        const int ihi = __builtin_amdgcn_readfirstlane(s_rng[1]);
        const int jlo = __builtin_amdgcn_readfirstlane(s_rng[2]);
        const int jhi = __builtin_amdgcn_readfirstlane(s_rng[3]);
        const int bn = __builtin_amdgcn_readfirstlane(s_tile);

        // Offspring.  Output i goes to the first parent j with c_j >= t_i
        // (lower_bound; every owned unwrapped target is <= c_last), and both
        // sequences are non-decreasing, so ancestors are too.
        // The outputs are taken in chunks of TILE.  Lane tid owns a
        // contiguous run of ceil(chunk / BLOCK) of them: it finds the first
        // one's parent by a binary search over the tile CDF and each next one
        // by galloping forward from the previous parent (O(log gap) whatever
        // the weight degeneracy), and stages the offspring in LDS, from where
        // the block stores the chunk with coalesced writes (lane-strided runs
        // stored directly cost 24 us more per launch at 51.2M particles).  The
        // Fisher variant keeps the merge path below, which moves the three
        // extra vectors.
        const int n_out = ihi - ilo;
        if (n_out > 0) {
            if constexpr (!FISH) {
                for (int c0 = 0; c0 < n_out; c0 += TILE) {
                    const int cnt = min(TILE, n_out - c0);
                    const int per = (cnt + BLOCK - 1) / BLOCK;
                    const int o0 = min(tid * per, cnt), o1 = min(o0 + per, cnt);
                    const int i0 = ilo + c0;
                    if (o0 < o1) {
                        double t = (double)target_at(u, i0 + o0, rn);
                        int lo = 0, hi = last;
                        while (lo < hi) {
                            const int mid = (lo + hi) >> 1;
                            if (s_c[pad16(mid)] < t) lo = mid + 1; else hi = mid;
                        }
                        int j = lo;
                        // c_j and c_(j+1) stay in registers: another child of
                        // j reads no LDS, a child of j + 1 waits on none (the
                        // refill of c_(j+2) is needed only by the next output)
                        double cj = s_c[pad16(j)];
                        double cn = s_c[pad16(min(j + 1, last))];
                        s_o[pad16(o0)] = s_x[pad16(j)];
                        for (int o = o0 + 1; o < o1; ++o) {
                            t = (double)target_at(u, i0 + o, rn);
                            if (cj < t) {
                                // t <= c_last, so j < last here
                                if (t <= cn) {
                                    ++j;
                                    cj = cn;
                                } else {
                                    // the parent lies in (j + 1, last]:
                                    // gallop, then search
                                    int g0 = j + 2, g1 = g0, step = 1;
                                    while (g1 < last && s_c[pad16(g1)] < t) {
                                        g0 = g1 + 1;
                                        step <<= 1;
                                        g1 = min(g0 + step - 1, last);
                                    }
                                    while (g0 < g1) {
                                        const int mid = (g0 + g1) >> 1;
                                        if (s_c[pad16(mid)] < t) g0 = mid + 1; else g1 = mid;
                                    }
                                    j = g0;
                                    cj = s_c[pad16(j)];
                                }
                                cn = s_c[pad16(min(j + 1, last))];
                            }
                            s_o[pad16(o)] = s_x[pad16(j)];
                        }
                    }
                    __syncthreads();
                    for (int o = tid; o < cnt; o += BLOCK) xd[i0 + o] = s_o[pad16(o)];
                    __syncthreads();    // s_o is refilled by the next chunk
                }
            } else {
                // merge path over the owned targets and the tile CDF:
                // emit output a when t_a <= c_j, else pass parent j; lane tid
                // takes the tid-th slice of the n_out + P steps and finds
                // where it starts by a binary search along the path
                const int P = last + 1;
                const int Lm = n_out + P;
                const int D = (Lm + BLOCK - 1) / BLOCK;
                const int d0 = min(tid * D, Lm), d1 = min(d0 + D, Lm);
                // outputs among the first d0 merge steps: output m comes before
                // parent d0 - m - 1 is passed iff t_m <= c_(d0-m-1)
                int lo = max(0, d0 - P), hi = min(d0, n_out);
                while (lo < hi) {
                    const int mid = (lo + hi) >> 1;
                    const double tm = (double)target_at(u, ilo + mid, rn);
                    if (tm <= s_c[pad16(d0 - mid - 1)]) lo = mid + 1; else hi = mid;
                }
                int oa = lo, j = d0 - lo;
                double cj = (j < P) ? s_c[pad16(j)] : __builtin_inf();
                double tv = (oa < n_out) ? (double)target_at(u, ilo + oa, rn) : 0.0;
                for (int d = d0; d < d1; ++d) {
                    if (oa < n_out && tv <= cj) {
                        const int jp = pad16(j);
                        xd[ilo + oa] = s_x[jp];
#pragma unroll
                        for (int c = 0; c < 3; ++c) ald[c][ilo + oa] = s_al[c][jp];
                        ++oa;
                        tv = (oa < n_out) ? (double)target_at(u, ilo + oa, rn) : 0.0;
                    } else {
                        ++j;
                        cj = (j < P) ? s_c[pad16(j)] : __builtin_inf();
                    }
                }
            }
        }
        // wrapped targets (a few at most)
        for (int i = jlo + tid; i < jhi; i += BLOCK) {
            double t = (double)(target_at(u, i, rn) - 1.0f);
            const int j = pad16(lds_lower_bound(s_c, last, t));
            xd[i] = s_x[j];
            if constexpr (FISH) {
#pragma unroll
                for (int c = 0; c < 3; ++c) ald[c][i] = s_al[c][j];
            }
        }
        __syncthreads();          // the next tile reuses the LDS arrays
        b = bn;
    }
}

// ---------------------------------------------------------------------------
// Adjoint of one always-resample tick (reverse mode, ancestors held fixed).
// With LL = sum_t [logsumexp_j lw(x_t,j) - log N] and
//   x_t,i = mu + phi (x_{t-1, a_i} - mu) + sigma z_t,i,
// the adjoint of particle j at tick t is
//   xbar_t,j = W_t,j lw'(x_t,j) + phi * sum_{i : a_t,i = j} xbar_{t+1,i},
// and the tile accumulates sum_j xbar_t,j * dx_t,j/d(mu, phi, sigma).  The
// children of j are the outputs of the tile's claimed target range with
// c_{j-1} < t_i <= c_j: a contiguous run, found by binary search over the
// non-decreasing targets.  The CDF is rebuilt bit for bit as in k_resample.
// ---------------------------------------------------------------------------

template <int BLOCK, int K>
__global__ __launch_bounds__(BLOCK) void k_adjoint(AdjArgs a) {
    constexpr int VEC = K < 4 ? K : 4;
    constexpr int NV = K / VEC;
    constexpr int TILE = BLOCK * K;
    constexpr int NW = BLOCK / kWave;
    __shared__ int s_bu[TILE];     // children run end (unwrapped targets) per parent
    __shared__ double s_c[TILE + TILE / 16];
    __shared__ double s_w[NW];
    __shared__ int s_rng[4];
    __shared__ double s_red[3][NW];
    float* s_q = reinterpret_cast<float*>(s_bu);   // normal-draw queue, used before s_bu

    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const int b = blockIdx.x;
    const int n = a.n;
    const long long base = (long long)b * TILE;
    const float mb = a.m[b];
    const bool dead = (mb == -__builtin_inff());

    float x[K], pre[K], z[K];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        long long off = base + (long long)(v * BLOCK + tid) * VEC;
        load_vec<VEC>(a.pre + off, &pre[v * VEC]);
    }
    // draw t of the streams (draw 0, the initial one, at tick 0)
    lane_normals<BLOCK, K>(a.pcg_a, a.pcg_g, base, tid, s_q, z);
#pragma unroll
    for (int k = 0; k < K; ++k) x[k] = (a.tick > 0) ? transition(pre[k], a.mu, a.phi, a.sig, z[k]) : pre[k];
    ObsModel obs = a.obs;
    obs.ysq = a.y * a.y;

    float e[K], gp[K];
#pragma unroll
    for (int v = 0; v < NV; ++v) {
#pragma unroll
        for (int c = 0; c < VEC; ++c) {
            const int k = v * VEC + c;
            const bool valid = base + (long long)(v * BLOCK + tid) * VEC + c < n;
            float lk = log_weight(x[k], obs);
            e[k] = (dead || !valid) ? 0.0f : expf_cw(lk - mb);
            float eta2 = obs.ysq * expf_cw(-x[k]);
            gp[k] = (obs.nu > 0.0f) ? obs.nph * eta2 / (obs.nu + eta2) - 0.5f : 0.5f * eta2 - 0.5f;
        }
    }
    const double ft = a.ftil[b];
    const float u = *a.u_ptr;
    const int W = *a.w_ptr;
    const int last = tile_cdf<BLOCK, K>(e, b, n, a.n_tiles, ft, a.btil[b], a.btil[b + 1], u, W,
                                 s_c, s_w, s_rng);
    const int ilo = s_rng[0], ihi = s_rng[1], jlo = s_rng[2], jhi = s_rng[3];

    // children run ends of every parent
    const double rn = target_recip(n);
    if (a.xbar_next) {
#pragma unroll
        for (int v = 0; v < NV; ++v) {
#pragma unroll
            for (int c = 0; c < VEC; ++c) {
                const int loc = (v * BLOCK + tid) * VEC + c;
                const double cj = s_c[pad16(loc)];
                s_bu[loc] = (loc >= last) ? ihi : count_targets_le(cj, u, 0.0f, ilo, ihi, n, rn);
            }
        }
    }
    __syncthreads();

    double g0 = 0.0, g1 = 0.0, g2 = 0.0;
    const double phi = (double)a.phi;
#pragma unroll
    for (int v = 0; v < NV; ++v) {
        float xb_out[VEC];
#pragma unroll
        for (int c = 0; c < VEC; ++c) {
            const int k = v * VEC + c;
            const int loc = (v * BLOCK + tid) * VEC + c;
            const bool valid = base + loc < n;
            double xb = (double)e[k] * ft * (double)gp[k];
            if (a.xbar_next && valid) {
                int u0 = (loc == 0) ? ilo : s_bu[loc - 1];
                int u1 = s_bu[loc];
                double cs = 0.0;
                for (int i = u0; i < u1; ++i) cs += (double)a.xbar_next[i];
                if (jlo < jhi) {   // wrapped targets (a few, usually tile 0's)
                    int w0 = (loc == 0) ? jlo : count_targets_le(s_c[pad16(loc - 1)], u, 1.0f, jlo, jhi, n, rn);
                    int w1 = (loc >= last) ? jhi : count_targets_le(s_c[pad16(loc)], u, 1.0f, jlo, jhi, n, rn);
                    for (int i = w0; i < w1; ++i) cs += (double)a.xbar_next[i];
                }
                xb += phi * cs;
            }
            xb = valid ? xb : 0.0;
            xb_out[c] = (float)xb;
            if (a.tick > 0) {
                g0 += xb * (1.0 - phi);
                g1 += xb * ((double)pre[k] - (double)a.mu);
                g2 += xb * (double)z[k];
            } else if (a.x0_reparam) {
                // x_0 = mu + sd(phi, sigma) z_0
                g0 += xb;
                g1 += xb * (double)z[k] * a.dsd_dphi;
                g2 += xb * (double)z[k] * a.dsd_dsig;
            }
        }
        long long off = base + (long long)(v * BLOCK + tid) * VEC;
        store_vec<VEC>(a.xbar_out + off, xb_out);
    }
    g0 = wave_sum(g0);
    g1 = wave_sum(g1);
    g2 = wave_sum(g2);
    if (lane == 0) {
        s_red[0][wid] = g0;
        s_red[1][wid] = g1;
        s_red[2][wid] = g2;
    }
    __syncthreads();
    if (tid == 0) {
        double r0 = s_red[0][0], r1 = s_red[1][0], r2 = s_red[2][0];
        for (int w = 1; w < NW; ++w) {
            r0 += s_red[0][w];
            r1 += s_red[1][w];
            r2 += s_red[2][w];
        }
        a.theta_acc[3 * b + 0] += r0;
        a.theta_acc[3 * b + 1] += r1;
        a.theta_acc[3 * b + 2] += r2;
    }
}

// ---------------------------------------------------------------------------
// Fisher-identity score of the latest tick, sum_i w_i alpha_i: tile b's
// normalised weight scale is tf_b * exp(gm_g - M) / S (as in tile_scale),
// applied to the tile sums of e * alpha.  One block, fixed order.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_fisher_score(FisherArgs a) {
    constexpr int NW = kBlock / kWave;
    constexpr float NEG_INF = -__builtin_inff();
    __shared__ double s_r[3][NW];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const float M = a.ctl->M;
    const double S = a.ctl->S;
    const int G = a.part.n_tiles;
    double acc[3] = {0.0, 0.0, 0.0};
    for (int b = tid; b < G; b += kBlock) {
        const float gm = a.part.gm[b / kGroup];
        const double fg = (gm == NEG_INF) ? 0.0 : (double)expf_cw_nonpos(gm - M);
        const double f = a.part.tf[b] * fg;
#pragma unroll
        for (int c = 0; c < 3; ++c) acc[c] = fma(f, a.fsum[(size_t)c * G + b], acc[c]);
    }
#pragma unroll
    for (int c = 0; c < 3; ++c) {
        acc[c] = wave_sum(acc[c]);
        if (lane == 0) s_r[c][wid] = acc[c];
    }
    __syncthreads();
    if (tid < 3) {
        double r = s_r[tid][0];
#pragma unroll
        for (int w = 1; w < NW; ++w) r += s_r[tid][w];
        a.out[tid] = r / S;
    }
}

// ---------------------------------------------------------------------------
// Silverman jitter after a resample (gpu_bpf/gpu_bpf_ptx_full.cu:708-739,
// ptx_assembly/bpf_kernels_full.ptx:852-1113): the spread of the resampled
// particles about the tick's mean (FP64 per-block sums, added on the host in
// block order), then x += bw * z with z from one more draw of the particle's
// stream, clamped to [mu - 8, mu + 8].
// ---------------------------------------------------------------------------
constexpr int kJitterGrid = 1024;

__global__ __launch_bounds__(kBlock) void k_jitter_var(JitterArgs a) {
    constexpr int NW = kBlock / kWave;
    __shared__ double s_r[NW];
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    double acc = 0.0;
    for (long long i = (long long)blockIdx.x * kBlock + tid; i < a.n; i += (long long)gridDim.x * kBlock) {
        const double d = (double)(a.x[i] - a.mean);
        acc = fma(d, d, acc);
    }
    acc = wave_sum(acc);
    if (lane == 0) s_r[wid] = acc;
    __syncthreads();
    if (tid == 0) {
        double r = s_r[0];
#pragma unroll
        for (int w = 1; w < NW; ++w) r += s_r[w];
        a.part[blockIdx.x] = r;
    }
}

__global__ __launch_bounds__(kBlock) void k_jitter(JitterArgs a) {
    const long long i = (long long)blockIdx.x * kBlock + threadIdx.x;
    if (i >= a.n) return;
    const uint64_t st = a.pcg_a + (uint64_t)(2 * i + 1) * a.pcg_g;
    const float z = icdf(pcg_uniform(pcg_output(st)));
    const float h = __builtin_fmaf(a.bw, z, a.x[i]);
    a.x[i] = fminf(fmaxf(h, a.lo), a.hi);
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

hipError_t launch_jitter_var(const JitterArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_jitter_var, dim3(kJitterGrid), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

int jitter_var_blocks() { return kJitterGrid; }

hipError_t launch_jitter(const JitterArgs& a, hipStream_t st) {
    const int grid = (int)(((long long)a.n + kBlock - 1) / kBlock);
    hipLaunchKernelGGL(k_jitter, dim3(grid), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}


hipError_t launch_fisher_score(const FisherArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_fisher_score, dim3(1), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

// Element-wise evaluation of the per-particle arithmetic on given inputs:
// fn 0 = inverse normal CDF of a uniform (central rational or tail branch, as
// the propagate pass selects them), 1 = expf_cw, 2 = expf_cw_nonpos,
// 3 = logf_cw.
__global__ __launch_bounds__(kBlock) void k_device_math(int fn, const float* in, float* out, int n) {
    const int i = blockIdx.x * kBlock + threadIdx.x;
    if (i >= n) return;
    const float v = in[i];
    float r;
    switch (fn) {
        case 0: r = icdf_is_central(v) ? icdf_central(v - 0.5f) : icdf_tail(v); break;
        case 1: r = expf_cw(v); break;
        case 2: r = expf_cw_nonpos(v); break;
        default: r = logf_cw(v); break;
    }
    out[i] = r;
}

hipError_t launch_device_math(int fn, const float* in, float* out, int n, hipStream_t st) {
    hipLaunchKernelGGL(k_device_math, dim3((n + kBlock - 1) / kBlock), dim3(kBlock), 0, st, fn, in,
                       out, n);
    return hipGetLastError();
}

// Points the control block at an observation series whose element 0 is the
// observation of the next tick (one thread; a kernel, so that it may sit
// in front of captured ticks).
__global__ void k_set_observations(Ctl* ctl, const float* y_base) {
    ctl->y_base = y_base;
    ctl->y_tick0 = ctl->tick;
}

hipError_t launch_set_observations(Ctl* ctl, const float* y_base, hipStream_t st) {
    hipLaunchKernelGGL(k_set_observations, dim3(1), dim3(1), 0, st, ctl, y_base);
    return hipGetLastError();
}

hipError_t launch_init(const InitArgs& a, hipStream_t st) {
    int grid = (a.n_pad + kBlock - 1) / kBlock;
    hipLaunchKernelGGL(k_init, dim3(grid), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

// Launch on st; with events, the start and stop events are tied to the
// kernel's own dispatch (no marker packets between the kernels).
template <typename Kern, typename Arg>
static void launch_timed(Kern k, dim3 grid, dim3 block, hipStream_t st, const hipEvent_t* ev,
                         const Arg& a) {
    if (ev) hipExtLaunchKernelGGL(k, grid, block, 0, st, ev[0], ev[1], 0, a);
    else hipLaunchKernelGGL(k, grid, block, 0, st, a);
}

// Threads of a resample (and adjoint) block for a tile, measured on MI355X
// (always-resampling ticks): 512 lanes for 4096-particle tiles (half the
// per-lane chain of 256 lanes), 256 for 2048, 128 for 512-1024 (small blocks,
// more tiles in flight per CU), and one lane per particle for 256-particle
// tiles, which keeps their CDF arithmetic that of the batched filters.
constexpr int resample_block(int tile) {
    return tile >= 4096 ? 512 : (tile == 2048 ? 256 : (tile >= 512 ? 128 : tile));
}

// ev (nullable): start/stop event pairs of propagate, finalise, resample.
template <int K, int MODE>
static hipError_t launch_tick_k(const PropArgs& p, const FinArgs& f, const ResArgs& r,
                                hipStream_t st, hipEvent_t* ev) {
    constexpr bool FISH = MODE & kFish;
    launch_timed(k_propagate<K, MODE>, dim3(p.part.n_tiles), dim3(kBlock), st, ev, p);
    launch_timed(k_finalize, dim3(f.part.n_groups), dim3(kFinBlock), st, ev ? ev + 2 : nullptr, f);
    constexpr int TILE = kBlock * K;
    constexpr int RB = resample_block(TILE);
    static int resident = 0;     // resample blocks the device holds at once
    if (resident == 0) {
        int per_cu = 0, cus = 0, dev = 0;
        hipError_t e = hipGetDevice(&dev);
        if (e == hipSuccess) e = hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        if (e == hipSuccess)
            e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_resample<RB, TILE / RB, FISH>, RB, 0);
        if (e != hipSuccess) return e;
        resident = per_cu > 0 && cus > 0 ? per_cu * cus : 1024;
    }
    // Up to four blocks per resident slot: one block per tile, no work
    // counters.  Beyond: a persistent grid of two blocks per resident slot
    // (one to three measure within 2 % of each other at 51.2M particles; a
    // tick without resampling costs the same).
    ResArgs ra = r;
    ra.one_tile = r.part.n_tiles <= 4 * resident;
    const int grid = ra.one_tile ? r.part.n_tiles : 2 * resident;
    launch_timed(k_resample<RB, TILE / RB, FISH>, dim3(grid), dim3(RB), st, ev ? ev + 4 : nullptr, ra);
    return hipGetLastError();
}

template <int K>
static hipError_t launch_tick_v(int var, const PropArgs& p, const FinArgs& f, const ResArgs& r,
                                hipStream_t st, hipEvent_t* ev) {
    switch (var) {
        case 0: return launch_tick_k<K, 0>(p, f, r, st, ev);
        case kFish: return launch_tick_k<K, kFish>(p, f, r, st, ev);
        case kLearn: return launch_tick_k<K, kLearn>(p, f, r, st, ev);
        case kTState: return launch_tick_k<K, kTState>(p, f, r, st, ev);
        case kLearn | kTState: return launch_tick_k<K, kLearn | kTState>(p, f, r, st, ev);
        case kApf1: return launch_tick_k<K, kApf1>(p, f, r, st, ev);
        case kApf2: return launch_tick_k<K, kApf2>(p, f, r, st, ev);
        case kApf2 | kTState: return launch_tick_k<K, kApf2 | kTState>(p, f, r, st, ev);
        default: return hipErrorInvalidValue;
    }
}

// The Fisher-identity variants are launched when PropArgs.alpha is set (the
// resample pass must then carry the same buffers), the learning variants
// when PropArgs.lsum is set, the Student-t state-noise variants when
// PropArgs.nu_int > 0; Fisher with neither of the others.
hipError_t launch_tick(int K, const PropArgs& p, const FinArgs& f, const ResArgs& r,
                       hipStream_t st, hipEvent_t* ev) {
    const bool fish = p.alpha[0][0] != nullptr;
    const bool learn = p.lsum != nullptr;
    const bool tstate = p.nu_int > 0;
    if (fish != (r.alpha[0][0] != nullptr)) return hipErrorInvalidValue;
    const int var = (fish ? kFish : 0) | (learn ? kLearn : 0) | (tstate ? kTState : 0) |
                    (p.apf_stage == 1 ? kApf1 : 0) | (p.apf_stage == 2 ? kApf2 : 0);
    switch (K) {
        case 1: return launch_tick_v<1>(var, p, f, r, st, ev);
        case 2: return launch_tick_v<2>(var, p, f, r, st, ev);
        case 4: return launch_tick_v<4>(var, p, f, r, st, ev);
        case 8: return launch_tick_v<8>(var, p, f, r, st, ev);
        case 16: return launch_tick_v<16>(var, p, f, r, st, ev);
        default: return hipErrorInvalidValue;
    }
}

// The adjoint runs each tile with the resample pass's lane layout, so that
// its in-tile FP64 CDF (tile_cdf) has the same association, bit for bit.
template <int K>
static hipError_t launch_adj_k(const AdjArgs& a, hipStream_t st) {
    constexpr int TILE = kBlock * K;
    constexpr int RB = resample_block(TILE);
    hipLaunchKernelGGL((k_adjoint<RB, TILE / RB>), dim3(a.n_tiles), dim3(RB), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_adjoint(int K, const AdjArgs& a, hipStream_t st) {
    switch (K) {
        case 1: return launch_adj_k<1>(a, st);
        case 2: return launch_adj_k<2>(a, st);
        case 4: return launch_adj_k<4>(a, st);
        case 8: return launch_adj_k<8>(a, st);
        case 16: return launch_adj_k<16>(a, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace bpf
