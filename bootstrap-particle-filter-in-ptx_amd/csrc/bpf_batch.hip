// bpf_batch.hip -- batched inner bootstrap filters (the SMC^2 inner layer)
// for gfx950.
//
// One launch per tick advances every inner filter: workgroup g runs the whole
// tick of slot order[g] (propagate, weights, FP64 statistics and
// log-likelihood, the resampling decision, FP64 CDF and systematic
// resampling), so filters of uneven size need no grid-wide step.  Each wave
// takes whole tiles of 256 KB particles (KB = 1, 2 or 4), 4 KB per lane.  The
// arithmetic restates the single filter's kernels at KB particles per lane --
// lane group q (64 lanes x KB particles) of a tile plays the single filter's
// wave q -- so an inner filter with the same parameters, particle count, seed
// and tile produces the single filter's particles, ancestors and statistics.
//
// Noise: the PCG32 streams of the single filter (per slot seed), optionally
// stored for later replays, or read back from the store, or proposed by a
// Crank-Nicolson step around the store (correlated pseudo-marginal moves).
#include <hip/hip_runtime.h>

#include <type_traits>

#include "bpf_common.hpp"

namespace bpf {

// Item p = q KB + c of a lane is tile particle (64 q + lane) KB + c: the
// single filter's tile of 256 KB particles at KB particles per lane, with
// 64-particle lane group q playing that filter's wave q.
template <int KB>
__device__ inline int tile_item(int q, int c, int lane) {
    return (64 * q + lane) * KB + c;
}

// Standard normals of the lane's 4 KB particles (tile at `base`) for the
// draw with jump constants (pa, pg).  Every item is evaluated as central;
// tail uniforms (4.85%) are compacted through the wave's LDS queue s_q
// (256 KB floats) by ballot rank, evaluated densely and read back over the
// central value.
template <int KB>
__device__ inline void wave_normals(uint64_t pa, uint64_t pg, int base, int lane, float* s_q,
                                    float (&z)[4 * KB]) {
    bool tail[4 * KB];
    uint64_t bal[4 * KB];
    int tot = 0;
    const uint64_t g2 = pg * 2u;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        uint64_t st = pa + (uint64_t)(2 * (base + tile_item<KB>(q, 0, lane)) + 1) * pg;
#pragma unroll
        for (int c = 0; c < KB; ++c) {
            const int p = q * KB + c;
            const float h = pcg_half(pcg_output(st));
            st += g2;
            tail[p] = !icdf_is_central_half(h);
            z[p] = icdf_central(h);
            bal[p] = __ballot(tail[p]);
            if (tail[p]) s_q[tot + lanes_below(bal[p], 0)] = h;
            tot += __popcll(bal[p]);
        }
    }
    if (tot == 0) return;
    wave_lds_sync();
    for (int e = lane; e < tot; e += kWave) s_q[e] = icdf_tail(uniform_from_half(s_q[e]));
    wave_lds_sync();
    int pos = 0;
#pragma unroll
    for (int p = 0; p < 4 * KB; ++p) {
        if (tail[p]) z[p] = s_q[pos + lanes_below(bal[p], 0)];
        pos += __popcll(bal[p]);
    }
    wave_lds_sync();
}

// The normals of one tick of slot m under the batch's noise source.
template <int KB>
__device__ inline void batch_normals(const BatchArgs& a, int m, uint64_t pa, uint64_t jm,
                                     uint64_t jg, int base, int lane, int zc, size_t zoff,
                                     float* s_q, float (&z)[4 * KB]) {
    if (a.noise == kNoiseStored || a.noise == kNoisePropose) {
        const float* zs = a.zpool[zc] + zoff + base;
#pragma unroll
        for (int q = 0; q < 4; ++q) load_vec<KB>(zs + tile_item<KB>(q, 0, lane), &z[q * KB]);
        if (a.noise == kNoisePropose) {
            float xi[4 * KB];
            const uint64_t pp = (uint64_t)(int64_t)a.prop_seed[m] * jm;
            wave_normals<KB>(pp, jg, base, lane, s_q, xi);
            float* zd = a.zpool[1 - zc] + zoff + base;
#pragma unroll
            for (int p = 0; p < 4 * KB; ++p) z[p] = __builtin_fmaf(a.cn_s, xi[p], a.cn_rho * z[p]);
#pragma unroll
            for (int q = 0; q < 4; ++q) store_vec<KB>(zd + tile_item<KB>(q, 0, lane), &z[q * KB]);
        }
        return;
    }
    wave_normals<KB>(pa, jg, base, lane, s_q, z);
    if (a.noise == kNoisePcgStore) {
        float* zd = a.zpool[zc] + zoff + base;
#pragma unroll
        for (int q = 0; q < 4; ++q) store_vec<KB>(zd + tile_item<KB>(q, 0, lane), &z[q * KB]);
    }
}

// ---------------------------------------------------------------------------
// Initialisation: x_0 = mu + sd z with z from draw 0 (as k_init), padding
// (i >= n) set to mu.
// ---------------------------------------------------------------------------

__global__ __launch_bounds__(kBlock) void k_batch_init(BatchInitArgs a) {
    const int m = blockIdx.x;
    const int n = a.n[m];
    const float mu = a.mu[m], phi = a.phi[m], sig = a.sig[m];
    const float sd = __builtin_sqrtf((sig * sig) / fmaxf(1.0f - phi * phi, 1e-6f));
    // derivatives of sd = sigma / sqrt(max(1 - phi^2, 1e-6)) (FP64)
    const double om = 1.0 - (double)phi * (double)phi;
    const float dsd_dsig = (float)(om > 1e-6 ? 1.0 / sqrt(om) : 1.0 / sqrt(1e-6));
    const float dsd_dphi = (float)(om > 1e-6 ? (double)sig * (double)phi / (om * sqrt(om)) : 0.0);
    const uint64_t pa = (uint64_t)(int64_t)a.seed[m] * a.jm;
    const int zc = (a.noise >= kNoisePcgStore) ? a.zcur[m] : 0;
    const size_t off = (size_t)m * a.cap;     // tick 0 of the pools
    float* x = a.x0 + off;
    for (int i = threadIdx.x; i < a.cap; i += kBlock) {
        float v = mu;
        if (i < n) {
            float z;
            if (a.noise == kNoiseStored || a.noise == kNoisePropose) {
                z = a.zpool[zc][off + i];
                if (a.noise == kNoisePropose) {
                    const uint64_t pp = (uint64_t)(int64_t)a.prop_seed[m] * a.jm;
                    const float xi = icdf(pcg_uniform(pcg_output(pp + (uint64_t)(2 * i + 1) * a.jg)));
                    z = __builtin_fmaf(a.cn_s, xi, a.cn_rho * z);
                    a.zpool[1 - zc][off + i] = z;
                }
            } else {
                z = icdf(pcg_uniform(pcg_output(pa + (uint64_t)(2 * i + 1) * a.jg)));
                if (a.noise == kNoisePcgStore) a.zpool[zc][off + i] = z;
            }
            v = __builtin_fmaf(sd, z, mu);
            if (a.dx0[0] && a.fisher) {   // grad log N(x_0; mu, sd^2), sd = sd(phi, sigma)
                const float z2m1 = __builtin_fmaf(z, z, -1.0f);
                a.dx0[0][off + i] = z / sd;
                a.dx0[1][off + i] = z2m1 * (dsd_dphi / sd);
                a.dx0[2][off + i] = z2m1 * (dsd_dsig / sd);
            } else if (a.dx0[0]) {
                a.dx0[0][off + i] = 1.0f;
                a.dx0[1][off + i] = z * dsd_dphi;
                a.dx0[2][off + i] = z * dsd_dsig;
            }
        }
        x[i] = v;
    }
}

// ---------------------------------------------------------------------------
// One tick of every inner filter.
// ---------------------------------------------------------------------------

// d log g(y | x) / dx of the observation density (Gaussian or Student-t).
template <bool STUDENT>
__device__ inline float dlw_dx(float x, const ObsModel& o) {
    const float eta2 = o.ysq * expf_cw(-x);
    if (STUDENT) return o.nph * eta2 / (o.nu + eta2) - 0.5f;   // as the reverse sweep
    return __builtin_fmaf(0.5f, eta2, -0.5f);
}

// TAN: also carry a per-particle 3-vector d through propagation and
// resampling.  FISHER false: d = dx/d(mu, phi, sigma), the tangents, and the
// filter accumulates the pathwise score of its log-likelihood,
// sum_t sum_i w_ti g_ti d_ti with g = d lw / dx (forward mode of the single
// filter's reverse-mode score; ancestors held fixed).  FISHER true: d is the
// path statistic alpha = grad_theta log p(x_0:t) along each particle's
// ancestry (the observation density does not depend on theta), and the
// filter's score is the Fisher-identity estimate sum_i w_ti alpha_ti of the
// latest tick.
template <int BLOCK, int KB, bool STUDENT, bool TAN, bool FISHER>
__global__ __launch_bounds__(BLOCK) void k_batch_tick(BatchArgs a) {
    constexpr int NW = BLOCK / kWave;
    constexpr int TB = kBatchTile * KB;                   // particles per tile
    constexpr int P = 4 * KB;                             // particles per lane per tile
    constexpr int TP = TB + TB / 16;                      // padded tile
    constexpr float NEG_INF = -__builtin_inff();
    __shared__ float s_q[NW][TB];             // tail queue of each wave
    __shared__ double s_c[NW][TP];            // in-tile CDF of each wave
    __shared__ float s_x[NW][TP];             // tile particles of each wave
    __shared__ float s_tmax[kBatchMaxTiles];
    __shared__ double s_ts[TAN ? 7 : 4][kBatchMaxTiles];
    __shared__ float s_dx[TAN ? NW : 1][TAN ? 3 : 1][TAN ? TP : 1];   // tile tangents of each wave
    __shared__ double s_tf[kBatchMaxTiles];
    __shared__ double s_tp[kBatchMaxTiles];
    __shared__ int s_flag, s_W;
    __shared__ float s_u, s_M;
    __shared__ double s_S;
    __shared__ double s_sc[3];
    __shared__ Ctl s_ctl;                 // the fold's inputs, loaded by wave 0 at the start
    __shared__ double s_fin[4];           // outer log-weight, accumulated score
    __shared__ float s_uin;               // stored resampling uniform

    if (a.halt && *a.halt) return;       // the outer layer stopped the run of ticks
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const int m = a.order ? a.order[blockIdx.x] : (int)blockIdx.x;
    const int n = a.n[m];
    const int nt = (n + TB - 1) / TB;
    Ctl* cp = a.ctl + m;
    const int cur = cp->cur, carry = cp->carry;
    const float moff = carry ? cp->m_prev : 0.0f;
    const size_t off = (size_t)m * a.cap;
    float* __restrict__ xb = (cur ? a.x1 : a.x0) + off;
    float* __restrict__ xo = (cur ? a.x0 : a.x1) + off;
    float* __restrict__ lb = a.lw + off;
    float* dxb[3];
    float* dxo[3];
    if (TAN) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            dxb[k] = (cur ? a.dx1[k] : a.dx0[k]) + off;
            dxo[k] = (cur ? a.dx0[k] : a.dx1[k]) + off;
        }
    }
    const float mu = a.mu[m], phi = a.phi[m], sig = a.sig[m];
    const int tick = cp->tick;
    uint64_t jm, jg;
    pcg_draw_constants((uint32_t)tick, jm, jg);
    const uint64_t pa = (uint64_t)(int64_t)a.seed[m] * jm;
    ObsModel obs = a.obs;
    {
        const float yv = a.y[tick];
        obs.ysq = yv * yv;
    }
    const int zc = (a.noise >= kNoisePcgStore) ? a.zcur[m] : 0;
    const size_t zoff = ((size_t)tick * a.n_slots + m) * a.cap;
    // The mean and variance are formed only for the statistics ring, and the
    // sum of squared weights only for it or an ESS threshold; the
    // log-likelihood, the decision and the CDF need the weight sum alone.
    const bool moments = a.ring != nullptr;
    const bool squares = moments || a.threshold > 0.0f;

    // The fold between the two phases reads the slot's control block, outer
    // log-weight, accumulated score and stored resampling uniform.  Wave 0
    // loads them now, one word per lane, and leaves them in LDS at the end of
    // the propagate phase, so that the fold waits on no global load.
    constexpr int kCtlW = (int)(sizeof(Ctl) / 4);
    static_assert(sizeof(Ctl) % 8 == 0 && kCtlW + 9 <= kWave, "fold inputs: one word per lane of a wave");
    const size_t ui = (size_t)tick * a.n_slots + m;
    const bool stored_u = a.noise == kNoiseStored || a.noise == kNoisePropose;
    uint32_t fold_word = 0;
    if (wid == 0) {
        const uint32_t* src = nullptr;
        if (lane < kCtlW) src = reinterpret_cast<const uint32_t*>(cp) + lane;
        else if (lane < kCtlW + 2) src = a.outer_logw ? reinterpret_cast<const uint32_t*>(a.outer_logw + m) + (lane - kCtlW) : nullptr;
        else if (lane < kCtlW + 8) src = (TAN && !FISHER) ? reinterpret_cast<const uint32_t*>(a.score + 3 * m) + (lane - kCtlW - 2) : nullptr;
        else if (lane == kCtlW + 8) src = stored_u ? reinterpret_cast<const uint32_t*>(a.upool[zc] + ui) : nullptr;
        if (src) fold_word = *src;
    }

    // Propagate, log-weights (carry as the single filter), tile maxima and
    // the tile sums of the single filter's propagate pass: the lane's KB
    // particles of group q summed in order, one wave sum per group, the four
    // groups added in order.
    // x and l outlive the loop: without tangents, the wave's last tile is
    // resampled first, from these registers, in the second phase (with them,
    // the held registers would cost more than the reload they save)
    float x[P], l[P];
    for (int b = wid; b < nt; b += NW) {
        const int base = b * TB;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            load_vec<KB>(xb + base + tile_item<KB>(q, 0, lane), &x[q * KB]);
            if (carry) load_vec<KB>(lb + base + tile_item<KB>(q, 0, lane), &l[q * KB]);
            else {
#pragma unroll
                for (int c = 0; c < KB; ++c) l[q * KB + c] = 0.0f;
            }
        }
        float d[TAN ? 3 : 1][TAN ? P : 1];
        if (TAN) {
#pragma unroll
            for (int k = 0; k < 3; ++k)
#pragma unroll
                for (int q = 0; q < 4; ++q) load_vec<KB>(dxb[k] + base + tile_item<KB>(q, 0, lane), &d[k][q * KB]);
        }
        if (tick > 0) {      // tick 0 weights the stationary draw
            float z[P];
            batch_normals<KB>(a, m, pa, jm, jg, base, lane, zc, zoff, s_q[wid], z);
#pragma unroll
            for (int p = 0; p < P; ++p) {
                if (TAN && FISHER) {   // grad log N(x'; mu + phi (x - mu), sigma^2)
                    const float zs = z[p] / sig;
                    d[0][p] = __builtin_fmaf(zs, 1.0f - phi, d[0][p]);
                    d[1][p] = __builtin_fmaf(zs, x[p] - mu, d[1][p]);
                    d[2][p] += __builtin_fmaf(z[p], z[p], -1.0f) / sig;
                } else if (TAN) {    // x' = mu + phi (x - mu) + sigma z
                    d[0][p] = __builtin_fmaf(phi, d[0][p], 1.0f - phi);
                    d[1][p] = __builtin_fmaf(phi, d[1][p], x[p] - mu);
                    d[2][p] = __builtin_fmaf(phi, d[2][p], z[p]);
                }
                x[p] = transition(x[p], mu, phi, sig, z[p]);
            }
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                store_vec<KB>(xb + base + tile_item<KB>(q, 0, lane), &x[q * KB]);
                if (TAN) {
#pragma unroll
                    for (int k = 0; k < 3; ++k) store_vec<KB>(dxb[k] + base + tile_item<KB>(q, 0, lane), &d[k][q * KB]);
                }
            }
        }
        float mloc = NEG_INF;
#pragma unroll
        for (int q = 0; q < 4; ++q) {
#pragma unroll
            for (int c = 0; c < KB; ++c) {
                const int p = q * KB + c;
                const int j = base + tile_item<KB>(q, c, lane);
                const float lw = STUDENT ? log_weight_t(x[p], obs) : log_weight_gauss(x[p], obs.ysq);
                const float lk = lw + (l[p] - moff);
                l[p] = (j < n) ? lk : NEG_INF;
                mloc = fmaxf(mloc, l[p]);
            }
            store_vec<KB>(lb + base + tile_item<KB>(q, 0, lane), &l[q * KB]);
        }
        const float mt = wave_max(mloc);
        constexpr int NS = TAN ? 7 : 4;
        double acc[NS];
#pragma unroll
        for (int k = 0; k < NS; ++k) acc[k] = 0.0;
        double ls[3] = {0.0, 0.0, 0.0};    // lane sums of the score terms
        if (mt != NEG_INF) {
            // MOMENTS: the sums of the mean and variance; SQUARES: the sum of
            // squared weights (ESS).  Sums that are not needed stay 0.
            auto tile_sums = [&](auto moments_c, auto squares_c) {
                constexpr bool MOMENTS = decltype(moments_c)::value;
                constexpr bool SQUARES = decltype(squares_c)::value;
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    double S = 0.0, SX = 0.0, SXX = 0.0, Q = 0.0;
#pragma unroll
                    for (int c = 0; c < KB; ++c) {
                        const int p = q * KB + c;
                        const int j = base + tile_item<KB>(q, c, lane);
                        const double de = (double)expf_cw_nonpos(l[p] - mt);
                        S += de;
                        if (MOMENTS) {
                            const double dx = (j < n) ? (double)x[p] : 0.0;
                            const double t = de * dx;
                            SX += t;
                            SXX = fma(t, dx, SXX);
                        }
                        if (SQUARES) Q = fma(de, de, Q);
                        if (TAN) {
                            const double eg =
                                (j < n) ? (FISHER ? de : de * (double)dlw_dx<STUDENT>(x[p], obs)) : 0.0;
#pragma unroll
                            for (int k = 0; k < 3; ++k) ls[k] = fma(eg, (double)d[k][p], ls[k]);
                        }
                    }
                    const double ws = wave_sum(S);
                    const double wx = MOMENTS ? wave_sum(SX) : 0.0;
                    const double wxx = MOMENTS ? wave_sum(SXX) : 0.0;
                    const double wq = SQUARES ? wave_sum(Q) : 0.0;
                    if (q == 0) {
                        acc[0] = ws; acc[1] = wx; acc[2] = wxx; acc[3] = wq;
                    } else {
                        acc[0] += ws; acc[1] += wx; acc[2] += wxx; acc[3] += wq;
                    }
                }
            };
            if (moments) tile_sums(std::true_type(), std::true_type());
            else if (squares) tile_sums(std::false_type(), std::true_type());
            else tile_sums(std::false_type(), std::false_type());
            if (TAN) {
#pragma unroll
                for (int k = 0; k < 3; ++k) acc[4 + k] = wave_sum(ls[k]);
            }
        } else {
            // the single filter's zero sums, added over its four waves
#pragma unroll
            for (int k = 0; k < 4; ++k) acc[k] = ((0.0 + 0.0) + 0.0) + 0.0;
        }
        if (lane == 0) {
            s_tmax[b] = mt;
#pragma unroll
            for (int k = 0; k < NS; ++k) s_ts[k][b] = acc[k];
        }
    }
    if (wid == 0) {
        uint32_t* dst = nullptr;
        if (lane < kCtlW) dst = reinterpret_cast<uint32_t*>(&s_ctl) + lane;
        else if (lane < kCtlW + 8) dst = reinterpret_cast<uint32_t*>(s_fin) + (lane - kCtlW);
        else if (lane == kCtlW + 8) dst = reinterpret_cast<uint32_t*>(&s_uin);
        if (dst) *dst = fold_word;
    }
    __syncthreads();
    // Fold of the filter's tiles (its one reduction group), statistics,
    // log-likelihood and decision -- the single filter's finalise pass with
    // one group: its empty waves and lanes contribute exact zeros.
    if (wid == 0) {
        const bool in = lane < nt;
        const float tm = in ? s_tmax[lane] : NEG_INF;
        const double t0 = in ? s_ts[0][lane] : 0.0;
        const double t1 = in ? s_ts[1][lane] : 0.0;
        const double t2 = in ? s_ts[2][lane] : 0.0;
        const double t3 = in ? s_ts[3][lane] : 0.0;
        const float gmax = wave_max(tm);
        const double f = (tm == NEG_INF) ? 0.0 : (double)expf_cw_nonpos(tm - gmax);
        double vs = f * t0, vx = f * t1, vxx = f * t2, vq = (f * f) * t3;
        const double pin = wave_incl_max(wave_incl_scan(vs, lane), lane);
        if (in) {
            s_tf[lane] = f;
            s_tp[lane] = 0.0 + pin;
        }
        vs = wave_sum(vs);
        vx = moments ? wave_sum(vx) : 0.0;
        vxx = moments ? wave_sum(vxx) : 0.0;
        vq = squares ? wave_sum(vq) : 0.0;
        if (TAN) {
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const double v = wave_sum(f * (in ? s_ts[4 + k][lane] : 0.0));
                if (lane == 0) s_sc[k] = v;
            }
        }
        if (lane == 0) {
            const double gs = ((vs + 0.0) + 0.0) + 0.0;
            const double gsx = ((vx + 0.0) + 0.0) + 0.0;
            const double gsxx = ((vxx + 0.0) + 0.0) + 0.0;
            const double gsq = ((vq + 0.0) + 0.0) + 0.0;
            const double fg = (gmax == NEG_INF) ? 0.0 : (double)expf_cw_nonpos(gmax - gmax);
            const double v0 = fg * gs;
            const double S = (((((((0.0 + v0) + 0.0) + 0.0) + 0.0) + 0.0) + 0.0) + 0.0) + 0.0;
            const double SX = (((((((0.0 + fg * gsx) + 0.0) + 0.0) + 0.0) + 0.0) + 0.0) + 0.0) + 0.0;
            const double SXX = (((((((0.0 + fg * gsxx) + 0.0) + 0.0) + 0.0) + 0.0) + 0.0) + 0.0) + 0.0;
            const double Q = (((((((0.0 + (fg * fg) * gsq) + 0.0) + 0.0) + 0.0) + 0.0) + 0.0) + 0.0) + 0.0;
            Ctl c = s_ctl;
            float ur_in = -1.0f;
            if (a.noise == kNoiseStored) {
                ur_in = s_uin;
            } else if (a.noise == kNoisePropose) {
                const uint64_t pp = (uint64_t)(int64_t)a.prop_seed[m] * jm;
                const float v = pcg_uniform(pcg_output(pp + (uint64_t)(2 * kBatchMaxN + 1) * jg));
                float w = s_uin + a.u_step * (2.0f * v - 1.0f);
                w = w - floorf(w);
                ur_in = (w < 1.0f) ? w : 0.0f;
            }
            finalize_tick(c, n, a.threshold, gmax, S, SX, SXX, Q, ur_in);
            if (TAN && c.last.mean == c.last.mean) {     // finite statistics
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    if (FISHER) a.score[3 * m + k] = s_sc[k] / S;
                    else a.score[3 * m + k] = s_fin[1 + k] + s_sc[k] / S;
                }
            }
            if (c.resample) {
                if (a.noise == kNoisePcgStore) a.upool[zc][ui] = c.u_raw;
                if (a.noise == kNoisePropose) a.upool[1 - zc][ui] = c.u_raw;
            }
            *cp = c;
            if (a.ring) a.ring[(size_t)(c.last.tick % a.ring_cap) * a.n_slots + m] = c.last;
            if (a.outer_logw) a.outer_logw[m] = s_fin[0] + c.last.ll_incr;
            s_flag = c.resample;
            s_S = c.S;
            s_u = c.u;
            s_W = c.wrap_start;
            s_M = gmax;
        }
    }
    __syncthreads();
    if (!s_flag) return;

    // Systematic resampling into the other buffer.  Tile b owns the CDF
    // interval (b0, b1]; its CDF is the single filter's tile CDF (the lane's
    // KB particles of group q in order, group q as wave q), and every target
    // the tile owns takes the first particle whose CDF value reaches it.
    const double rS = 1.0 / s_S;
    const float u = s_u;
    const int W = s_W;
    const float M = s_M;
    const double fg = (M == NEG_INF) ? 0.0 : (double)expf_cw_nonpos(M - M);
    double* cw = s_c[wid];
    float* xw = s_x[wid];
    const double rn = target_recip(n);
    // the wave's tiles in reverse order, so that the first is the one whose
    // states and log-weights the first phase left in registers; a tile's
    // outputs do not depend on the order
    const int wt = (wid < nt) ? (nt - 1 - wid) / NW + 1 : 0;
    for (int i = wt - 1; i >= 0; --i) {
        const int b = wid + i * NW;
        const bool held = !TAN && i == wt - 1;
        const int base = b * TB;
        const float mb = s_tmax[b];
        const double pex = (b == 0) ? 0.0 : s_tp[b - 1];
        const double pin = s_tp[b];
        const double ft = (s_tf[b] * fg) * rS;
        const double b0 = fmin(fma(fg * pex, rS, 0.0), 1.0);
        const double b1 = (b == kGroup - 1 || b == nt - 1) ? 1.0 : fmin(fma(fg * pin, rS, 0.0), 1.0);
        const int last = (b == nt - 1) ? (n - 1 - base) : (TB - 1);
        float xv[P], e[P];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            float lk[KB];
            if (held) {
#pragma unroll
                for (int c = 0; c < KB; ++c) {
                    xv[q * KB + c] = x[q * KB + c];
                    lk[c] = l[q * KB + c];
                }
            } else {
                load_vec<KB>(xb + base + tile_item<KB>(q, 0, lane), &xv[q * KB]);
                load_vec<KB>(lb + base + tile_item<KB>(q, 0, lane), lk);
            }
#pragma unroll
            for (int c = 0; c < KB; ++c) {
                const int j = base + tile_item<KB>(q, c, lane);
                e[q * KB + c] = (mb == NEG_INF || j >= n) ? 0.0f : expf_cw_nonpos(lk[c] - mb);
            }
        }
        double excl[4], tot4[4], cl[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            double tot = 0.0;
#pragma unroll
            for (int c = 0; c < KB; ++c) tot += (double)e[q * KB + c];
            const double incl = wave_incl_scan(tot, lane);
            tot4[q] = wave_last(incl);
            double ex = incl - tot;
            for (int r = 0; r < q; ++r) ex += tot4[r];
            ex = (q == 0 && lane == 0) ? 0.0 : ex;
            excl[q] = ex;
            double L = ex;                                 // the lane's last CDF value
#pragma unroll
            for (int c = 0; c < KB; ++c) L += (double)e[q * KB + c];
            cl[q] = fmin(fma(ft, L, b0), b1);
        }
        double pm4[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const double pm = wave_incl_max(cl[q], lane);
            double prev = wave_prev(pm, b0);
            pm4[q] = wave_last(pm);
            for (int r = 0; r < q; ++r) prev = fmax(prev, pm4[r]);
            double L = excl[q];
#pragma unroll
            for (int c = 0; c < KB; ++c) {
                const int p = q * KB + c;
                const int jl = tile_item<KB>(q, c, lane);
                L += (double)e[p];
                double cv = fmax(fmin(fma(ft, L, b0), b1), prev);
                cv = (jl == last) ? b1 : cv;
                cv = (jl > last) ? __builtin_inf() : cv;
                cw[pad16(jl)] = cv;
                xw[pad16(jl)] = xv[p];
                if (TAN) {
#pragma unroll
                    for (int k = 0; k < 3; ++k) s_dx[wid][k][pad16(jl)] = dxb[k][base + jl];
                }
            }
        }
        // lanes 0, 1: unwrapped targets [0, W) at or below b0, b1; lanes 2,
        // 3: wrapped targets [W, n) likewise
        int r = 0;
        if (lane < 4) {
            const bool up = lane & 1, wr = lane >= 2;
            const int lo = wr ? W : 0, hi = wr ? n : W;
            const bool end = up ? (b == nt - 1) : (b == 0);
            r = end ? (up ? hi : lo) : count_targets_le(up ? b1 : b0, u, wr ? 1.0f : 0.0f, lo, hi, n, rn);
        }
        const int ilo = lane_i32(r, 0), ihi = lane_i32(r, 1);
        const int jlo = lane_i32(r, 2), jhi = lane_i32(r, 3);
        wave_lds_sync();
        // Unwrapped outputs in chunks of TB: lane `lane` owns a
        // contiguous run of at most 4 KB, finds the first one's parent by a
        // binary search over the tile CDF and the next ones by stepping or
        // galloping forward, and stages the parent indices in the wave's tail
        // queue (idle until the next tick); the wave then gathers the chunk
        // with coalesced stores.
        int* par = reinterpret_cast<int*>(s_q[wid]);
        for (int c0 = 0; c0 < ihi - ilo; c0 += TB) {
            const int i0 = ilo + c0;
            const int cnt = min(TB, ihi - i0);
            const int per = (cnt + kWave - 1) / kWave;
            const int o0 = min(lane * per, cnt), o1 = min(o0 + per, cnt);
            if (o0 < o1) {
                int j = lds_lower_bound(cw, last, (double)target_at(u, i0 + o0, rn));
                double cj = cw[pad16(j)];
                par[o0] = j;
                for (int o = o0 + 1; o < o1; ++o) {
                    const double t = (double)target_at(u, i0 + o, rn);
                    if (cj < t) {
                        // t <= c_last, so j < last here: gallop from j + 1
                        int g0 = j + 1, g1 = g0, step = 1;
                        while (g1 < last && cw[pad16(g1)] < t) {
                            g0 = g1 + 1;
                            step <<= 1;
                            g1 = min(g0 + step - 1, last);
                        }
                        while (g0 < g1) {
                            const int mid = (g0 + g1) >> 1;
                            if (cw[pad16(mid)] < t) g0 = mid + 1; else g1 = mid;
                        }
                        j = g0;
                        cj = cw[pad16(j)];
                    }
                    par[o] = j;
                }
            }
            wave_lds_sync();
            for (int o = lane; o < cnt; o += kWave) {
                const int p = pad16(par[o]);
                xo[i0 + o] = xw[p];
                if (TAN) {
#pragma unroll
                    for (int k = 0; k < 3; ++k) dxo[k][i0 + o] = s_dx[wid][k][p];
                }
            }
            wave_lds_sync();      // par is refilled by the next chunk
        }
        for (int i = jlo + lane; i < jhi; i += kWave) {
            const double t = (double)(target_at(u, i, rn) - 1.0f);
            const int p = pad16(lds_lower_bound(cw, last, t));
            xo[i] = xw[p];
            if (TAN) {
#pragma unroll
                for (int k = 0; k < 3; ++k) dxo[k][i] = s_dx[wid][k][p];
            }
        }
        wave_lds_sync();
    }
}

// ---------------------------------------------------------------------------
// SMC^2 outer layer
// ---------------------------------------------------------------------------

// Outer ESS of the slots' log-weights after a tick (one block of 1024):
// ESS = (sum w)^2 / sum w^2, w = exp(logw - max), in FP64.  Below
// threshold * M (or non-finite) it raises the halt flag and records the tick.
constexpr int kOuterBlock = 1024;

__global__ __launch_bounds__(kOuterBlock) void k_outer_check(const double* logw, const Ctl* ctl, int M,
                                                             float threshold, int* halt, double* ess_out) {
    constexpr int NW = kOuterBlock / kWave;
    __shared__ double s_a[NW], s_b[NW];
    if (*halt) return;
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;
    double mx = -__builtin_inf();
    for (int m = tid; m < M; m += kOuterBlock) mx = fmax(mx, logw[m]);
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, kWave));
    if (lane == 0) s_a[wid] = mx;
    __syncthreads();
    mx = s_a[0];
    for (int w = 1; w < NW; ++w) mx = fmax(mx, s_a[w]);
    __syncthreads();
    double S = 0.0, Q = 0.0;
    for (int m = tid; m < M; m += kOuterBlock) {
        const double w = exp(logw[m] - mx);
        S += w;
        Q += w * w;
    }
    S = wave_sum(S);
    Q = wave_sum(Q);
    if (lane == 0) {
        s_a[wid] = S;
        s_b[wid] = Q;
    }
    __syncthreads();
    if (tid == 0) {
        double St = s_a[0], Qt = s_b[0];
        for (int w = 1; w < NW; ++w) {
            St += s_a[w];
            Qt += s_b[w];
        }
        const double ess = (St * St) / Qt;
        *ess_out = ess;
        if (!(ess >= (double)threshold * (double)M)) {
            halt[0] = 1;
            halt[1] = ctl[0].tick;
        }
    }
}

// Slot m receives slot anc[m]'s whole inner filter (GatherArgs).
__global__ __launch_bounds__(kBlock) void k_smc2_gather(GatherArgs a) {
    const int m = blockIdx.x, src = a.anc[m];
    const int n = a.n_src[src];
    const Ctl c = a.ctl_src[src];
    if (threadIdx.x == 0) {
        a.n_dst[m] = n;
        a.mu_dst[m] = a.mu_src[src];
        a.phi_dst[m] = a.phi_src[src];
        a.sig_dst[m] = a.sig_src[src];
        Ctl d = c;
        d.cur = 0;
        a.ctl_dst[m] = d;
    }
    const float* xs = a.x_src[c.cur] + (size_t)src * a.cap;
    const float* ls = a.l_src + (size_t)src * a.cap;
    float* xd = a.x_dst + (size_t)m * a.cap;
    float* ld = a.l_dst + (size_t)m * a.cap;
    for (int i = threadIdx.x; i < a.cap; i += kBlock) {
        xd[i] = xs[i];
        ld[i] = ls[i];
    }
    if (a.score_src) {
        for (int k = 0; k < 3; ++k) {
            const float* ds = a.dx_src[c.cur][k] + (size_t)src * a.cap;
            float* dd = a.dx_dst[k] + (size_t)m * a.cap;
            for (int i = threadIdx.x; i < a.cap; i += kBlock) dd[i] = ds[i];
        }
        if (threadIdx.x < 3) a.score_dst[3 * m + threadIdx.x] = a.score_src[3 * src + threadIdx.x];
    }
    if (a.rows > 0) {
        const float* zs = a.zpool[a.zcur[src]];
        float* zd = a.zpool[1 - a.zcur[m]];
        for (int t = 0; t < a.rows; ++t) {
            const size_t rs = ((size_t)t * a.n_slots + src) * a.cap, rd = ((size_t)t * a.n_slots + m) * a.cap;
            for (int i = threadIdx.x; i < n; i += kBlock) zd[rd + i] = zs[rs + i];
        }
        const float* us = a.upool[a.zcur[src]];
        float* ud = a.upool[1 - a.zcur[m]];
        for (int t = threadIdx.x; t < a.rows; t += kBlock)
            ud[(size_t)t * a.n_slots + m] = us[(size_t)t * a.n_slots + src];
    }
}

// Accepted moves: slot m takes the replayed filter where acc[m] is set.
__global__ __launch_bounds__(kBlock) void k_smc2_accept(AcceptArgs a) {
    const int m = blockIdx.x;
    if (!a.acc[m]) return;
    const Ctl c = a.ctl_src[m];
    if (threadIdx.x == 0) {
        a.mu_dst[m] = a.mu_src[m];
        a.phi_dst[m] = a.phi_src[m];
        a.sig_dst[m] = a.sig_src[m];
        a.ctl_dst[m] = c;
    }
    const float* xs = a.x_src[c.cur] + (size_t)m * a.cap;
    float* xd = a.x_dst[c.cur] + (size_t)m * a.cap;
    const float* ls = a.l_src + (size_t)m * a.cap;
    float* ld = a.l_dst + (size_t)m * a.cap;
    for (int i = threadIdx.x; i < a.cap; i += kBlock) {
        xd[i] = xs[i];
        ld[i] = ls[i];
    }
    if (a.score_src) {
        for (int k = 0; k < 3; ++k) {
            const float* ds = a.dx_src[c.cur][k] + (size_t)m * a.cap;
            float* dd = a.dx_dst[c.cur][k] + (size_t)m * a.cap;
            for (int i = threadIdx.x; i < a.cap; i += kBlock) dd[i] = ds[i];
        }
        if (threadIdx.x < 3) a.score_dst[3 * m + threadIdx.x] = a.score_src[3 * m + threadIdx.x];
    }
    if (a.seed_src && threadIdx.x == 0) a.seed_dst[m] = a.seed_src[m];
}

// ---------------------------------------------------------------------------
// Launchers
// ---------------------------------------------------------------------------

hipError_t launch_outer_check(const double* logw, const Ctl* ctl, int M, float threshold, int* halt,
                              double* ess_out, hipStream_t st) {
    hipLaunchKernelGGL(k_outer_check, dim3(1), dim3(kOuterBlock), 0, st, logw, ctl, M, threshold, halt,
                       ess_out);
    return hipGetLastError();
}

hipError_t launch_smc2_gather(const GatherArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_smc2_gather, dim3(a.n_slots), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_smc2_accept(const AcceptArgs& a, int n_slots, hipStream_t st) {
    hipLaunchKernelGGL(k_smc2_accept, dim3(n_slots), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

hipError_t launch_batch_init(const BatchInitArgs& a, hipStream_t st) {
    hipLaunchKernelGGL(k_batch_init, dim3(a.n_slots), dim3(kBlock), 0, st, a);
    return hipGetLastError();
}

// Static LDS of k_batch_tick<BLOCK, KB, *, TAN, *> (the arrays above).
constexpr size_t batch_tick_lds(int block, int kb, bool tan) {
    const size_t nw = (size_t)block / kWave, tb = (size_t)kBatchTile * kb, tp = tb + tb / 16;
    return nw * tb * 4 + nw * tp * 8 + nw * tp * 4 + kBatchMaxTiles * 4 + (tan ? 7 : 4) * kBatchMaxTiles * 8 +
           (tan ? nw * 3 * tp * 4 : 4) + 2 * kBatchMaxTiles * 8 + 64 + sizeof(Ctl) + 5 * 8;
}
constexpr size_t kLdsPerWorkgroup = 160 * 1024;

template <int BLOCK, int KB, bool STUDENT, bool TAN, bool FISHER>
static hipError_t launch_batch_tick_v(const BatchArgs& a, int n_blocks, hipStream_t st) {
    if constexpr (batch_tick_lds(BLOCK, KB, TAN) <= kLdsPerWorkgroup) {
        hipLaunchKernelGGL((k_batch_tick<BLOCK, KB, STUDENT, TAN, FISHER>), dim3(n_blocks), dim3(BLOCK), 0, st, a);
        return hipGetLastError();
    } else {
        return hipErrorInvalidConfiguration;    // the tile's LDS does not fit this many waves
    }
}

template <int BLOCK, int KB>
static hipError_t launch_batch_tick_k(const BatchArgs& a, int n_blocks, hipStream_t st) {
    const bool student = a.obs.nu > 0.0f;
    if (!a.score) {
        return student ? launch_batch_tick_v<BLOCK, KB, true, false, false>(a, n_blocks, st)
                       : launch_batch_tick_v<BLOCK, KB, false, false, false>(a, n_blocks, st);
    }
    if (a.fisher) {
        return student ? launch_batch_tick_v<BLOCK, KB, true, true, true>(a, n_blocks, st)
                       : launch_batch_tick_v<BLOCK, KB, false, true, true>(a, n_blocks, st);
    }
    return student ? launch_batch_tick_v<BLOCK, KB, true, true, false>(a, n_blocks, st)
                   : launch_batch_tick_v<BLOCK, KB, false, true, false>(a, n_blocks, st);
}

template <int BLOCK>
static hipError_t launch_batch_tick_b(const BatchArgs& a, int n_blocks, int tile, hipStream_t st) {
    switch (tile) {
        case kBatchTile: return launch_batch_tick_k<BLOCK, 1>(a, n_blocks, st);
        case 2 * kBatchTile: return launch_batch_tick_k<BLOCK, 2>(a, n_blocks, st);
        case 4 * kBatchTile: return launch_batch_tick_k<BLOCK, 4>(a, n_blocks, st);
        default: return hipErrorInvalidValue;
    }
}

bool batch_tick_fits(int block, int tile, bool tan) {
    return tile % kBatchTile == 0 && batch_tick_lds(block, tile / kBatchTile, tan) <= kLdsPerWorkgroup;
}

// block: 64, 128, 256, 512 or 1024 threads per inner filter; tile: 256, 512
// or 1024 particles (each wave takes whole tiles)
hipError_t launch_batch_tick(const BatchArgs& a, int n_blocks, int block, int tile, hipStream_t st) {
    switch (block) {
        case 64: return launch_batch_tick_b<64>(a, n_blocks, tile, st);
        case 128: return launch_batch_tick_b<128>(a, n_blocks, tile, st);
        case 256: return launch_batch_tick_b<256>(a, n_blocks, tile, st);
        case 512: return launch_batch_tick_b<512>(a, n_blocks, tile, st);
        case 1024: return launch_batch_tick_b<1024>(a, n_blocks, tile, st);
        default: return hipErrorInvalidValue;
    }
}

}  // namespace bpf
