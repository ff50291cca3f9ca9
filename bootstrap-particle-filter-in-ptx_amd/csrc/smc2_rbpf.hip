// smc2_rbpf.hip -- SMC^2 over (rho, sigma_z, mu_base, sigma_base) with a
// Rao-Blackwellised inner filter per theta-particle (include/smc2_rbpf.h).
//
// Reference: smc2/smc2_rbpf_batch.cu (kernels :156-1003, host API
// :1028-1557), smc2/smc2_sorting.cuh (sort coupling), smc2/
// smc2_noise_precision.cuh (FP32 noise storage, the reference default).
//
// Layout in HBM (theta-major, one workgroup of N_inner lanes per theta):
//   theta arrays   rho, sigma_z, mu_base, sigma_base  f32[N_theta]   (x2)
//                  log-likelihood f64[N_theta], inner ESS f32[N_theta] (x2)
//   inner state    z~, mu_h, var_h, log_w   f32[N_theta * N_inner]    (x2)
//   noise history  z  f32[N_theta][cap + 1][N_inner],  u0 f32[N_theta][cap + 1]
//                  (x2: current and proposed / gathered)
//   checkpoint     inner state + log-likelihood of tick t_cp          (x2)
// The doubled arrays are the source and destination of an outer resample
// (gathers by ancestor) and are swapped after it.
//
// Kernels (all one workgroup per theta-particle unless noted):
//   k_rb_init<NI>   prior draw (lane 0, rejection) + stationary inner draw
//   k_rb_step<NI>   one RBPF tick: resample (every tick), sort by mu_h every
//                   4th tick, regime transition, Kalman predict, OCSN update,
//                   FP64 normalisation, inner ESS and log-likelihood increment
//   k_rb_move<NI>   one CPMMH move: propose theta*, replay the history (or the
//                   fixed-lag window) with correlated noise, MH accept/reject
//   k_rows_*        gathers / masked copies of rows (outer resample, commit)
// The outer layer (ESS, systematic resample, proposal covariance) runs on
// the host in FP64 over N_theta values per tick.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/smc2_rbpf.h"
#include "api_util.hpp"
#include "bpf_common.hpp"
#include "rbpf_device.hpp"

namespace bpf {
namespace rb {

// ---------------------------------------------------------------------------
// Block helpers: FP32 max, FP64 sums (wave butterfly, then the wave totals in
// wave order) and FP64 inclusive scan (per-wave Hillis-Steele, then the
// running sum of the wave totals in wave order).  The CPU oracle repeats
// these association orders.
// ---------------------------------------------------------------------------
template <int NI>
struct Lds {
    float z[NI], mu[NI], var[NI];
    double c[NI];
    uint64_t k[2][NI];
    double red[2][NI / kWave];
    float fred[NI / kWave];
    float u0;
};

template <int NI>
__device__ inline float blk_max(float v, float* s, int lane, int wid) {
    v = wave_max(v);
    if (lane == 0) s[wid] = v;
    __syncthreads();
    float m = s[0];
#pragma unroll
    for (int w = 1; w < NI / kWave; ++w) m = fmaxf(m, s[w]);
    __syncthreads();
    return m;
}

template <int NI>
__device__ inline void blk_sum2(double& a, double& b, double (*s)[NI / kWave], int lane, int wid) {
    a = wave_sum_bfly(a);
    b = wave_sum_bfly(b);
    if (lane == 0) {
        s[0][wid] = a;
        s[1][wid] = b;
    }
    __syncthreads();
    double x = s[0][0], y = s[1][0];
#pragma unroll
    for (int w = 1; w < NI / kWave; ++w) {
        x += s[0][w];
        y += s[1][w];
    }
    __syncthreads();
    a = x;
    b = y;
}

template <int NI>
__device__ inline double blk_scan(double v, double* s, int lane, int wid, double& total) {
    const double incl = wave_incl_scan_hs(v, lane);
    if (lane == kWave - 1) s[wid] = incl;
    __syncthreads();
    double P = 0.0, T = 0.0;
#pragma unroll
    for (int w = 0; w < NI / kWave; ++w) {
        if (w == wid) P = T;
        T += s[w];
    }
    __syncthreads();
    total = T;
    return incl + P;
}

// Ascending bitonic sort of NI distinct 64-bit keys, one per lane, held in
// a register: stages whose partner lane (tid ^ j) lies in the same wave
// exchange keys by lane shuffles, the others through LDS.  Lane tid ends with
// the tid-th smallest key.  The keys are distinct, so the result is the
// sorted order whatever the network.
__device__ inline uint64_t shfl_xor_u64(uint64_t v, int m) {
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, m, kWave);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), m, kWave);
    return ((uint64_t)hi << 32) | lo;
}

template <int NI>
__device__ inline uint64_t bitonic(uint64_t key, uint64_t (*k)[NI], int tid) {
    int buf = 0;     // LDS stages alternate buffers: one barrier per stage
#pragma unroll
    for (int s = 2; s <= NI; s <<= 1) {
#pragma unroll
        for (int j = s >> 1; j > 0; j >>= 1) {
            uint64_t p;
            if (j < kWave) {
                p = shfl_xor_u64(key, j);
            } else {
                k[buf][tid] = key;
                __syncthreads();
                p = k[buf][tid ^ j];
                buf ^= 1;
            }
            const bool asc = (tid & s) == 0, lower = (tid & j) == 0;
            const uint64_t mn = key < p ? key : p, mx = key < p ? p : key;
            key = (asc == lower) ? mn : mx;
        }
    }
    return key;
}

// Noise history elements: FP32 (default) or IEEE half (the reference's
// SMC2_NOISE_FP16, smc2/smc2_noise_precision.cuh:23-58).  A stored noise is
// always used as stored, so a replay repeats the forward tick.
template <bool H>
struct NoiseT {
    using type = float;
};
template <>
struct NoiseT<true> {
    using type = _Float16;
};
template <bool H>
__device__ inline float nstore(void* p, long long i, float v) {
    using T = typename NoiseT<H>::type;
    const T h = (T)v;
    static_cast<T*>(p)[i] = h;
    return (float)h;
}
template <bool H>
__device__ inline float nload(const void* p, long long i) {
    using T = typename NoiseT<H>::type;
    return (float)static_cast<const T*>(p)[i];
}

// One RBPF tick of one theta-particle (smc2/smc2_rbpf_batch.cu:305-393; the
// same tick is replayed by CPMMH, :786-875).  Lane tid holds inner particle
// tid; zn is its regime noise for this tick and u0 the block's resampling
// offset.  Returns the tick's log-likelihood increment and inner ESS (every
// lane gets the same values).
template <int NI>
__device__ inline void tick(Lds<NI>& L, float& zt, float& mu, float& var, float& lw, float zn,
                            float u0, float y, int t, float rho, float sz, float mub, float sgb,
                            const Model& md, double& ll_incr, double& ess) {
    const int tid = threadIdx.x, lane = tid & (kWave - 1), wid = tid / kWave;

    // systematic resample on the incoming weights, every tick (:308-330):
    // FP64 CDF of exp(lw - max), normalised, last entry 1; output i takes
    // the first parent with cdf >= (u0 + i) / NI
    const float m0 = blk_max<NI>(lw, L.fred, lane, wid);
    const float w0 = expf_cw(lw - m0);
    double S0;
    const double c = blk_scan<NI>((double)w0, L.red[0], lane, wid, S0);
    L.z[tid] = zt;
    L.mu[tid] = mu;
    L.var[tid] = var;
    L.c[tid] = (tid == NI - 1) ? 1.0 : c / S0;
    __syncthreads();
    const float u = (u0 + (float)tid) / (float)NI;
    const double uu = (double)u;
    int lo = 0, hi = NI - 1;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (L.c[mid] < uu) lo = mid + 1; else hi = mid;
    }
    zt = L.z[lo];
    mu = L.mu[lo];
    var = L.var[lo];
    lw = md.lw0;
    __syncthreads();

    // CPMMH coupling: order the particles by mu_h every 4th tick
    // (smc2/smc2_sorting.cuh:47-106); ties keep their resampled order
    if (t % kSortEvery == 0) {
        L.z[tid] = zt;
        L.mu[tid] = mu;
        L.var[tid] = var;
        __syncthreads();
        const uint64_t kk = bitonic<NI>(((uint64_t)sort_key(mu) << 32) | (uint32_t)tid, L.k, tid);
        const int src = (int)(uint32_t)kk;
        zt = L.z[src];
        mu = L.mu[src];
        var = L.var[src];
        __syncthreads();
    }

    // regime transition, Kalman prediction, OCSN update (:358-372)
    float mp, vp, lnorm;
    predict(rho, sz, mub, sgb, zn, md, zt, mu, var, mp, vp);
    ocsn_update(y, mp, vp, md, mu, var, lnorm);
    lw = lw + lnorm;

    // normalisation, inner ESS, log-likelihood increment (:374-393), FP64.
    // The log-weights carry the 1/NI of the resample, so the increment is
    // log sum exp(lw); the reference subtracts log NI once more (a constant
    // offset per tick, the same for every theta-particle)
    const float m1 = blk_max<NI>(lw, L.fred, lane, wid);
    const float w1 = expf_cw(lw - m1);
    double S = (double)w1, Q = (double)w1 * (double)w1;
    blk_sum2<NI>(S, Q, L.red, lane, wid);
    ess = (S * S) / Q;
    ll_incr = (double)m1 + log(S);
}

// log prior of smc2_rbpf_batch.cu:139-151 in FP64: -inf outside the bounds.
__device__ __host__ inline double log_prior(const float* th, const SVPrior& p, const SVBounds& b) {
    if (th[0] < b.rho_min || th[0] > b.rho_max) return -HUGE_VAL;
    if (th[1] < b.sigma_z_min || th[1] > b.sigma_z_max) return -HUGE_VAL;
    if (th[2] < b.mu_base_min || th[2] > b.mu_base_max) return -HUGE_VAL;
    if (th[3] < b.sigma_base_min || th[3] > b.sigma_base_max) return -HUGE_VAL;
    const double d0 = ((double)th[0] - (double)p.rho_mean) / (double)p.rho_std;
    const double d1 = ((double)th[1] - (double)p.sigma_z_mean) / (double)p.sigma_z_std;
    const double d2 = ((double)th[2] - (double)p.mu_base_mean) / (double)p.mu_base_std;
    const double d3 = ((double)th[3] - (double)p.sigma_base_mean) / (double)p.sigma_base_std;
    return -0.5 * (d0 * d0 + d1 * d1 + d2 * d2 + d3 * d3);
}

struct Theta {            // one buffer of theta-particle arrays
    float* p[kParams];
    double* ll;
    float* essi;
};
struct Inner {
    float* a[4];          // z~, mu_h, var_h, log_w
};

struct RbInitArgs {
    Theta th;
    Inner in;
    void* zn;
    void* un;
    uint64_t kprior, kz, ku;
    long long zrow;
    int ucap1;
    SVPrior prior;
    SVBounds bounds;
    Model md;
};

// smc2_rbpf_batch.cu:166-230.  Lane 0 draws theta from the prior, up to 1000
// attempts until all four lie within the bounds (the last attempt is kept
// otherwise); every lane draws its regime from the stationary law.
template <int NI, bool H>
__global__ __launch_bounds__(NI) void k_rb_init(RbInitArgs a) {
    __shared__ float s_th[kParams];
    const int m = blockIdx.x, tid = threadIdx.x;
    const long long g = (long long)m * NI + tid;
    if (tid == 0) {
        const uint64_t kp = key(a.kprior, (uint64_t)m);
        const float mean[4] = {a.prior.rho_mean, a.prior.sigma_z_mean, a.prior.mu_base_mean,
                               a.prior.sigma_base_mean};
        const float sd[4] = {a.prior.rho_std, a.prior.sigma_z_std, a.prior.mu_base_std,
                             a.prior.sigma_base_std};
        const float lo[4] = {a.bounds.rho_min, a.bounds.sigma_z_min, a.bounds.mu_base_min,
                             a.bounds.sigma_base_min};
        const float hi[4] = {a.bounds.rho_max, a.bounds.sigma_z_max, a.bounds.mu_base_max,
                             a.bounds.sigma_base_max};
        float th[4];
        for (int at = 0; at < 1000; ++at) {
            bool ok = true;
            for (int q = 0; q < kParams; ++q) {
                th[q] = mean[q] + sd[q] * normal(kp, (uint32_t)(4 * at + q));
                ok = ok && th[q] >= lo[q] && th[q] <= hi[q];
            }
            if (ok) break;
        }
        for (int q = 0; q < kParams; ++q) {
            s_th[q] = th[q];
            a.th.p[q][m] = th[q];
        }
        a.th.ll[m] = 0.0;
        a.th.essi[m] = (float)NI;
        nstore<H>(a.un, (long long)m * a.ucap1, normal(key(a.ku, (uint64_t)m), 0u));
    }
    __syncthreads();
    const float n0 = nstore<H>(a.zn, (long long)m * a.zrow + tid, normal(key(a.kz, (uint64_t)g), 0u));
    float zt, mu, var;
    init_particle(s_th[0], s_th[1], s_th[2], s_th[3], n0, a.md, zt, mu, var);
    a.in.a[0][g] = zt;
    a.in.a[1][g] = mu;
    a.in.a[2][g] = var;
    a.in.a[3][g] = a.md.lw0;
}

struct RbStepArgs {
    Theta th;
    Inner in;
    void* zn;
    void* un;
    float* y_hist;            // the observation is stored at index t
    double* ll_incr;          // mapped pinned host memory
    uint64_t kz, ku;
    long long zrow;
    int ucap1;
    int t;
    float y;
    Model md;
};

// One forward tick at observation t (smc2_rbpf_batch.cu:245-398): draws and
// stores this tick's noise (index t + 1), runs the tick in place, adds the
// increment to the pseudo-marginal log-likelihood.
template <int NI, bool H>
__global__ __launch_bounds__(NI) void k_rb_step(RbStepArgs a) {
    __shared__ Lds<NI> L;
    const int m = blockIdx.x, tid = threadIdx.x;
    const long long g = (long long)m * NI + tid;
    const float rho = a.th.p[0][m], sz = a.th.p[1][m], mub = a.th.p[2][m], sgb = a.th.p[3][m];
    float zt = a.in.a[0][g], mu = a.in.a[1][g], var = a.in.a[2][g], lw = a.in.a[3][g];
    const float zn = nstore<H>(a.zn, (long long)m * a.zrow + (long long)(a.t + 1) * NI + tid,
                               normal(key(a.kz, (uint64_t)g), (uint32_t)(a.t + 1)));
    if (tid == 0) {
        const float n = nstore<H>(a.un, (long long)m * a.ucap1 + a.t + 1,
                                  normal(key(a.ku, (uint64_t)m), (uint32_t)(a.t + 1)));
        L.u0 = u0_of(n);
        if (m == 0) a.y_hist[a.t] = a.y;
    }
    __syncthreads();
    double lli, ess;
    tick<NI>(L, zt, mu, var, lw, zn, L.u0, a.y, a.t, rho, sz, mub, sgb, a.md, lli, ess);
    a.in.a[0][g] = zt;
    a.in.a[1][g] = mu;
    a.in.a[2][g] = var;
    a.in.a[3][g] = lw;
    if (tid == 0) {
        a.ll_incr[m] = lli;
        a.th.ll[m] += lli;
        a.th.essi[m] = (float)ess;
    }
}

struct RbMoveArgs {
    Theta th;
    Inner in;
    const void* zc;
    void* zo;
    const void* uc;
    void* uo;
    const float* y;
    int* acc;
    Inner cp;                 // checkpoint state (used when t_cp >= 0)
    const double* cp_ll;
    int t_cp;
    uint64_t kmove, kfz, kfu;
    long long zrow;
    int ucap1;
    int t_cur;
    float crho, cscale;
    float chol[kParams * kParams];
    float pstd[kParams];
    SVPrior prior;
    SVBounds bounds;
    Model md;
};

// One CPMMH move (smc2_rbpf_batch.cu:651-964): lane 0 proposes theta* (95 %
// adaptive Cholesky random walk, 5 % fixed diagonal); the block replays
// ticks t0..t_cur under theta* with noise rho_c * current + sqrt(1-rho_c^2)
// * fresh (stored into the proposal buffers) from the stationary start or
// the fixed-lag checkpoint; lane 0 accepts with probability
// min(1, exp(log-likelihood and prior difference)).  An accepted move
// overwrites theta, its log-likelihood, inner ESS and inner state.
template <int NI, bool H>
__global__ __launch_bounds__(NI) void k_rb_move(RbMoveArgs a) {
    __shared__ Lds<NI> L;
    __shared__ float s_th[kParams];
    __shared__ double s_lp[2];
    __shared__ int s_flag;
    const int m = blockIdx.x, tid = threadIdx.x;
    const long long g = (long long)m * NI + tid;
    const uint64_t km = key(a.kmove, (uint64_t)m);
    if (tid == 0) {
        float th[kParams], z[kParams], pert[kParams], tp[kParams];
        for (int q = 0; q < kParams; ++q) th[q] = a.th.p[q][m];
        const double lpc = log_prior(th, a.prior, a.bounds);
        for (int q = 0; q < kParams; ++q) z[q] = normal(km, (uint32_t)q);
        const float umix = uniform(km, 4u);
        if (umix > 0.05f) {
            for (int i = 0; i < kParams; ++i) {
                float s = 0.0f;
                for (int j = 0; j <= i; ++j) s += a.chol[i * kParams + j] * z[j];
                pert[i] = s;
            }
        } else {
            for (int i = 0; i < kParams; ++i) pert[i] = a.pstd[i] * z[i];
        }
        for (int q = 0; q < kParams; ++q) {
            tp[q] = th[q] + pert[q];
            s_th[q] = tp[q];
        }
        const double lpp = log_prior(tp, a.prior, a.bounds);
        s_lp[0] = lpc;
        s_lp[1] = lpp;
        s_flag = __builtin_isfinite(lpp) ? 1 : 0;
    }
    __syncthreads();
    if (!s_flag) {
        if (tid == 0) a.acc[m] = 0;
        return;
    }
    const float rho = s_th[0], sz = s_th[1], mub = s_th[2], sgb = s_th[3];
    const uint64_t kz = key(a.kfz, (uint64_t)g);
    const uint64_t ku = key(a.kfu, (uint64_t)m);
    const long long zb = (long long)m * a.zrow;
    const long long ub = (long long)m * a.ucap1;
    float zt, mu, var, lw;
    int t0;
    if (a.t_cp >= 0) {
        zt = a.cp.a[0][g];
        mu = a.cp.a[1][g];
        var = a.cp.a[2][g];
        lw = a.cp.a[3][g];
        t0 = a.t_cp + 1;
    } else {
        const float zp = nstore<H>(a.zo, zb + tid, a.crho * nload<H>(a.zc, zb + tid) + a.cscale * normal(kz, 0u));
        init_particle(rho, sz, mub, sgb, zp, a.md, zt, mu, var);
        lw = a.md.lw0;
        t0 = 0;
    }
    double llacc = 0.0, ess = (double)NI;
    for (int t = t0; t <= a.t_cur; ++t) {
        const long long zi = zb + (long long)(t + 1) * NI + tid;
        const float zp = nstore<H>(a.zo, zi, a.crho * nload<H>(a.zc, zi) + a.cscale * normal(kz, (uint32_t)(t + 1)));
        if (tid == 0) {
            const float up = nstore<H>(a.uo, ub + t + 1,
                                       a.crho * nload<H>(a.uc, ub + t + 1) + a.cscale * normal(ku, (uint32_t)(t + 1)));
            L.u0 = u0_of(up);
        }
        __syncthreads();
        double lli;
        tick<NI>(L, zt, mu, var, lw, zp, L.u0, a.y[t], t, rho, sz, mub, sgb, a.md, lli, ess);
        llacc += lli;
    }
    if (tid == 0) {
        const double llb = (a.t_cp >= 0) ? a.cp_ll[m] : 0.0;
        const double llc = a.th.ll[m] - llb;
        const double la = (llacc + s_lp[1]) - (llc + s_lp[0]);
        const float u = uniform(km, 5u);
        const int ok = log((double)u) < la;
        s_flag = ok;
        a.acc[m] = ok;
        if (ok) {
            for (int q = 0; q < kParams; ++q) a.th.p[q][m] = s_th[q];
            a.th.ll[m] = llb + llacc;
            a.th.essi[m] = (float)ess;
        }
    }
    __syncthreads();
    if (s_flag) {
        a.in.a[0][g] = zt;
        a.in.a[1][g] = mu;
        a.in.a[2][g] = var;
        a.in.a[3][g] = lw;
    }
}

// dst row m <- src row anc[m] (row length len elements; 16-byte vectors
// when VEC).
template <class T, bool VEC>
__global__ void k_rows_gather(const T* __restrict__ src, T* __restrict__ dst,
                              const int* __restrict__ anc, long long stride, long long len) {
    constexpr int W = 16 / (int)sizeof(T);
    const int m = blockIdx.y;
    const long long s = (long long)anc[m] * stride, d = (long long)m * stride;
    const long long step = (long long)gridDim.x * blockDim.x;
    if constexpr (VEC) {
        const long long nv = len / W;
        for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < nv; j += step)
            reinterpret_cast<float4*>(dst + d)[j] = reinterpret_cast<const float4*>(src + s)[j];
    } else {
        for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < len; j += step)
            dst[d + j] = src[s + j];
    }
}

// dst row m <- src row m over [off, off + len) where acc[m] (commit of the
// accepted proposals' noise).
template <class T, bool VEC>
__global__ void k_rows_commit(const T* __restrict__ src, T* __restrict__ dst,
                              const int* __restrict__ acc, long long stride, long long off,
                              long long len) {
    constexpr int W = 16 / (int)sizeof(T);
    const int m = blockIdx.y;
    if (!acc[m]) return;
    const long long b = (long long)m * stride + off;
    const long long step = (long long)gridDim.x * blockDim.x;
    if constexpr (VEC) {
        const long long nv = len / W;
        for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < nv; j += step)
            reinterpret_cast<float4*>(dst + b)[j] = reinterpret_cast<const float4*>(src + b)[j];
    } else {
        for (long long j = (long long)blockIdx.x * blockDim.x + threadIdx.x; j < len; j += step)
            dst[b + j] = src[b + j];
    }
}

__global__ void k_theta_gather(Theta src, Theta dst, const int* __restrict__ anc, int n) {
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const int s = anc[m];
    for (int q = 0; q < kParams; ++q) dst.p[q][m] = src.p[q][s];
    dst.ll[m] = src.ll[s];
    dst.essi[m] = src.essi[s];
}

}  // namespace rb
}  // namespace bpf

using namespace bpf;
using namespace bpf::rb;

// ---------------------------------------------------------------------------
// Host runtime
// ---------------------------------------------------------------------------
struct SMC2StateCUDA {
    int nt = 0, ni = 0;
    hipStream_t st = nullptr;
    hipEvent_t ev_tick = nullptr;
    Theta th[2] = {};
    int tb = 0;
    Inner in[2] = {};
    int ib = 0;
    void* zn[2] = {nullptr, nullptr};   // noise histories, elements of es bytes
    void* un[2] = {nullptr, nullptr};
    int es = 4;                          // 4: FP32, 2: IEEE half
    int nb = 0;
    int cap = 0;
    Inner cp[2] = {};
    double* cpll[2] = {nullptr, nullptr};
    int cb = 0;
    float* d_y = nullptr;
    int y_cap = 0;
    int* d_anc = nullptr;
    int* d_acc = nullptr;
    double* d_lli = nullptr;   // device view of h_lli
    // pinned host staging
    double* h_lli = nullptr;
    int* h_anc = nullptr;
    int* h_acc = nullptr;
    float* h_th = nullptr;     // kParams * nt
    // outer state (host, FP64)
    std::vector<double> lw, w;
    int t_cur = -1, t_cp = -1, lag = 0, K = 5;
    float ess_thr = 0.5f, crho = 0.99f, ess = 0.0f;
    float pstd[kParams] = {0.01f, 0.02f, 0.1f, 0.02f};
    float chol[kParams * kParams] = {};
    SVPrior prior{};
    SVBounds bounds{};
    SVCurve tcurve{};
    SVFixedCurves fixed{};
    SVPrior run_prior{};      // the prior and bounds of the current run (copied at init)
    SVBounds run_bounds{};
    bool adaptive = true;
    uint64_t seed = 12345;
    uint64_t host_rng = 0;
    int n_res = 0, n_acc = 0, n_tot = 0;
    bool inited = false;
    Model md{};
};

namespace {

constexpr float kAdaptScale2 = 2.38f * 2.38f / (float)kParams;

template <class T>
int dmalloc(T** p, size_t n) {
    HIP_TRY(hipMalloc((void**)p, (n ? n : 1) * sizeof(T)));
    return 0;
}

// xorshift64* host stream of smc2_rbpf_batch.cu:1009-1018.
uint64_t xs64(uint64_t* s) {
    uint64_t x = *s;
    x ^= x >> 12;
    x ^= x << 25;
    x ^= x >> 27;
    *s = x;
    return x * 0x2545F4914F6CDD1DULL;
}
float xs64_uniform(uint64_t* s) {
    return (float)((xs64(s) >> 11) + 1) * (1.0f / 9007199254740994.0f);
}

void default_model(SMC2StateCUDA* s) {
    // smc2_rbpf_batch.cu:1080-1110
    s->prior = {0.95f, 0.02f, 0.1f, 0.1f, -1.0f, 0.5f, 0.15f, 0.05f};
    s->bounds = {0.8f, 0.999f, 0.01f, 1.0f, -10.0f, 5.0f, 0.01f, 1.0f};
    s->fixed = {0.5f, 1.0f, 0.1f, 1.0f};
    s->tcurve = {0.02f, 0.08f, 1.5f};
}

int build_model(SMC2StateCUDA* s) {
    Model& m = s->md;
    m.tb = s->tcurve.base;
    m.ts = s->tcurve.scale;
    m.tr = s->tcurve.rate;
    m.ms = s->fixed.mu_scale;
    m.mr = s->fixed.mu_rate;
    m.ss = s->fixed.sigma_scale;
    m.sr = s->fixed.sigma_rate;
    const double half_log_2pi = 0.5 * std::log(2.0 * 3.14159265358979323846);
    for (int k = 0; k < kOcsnK; ++k) {
        if ((float)(std::log((double)kOcsnP[k]) - half_log_2pi) != kOcsnLogW[k]) {
            set_err("OCSN log-weight %d differs from this host's libm", k);
            return -1;
        }
    }
    // log-weight of every inner particle after a resample
    m.lw0 = -(float)std::log((double)s->ni);
    return 0;
}

uint64_t base_key(const SMC2StateCUDA* s, Purpose p) { return key(s->seed, (uint64_t)p); }

// Outer weights and ESS from the log-weights (FP64, sequential order).
float outer_ess(SMC2StateCUDA* s) {
    const int n = s->nt;
    double mx = -HUGE_VAL;
    for (int m = 0; m < n; ++m) mx = std::fmax(mx, s->lw[m]);
    double S = 0.0;
    for (int m = 0; m < n; ++m) {
        s->w[m] = std::exp(s->lw[m] - mx);
        S += s->w[m];
    }
    double Q = 0.0;
    for (int m = 0; m < n; ++m) {
        s->w[m] /= S;
        Q += s->w[m] * s->w[m];
    }
    s->ess = (float)(1.0 / std::fmax(Q, 1e-300));
    return s->ess;
}

// Adaptive random-walk Cholesky factor (smc2_rbpf_batch.cu:1248-1294) from
// the unweighted moments of the (just resampled) theta-particles, FP64.
void adapt_chol(SMC2StateCUDA* s) {
    const int n = s->nt;
    double mean[kParams] = {}, cov[kParams][kParams] = {}, L[kParams][kParams] = {};
    for (int q = 0; q < kParams; ++q) {
        double a = 0.0;
        for (int m = 0; m < n; ++m) a += (double)s->h_th[q * n + m];
        mean[q] = a / n;
    }
    for (int i = 0; i < kParams; ++i)
        for (int j = 0; j <= i; ++j) {
            double a = 0.0;
            for (int m = 0; m < n; ++m)
                a += ((double)s->h_th[i * n + m] - mean[i]) * ((double)s->h_th[j * n + m] - mean[j]);
            cov[i][j] = cov[j][i] = a / (n - 1);
        }
    for (int i = 0; i < kParams; ++i) cov[i][i] += 1e-6;
    bool ok = true;
    for (int i = 0; i < kParams; ++i)
        for (int j = 0; j <= i; ++j) {
            double a = 0.0;
            for (int k = 0; k < j; ++k) a += L[i][k] * L[j][k];
            if (i == j) {
                double v = cov[i][i] - a;
                if (v <= 0.0) {
                    ok = false;
                    v = 1e-8;
                }
                L[i][j] = std::sqrt(v);
            } else {
                L[i][j] = (L[j][j] > 1e-10) ? (cov[i][j] - a) / L[j][j] : 0.0;
            }
        }
    if (!ok) {
        for (int i = 0; i < kParams; ++i)
            for (int j = 0; j < kParams; ++j) L[i][j] = (i == j) ? std::sqrt(std::fmax(cov[i][i], 1e-8)) : 0.0;
    }
    const double sc = std::sqrt((double)kAdaptScale2);
    for (int i = 0; i < kParams; ++i)
        for (int j = 0; j < kParams; ++j) s->chol[i * kParams + j] = (float)(L[i][j] * sc);
}

void initial_chol(SMC2StateCUDA* s) {
    const double sc = std::sqrt((double)kAdaptScale2);
    for (int i = 0; i < kParams; ++i)
        for (int j = 0; j < kParams; ++j)
            s->chol[i * kParams + j] = (i == j) ? (float)((double)s->pstd[i] * sc) : 0.0f;
}

int alloc_noise(SMC2StateCUDA* s, int cap, void* z[2], void* u[2]) {
    const size_t nz = (size_t)s->nt * (size_t)(cap + 1) * (size_t)s->ni * (size_t)s->es;
    const size_t nu = (size_t)s->nt * (size_t)(cap + 1) * (size_t)s->es;
    for (int b = 0; b < 2; ++b) {
        HIP_TRY(hipMalloc(&z[b], nz));
        HIP_TRY(hipMalloc(&u[b], nu));
    }
    return 0;
}

void free_noise(void* z[2], void* u[2]) {
    for (int b = 0; b < 2; ++b) {
        if (z[b]) (void)hipFree(z[b]);
        if (u[b]) (void)hipFree(u[b]);
        z[b] = u[b] = nullptr;
    }
}

int grow_noise(SMC2StateCUDA* s, int cap) {
    void* z[2] = {nullptr, nullptr};
    void* u[2] = {nullptr, nullptr};
    if (alloc_noise(s, cap, z, u)) {
        free_noise(z, u);
        return -1;
    }
    // keep every theta's rows of the current buffer (strides change)
    const size_t zr = (size_t)s->ni * (size_t)s->es, ur = (size_t)s->es;
    HIP_TRY(hipMemcpy2DAsync(z[s->nb], (size_t)(cap + 1) * zr, s->zn[s->nb], (size_t)(s->cap + 1) * zr,
                             (size_t)(s->cap + 1) * zr, (size_t)s->nt, hipMemcpyDeviceToDevice, s->st));
    HIP_TRY(hipMemcpy2DAsync(u[s->nb], (size_t)(cap + 1) * ur, s->un[s->nb], (size_t)(s->cap + 1) * ur,
                             (size_t)(s->cap + 1) * ur, (size_t)s->nt, hipMemcpyDeviceToDevice, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    free_noise(s->zn, s->un);
    for (int b = 0; b < 2; ++b) {
        s->zn[b] = z[b];
        s->un[b] = u[b];
    }
    s->cap = cap;
    return 0;
}

int ensure_y(SMC2StateCUDA* s, int need) {
    if (need <= s->y_cap) return 0;
    int cap = s->y_cap ? s->y_cap : 8000;
    while (cap < need) cap *= 2;
    float* d = nullptr;
    if (dmalloc(&d, (size_t)cap)) return -1;
    if (s->t_cur >= 0) {
        hipError_t e = hipMemcpyAsync(d, s->d_y, (size_t)(s->t_cur + 1) * 4, hipMemcpyDeviceToDevice, s->st);
        if (e == hipSuccess) e = hipStreamSynchronize(s->st);
        if (e != hipSuccess) {
            (void)hipFree(d);
            set_err("history copy: %s", hipGetErrorString(e));
            return -1;
        }
    }
    if (s->d_y) (void)hipFree(s->d_y);
    s->d_y = d;
    s->y_cap = cap;
    return 0;
}

template <int NI, bool H>
int launch_init_t(SMC2StateCUDA* s) {
    RbInitArgs a{};
    a.th = s->th[s->tb];
    a.in = s->in[s->ib];
    a.zn = s->zn[s->nb];
    a.un = s->un[s->nb];
    a.kprior = base_key(s, kPrior);
    a.kz = key(base_key(s, kZ), 0);
    a.ku = key(base_key(s, kU0), 0);
    a.zrow = (long long)(s->cap + 1) * NI;
    a.ucap1 = s->cap + 1;
    a.prior = s->prior;
    a.bounds = s->bounds;
    a.md = s->md;
    hipLaunchKernelGGL((k_rb_init<NI, H>), dim3(s->nt), dim3(NI), 0, s->st, a);
    HIP_TRY(hipGetLastError());
    return 0;
}

template <int NI, bool H>
int launch_step_t(SMC2StateCUDA* s, float y) {
    RbStepArgs a{};
    a.th = s->th[s->tb];
    a.in = s->in[s->ib];
    a.zn = s->zn[s->nb];
    a.un = s->un[s->nb];
    a.y_hist = s->d_y;
    a.ll_incr = s->d_lli;
    a.kz = key(base_key(s, kZ), (uint64_t)s->n_res);
    a.ku = key(base_key(s, kU0), (uint64_t)s->n_res);
    a.zrow = (long long)(s->cap + 1) * NI;
    a.ucap1 = s->cap + 1;
    a.t = s->t_cur;
    a.y = y;
    a.md = s->md;
    hipLaunchKernelGGL((k_rb_step<NI, H>), dim3(s->nt), dim3(NI), 0, s->st, a);
    HIP_TRY(hipGetLastError());
    return 0;
}

template <int NI, bool H>
int launch_move_t(SMC2StateCUDA* s, int k, bool use_cp) {
    RbMoveArgs a{};
    a.th = s->th[s->tb];
    a.in = s->in[s->ib];
    a.zc = s->zn[s->nb];
    a.zo = s->zn[1 - s->nb];
    a.uc = s->un[s->nb];
    a.uo = s->un[1 - s->nb];
    a.y = s->d_y;
    a.acc = s->d_acc;
    a.cp = s->cp[s->cb];
    a.cp_ll = s->cpll[s->cb];
    a.t_cp = use_cp ? s->t_cp : -1;
    a.kmove = key(key(base_key(s, kMove), (uint64_t)s->n_res), (uint64_t)k);
    a.kfz = key(key(base_key(s, kFreshZ), (uint64_t)s->n_res), (uint64_t)k);
    a.kfu = key(key(base_key(s, kFreshU), (uint64_t)s->n_res), (uint64_t)k);
    a.zrow = (long long)(s->cap + 1) * NI;
    a.ucap1 = s->cap + 1;
    a.t_cur = s->t_cur;
    a.crho = s->crho;
    a.cscale = __builtin_sqrtf(1.0f - s->crho * s->crho);
    std::memcpy(a.chol, s->chol, sizeof(a.chol));
    std::memcpy(a.pstd, s->pstd, sizeof(a.pstd));
    a.prior = s->run_prior;
    a.bounds = s->run_bounds;
    a.md = s->md;
    hipLaunchKernelGGL((k_rb_move<NI, H>), dim3(s->nt), dim3(NI), 0, s->st, a);
    HIP_TRY(hipGetLastError());
    return 0;
}

#define RB_DISPATCH_H(ni, H, fn, ...)                      \
    switch (ni) {                                          \
        case 64: return fn<64, H>(__VA_ARGS__);            \
        case 128: return fn<128, H>(__VA_ARGS__);          \
        case 256: return fn<256, H>(__VA_ARGS__);          \
        case 512: return fn<512, H>(__VA_ARGS__);          \
        case 1024: return fn<1024, H>(__VA_ARGS__);        \
        default: set_err("N_inner %d", ni); return -1;     \
    }
#define RB_DISPATCH(s, fn, ...)                                      \
    do {                                                             \
        if ((s)->es == 2) { RB_DISPATCH_H((s)->ni, true, fn, __VA_ARGS__) } \
        RB_DISPATCH_H((s)->ni, false, fn, __VA_ARGS__)               \
    } while (0)

int launch_init(SMC2StateCUDA* s) { RB_DISPATCH(s, launch_init_t, s); }
int launch_step(SMC2StateCUDA* s, float y) { RB_DISPATCH(s, launch_step_t, s, y); }
int launch_move(SMC2StateCUDA* s, int k, bool cp) { RB_DISPATCH(s, launch_move_t, s, k, cp); }

dim3 row_grid(long long len, int rows) {
    long long b = (len / 4 + 255) / 256;
    if (b < 1) b = 1;
    if (b > 64) b = 64;
    return dim3((unsigned)b, (unsigned)rows);
}

template <class T>
int gather_rows_t(SMC2StateCUDA* s, const void* src, void* dst, long long stride, long long len) {
    if (len <= 0) return 0;
    constexpr int W = 16 / (int)sizeof(T);
    const T* a = static_cast<const T*>(src);
    T* b = static_cast<T*>(dst);
    if (stride % W == 0 && len % W == 0)
        hipLaunchKernelGGL((k_rows_gather<T, true>), row_grid(len * sizeof(T) / 4, s->nt), dim3(256), 0, s->st, a, b,
                           s->d_anc, stride, len);
    else
        hipLaunchKernelGGL((k_rows_gather<T, false>), row_grid(len * 4, s->nt), dim3(256), 0, s->st, a, b, s->d_anc,
                           stride, len);
    HIP_TRY(hipGetLastError());
    return 0;
}

template <class T>
int commit_rows_t(SMC2StateCUDA* s, const void* src, void* dst, long long stride, long long off, long long len) {
    if (len <= 0) return 0;
    constexpr int W = 16 / (int)sizeof(T);
    const T* a = static_cast<const T*>(src);
    T* b = static_cast<T*>(dst);
    if (stride % W == 0 && off % W == 0 && len % W == 0)
        hipLaunchKernelGGL((k_rows_commit<T, true>), row_grid(len * sizeof(T) / 4, s->nt), dim3(256), 0, s->st, a, b,
                           s->d_acc, stride, off, len);
    else
        hipLaunchKernelGGL((k_rows_commit<T, false>), row_grid(len * 4, s->nt), dim3(256), 0, s->st, a, b,
                           s->d_acc, stride, off, len);
    HIP_TRY(hipGetLastError());
    return 0;
}

// rows of 4-byte elements (inner states, checkpoint log-likelihood halves)
int gather_rows(SMC2StateCUDA* s, const void* src, void* dst, long long stride, long long len) {
    return gather_rows_t<float>(s, src, dst, stride, len);
}
// rows of noise-history elements
int gather_noise(SMC2StateCUDA* s, const void* src, void* dst, long long stride, long long len) {
    return s->es == 2 ? gather_rows_t<uint16_t>(s, src, dst, stride, len)
                      : gather_rows_t<float>(s, src, dst, stride, len);
}
int commit_noise(SMC2StateCUDA* s, const void* src, void* dst, long long stride, long long off, long long len) {
    return s->es == 2 ? commit_rows_t<uint16_t>(s, src, dst, stride, off, len)
                      : commit_rows_t<float>(s, src, dst, stride, off, len);
}

int download_theta(SMC2StateCUDA* s) {
    const Theta& t = s->th[s->tb];
    for (int q = 0; q < kParams; ++q)
        HIP_TRY(hipMemcpyAsync(s->h_th + (size_t)q * s->nt, t.p[q], (size_t)s->nt * 4, hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    return 0;
}

// Outer resample + rejuvenation (smc2_rbpf_batch.cu:1376-1481).
int resample_and_move(SMC2StateCUDA* s) {
    const int n = s->nt, ni = s->ni, t = s->t_cur;
    s->n_res += 1;
    // systematic resample of the theta-particles (host, FP64 CDF)
    const float u = xs64_uniform(&s->host_rng);
    std::vector<double> cdf(n);
    double c = 0.0;
    for (int m = 0; m < n; ++m) {
        c += s->w[m];
        cdf[m] = c;
    }
    cdf[n - 1] = 1.0;
    for (int i = 0; i < n; ++i) {
        const double ui = (double)((u + (float)i) / (float)n);
        int lo = 0, hi = n - 1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (cdf[mid] < ui) lo = mid + 1; else hi = mid;
        }
        s->h_anc[i] = lo;
    }
    HIP_TRY(hipMemcpyAsync(s->d_anc, s->h_anc, (size_t)n * 4, hipMemcpyHostToDevice, s->st));
    hipLaunchKernelGGL(k_theta_gather, dim3((n + 255) / 256), dim3(256), 0, s->st, s->th[s->tb],
                       s->th[1 - s->tb], s->d_anc, n);
    HIP_TRY(hipGetLastError());
    for (int q = 0; q < 4; ++q)
        if (gather_rows(s, s->in[s->ib].a[q], s->in[1 - s->ib].a[q], ni, ni)) return -1;
    const long long zrow = (long long)(s->cap + 1) * ni;
    if (gather_noise(s, s->zn[s->nb], s->zn[1 - s->nb], zrow, (long long)(t + 2) * ni)) return -1;
    if (gather_noise(s, s->un[s->nb], s->un[1 - s->nb], s->cap + 1, t + 2)) return -1;
    s->tb = 1 - s->tb;
    s->ib = 1 - s->ib;
    s->nb = 1 - s->nb;
    if (s->lag > 0 && s->t_cp >= 0) {
        // the checkpoint (state and log-likelihood) follows its theta-particle
        for (int q = 0; q < 4; ++q)
            if (gather_rows(s, s->cp[s->cb].a[q], s->cp[1 - s->cb].a[q], ni, ni)) return -1;
        if (gather_rows(s, reinterpret_cast<const float*>(s->cpll[s->cb]),
                        reinterpret_cast<float*>(s->cpll[1 - s->cb]), 2, 2))
            return -1;
        s->cb = 1 - s->cb;
    }
    for (int m = 0; m < n; ++m) s->lw[m] = 0.0;

    bool use_cp = false;
    if (s->lag > 0 && s->t_cp >= 0) {
        const int steps = t - s->t_cp;
        use_cp = steps > 0 && steps <= 2 * s->lag;
    }
    if (download_theta(s)) return -1;
    if (s->adaptive) adapt_chol(s);
    else initial_chol(s);
    const long long t_start = use_cp ? (long long)s->t_cp + 1 : 0;
    for (int k = 0; k < s->K; ++k) {
        if (launch_move(s, k, use_cp)) return -1;
        if (commit_noise(s, s->zn[1 - s->nb], s->zn[s->nb], zrow, t_start * ni, ((long long)t + 2 - t_start) * ni))
            return -1;
        if (commit_noise(s, s->un[1 - s->nb], s->un[s->nb], s->cap + 1, t_start, (long long)t + 2 - t_start))
            return -1;
        HIP_TRY(hipMemcpyAsync(s->h_acc, s->d_acc, (size_t)n * 4, hipMemcpyDeviceToHost, s->st));
        HIP_TRY(hipStreamSynchronize(s->st));
        int a = 0;
        for (int m = 0; m < n; ++m) a += s->h_acc[m];
        s->n_acc += a;
        s->n_tot += n;
    }
    if (s->K > 0 && download_theta(s)) return -1;
    for (int m = 0; m < n; ++m) s->lw[m] = 0.0;
    outer_ess(s);
    return 0;
}

int save_checkpoint(SMC2StateCUDA* s) {
    const size_t bytes = (size_t)s->nt * s->ni * 4;
    for (int q = 0; q < 4; ++q)
        HIP_TRY(hipMemcpyAsync(s->cp[s->cb].a[q], s->in[s->ib].a[q], bytes, hipMemcpyDeviceToDevice, s->st));
    HIP_TRY(hipMemcpyAsync(s->cpll[s->cb], s->th[s->tb].ll, (size_t)s->nt * 8, hipMemcpyDeviceToDevice, s->st));
    return 0;
}

void release(SMC2StateCUDA* s) {
    for (int b = 0; b < 2; ++b) {
        for (int q = 0; q < kParams; ++q)
            if (s->th[b].p[q]) (void)hipFree(s->th[b].p[q]);
        if (s->th[b].ll) (void)hipFree(s->th[b].ll);
        if (s->th[b].essi) (void)hipFree(s->th[b].essi);
        for (int q = 0; q < 4; ++q) {
            if (s->in[b].a[q]) (void)hipFree(s->in[b].a[q]);
            if (s->cp[b].a[q]) (void)hipFree(s->cp[b].a[q]);
        }
        if (s->cpll[b]) (void)hipFree(s->cpll[b]);
    }
    free_noise(s->zn, s->un);
    if (s->d_y) (void)hipFree(s->d_y);
    if (s->d_anc) (void)hipFree(s->d_anc);
    if (s->d_acc) (void)hipFree(s->d_acc);
    if (s->h_lli) (void)hipHostFree(s->h_lli);
    if (s->h_anc) (void)hipHostFree(s->h_anc);
    if (s->h_acc) (void)hipHostFree(s->h_acc);
    if (s->h_th) (void)hipHostFree(s->h_th);
    if (s->ev_tick) (void)hipEventDestroy(s->ev_tick);
    if (s->st) (void)hipStreamDestroy(s->st);
    delete s;
}

int alloc_all(SMC2StateCUDA* s) {
    const size_t n = (size_t)s->nt, tot = n * (size_t)s->ni;
    HIP_TRY(hipStreamCreateWithFlags(&s->st, hipStreamNonBlocking));
    HIP_TRY(hipEventCreateWithFlags(&s->ev_tick, hipEventDisableTiming));
    for (int b = 0; b < 2; ++b) {
        for (int q = 0; q < kParams; ++q)
            if (dmalloc(&s->th[b].p[q], n)) return -1;
        if (dmalloc(&s->th[b].ll, n) || dmalloc(&s->th[b].essi, n)) return -1;
        for (int q = 0; q < 4; ++q)
            if (dmalloc(&s->in[b].a[q], tot) || dmalloc(&s->cp[b].a[q], tot)) return -1;
        if (dmalloc(&s->cpll[b], n)) return -1;
    }
    if (alloc_noise(s, s->cap, s->zn, s->un)) return -1;
    if (dmalloc(&s->d_anc, n) || dmalloc(&s->d_acc, n)) return -1;
    // the step kernel writes the per-theta increments straight to host memory
    HIP_TRY(hipHostMalloc((void**)&s->h_lli, n * 8, hipHostMallocMapped));
    HIP_TRY(hipHostGetDevicePointer((void**)&s->d_lli, s->h_lli, 0));
    HIP_TRY(hipHostMalloc((void**)&s->h_anc, n * 4, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void**)&s->h_acc, n * 4, hipHostMallocDefault));
    HIP_TRY(hipHostMalloc((void**)&s->h_th, n * 4 * kParams, hipHostMallocDefault));
    if (ensure_y(s, 8000)) return -1;
    return 0;
}

}  // namespace

extern "C" {

SMC2StateCUDA* smc2_cuda_alloc(int N_theta, int N_inner) {
    clear_err();
    if (N_theta < 2) {
        set_err("smc2_cuda_alloc: N_theta must be >= 2 (got %d)", N_theta);
        return nullptr;
    }
    if (N_inner != 64 && N_inner != 128 && N_inner != 256 && N_inner != 512 && N_inner != 1024) {
        set_err("smc2_cuda_alloc: N_inner must be 64, 128, 256, 512 or 1024 (got %d)", N_inner);
        return nullptr;
    }
    SMC2StateCUDA* s = new SMC2StateCUDA();
    s->nt = N_theta;
    s->ni = N_inner;
    s->cap = 2048;
    s->lw.assign(N_theta, 0.0);
    s->w.assign(N_theta, 1.0 / N_theta);
    default_model(s);
    s->host_rng = 0x853C49E6748FEA9BULL ^ s->seed;
    if (alloc_all(s)) {
        release(s);
        return nullptr;
    }
    if (build_model(s)) {
        release(s);
        return nullptr;
    }
    initial_chol(s);
    return s;
}

void smc2_cuda_free(SMC2StateCUDA* state) {
    if (!state) return;
    (void)hipStreamSynchronize(state->st);
    release(state);
}

void smc2_cuda_set_seed(SMC2StateCUDA* state, uint64_t seed) {
    if (!state) return;
    state->seed = seed ? seed : 12345;
    state->host_rng = 0x853C49E6748FEA9BULL ^ state->seed;
}

void smc2_cuda_set_noise_capacity(SMC2StateCUDA* state, int capacity) {
    if (!state || capacity <= state->cap) return;
    clear_err();
    (void)grow_noise(state, capacity);
}

void smc2_cuda_set_fixed_lag(SMC2StateCUDA* state, int L) {
    if (!state) return;
    state->lag = L > 0 ? L : 0;
    state->t_cp = -1;
}

void smc2_cuda_set_cpmmh_rho(SMC2StateCUDA* state, float rho) {
    if (state) state->crho = rho;
}

void smc2_cuda_set_proposal_std(SMC2StateCUDA* state, const float* std) {
    if (!state) return;
    static const float def[kParams] = {0.01f, 0.02f, 0.1f, 0.02f};
    std::memcpy(state->pstd, std ? std : def, sizeof(state->pstd));
}

void smc2_cuda_init_from_prior(SMC2StateCUDA* state) {
    if (!state) return;
    clear_err();
    SMC2StateCUDA* s = state;
    if (build_model(s)) return;
    initial_chol(s);
    s->run_prior = s->prior;
    s->run_bounds = s->bounds;
    s->tb = s->ib = s->nb = s->cb = 0;
    s->n_res = s->n_acc = s->n_tot = 0;
    s->t_cur = -1;
    s->t_cp = -1;
    for (int m = 0; m < s->nt; ++m) s->lw[m] = 0.0;
    if (launch_init(s)) return;
    if (hipStreamSynchronize(s->st) != hipSuccess) {
        set_err("smc2_cuda_init_from_prior: %s", hipGetErrorString(hipGetLastError()));
        return;
    }
    if (download_theta(s)) return;
    outer_ess(s);
    s->inited = true;
}

float smc2_cuda_update(SMC2StateCUDA* state, float y_obs) {
    const float nan = __builtin_nanf("");
    if (!state) return nan;
    clear_err();
    SMC2StateCUDA* s = state;
    if (!s->inited) {
        set_err("smc2_cuda_update: call smc2_cuda_init_from_prior first");
        return nan;
    }
    const int t = s->t_cur + 1;
    if (ensure_y(s, t + 1)) return nan;
    if (t >= s->cap && grow_noise(s, 2 * s->cap)) return nan;
    s->t_cur = t;
    if (launch_step(s, y_obs)) return nan;
    // the tick's increments are read right away: wait by polling an event
    // (a blocking synchronisation adds the host thread's wake-up latency to
    // every update)
    hipError_t we = hipEventRecord(s->ev_tick, s->st);
    if (we == hipSuccess) {
        do {
            we = hipEventQuery(s->ev_tick);
        } while (we == hipErrorNotReady);
    }
    if (we != hipSuccess) {
        set_err("smc2_cuda_update: tick failed: %s", hipGetErrorString(we));
        return nan;
    }
    for (int m = 0; m < s->nt; ++m) s->lw[m] += s->h_lli[m];
    float ess = outer_ess(s);
    if (ess < s->ess_thr * (float)s->nt) {
        if (resample_and_move(s)) return nan;
        ess = s->ess;
    }
    // fixed-lag checkpoint at every multiple of L (smc2_rbpf_batch.cu:1484-1494
    // also saves at t = 1..L-1 under the label 0, so a replay from it would
    // apply tick 1 twice; here a checkpoint is always labelled with its tick)
    if (s->lag > 0 && t > 0 && t % s->lag == 0 && t > s->t_cp) {
        if (save_checkpoint(s)) return nan;
        s->t_cp = t;
        if (hipStreamSynchronize(s->st) != hipSuccess) {
            set_err("smc2_cuda_update: checkpoint: %s", hipGetErrorString(hipGetLastError()));
            return nan;
        }
    }
    return ess;
}

void smc2_cuda_get_theta_mean(SMC2StateCUDA* state, float* theta_mean) {
    if (!state || !theta_mean) return;
    const int n = state->nt;
    for (int q = 0; q < kParams; ++q) {
        double a = 0.0;
        for (int m = 0; m < n; ++m) a += state->w[m] * (double)state->h_th[q * n + m];
        theta_mean[q] = (float)a;
    }
}

void smc2_cuda_get_theta_std(SMC2StateCUDA* state, float* theta_std) {
    if (!state || !theta_std) return;
    const int n = state->nt;
    for (int q = 0; q < kParams; ++q) {
        double mean = 0.0;
        for (int m = 0; m < n; ++m) mean += state->w[m] * (double)state->h_th[q * n + m];
        double v = 0.0;
        for (int m = 0; m < n; ++m) {
            const double d = (double)state->h_th[q * n + m] - mean;
            v += state->w[m] * d * d;
        }
        theta_std[q] = (float)std::sqrt(v);
    }
}

float smc2_cuda_get_outer_ess(SMC2StateCUDA* state) { return state ? state->ess : __builtin_nanf(""); }

int smc2_cuda_set_model(SMC2StateCUDA* state, const SVPrior* prior, const SVBounds* bounds,
                        const SVCurve* theta_curve, const SVFixedCurves* fixed_curves) {
    clear_err();
    if (!state) {
        set_err("smc2_cuda_set_model: null state");
        return -1;
    }
    if (prior) state->prior = *prior;
    if (bounds) state->bounds = *bounds;
    if (theta_curve) state->tcurve = *theta_curve;
    if (fixed_curves) state->fixed = *fixed_curves;
    return 0;
}

int smc2_cuda_get_model(SMC2StateCUDA* state, SVPrior* prior, SVBounds* bounds, SVCurve* theta_curve,
                        SVFixedCurves* fixed_curves) {
    clear_err();
    if (!state) {
        set_err("smc2_cuda_get_model: null state");
        return -1;
    }
    if (prior) *prior = state->prior;
    if (bounds) *bounds = state->bounds;
    if (theta_curve) *theta_curve = state->tcurve;
    if (fixed_curves) *fixed_curves = state->fixed;
    return 0;
}

int smc2_cuda_set_rejuv(SMC2StateCUDA* state, int K_rejuv, float ess_threshold_outer) {
    clear_err();
    if (!state || K_rejuv < 0 || !(ess_threshold_outer >= 0.0f && ess_threshold_outer <= 1.0f)) {
        set_err("smc2_cuda_set_rejuv: need K_rejuv >= 0 and a threshold in [0, 1]");
        return -1;
    }
    state->K = K_rejuv;
    state->ess_thr = ess_threshold_outer;
    return 0;
}

int smc2_cuda_set_noise_precision(SMC2StateCUDA* state, int bits) {
    clear_err();
    if (!state || (bits != 16 && bits != 32)) {
        set_err("smc2_cuda_set_noise_precision: bits must be 16 or 32");
        return -1;
    }
    SMC2StateCUDA* s = state;
    const int es = bits / 8;
    if (es == s->es) return 0;
    if (hipStreamSynchronize(s->st) != hipSuccess) {
        set_err("smc2_cuda_set_noise_precision: %s", hipGetErrorString(hipGetLastError()));
        return -1;
    }
    free_noise(s->zn, s->un);
    s->es = es;
    if (alloc_noise(s, s->cap, s->zn, s->un)) return -1;
    s->inited = false;       // the histories restart with the next init
    return 0;
}

int smc2_cuda_set_adaptive(SMC2StateCUDA* state, int on) {
    clear_err();
    if (!state) {
        set_err("smc2_cuda_set_adaptive: null state");
        return -1;
    }
    state->adaptive = on != 0;
    return 0;
}

int smc2_cuda_get_counters(SMC2StateCUDA* state, SMC2Counters* out) {
    clear_err();
    if (!state || !out) {
        set_err("smc2_cuda_get_counters: null argument");
        return -1;
    }
    out->n_resamples = state->n_res;
    out->n_rejuv_accepts = state->n_acc;
    out->n_rejuv_total = state->n_tot;
    out->t_current = state->t_cur;
    out->t_checkpoint = state->t_cp;
    out->noise_capacity = state->cap;
    return 0;
}

int smc2_cuda_get_particles(SMC2StateCUDA* state, float* rho, float* sigma_z, float* mu_base,
                            float* sigma_base, double* log_weight, double* log_likelihood,
                            float* ess_inner) {
    clear_err();
    if (!state) {
        set_err("smc2_cuda_get_particles: null state");
        return -1;
    }
    SMC2StateCUDA* s = state;
    const Theta& t = s->th[s->tb];
    float* outs[kParams] = {rho, sigma_z, mu_base, sigma_base};
    for (int q = 0; q < kParams; ++q)
        if (outs[q]) HIP_TRY(hipMemcpyAsync(outs[q], t.p[q], (size_t)s->nt * 4, hipMemcpyDeviceToHost, s->st));
    if (log_likelihood) HIP_TRY(hipMemcpyAsync(log_likelihood, t.ll, (size_t)s->nt * 8, hipMemcpyDeviceToHost, s->st));
    if (ess_inner) HIP_TRY(hipMemcpyAsync(ess_inner, t.essi, (size_t)s->nt * 4, hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    if (log_weight) std::memcpy(log_weight, s->lw.data(), (size_t)s->nt * 8);
    return 0;
}

int smc2_cuda_get_inner(SMC2StateCUDA* state, float* z_tilde, float* mu_h, float* var_h, float* log_w) {
    clear_err();
    if (!state) {
        set_err("smc2_cuda_get_inner: null state");
        return -1;
    }
    SMC2StateCUDA* s = state;
    float* outs[4] = {z_tilde, mu_h, var_h, log_w};
    const size_t bytes = (size_t)s->nt * s->ni * 4;
    for (int q = 0; q < 4; ++q)
        if (outs[q]) HIP_TRY(hipMemcpyAsync(outs[q], s->in[s->ib].a[q], bytes, hipMemcpyDeviceToHost, s->st));
    HIP_TRY(hipStreamSynchronize(s->st));
    return 0;
}

}  // extern "C"
