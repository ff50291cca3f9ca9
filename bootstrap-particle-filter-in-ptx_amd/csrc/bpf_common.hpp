// bpf_common.hpp -- device helpers shared by the single-filter kernels and
// the batched inner-filter kernels: wave64 reductions and scans, vector
// loads, the host-stream uniform, the systematic-target counts, and the
// per-tick statistics and resampling decision.
#pragma once

#include "bpf_types.hpp"

namespace bpf {

// ---------------------------------------------------------------------------
// Wavefront helpers (wave64) on DPP cross-lane moves -- quad permutations,
// shifts and mirrors within 16-lane rows, the row broadcasts of lanes 15 and
// 31 -- and v_readlane, so no step goes through the LDS crossbar.  FP64
// values move as two 32-bit halves.  Each helper has one fixed association
// order, so every kernel that uses it forms the same sums bit for bit.  Call
// them with the whole wave active.
// ---------------------------------------------------------------------------

namespace dpp {
constexpr int kQuad1032 = 0xB1;      // quad_perm [1, 0, 3, 2]
constexpr int kQuad2301 = 0x4E;      // quad_perm [2, 3, 0, 1]
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr3 = 0x113, kRowShr4 = 0x114, kRowShr8 = 0x118;
constexpr int kWaveShr1 = 0x138;     // lane i takes lane i - 1 across rows
constexpr int kRowMirror = 0x140;    // lane i of a row takes lane 15 - i
constexpr int kRowHalfMirror = 0x141;  // lane i of a half row takes lane 7 - i
constexpr int kRowBcast15 = 0x142;   // rows take lane 15 of the row before
constexpr int kRowBcast31 = 0x143;   // rows 2 and 3 take lane 31
}  // namespace dpp

// The value from the lane that CTRL selects; lanes outside the ROWS / BANKS
// masks, and lanes whose source lies outside their row, receive `id`.
template <int CTRL, int ROWS = 0xf, int BANKS = 0xf>
__device__ inline int dpp_i32(int v, int id) {
    return __builtin_amdgcn_update_dpp(id, v, CTRL, ROWS, BANKS, false);
}

template <int CTRL, int ROWS = 0xf, int BANKS = 0xf>
__device__ inline float dpp_f32(float v, float id) {
    return __int_as_float(dpp_i32<CTRL, ROWS, BANKS>(__float_as_int(v), __float_as_int(id)));
}

template <int CTRL, int ROWS = 0xf, int BANKS = 0xf>
__device__ inline double dpp_f64(double v, double id) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned long long ib = (unsigned long long)__double_as_longlong(id);
    const unsigned lo = (unsigned)dpp_i32<CTRL, ROWS, BANKS>((int)(unsigned)b, (int)(unsigned)ib);
    const unsigned hi = (unsigned)dpp_i32<CTRL, ROWS, BANKS>((int)(unsigned)(b >> 32), (int)(unsigned)(ib >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// The same moves with nothing preset in the destination: lanes that the
// masks disable keep whatever the register held (callers read none of
// them), and with ZERO a lane whose source lies outside its row receives
// +0.0 / 0.  They save the move of the identity that the forms above need.
template <int CTRL, int ROWS = 0xf, int BANKS = 0xf, bool ZERO = false>
__device__ inline int dpp_raw_i32(int v) {
    return __builtin_amdgcn_mov_dpp(v, CTRL, ROWS, BANKS, ZERO);
}

template <int CTRL, int ROWS = 0xf, int BANKS = 0xf, bool ZERO = false>
__device__ inline float dpp_raw_f32(float v) {
    return __int_as_float(dpp_raw_i32<CTRL, ROWS, BANKS, ZERO>(__float_as_int(v)));
}

template <int CTRL, int ROWS = 0xf, int BANKS = 0xf, bool ZERO = false>
__device__ inline double dpp_raw_f64(double v) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)dpp_raw_i32<CTRL, ROWS, BANKS, ZERO>((int)(unsigned)b);
    const unsigned hi = (unsigned)dpp_raw_i32<CTRL, ROWS, BANKS, ZERO>((int)(unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Orders the wave's own LDS writes before its later LDS reads.
__device__ inline void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Rank of the calling lane among the lanes of `mask` below it, plus `base`.
__device__ inline int lanes_below(uint64_t mask, int base) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, base));
}

// Lane L's value in every lane (L a constant).
__device__ inline double lane_f64(double v, int L) {
    const unsigned long long b = (unsigned long long)__double_as_longlong(v);
    const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, L);
    const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), L);
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

__device__ inline int lane_i32(int v, int L) { return __builtin_amdgcn_readlane(v, L); }

__device__ inline float lane_f32(float v, int L) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), L));
}

// The last lane's value, and the previous lane's (lane 0: fill).
__device__ inline double wave_last(double v) { return lane_f64(v, kWave - 1); }
__device__ inline double wave_prev(double v, double fill) { return dpp_f64<dpp::kWaveShr1>(v, fill); }

// Reductions: quads, half rows, rows, then rows 0 + 1 into row 1 and rows
// 2 + 3 into row 3, and the two into lane 63, which every lane reads.  The
// quad and mirror moves give every lane a source; after the two row
// broadcasts only rows 1 and 3 hold defined values, and only lane 63 is read.
__device__ inline float wave_max(float v) {
    v = fmaxf(v, dpp_raw_f32<dpp::kQuad1032>(v));
    v = fmaxf(v, dpp_raw_f32<dpp::kQuad2301>(v));
    v = fmaxf(v, dpp_raw_f32<dpp::kRowHalfMirror>(v));
    v = fmaxf(v, dpp_raw_f32<dpp::kRowMirror>(v));
    v = fmaxf(v, dpp_raw_f32<dpp::kRowBcast15, 0xa>(v));
    v = fmaxf(v, dpp_raw_f32<dpp::kRowBcast31, 0xc>(v));
    return lane_f32(v, kWave - 1);
}

__device__ inline double wave_sum(double v) {
    v += dpp_raw_f64<dpp::kQuad1032>(v);
    v += dpp_raw_f64<dpp::kQuad2301>(v);
    v += dpp_raw_f64<dpp::kRowHalfMirror>(v);
    v += dpp_raw_f64<dpp::kRowMirror>(v);
    v += dpp_raw_f64<dpp::kRowBcast15, 0xa>(v);
    v += dpp_raw_f64<dpp::kRowBcast31, 0xc>(v);
    return wave_last(v);
}

// Inclusive scans: each lane adds its three row predecessors, then the
// partial sums 4 and 8 lanes back, then the totals of the rows before (row
// broadcasts of lanes 15 and 31).  A lane with no predecessor that far back
// in its row adds +0 (the row shifts' zero fill); rows 0 and 2, and rows 0
// and 1, add +0 in the two broadcasts.
__device__ inline double wave_incl_scan(double v, int /*lane*/) {
    double s = v + dpp_raw_f64<dpp::kRowShr1, 0xf, 0xf, true>(v);
    s += dpp_raw_f64<dpp::kRowShr2, 0xf, 0xf, true>(v);
    s += dpp_raw_f64<dpp::kRowShr3, 0xf, 0xf, true>(v);
    s += dpp_raw_f64<dpp::kRowShr4, 0xf, 0xf, true>(s);
    s += dpp_raw_f64<dpp::kRowShr8, 0xf, 0xf, true>(s);
    s += dpp_f64<dpp::kRowBcast15, 0xa>(s, 0.0);
    s += dpp_f64<dpp::kRowBcast31, 0xc>(s, 0.0);
    return s;
}

__device__ inline int wave_incl_scan(int v, int /*lane*/) {
    int s = v + dpp_raw_i32<dpp::kRowShr1, 0xf, 0xf, true>(v);
    s += dpp_raw_i32<dpp::kRowShr2, 0xf, 0xf, true>(v);
    s += dpp_raw_i32<dpp::kRowShr3, 0xf, 0xf, true>(v);
    s += dpp_raw_i32<dpp::kRowShr4, 0xf, 0xf, true>(s);
    s += dpp_raw_i32<dpp::kRowShr8, 0xf, 0xf, true>(s);
    s += dpp_i32<dpp::kRowBcast15, 0xa>(s, 0);
    s += dpp_i32<dpp::kRowBcast31, 0xc>(s, 0);
    return s;
}

// Inclusive prefix maximum of non-negative values (CDF values and prefix
// sums of weights), in the scan's order; +0 stands in for a missing
// predecessor, which leaves any value >= +0 unchanged.
__device__ inline double wave_incl_max(double v, int /*lane*/) {
    double s = fmax(v, dpp_raw_f64<dpp::kRowShr1, 0xf, 0xf, true>(v));
    s = fmax(s, dpp_raw_f64<dpp::kRowShr2, 0xf, 0xf, true>(v));
    s = fmax(s, dpp_raw_f64<dpp::kRowShr3, 0xf, 0xf, true>(v));
    s = fmax(s, dpp_raw_f64<dpp::kRowShr4, 0xf, 0xf, true>(s));
    s = fmax(s, dpp_raw_f64<dpp::kRowShr8, 0xf, 0xf, true>(s));
    s = fmax(s, dpp_f64<dpp::kRowBcast15, 0xa>(s, 0.0));
    s = fmax(s, dpp_f64<dpp::kRowBcast31, 0xc>(s, 0.0));
    return s;
}

// Lane-shuffle forms with the butterfly (offsets 32, 16, ..., 1) and
// Hillis-Steele (offsets 1, 2, ..., 32) association orders, for the kernels
// whose CPU restatement repeats those orders bit for bit (the RBPF inner
// filter).
__device__ inline double wave_sum_bfly(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, kWave);
    return v;
}

__device__ inline double wave_incl_scan_hs(double v, int lane) {
#pragma unroll
    for (int off = 1; off < kWave; off <<= 1) {
        double o = __shfl_up(v, off, kWave);
        v += (lane >= off) ? o : 0.0;
    }
    return v;
}

// LDS index padded by one element per 16, so lane-contiguous runs of 16
// (stride 16 elements between lanes) fall into distinct banks.
__device__ inline int pad16(int j) { return j + (j >> 4); }

typedef float v4f __attribute__((ext_vector_type(4)));

// Streaming (non-temporal) variants for particle arrays far larger than the
// caches: each element is read once and written once per tick.
template <int VEC>
__device__ inline void load_vec_nt(const float* p, float* v) {
    if constexpr (VEC == 4) {
        v4f t = __builtin_nontemporal_load(reinterpret_cast<const v4f*>(p));
        v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
    } else {
        for (int c = 0; c < VEC; ++c) v[c] = __builtin_nontemporal_load(p + c);
    }
}

template <int VEC>
__device__ inline void store_vec_nt(float* p, const float* v) {
    if constexpr (VEC == 4) {
        v4f t = {v[0], v[1], v[2], v[3]};
        __builtin_nontemporal_store(t, reinterpret_cast<v4f*>(p));
    } else {
        for (int c = 0; c < VEC; ++c) __builtin_nontemporal_store(v[c], p + c);
    }
}

template <int VEC>
__device__ inline void load_vec(const float* p, float* v) {
    if constexpr (VEC == 4) {
        float4 t = *reinterpret_cast<const float4*>(p);
        v[0] = t.x; v[1] = t.y; v[2] = t.z; v[3] = t.w;
    } else if constexpr (VEC == 2) {
        float2 t = *reinterpret_cast<const float2*>(p);
        v[0] = t.x; v[1] = t.y;
    } else {
        v[0] = *p;
    }
}

template <int VEC>
__device__ inline void store_vec(float* p, const float* v) {
    if constexpr (VEC == 4) {
        *reinterpret_cast<float4*>(p) = make_float4(v[0], v[1], v[2], v[3]);
    } else if constexpr (VEC == 2) {
        *reinterpret_cast<float2*>(p) = make_float2(v[0], v[1]);
    } else {
        *p = v[0];
    }
}


// Jump constants of draw k (the single filter's tick-k draw): particle
// i's PCG32 state is seed * A + (2i + 1) * G with A = M^(k+2) and
// G = sum_{j < k+2} M^j (mod 2^64), by binary powering of the LCG step.
__device__ inline void pcg_draw_constants(uint32_t k, uint64_t& A, uint64_t& G) {
    uint64_t am = 1, ap = 0, cm = kPcgMult, cpl = 1;
    for (uint32_t d = k + 2; d > 0; d >>= 1) {
        if (d & 1u) {
            am *= cm;
            ap = ap * cm + cpl;
        }
        cpl = (cm + 1) * cpl;
        cm *= cm;
    }
    A = am;
    G = ap;
}

__device__ inline float host_pcg_float(unsigned long long* st) {
    unsigned long long old = *st;
    *st = old * kPcgMult + kPcgHostInc;
    return (float)(pcg_output(old) >> 9) * (1.0f / 8388608.0f);
}

// W = first i in [0, n] with u + i/n >= 1 (FP32, gpu_bpf/gpu_bpf.cu:300-302);
// the targets are non-decreasing in i.  Starts from the real-number estimate
// ceil((1 - u) n), which the float rounding moves by a few positions at most,
// and steps to the exact boundary; bisects [0, n] only if that takes more
// than 16 steps down or 32 in all.
__device__ inline int wrap_start(float u, int n) {
    const double rn = target_recip(n);
    const long long est = (long long)ceil((1.0 - (double)u) * (double)n);
    int g = (int)max(0LL, min((long long)n, est));
    int steps = 0;
    bool ok = true;
    while (g > 0 && target_at(u, g - 1, rn) >= 1.0f) {
        --g;
        if (++steps > 16) {
            ok = false;
            break;
        }
    }
    if (ok) {
        while (g < n && target_at(u, g, rn) < 1.0f) {
            ++g;
            if (++steps > 32) {
                ok = false;
                break;
            }
        }
    }
    if (!ok) {
        int lo = 0, hi = n;
        while (lo < hi) {
            const int mid = (int)(((unsigned)lo + (unsigned)hi) >> 1);
            if (target_at(u, mid, rn) >= 1.0f) hi = mid; else lo = mid + 1;
        }
        g = lo;
    }
    return g;
}

// The tick's statistics and decisions from the global sums (one thread, on
// a register copy of the control block).
// u_raw_in >= 0 replaces the host-stream draw (replays of stored noise).
__device__ inline void finalize_tick(Ctl& c, int n_particles, float threshold, float M, double S,
                                     double SX, double SXX, double Q, float u_raw_in = -1.0f) {
    const int carry = c.carry;
    const double n = (double)n_particles;
    const double Md = (double)M;
    BpfTickStats r;
    r.tick = c.tick;
    const bool ok = isfinite(S) && S > 0.0 && isfinite(Md);
    const double ess = (S * S) / Q;
    const int resample = ok && (threshold <= 0.0f || (ess / n) < (double)threshold);
    // log(S): the increment of a tick whose weights carried in, and the next
    // tick's when they carry on (kept in the control block, so each tick
    // takes one log of S rather than logs of S and of the previous S)
    const double log_s = (carry || !resample) ? log(S) : 0.0;
    if (ok) {
        r.mean = SX / S;
        r.var = SXX / S - r.mean * r.mean;
        r.ess = ess;
        // one log of S / n serves both log-likelihoods
        const double sn = S / n;
        const double lsn = log(sn);
        const double fl = (double)bits_f32(0x0D6BF94Eu);
        double lref;
        if (sn >= fl) lref = lsn;
        else lref = log(fl);
        r.ll_ref = c.off + Md + lref;
        r.ll_incr = carry ? (Md + log_s) - c.log_s_prev : Md + lsn;
    } else {
        const double qn = __builtin_nan("");
        r.mean = r.var = r.ess = r.ll_ref = r.ll_incr = qn;
    }
    c.ll_cum += r.ll_incr;
    r.ll_cum = c.ll_cum;
    r.h_mean = (float)r.mean;
    r.log_lik = (float)r.ll_ref;
    r.resampled = resample;
    if (resample) {
        const float ur = (u_raw_in >= 0.0f) ? u_raw_in : host_pcg_float(&c.host_rng);
        const float u = ur / (float)n_particles;
        c.u_raw = ur;
        c.u = u;
        c.wrap_start = wrap_start(u, n_particles);
        c.src = c.cur;
        c.dst = 1 - c.cur;
        c.cur = c.dst;
        c.carry = 0;
        c.off = 0.0;
        c.resample_count += 1;
    } else {
        c.carry = 1;
        c.off += Md;
        c.s_prev = S;
        c.log_s_prev = log_s;
    }
    c.m_prev = M;
    c.M = M;
    c.S = S;
    c.resample = resample;
    c.last = r;
    c.tick += 1;
}

// ---------------------------------------------------------------------------
// Systematic targets t_i = u + i/N (wrapped once at 1 from i = W on).
// ---------------------------------------------------------------------------

// Number of i in [lo, hi) whose resampling target lies at or below v:
// (double)(t_i - sub) <= v with t_i = target_at(u, i, rn), sub = 0 for the
// unwrapped targets and 1 for the wrapped ones (i >= W, t_i in [1, 2), so
// t_i - 1 is exact).  The targets are non-decreasing in i, so the count is
// where a monotone predicate flips.  It starts from the real-number estimate
// floor((v + sub - u) N) + 1, which the float rounding of t_i moves by at most
// about N 2^-23 positions, steps to the exact flip, and bisects [lo, hi)
// only if that takes more than 16 steps down or 32 in all.  Inline and
// call-free, so the four range counts of a tile run side by side in four
// lanes and a lane's per-parent counts interleave.
__device__ inline int count_targets_le(double v, float u, float sub, int lo, int hi, int n, double rn) {
    if (lo >= hi) return lo;
    double est = floor((v + (double)sub - (double)u) * (double)n) + 1.0;
    est = fmin(fmax(est, (double)lo), (double)hi);
    int g = (int)est;
    int steps = 0;
    bool ok = true;
    while (g > lo && (double)(target_at(u, g - 1, rn) - sub) > v) {
        --g;
        if (++steps > 16) {
            ok = false;
            break;
        }
    }
    if (ok) {
        while (g < hi && (double)(target_at(u, g, rn) - sub) <= v) {
            ++g;
            if (++steps > 32) {
                ok = false;
                break;
            }
        }
    }
    if (!ok) {
        int a = lo, b = hi;
        while (a < b) {
            const int mid = (int)(((unsigned)a + (unsigned)b) >> 1);
            if ((double)(target_at(u, mid, rn) - sub) <= v) a = mid + 1; else b = mid;
        }
        g = a;
    }
    return g;
}

// First j in [0, last] with c_j >= t over a non-decreasing padded CDF
// (last when none is).
__device__ inline int lds_lower_bound(const double* c, int last, double t) {
    int lo = 0, hi = last;
    while (lo < hi) {
        int mid = (lo + hi) >> 1;
        if (c[pad16(mid)] < t) lo = mid + 1; else hi = mid;
    }
    return lo;
}

__device__ inline int wave_incl_max_int(int v, int /*lane*/) {
    constexpr int NI = -0x7fffffff - 1;
    int s = max(v, dpp_i32<dpp::kRowShr1>(v, NI));
    s = max(s, dpp_i32<dpp::kRowShr2>(v, NI));
    s = max(s, dpp_i32<dpp::kRowShr3>(v, NI));
    s = max(s, dpp_i32<dpp::kRowShr4, 0xf, 0xe>(s, NI));
    s = max(s, dpp_i32<dpp::kRowShr8, 0xf, 0xc>(s, NI));
    s = max(s, dpp_i32<dpp::kRowBcast15, 0xa>(s, NI));
    s = max(s, dpp_i32<dpp::kRowBcast31, 0xc>(s, NI));
    return s;
}

}  // namespace bpf
