"""ctypes binding of the CPU oracle (oracle/_build/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline.
The shipped GPU library never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "liboracle.so")

MODE_F32 = 0
MODE_F64 = 1
MODE_F64_COUPLED = 2


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("n_particles", ctypes.c_int),
        ("rho", ctypes.c_float),
        ("sigma_z", ctypes.c_float),
        ("mu", ctypes.c_float),
        ("nu_obs", ctypes.c_float),
        ("ess_threshold", ctypes.c_float),
        ("seed", ctypes.c_int),
        ("tile", ctypes.c_int),
    ]


class OrcLearn(ctypes.Structure):
    _fields_ = [
        ("update_K", ctypes.c_int),
        ("learn_rho", ctypes.c_int),
        ("c", ctypes.c_float),
        ("t0", ctypes.c_float),
        ("gamma", ctypes.c_float),
        ("grad_clip", ctypes.c_float),
    ]


class OrcBands(ctypes.Structure):
    _fields_ = [
        ("adaptive", ctypes.c_int),
        ("n_particles", ctypes.c_int),
        ("n_bands", ctypes.c_int * 3),
        ("fracs", (ctypes.c_float * 4) * 3),
        ("scales", (ctypes.c_float * 4) * 3),
        ("thr_alert", ctypes.c_float),
        ("thr_panic", ctypes.c_float),
    ]


TICK_DTYPE = np.dtype(
    [
        ("mean", "f8"),
        ("var", "f8"),
        ("ess", "f8"),
        ("ll_incr", "f8"),
        ("ll_cum", "f8"),
        ("ll_ref", "f8"),
        ("h_mean", "f4"),
        ("log_lik", "f4"),
        ("resampled", "i4"),
        ("tick", "i4"),
    ]
)


class OrcRbCfg(ctypes.Structure):
    _fields_ = [("n_theta", ctypes.c_int), ("n_inner", ctypes.c_int), ("K_rejuv", ctypes.c_int),
                ("lag", ctypes.c_int), ("cap", ctypes.c_int), ("adaptive", ctypes.c_int),
                ("noise_bits", ctypes.c_int),
                ("ess_thr", ctypes.c_float), ("crho", ctypes.c_float), ("seed", ctypes.c_uint64),
                ("pstd", ctypes.c_float * 4), ("prior", ctypes.c_float * 8),
                ("bounds", ctypes.c_float * 8), ("tcurve", ctypes.c_float * 3),
                ("fixed", ctypes.c_float * 4)]


class OrcRbCount(ctypes.Structure):
    _fields_ = [("n_resamples", ctypes.c_int), ("n_acc", ctypes.c_int), ("n_tot", ctypes.c_int),
                ("t_cp", ctypes.c_int)]


class OrcSmc2Cfg(ctypes.Structure):
    _fields_ = [("n_theta", ctypes.c_int), ("n_min", ctypes.c_int), ("n_max", ctypes.c_int),
                ("seed", ctypes.c_int), ("kind", ctypes.c_int), ("moves", ctypes.c_int),
                ("nu_obs", ctypes.c_float), ("cn_rho", ctypes.c_float), ("u_step", ctypes.c_float),
                ("prop_scale", ctypes.c_float), ("mala_h", ctypes.c_float),
                ("outer_thr", ctypes.c_float), ("lo", ctypes.c_float * 3), ("hi", ctypes.c_float * 3),
                ("tile", ctypes.c_int)]


class OrcSmc2Count(ctypes.Structure):
    _fields_ = [(k, ctypes.c_longlong) for k in ("ticks", "outer_resamples", "proposed", "accepted",
                                                 "replay_ticks", "pt_forward", "pt_replay")]


# Defaults of smc2/smc2_rbpf_batch.cu:1080-1110 (prior, bounds, curves, proposal std)
RB_PRIOR = (0.95, 0.02, 0.1, 0.1, -1.0, 0.5, 0.15, 0.05)
RB_BOUNDS = (0.8, 0.999, 0.01, 1.0, -10.0, 5.0, 0.01, 1.0)
RB_TCURVE = (0.02, 0.08, 1.5)
RB_FIXED = (0.5, 1.0, 0.1, 1.0)
RB_PSTD = (0.01, 0.02, 0.1, 0.02)


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        u64, u32, f32, f64, i32 = (
            ctypes.c_uint64,
            ctypes.c_uint32,
            ctypes.c_float,
            ctypes.c_double,
            ctypes.c_int,
        )
        P = ctypes.c_void_p
        L.orc_pcg_state.restype = u64
        L.orc_pcg_state.argtypes = [i32, u32, u32]
        L.orc_pcg_state_naive.restype = u64
        L.orc_pcg_state_naive.argtypes = [i32, u32, u32]
        L.orc_pcg_output.restype = u32
        L.orc_pcg_output.argtypes = [u64]
        L.orc_pcg_host_seed.restype = u64
        L.orc_pcg_host_seed.argtypes = [i32]
        L.orc_pcg_host_next.restype = u32
        L.orc_pcg_host_next.argtypes = [ctypes.POINTER(u64)]
        L.orc_uniform.restype = f32
        L.orc_uniform.argtypes = [u32]
        L.orc_expf.restype = f32
        L.orc_expf.argtypes = [f32]
        L.orc_logf.restype = f32
        L.orc_logf.argtypes = [f32]
        L.orc_icdf_f32.restype = f32
        L.orc_icdf_f32.argtypes = [f32]
        L.orc_icdf_f64.restype = f64
        L.orc_icdf_f64.argtypes = [f64]
        L.orc_normal.restype = f32
        L.orc_normal.argtypes = [i32, u32, u32]
        L.orc_dgp.restype = None
        L.orc_dgp.argtypes = [i32, f64, f64, f64, f64, f64, u64, P, P]
        L.orc_dgp_regimes.restype = None
        L.orc_dgp_regimes.argtypes = [i32, f64, f64, P, f64, u64, P, P]
        L.orc_bpf_run.restype = i32
        L.orc_bpf_run.argtypes = [ctypes.POINTER(OrcParams), i32, P, P, i32, P, P, P, P, i32]
        L.orc_bpf_run_learn.restype = i32
        L.orc_bpf_run_learn.argtypes = [ctypes.POINTER(OrcParams), ctypes.POINTER(OrcLearn), P, i32,
                                        P, P, P, P, i32]
        L.orc_bpf_run_bands.restype = i32
        L.orc_bpf_run_bands.argtypes = [ctypes.POINTER(OrcParams), ctypes.POINTER(OrcBands), P, i32,
                                        P, P, P, i32]
        L.orc_bpf_run_ext.restype = i32
        L.orc_bpf_run_ext.argtypes = [ctypes.POINTER(OrcParams), f32, f32, P, i32, P, P, i32]
        L.orc_apf_run.restype = i32
        L.orc_apf_run.argtypes = [ctypes.POINTER(OrcParams), f32, P, i32, P, P, i32]
        L.orc_bpf_run_all.restype = i32
        L.orc_bpf_run_all.argtypes = [ctypes.POINTER(OrcParams), ctypes.POINTER(OrcLearn),
                                      ctypes.POINTER(OrcBands), f32, f32, P, i32, P, P, P, P, i32]
        L.orc_learn_update.restype = None
        L.orc_learn_update.argtypes = [ctypes.POINTER(OrcLearn), P, i32, ctypes.POINTER(i32),
                                       ctypes.POINTER(f32), ctypes.POINTER(f32)]
        L.orc_rb_smc2_run.restype = i32
        L.orc_rb_smc2_run.argtypes = [ctypes.POINTER(OrcRbCfg), P, i32, P, P, P, P, P, P, P,
                                      ctypes.POINTER(OrcRbCount)]
        L.orc_rb_inner_run.restype = i32
        L.orc_rb_inner_run.argtypes = [ctypes.POINTER(OrcRbCfg), P, P, i32, P, P, P]
        L.orc_rb_ocsn.restype = None
        L.orc_rb_ocsn.argtypes = [f32, f32, f32, P]
        L.orc_rb_math.restype = None
        L.orc_rb_math.argtypes = [i32, P, P, i32]
        L.orc_rb_half_roundtrip.restype = None
        L.orc_rb_half_roundtrip.argtypes = [P, P, i32]
        L.orc_rb_ocsn_table.restype = None
        L.orc_rb_ocsn_table.argtypes = [P, P, P]
        L.orc_resample_ancestors.restype = None
        L.orc_resample_ancestors.argtypes = [P, i32, f32, P]
        L.orc_bpf_score.restype = i32
        L.orc_bpf_score.argtypes = [ctypes.POINTER(OrcParams), P, i32, i32, P, P, P, P, P, P]
        L.orc_set_fault.restype = None
        L.orc_set_fault.argtypes = [i32]
        L.orc_smc2_set_fault.restype = None
        L.orc_smc2_set_fault.argtypes = [i32]
        L.orc_sv_smc2_run.restype = i32
        L.orc_sv_smc2_run.argtypes = [ctypes.POINTER(OrcSmc2Cfg), P, i32, P, P, P, P, P, P, P, P, i32,
                                      ctypes.POINTER(OrcSmc2Count)]
        L.orc_set_threads.restype = None
        L.orc_set_threads.argtypes = [i32]
        L.orc_bpf_tick.restype = i32
        L.orc_bpf_tick.argtypes = [ctypes.POINTER(OrcParams), f32, i32, f32, i32, ctypes.POINTER(u64), P, P,
                                   P, P, P, ctypes.POINTER(f64), i32]
        L.orc_smc2_ancestors.restype = i32
        L.orc_smc2_ancestors.argtypes = [P, i32, f64, P]
        L.orc_smc2_selftest.restype = i32
        L.orc_smc2_selftest.argtypes = [i32, i32, i32, P]
        L.orc_sv_inner_run.restype = f64
        L.orc_sv_inner_run.argtypes = [i32, f32, f32, f32, f32, i32, P, i32, P, P]
        L.orc_math_array.restype = None
        L.orc_math_array.argtypes = [i32, P, P, i32]
        L.orc_grid_loglik.restype = f64
        L.orc_grid_loglik.argtypes = [P, i32, f64, f64, f64, i32, P]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def expf(x):
    return lib().orc_expf(float(x))


def logf(x):
    return lib().orc_logf(float(x))


def icdf_f32(u):
    return lib().orc_icdf_f32(float(u))


MATH_ICDF, MATH_EXP, MATH_EXP_NONPOS, MATH_LOG = 0, 1, 2, 3


def math_array(fn, x):
    """Element-wise oracle arithmetic over a float32 array (fn as above)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty_like(x)
    lib().orc_math_array(fn, _ptr(x), _ptr(out), x.size)
    return out


def icdf_f64(u):
    return lib().orc_icdf_f64(float(u))


def pcg_state(seed, i, k):
    return lib().orc_pcg_state(int(seed), int(i), int(k))


def pcg_state_naive(seed, i, k):
    return lib().orc_pcg_state_naive(int(seed), int(i), int(k))


def pcg_output(s):
    return lib().orc_pcg_output(int(s))


def uniform(r):
    return lib().orc_uniform(int(r))


def normal(seed, i, k):
    return lib().orc_normal(int(seed), int(i), int(k))


def host_stream(seed, n):
    st = ctypes.c_uint64(lib().orc_pcg_host_seed(int(seed)))
    return [lib().orc_pcg_host_next(ctypes.byref(st)) for _ in range(n)]


def dgp(n_ticks, rho=0.97, sigma_z=0.15, mu=-1.0, nu_state=0.0, nu_obs=0.0, seed=42):
    h = np.empty(n_ticks, np.float64)
    y = np.empty(n_ticks, np.float64)
    lib().orc_dgp(n_ticks, rho, sigma_z, mu, nu_state, nu_obs, seed, _ptr(h), _ptr(y))
    return h, y


def dgp_regimes(mu_t, rho=0.97, sigma_z=0.15, nu_obs=0.0, seed=42):
    mu_t = np.ascontiguousarray(mu_t, np.float64)
    n = len(mu_t)
    h = np.empty(n, np.float64)
    y = np.empty(n, np.float64)
    lib().orc_dgp_regimes(n, rho, sigma_z, _ptr(mu_t), nu_obs, seed, _ptr(h), _ptr(y))
    return h, y


def run_filter(y, n_particles, rho=0.97, sigma_z=0.15, mu=-1.0, nu_obs=0.0,
               ess_threshold=0.0, seed=42, tile=256, mode=MODE_F32, mu_t=None,
               want_x=False, want_anc=False, threads=1):
    """Run the oracle filter; returns (ticks structured array, resample count, x, anc)."""
    y = np.ascontiguousarray(y, np.float32)
    T = len(y)
    p = OrcParams(n_particles, rho, sigma_z, mu, nu_obs, ess_threshold, seed, tile)
    out = np.zeros(T, TICK_DTYPE)
    x = np.empty(n_particles, np.float32) if (want_x and mode == MODE_F32) else None
    x64 = np.empty(n_particles, np.float64) if (want_x and mode != MODE_F32) else None
    anc = np.empty(n_particles, np.int32) if want_anc else None
    mt = None if mu_t is None else np.ascontiguousarray(mu_t, np.float32)
    cnt = lib().orc_bpf_run(ctypes.byref(p), mode, _ptr(y), _ptr(mt), T, _ptr(out),
                            _ptr(x), _ptr(x64), _ptr(anc), int(threads))
    if cnt < 0:
        raise ValueError("oracle rejected the arguments")
    return out, cnt, (x if x is not None else x64), anc


def run_filter_learn(y, n_particles, rho=0.97, sigma_z=0.15, mu=-1.0, nu_obs=0.0,
                     ess_threshold=0.0, seed=42, tile=256, update_K=50, learn_rho=False,
                     c=0.1, t0=10.0, gamma=0.667, grad_clip=5.0, want_x=False, threads=1):
    """FP32 filter with the online mu/rho learner on from tick 0.
    Returns (ticks, per-tick learner sums [T, 4], (mu, rho) after each tick [T, 2], x)."""
    y = np.ascontiguousarray(y, np.float32)
    T = len(y)
    p = OrcParams(n_particles, rho, sigma_z, mu, nu_obs, ess_threshold, seed, tile)
    lp = OrcLearn(int(update_K), 1 if learn_rho else 0, c, t0, gamma, grad_clip)
    out = np.zeros(T, TICK_DTYPE)
    ls = np.zeros((T, 4), np.float64)
    par = np.zeros((T, 2), np.float32)
    x = np.empty(n_particles, np.float32) if want_x else None
    cnt = lib().orc_bpf_run_learn(ctypes.byref(p), ctypes.byref(lp), _ptr(y), T, _ptr(out),
                                  _ptr(ls), _ptr(par), _ptr(x), int(threads))
    if cnt < 0:
        raise ValueError("oracle rejected the arguments")
    return out, ls, par, x


def run_filter_bands(y, n_particles, configs, adaptive=False, band_n=None, rho=0.97, sigma_z=0.15,
                     mu=-1.0, nu_obs=0.0, ess_threshold=0.0, seed=42, tile=256, thr_alert=2.0,
                     thr_panic=4.0, want_x=False, threads=1):
    """FP32 filter with sigma_z bands.  configs: one (fracs, scales) pair (static) or
    three (calm, alert, panic; adaptive).  band_n: the particle count the
    fractions refer to (default n_particles).  Returns (ticks, regimes[T], x)."""
    y = np.ascontiguousarray(y, np.float32)
    T = len(y)
    p = OrcParams(n_particles, rho, sigma_z, mu, nu_obs, ess_threshold, seed, tile)
    b = OrcBands()
    b.adaptive = 1 if adaptive else 0
    b.n_particles = int(n_particles if band_n is None else band_n)
    cfgs = list(configs) if adaptive else [configs] * 3
    for c, (fr, sc) in enumerate(cfgs):
        b.n_bands[c] = len(sc)
        for q, v in enumerate(fr):
            b.fracs[c][q] = v
        for q, v in enumerate(sc):
            b.scales[c][q] = v
    b.thr_alert = thr_alert
    b.thr_panic = thr_panic
    out = np.zeros(T, TICK_DTYPE)
    reg = np.zeros(T, np.int32)
    x = np.empty(n_particles, np.float32) if want_x else None
    cnt = lib().orc_bpf_run_bands(ctypes.byref(p), ctypes.byref(b), _ptr(y), T, _ptr(out), _ptr(x),
                                  _ptr(reg), int(threads))
    if cnt < 0:
        raise ValueError("oracle rejected the arguments")
    return out, reg, x


def run_filter_ext(y, n_particles, rho=0.97, sigma_z=0.15, mu=-1.0, nu_obs=0.0, nu_state=0.0,
                   silverman=0.0, ess_threshold=0.0, seed=42, tile=256, want_x=False, threads=1):
    """FP32 filter with Student-t state noise and / or Silverman jitter.
    Returns (ticks, resample count, x)."""
    y = np.ascontiguousarray(y, np.float32)
    T = len(y)
    p = OrcParams(n_particles, rho, sigma_z, mu, nu_obs, ess_threshold, seed, tile)
    out = np.zeros(T, TICK_DTYPE)
    x = np.empty(n_particles, np.float32) if want_x else None
    cnt = lib().orc_bpf_run_ext(ctypes.byref(p), float(nu_state), float(silverman), _ptr(y), T,
                                _ptr(out), _ptr(x), int(threads))
    if cnt < 0:
        raise ValueError("oracle rejected the arguments")
    return out, cnt, x


def _bands_struct(configs, adaptive, band_n, thr_alert, thr_panic):
    b = OrcBands()
    b.adaptive = 1 if adaptive else 0
    b.n_particles = int(band_n)
    cfgs = list(configs) if adaptive else [configs] * 3
    for c, (fr, sc) in enumerate(cfgs):
        b.n_bands[c] = len(sc)
        for q, v in enumerate(fr):
            b.fracs[c][q] = v
        for q, v in enumerate(sc):
            b.scales[c][q] = v
    b.thr_alert = thr_alert
    b.thr_panic = thr_panic
    return b


def run_filter_all(y, n_particles, rho=0.97, sigma_z=0.15, mu=-1.0, nu_obs=0.0, ess_threshold=0.5,
                   seed=42, tile=256, learn=None, bands=None, adaptive=False, thr_alert=2.0,
                   thr_panic=4.0, nu_state=0.0, silverman=0.0, want_x=False, threads=1):
    """FP32 filter with any combination of options.  learn: dict(update_K, learn_rho, c, t0,
    gamma) or None; bands: a (fracs, scales) pair, or three of them with adaptive=True.
    Returns (ticks, (mu, rho) after each tick [T, 2], regimes [T], x)."""
    y = np.ascontiguousarray(y, np.float32)
    T = len(y)
    p = OrcParams(n_particles, rho, sigma_z, mu, nu_obs, ess_threshold, seed, tile)
    lp = None
    if learn is not None:
        lp = OrcLearn(int(learn.get("update_K", 50)), 1 if learn.get("learn_rho") else 0,
                      learn.get("c", 0.1), learn.get("t0", 10.0), learn.get("gamma", 0.667), 5.0)
    bd = None if bands is None else _bands_struct(bands, adaptive, n_particles, thr_alert, thr_panic)
    out = np.zeros(T, TICK_DTYPE)
    par = np.zeros((T, 2), np.float32)
    reg = np.zeros(T, np.int32)
    x = np.empty(n_particles, np.float32) if want_x else None
    rc = lib().orc_bpf_run_all(ctypes.byref(p), ctypes.byref(lp) if lp is not None else None,
                               ctypes.byref(bd) if bd is not None else None, float(nu_state),
                               float(silverman), _ptr(y), T, _ptr(out), _ptr(par), _ptr(reg), _ptr(x),
                               int(threads))
    if rc < 0:
        raise ValueError("oracle rejected the arguments")
    return out, par, reg, x


def run_apf(y, n_particles, rho=0.97, sigma_z=0.15, mu=-1.0, nu_obs=0.0, nu_state=0.0, seed=42,
            tile=256, want_x=False, threads=1):
    """FP32 auxiliary particle filter; returns (ticks, x)."""
    y = np.ascontiguousarray(y, np.float32)
    T = len(y)
    p = OrcParams(n_particles, rho, sigma_z, mu, nu_obs, 0.0, seed, tile)
    out = np.zeros(T, TICK_DTYPE)
    x = np.empty(n_particles, np.float32) if want_x else None
    if lib().orc_apf_run(ctypes.byref(p), float(nu_state), _ptr(y), T, _ptr(out), _ptr(x),
                         int(threads)) < 0:
        raise ValueError("oracle rejected the arguments")
    return out, x


def learn_update(sums, k, rm_step, mu, rho, learn_rho=True, c=0.1, t0=10.0, gamma=0.667,
                 grad_clip=5.0):
    """One learner update; returns (rm_step, mu, rho) after it."""
    lp = OrcLearn(1, 1 if learn_rho else 0, c, t0, gamma, grad_clip)
    sm = np.ascontiguousarray(sums, np.float64)
    st = ctypes.c_int(int(rm_step))
    m = ctypes.c_float(mu)
    r = ctypes.c_float(rho)
    lib().orc_learn_update(ctypes.byref(lp), _ptr(sm), int(k), ctypes.byref(st),
                           ctypes.byref(m), ctypes.byref(r))
    return st.value, m.value, r.value


def score(y, n_particles, rho=0.97, sigma_z=0.15, mu=-1.0, nu_obs=0.0, seed=42, tile=256,
          x0_reparam=True, theta64=None, anc_in=None, want_anc=False, threads=1):
    """Pathwise score of the always-resample filter with frozen ancestors.
    Returns (ll64, score[3], ll32, anc[T,N] or None)."""
    y = np.ascontiguousarray(y, np.float32)
    T = len(y)
    p = OrcParams(n_particles, rho, sigma_z, mu, nu_obs, 0.0, seed, tile)
    th = None if theta64 is None else np.ascontiguousarray(theta64, np.float64)
    ai = None if anc_in is None else np.ascontiguousarray(anc_in, np.int32)
    ao = np.empty((T, n_particles), np.int32) if want_anc else None
    ll = ctypes.c_double()
    ll32 = ctypes.c_double()
    sc = np.zeros(3, np.float64)
    lib().orc_set_threads(int(threads))
    try:
        rc = lib().orc_bpf_score(ctypes.byref(p), _ptr(y), T, 1 if x0_reparam else 0, _ptr(th),
                                 _ptr(ai), _ptr(ao), ctypes.byref(ll), _ptr(sc), ctypes.byref(ll32))
    finally:
        lib().orc_set_threads(1)
    if rc < 0:
        raise ValueError("oracle rejected the arguments")
    return ll.value, sc, ll32.value, ao


def resample_ancestors(cdf, u):
    cdf = np.ascontiguousarray(cdf, np.float64)
    anc = np.empty(len(cdf), np.int32)
    lib().orc_resample_ancestors(_ptr(cdf), len(cdf), float(u), _ptr(anc))
    return anc


class planted_fault:
    """Context manager planting mistake `code` in the oracle (see bpf_oracle.h)."""

    def __init__(self, code):
        self.code = code

    def __enter__(self):
        lib().orc_set_fault(self.code)

    def __exit__(self, *exc):
        lib().orc_set_fault(0)


def grid_loglik(y, rho=0.97, sigma_z=0.15, mu=-1.0, n_grid=2048):
    y = np.ascontiguousarray(y, np.float32)
    llt = np.empty(len(y), np.float64)
    tot = lib().orc_grid_loglik(_ptr(y), len(y), rho, sigma_z, mu, n_grid, _ptr(llt))
    return tot, llt


# ---- SMC^2 with the RBPF / OCSN inner filter (rbpf_oracle.c) -------------

def rb_cfg(n_theta, n_inner, seed=12345, K_rejuv=5, lag=0, cap=None, ess_thr=0.5, crho=0.99,
           adaptive=True, pstd=RB_PSTD, prior=RB_PRIOR, bounds=RB_BOUNDS, tcurve=RB_TCURVE,
           fixed=RB_FIXED, T=0, noise_bits=32):
    c = OrcRbCfg()
    c.noise_bits = noise_bits
    c.n_theta, c.n_inner, c.K_rejuv, c.lag = n_theta, n_inner, K_rejuv, lag
    c.cap = cap if cap is not None else max(2048, T)
    c.adaptive = 1 if adaptive else 0
    c.ess_thr, c.crho = ess_thr, crho
    c.seed = seed if seed else 12345
    c.pstd[:] = list(pstd)
    c.prior[:] = list(prior)
    c.bounds[:] = list(bounds)
    c.tcurve[:] = list(tcurve)
    c.fixed[:] = list(fixed)
    return c


def run_rb_smc2(cfg, y):
    """Whole SMC^2 run; returns a dict of the final state and per-tick ESS."""
    L = lib()
    y = np.ascontiguousarray(y, np.float32)
    T, nt, ni = len(y), cfg.n_theta, cfg.n_inner
    ess = np.zeros(T, np.float32)
    th = np.zeros((4, nt), np.float32)
    ll = np.zeros(nt, np.float64)
    lw = np.zeros(nt, np.float64)
    essi = np.zeros(nt, np.float32)
    inner = np.zeros((4, nt * ni), np.float32)
    rt = np.zeros(T + 1, np.int32)
    cnt = OrcRbCount()
    r = L.orc_rb_smc2_run(ctypes.byref(cfg), _ptr(y), T, _ptr(ess), _ptr(th), _ptr(ll), _ptr(lw),
                          _ptr(essi), _ptr(inner), _ptr(rt), ctypes.byref(cnt))
    if r != 0:
        raise ValueError("orc_rb_smc2_run rejected its arguments")
    return dict(ess=ess, theta=th, ll=ll, lw=lw, ess_inner=essi, inner=inner,
                resample_ticks=rt[:cnt.n_resamples].copy(), n_resamples=cnt.n_resamples,
                n_acc=cnt.n_acc, n_tot=cnt.n_tot, t_cp=cnt.t_cp)


def rb_inner_run(cfg, theta, y):
    """Inner RBPF of one theta-particle (index 0, first epoch), theta fixed."""
    L = lib()
    y = np.ascontiguousarray(y, np.float32)
    T, ni = len(y), cfg.n_inner
    th = np.ascontiguousarray(theta, np.float32)
    lli = np.zeros(T, np.float64)
    ess = np.zeros(T, np.float64)
    st = np.zeros((4, ni), np.float32)
    if L.orc_rb_inner_run(ctypes.byref(cfg), _ptr(th), _ptr(y), T, _ptr(lli), _ptr(ess), _ptr(st)) != 0:
        raise ValueError("orc_rb_inner_run rejected its arguments")
    return lli, ess, st


def rb_ocsn(y, mp, vp):
    out = np.zeros(3, np.float32)
    lib().orc_rb_ocsn(float(y), float(mp), float(vp), _ptr(out))
    return out


def rb_math(fn, x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros_like(x)
    lib().orc_rb_math(fn, _ptr(x), _ptr(out), len(x))
    return out


def rb_ocsn_table():
    p, m, v = (np.zeros(10, np.float32) for _ in range(3))
    lib().orc_rb_ocsn_table(_ptr(p), _ptr(m), _ptr(v))
    return p, m, v


def rb_dgp(T, rho=0.95, sigma_z=0.10, mu_base=-1.0, sigma_base=0.15, fixed=RB_FIXED,
           tcurve=RB_TCURVE, seed=12345):
    """Regime-switching SV generator of smc2/test_smc2_rbpf.cu:95-133 (FP64,
    NumPy normals): returns y = log r^2 (floored at log 1e-30), h and z."""
    rng = np.random.default_rng(seed)
    ms, mr, ss, sr = fixed
    tb, ts, tr = tcurve

    def curve(b, s, r, z):
        return b + s * (1.0 - np.exp(-r * z))

    zt = sigma_z / np.sqrt(max(1.0 - rho * rho, 1e-6)) * rng.standard_normal()
    z = 1.5 * (1.0 + np.tanh(zt))
    th, mu, sh = curve(tb, ts, tr, z), curve(mu_base, ms, mr, z), curve(sigma_base, ss, sr, z)
    phi = 1.0 - th
    h = mu + np.sqrt(sh * sh / max(1.0 - phi * phi, 1e-6)) * rng.standard_normal()
    y, hs, zs = np.zeros(T), np.zeros(T), np.zeros(T)
    for t in range(T):
        zt = rho * zt + sigma_z * rng.standard_normal()
        z = 1.5 * (1.0 + np.tanh(zt))
        th, mu, sh = curve(tb, ts, tr, z), curve(mu_base, ms, mr, z), curve(sigma_base, ss, sr, z)
        h = (1.0 - th) * h + th * mu + sh * rng.standard_normal()
        r2 = max((np.exp(h / 2.0) * rng.standard_normal()) ** 2, 1e-30)
        y[t], hs[t], zs[t] = np.log(r2), h, z
    return y.astype(np.float32), hs, zs


def rb_half_roundtrip(x):
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    lib().orc_rb_half_roundtrip(_ptr(x), _ptr(out), len(x))
    return out


def run_sv_smc2(y, n_theta, n_min, n_max, seed=42, kind="cpmmh", moves=5, nu_obs=0.0, cn_rho=0.99,
                u_step=0.05, theta_scale=1.0, mala_step=0.5, outer_threshold=0.5,
                prior=((-2.0, 0.90, 0.05), (0.0, 0.995, 0.40)), fault=0, tile=256):
    """SMC^2 over batched SV bootstrap filters (smc2_oracle.c), with the
    arguments of gpu_bpf_batch.GpuSmc2 (tile: particles per inner tile).  Returns a dict: theta [M, 3], n [M],
    logw [M], ll [M], score [M, 3], ess [T], history (ticks, sum_n), counters."""
    y = np.ascontiguousarray(y, np.float32)
    T, M = len(y), int(n_theta)
    c = OrcSmc2Cfg()
    c.n_theta, c.n_min, c.n_max, c.seed = M, int(n_min), int(n_max), int(seed)
    c.kind = 1 if kind == "mala" else 0
    c.moves = int(moves)
    c.nu_obs, c.cn_rho, c.u_step = float(nu_obs), float(cn_rho), float(u_step)
    c.prop_scale, c.mala_h, c.outer_thr = float(theta_scale), float(mala_step), float(outer_threshold)
    for k in range(3):
        c.lo[k], c.hi[k] = float(prior[0][k]), float(prior[1][k])
    c.tile = int(tile)
    th = np.zeros((M, 3), np.float32)
    n = np.zeros(M, np.int32)
    lw = np.zeros(M, np.float64)
    ll = np.zeros(M, np.float64)
    sc = np.zeros((M, 3), np.float64)
    ess = np.zeros(max(T, 1), np.float64)
    cap = max(T, 1)
    ht = np.zeros(cap, np.int32)
    hs = np.zeros(cap, np.int64)
    cnt = OrcSmc2Count()
    L = lib()
    L.orc_smc2_set_fault(int(fault))
    try:
        k = L.orc_sv_smc2_run(ctypes.byref(c), _ptr(y), T, _ptr(th), _ptr(n), _ptr(lw), _ptr(ll), _ptr(sc),
                              _ptr(ess), _ptr(ht), _ptr(hs), cap, ctypes.byref(cnt))
    finally:
        L.orc_smc2_set_fault(0)
    if k < 0:
        raise RuntimeError("orc_sv_smc2_run failed")
    return {"theta": th, "n": n, "logw": lw, "ll": ll, "score": sc, "ess": ess[:T],
            "history": (ht[:k], hs[:k]),
            "counters": {f: getattr(cnt, f) for f, _ in OrcSmc2Count._fields_}}


def sv_inner_run(y, n, mu, phi, sig, nu_obs=0.0, seed=42):
    """One batched inner filter (tiles of 256, every tick resampled, PCG32
    noise) with its FP32 tangents: (log-likelihood, pathwise score[3], x)."""
    y = np.ascontiguousarray(y, np.float32)
    sc = np.zeros(3, np.float64)
    x = np.zeros(int(n), np.float32)
    ll = lib().orc_sv_inner_run(int(n), float(mu), float(phi), float(sig), float(nu_obs), int(seed), _ptr(y),
                                len(y), _ptr(sc), _ptr(x))
    return ll, sc, x


def smc2_ancestors(log_w, u):
    lw = np.ascontiguousarray(log_w, np.float64)
    anc = np.zeros(lw.size, np.int32)
    if lib().orc_smc2_ancestors(_ptr(lw), lw.size, float(u), _ptr(anc)) < 0:
        raise ValueError("no finite outer log-weight")
    return anc


def smc2_selftest(kind, iters, seed=1, fault=0):
    """The oracle's move code on closed-form targets: kind 'rw' / 'mala' give
    (mean[3], cov[3, 3], acceptance) of a chain on N(m, C); kind 'cn' gives
    (mean, variance, lag-1 correlation) of the Crank-Nicolson noise chain."""
    out = np.zeros(13, np.float64)
    L = lib()
    L.orc_smc2_set_fault(int(fault))
    try:
        rc = L.orc_smc2_selftest({"rw": 0, "mala": 1, "cn": 2}[kind], int(iters), int(seed), _ptr(out))
    finally:
        L.orc_smc2_set_fault(0)
    if rc < 0:
        raise ValueError("bad selftest arguments")
    if kind == "cn":
        return out[0], out[1], out[2]
    return out[:3], out[3:12].reshape(3, 3), out[12]


def filter_tick(x, l, t, y, carry, hrng, n_particles, mu, rho=0.97, sigma_z=0.15, nu_obs=0.0,
                ess_threshold=0.5, seed=42, tile=4096, threads=1):
    """One tick of the FP32 restatement from the state (x, l, carry, host
    stream hrng).  x and l are updated in place; returns (stats record,
    resampled, x_pre, anc, FP64 ll_incr, new hrng)."""
    n = int(n_particles)
    assert x.dtype == np.float32 and l.dtype == np.float32 and x.size == n and l.size == n
    p = OrcParams(n, rho, sigma_z, mu, nu_obs, ess_threshold, seed, tile)
    x_pre = np.empty(n, np.float32)
    anc = np.empty(n, np.int32)
    out = np.zeros(1, TICK_DTYPE)
    h = ctypes.c_uint64(int(hrng))
    ll64 = ctypes.c_double()
    rc = lib().orc_bpf_tick(ctypes.byref(p), float(mu), int(t), float(y), int(carry), ctypes.byref(h), _ptr(x),
                            _ptr(l), _ptr(x_pre), _ptr(anc), _ptr(out), ctypes.byref(ll64), int(threads))
    if rc < 0:
        raise ValueError("oracle rejected the arguments")
    return out[0], bool(rc), x_pre, anc, ll64.value, h.value
