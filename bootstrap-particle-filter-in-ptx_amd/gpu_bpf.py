"""Python mirror of the filter's C ABI (include/gpu_bpf.h) over ctypes.

The class methods follow the reference's C API one for one
(gpu_bpf/gpu_bpf_full.cuh:144-301): create/step/step_async/get_result/destroy,
set_ess_threshold, get_resample_count, get/set mu and rho, run_rmse, plus the
build-owned batch and inspection entry points.  The compute path is the HIP
library _build/libgpubpf.so; if it is missing this module raises instead of
falling back to anything else.
"""
import ctypes
import os
import re

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
REPO_DIR = os.path.dirname(PKG_DIR)
# largest particle count of gpu_bpf_run_device_noalloc (GPU_BPF_NOALLOC_MAX_N)
NOALLOC_MAX_N = 4194304
LIB_PATH = os.path.join(PKG_DIR, "_build", "libgpubpf.so")
HEADER_PATH = os.path.join(REPO_DIR, "include", "gpu_bpf.h")


class BpfResult(ctypes.Structure):
    _fields_ = [("h_mean", ctypes.c_float), ("log_lik", ctypes.c_float)]


STATS_DTYPE = np.dtype(
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
RESULT_DTYPE = np.dtype([("h_mean", "f4"), ("log_lik", "f4")])


class ImmModelParams(ctypes.Structure):
    _fields_ = [("rho", ctypes.c_float), ("sigma_z", ctypes.c_float), ("mu", ctypes.c_float),
                ("nu_state", ctypes.c_float), ("nu_obs", ctypes.c_float)]


class ImmResult(ctypes.Structure):
    _fields_ = [("h_mean", ctypes.c_float), ("vol", ctypes.c_float), ("log_lik", ctypes.c_float),
                ("best_model", ctypes.c_int), ("best_prob", ctypes.c_float)]


class GpuBpfError(RuntimeError):
    pass


_lib = None


def header_symbols(path=HEADER_PATH):
    """Function names declared in a boundary header (default include/gpu_bpf.h)."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b((?:gpu_bpf|gpu_smc2|gpu_apf|gpu_imm|smc2_cuda)_[a-z_0-9]+)\s*\(", text)))


def load(path=LIB_PATH):
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise GpuBpfError(
            f"{path} is missing: build it with `make -C {PKG_DIR}/csrc` "
            "(or __graft_entry__.build()); there is no CPU fallback")
    L = ctypes.CDLL(path)
    P, F, I, D = ctypes.c_void_p, ctypes.c_float, ctypes.c_int, ctypes.c_double
    sig = {
        "gpu_bpf_create": (P, [I, F, F, F, F, F, I]),
        "gpu_bpf_destroy": (None, [P]),
        "gpu_bpf_step": (BpfResult, [P, F]),
        "gpu_bpf_step_async": (None, [P, F]),
        "gpu_bpf_get_result": (BpfResult, [P]),
        "gpu_bpf_run_rmse": (D, [P, P, I, I, F, F, F, F, F, I]),
        "gpu_bpf_set_ess_threshold": (None, [P, F]),
        "gpu_bpf_get_resample_count": (I, [P]),
        "gpu_bpf_get_mu": (F, [P]),
        "gpu_bpf_get_rho": (F, [P]),
        "gpu_bpf_set_mu": (None, [P, F]),
        "gpu_bpf_set_rho": (None, [P, F]),
        "gpu_bpf_get_learn_sums": (I, [P, P, P]),
        "gpu_bpf_get_sigma_z": (F, [P]),
        "gpu_bpf_set_sigma_z": (None, [P, F]),
        "gpu_bpf_enable_mu_learning": (None, [P, I, I, F, F, F]),
        "gpu_bpf_disable_mu_learning": (None, [P]),
        "gpu_bpf_enable_rho_learning": (None, [P, I]),
        "gpu_bpf_set_bands": (None, [I, I, P, P]),
        "gpu_bpf_set_adaptive_bands": (None, [I, P, P, I, P, P, I, P, P, I, F, F]),
        "gpu_bpf_get_band_regime": (I, [P]),
        "gpu_bpf_set_silverman_shrink": (None, [P, F]),
        "gpu_apf_create": (P, [I, F, F, F, F, F, I]),
        "gpu_apf_destroy": (None, [P]),
        "gpu_apf_step": (BpfResult, [P, F]),
        "gpu_apf_run_rmse": (D, [P, P, I, I, F, F, F, F, F, I]),
        "gpu_apf_run": (I, [P, P, I, P, P]),
        "gpu_apf_copy_particles": (I, [P, P]),
        "gpu_apf_get_layout": (I, [P, P, P]),
        "gpu_imm_create": (P, [P, I, I, P, I]),
        "gpu_imm_destroy": (None, [P]),
        "gpu_imm_step": (ImmResult, [P, F]),
        "gpu_imm_get_prob": (F, [P, I]),
        "gpu_imm_get_probs": (None, [P, P]),
        "gpu_imm_build_grid": (P, [P, I, P, I, P, I, F, F, P]),
        "gpu_imm_uses_batch": (I, [P]),
        "gpu_bpf_last_error": (ctypes.c_char_p, []),
        "gpu_bpf_run": (I, [P, P, I, P, P]),
        "gpu_bpf_run_device": (I, [P, P, I]),
        "gpu_bpf_copy_stats": (I, [P, P, I]),
        "gpu_bpf_get_stats": (I, [P, P]),
        "gpu_bpf_copy_particles": (I, [P, P, P]),
        "gpu_bpf_get_layout": (I, [P, ctypes.POINTER(I), ctypes.POINTER(I)]),
        "gpu_bpf_set_items_per_thread": (I, [P, I]),
        "gpu_bpf_synchronize": (I, [P]),
        "gpu_bpf_get_timestep": (I, [P]),
        "gpu_bpf_restart": (I, [P]),
        "gpu_bpf_device_upload": (P, [P, I]),
        "gpu_bpf_device_free": (None, [P]),
        "gpu_bpf_device_download": (I, [P, P, I]),
        "gpu_bpf_run_device_noalloc": (I, [I, F, F, F, F, I, F, P, I, P, P, P]),
        "gpu_bpf_set_stats_capacity": (I, [P, I]),
        "gpu_bpf_enable_kernel_timing": (I, [P, I]),
        "gpu_bpf_score_forward": (I, [P, P, I, I, ctypes.POINTER(D)]),
        "gpu_bpf_kernel_tick_times": (I, [P, P, I]),
        "gpu_bpf_score_backward": (I, [P, I, P]),
        "gpu_bpf_set_score_estimator": (I, [P, I]),
        "gpu_bpf_kernel_times": (I, [P, P, P, I]),
        "gpu_bpf_eval_device_math": (I, [I, P, P, I]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def last_error():
    return load().gpu_bpf_last_error().decode()


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


class GpuBpf:
    """One filter state (gpu_bpf_create ... gpu_bpf_destroy)."""

    def __init__(self, n_particles, rho, sigma_z, mu, nu_state=0.0, nu_obs=0.0, seed=42,
                 ess_threshold=None, items_per_thread=None):
        self.L = load()
        self.n = int(n_particles)
        self.h = self.L.gpu_bpf_create(self.n, rho, sigma_z, mu, nu_state, nu_obs, int(seed))
        if not self.h:
            raise GpuBpfError("gpu_bpf_create failed: " + last_error())
        if ess_threshold is not None:
            self.set_ess_threshold(ess_threshold)
        if items_per_thread is not None:
            self._check(self.L.gpu_bpf_set_items_per_thread(self.h, int(items_per_thread)))

    def _check(self, rc):
        if rc < 0:
            raise GpuBpfError(last_error())
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.L.gpu_bpf_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    # reference API
    def step(self, y):
        r = self.L.gpu_bpf_step(self.h, float(y))
        return r.h_mean, r.log_lik

    def step_async(self, y):
        self.L.gpu_bpf_step_async(self.h, float(y))

    def get_result(self):
        r = self.L.gpu_bpf_get_result(self.h)
        return r.h_mean, r.log_lik

    def set_ess_threshold(self, thr):
        self.L.gpu_bpf_set_ess_threshold(self.h, float(thr))

    def resample_count(self):
        return self._check(self.L.gpu_bpf_get_resample_count(self.h))

    def set_mu(self, mu):
        self.L.gpu_bpf_set_mu(self.h, float(mu))

    def set_rho(self, rho):
        self.L.gpu_bpf_set_rho(self.h, float(rho))

    def get_mu(self):
        return self.L.gpu_bpf_get_mu(self.h)

    def get_rho(self):
        return self.L.gpu_bpf_get_rho(self.h)

    def enable_mu_learning(self, mode=1, K=50, c=0.1, t0=10.0, gamma=0.667):
        self.L.gpu_bpf_enable_mu_learning(self.h, int(mode), int(K), float(c), float(t0), float(gamma))

    def disable_mu_learning(self):
        self.L.gpu_bpf_disable_mu_learning(self.h)

    def enable_rho_learning(self, enable=True):
        self.L.gpu_bpf_enable_rho_learning(self.h, int(enable))

    def set_silverman_shrink(self, shrink):
        self.L.gpu_bpf_set_silverman_shrink(self.h, float(shrink))

    def band_regime(self):
        return self._check(self.L.gpu_bpf_get_band_regime(self.h))

    def learn_sums(self):
        """(sums[4], ticks): the learner's accumulators since its last update."""
        out = np.zeros(4, np.float64)
        n = ctypes.c_int(0)
        self._check(self.L.gpu_bpf_get_learn_sums(self.h, _ptr(out), ctypes.byref(n)))
        return out, n.value

    # build-owned
    def run(self, y, stats=True):
        y = np.ascontiguousarray(y, np.float32)
        res = np.zeros(len(y), RESULT_DTYPE)
        st = np.zeros(len(y), STATS_DTYPE) if stats else None
        self._check(self.L.gpu_bpf_run(self.h, _ptr(y), len(y), _ptr(res), _ptr(st)))
        return (res, st) if stats else res

    def stats(self):
        st = np.zeros(1, STATS_DTYPE)
        self._check(self.L.gpu_bpf_get_stats(self.h, _ptr(st)))
        return st[0]

    def particles(self, with_logw=False):
        x = np.empty(self.n, np.float32)
        lw = np.empty(self.n, np.float32) if with_logw else None
        self._check(self.L.gpu_bpf_copy_particles(self.h, _ptr(x), _ptr(lw)))
        return (x, lw) if with_logw else x

    def layout(self):
        t, g = ctypes.c_int(), ctypes.c_int()
        self._check(self.L.gpu_bpf_get_layout(self.h, ctypes.byref(t), ctypes.byref(g)))
        return t.value, g.value

    def synchronize(self):
        self._check(self.L.gpu_bpf_synchronize(self.h))

    def timestep(self):
        return self.L.gpu_bpf_get_timestep(self.h)

    def set_stats_capacity(self, ticks):
        self._check(self.L.gpu_bpf_set_stats_capacity(self.h, int(ticks)))

    def stats_tail(self, n):
        """Statistics of the last n ticks (n <= the ring capacity)."""
        st = np.zeros(n, STATS_DTYPE)
        self._check(self.L.gpu_bpf_copy_stats(self.h, _ptr(st), int(n)))
        return st

    def restart(self):
        """Back to tick 0 with the current parameters (no allocation)."""
        self._check(self.L.gpu_bpf_restart(self.h))

    def set_mu_schedule_run(self, y, mu_t, stats=False):
        """Run y with mu changed (set_mu) wherever mu_t changes value."""
        y = np.ascontiguousarray(y, np.float32)
        mu_t = np.asarray(mu_t, np.float32)
        cuts = [0] + [t for t in range(1, len(y)) if mu_t[t] != mu_t[t - 1]] + [len(y)]
        out = []
        for a, b in zip(cuts[:-1], cuts[1:]):
            self.set_mu(float(mu_t[a]))
            if stats:
                out.append(self.run(y[a:b])[1])
            else:
                yd = y[a:b]
                self._check(self.L.gpu_bpf_run(self.h, _ptr(yd), len(yd), None, None))
        return np.concatenate(out) if stats else None

    def run_device(self, y_dev_ptr, n_ticks):
        """Enqueue n_ticks ticks on observations already in device memory."""
        self._check(self.L.gpu_bpf_run_device(self.h, ctypes.c_void_p(int(y_dev_ptr)), int(n_ticks)))

    def enqueue(self, y):
        """Enqueue len(y) ticks without waiting (observations staged once)."""
        y = np.ascontiguousarray(y, np.float32)
        self._y_keep = y
        self._check(self.L.gpu_bpf_run(self.h, _ptr(y), len(y), None, None))

    def set_score_estimator(self, estimator):
        """'pathwise' (default) or 'fisher' for the next score_forward."""
        code = {"pathwise": 0, "fisher": 1}[estimator]
        self._check(self.L.gpu_bpf_set_score_estimator(self.h, code))

    def score_forward(self, y, checkpoint_every=64):
        """Forward pass of the score (checkpointed for the pathwise
        estimator, complete for the Fisher-identity one); returns LL (FP64)."""
        y = np.ascontiguousarray(y, np.float32)
        ll = ctypes.c_double()
        self._check(self.L.gpu_bpf_score_forward(self.h, _ptr(y), len(y), int(checkpoint_every),
                                                 ctypes.byref(ll)))
        return ll.value

    def score_backward(self, x0_reparam=True):
        """Reverse sweep (pathwise) or the forward pass's result (Fisher);
        returns dLL/d(mu, phi, sigma_z)."""
        g = np.zeros(3, np.float64)
        self._check(self.L.gpu_bpf_score_backward(self.h, 1 if x0_reparam else 0, _ptr(g)))
        return g

    def enable_kernel_timing(self, on=True, stride=1):
        """Event timing of the three kernels of every stride-th tick."""
        self._check(self.L.gpu_bpf_enable_kernel_timing(self.h, int(stride) if on else 0))

    def kernel_tick_times(self, n_ticks):
        """[n, 3] per-tick milliseconds (propagate, finalise, resample)."""
        ms = np.zeros((n_ticks, 3), np.float32)
        got = self._check(self.L.gpu_bpf_kernel_tick_times(self.h, _ptr(ms), n_ticks))
        return ms[:got]

    def kernel_times(self, reset=True):
        """(ms[3], launches[3]) for propagate, finalise, resample."""
        ms = np.zeros(3, np.float64)
        cnt = np.zeros(3, np.int32)
        self._check(self.L.gpu_bpf_kernel_times(self.h, _ptr(ms), _ptr(cnt), 1 if reset else 0))
        return ms, cnt


class DeviceArray:
    """Floats copied to device memory by the library (gpu_bpf_device_upload);
    .ptr is the device address for the *_run_device entry points."""

    def __init__(self, host):
        self.L = load()
        a = np.ascontiguousarray(host, np.float32)
        self.n = a.size
        self.ptr = self.L.gpu_bpf_device_upload(_ptr(a), self.n)
        if not self.ptr:
            raise GpuBpfError(last_error())

    @classmethod
    def zeros(cls, n):
        return cls(np.zeros(int(n), np.float32))

    def download(self):
        out = np.empty(self.n, np.float32)
        if self.L.gpu_bpf_device_download(self.ptr, _ptr(out), self.n) < 0:
            raise GpuBpfError(last_error())
        return out

    def close(self):
        if getattr(self, "ptr", None):
            self.L.gpu_bpf_device_free(self.ptr)
            self.ptr = None

    __del__ = close


class GpuApf:
    """Auxiliary particle filter (gpu_apf_create ... gpu_apf_destroy)."""

    def __init__(self, n_particles, rho, sigma_z, mu, nu_state=0.0, nu_obs=0.0, seed=42):
        self.L = load()
        self.n = int(n_particles)
        self.h = self.L.gpu_apf_create(self.n, rho, sigma_z, mu, nu_state, nu_obs, int(seed))
        if not self.h:
            raise GpuBpfError(last_error())

    def close(self):
        if self.h:
            self.L.gpu_apf_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _check(self, rc):
        if rc < 0:
            raise GpuBpfError(last_error())
        return rc

    def step(self, y):
        r = self.L.gpu_apf_step(self.h, float(y))
        return r.h_mean, r.log_lik

    def run(self, y):
        y = np.ascontiguousarray(y, np.float32)
        res = np.zeros(len(y), RESULT_DTYPE)
        st = np.zeros(len(y), STATS_DTYPE)
        self._check(self.L.gpu_apf_run(self.h, _ptr(y), len(y), _ptr(res), _ptr(st)))
        return res, st

    def particles(self):
        x = np.empty(self.n, np.float32)
        self._check(self.L.gpu_apf_copy_particles(self.h, _ptr(x)))
        return x

    def layout(self):
        t, g = ctypes.c_int(), ctypes.c_int()
        self._check(self.L.gpu_apf_get_layout(self.h, ctypes.byref(t), ctypes.byref(g)))
        return t.value, g.value


class GpuImm:
    """Interacting multiple models (gpu_imm_create ... gpu_imm_destroy).
    models: sequence of (rho, sigma_z, mu, nu_state, nu_obs)."""

    def __init__(self, models, n_particles, transition=None, seed=42):
        self.L = load()
        self.K = len(models)
        arr = (ImmModelParams * self.K)(*[ImmModelParams(*map(float, m)) for m in models])
        tm = None if transition is None else np.ascontiguousarray(transition, np.float32)
        self.h = self.L.gpu_imm_create(arr, self.K, int(n_particles), _ptr(tm), int(seed))
        if not self.h:
            raise GpuBpfError(last_error())

    def close(self):
        if self.h:
            self.L.gpu_imm_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def step(self, y):
        r = self.L.gpu_imm_step(self.h, float(y))
        return r.h_mean, r.vol, r.log_lik, r.best_model, r.best_prob

    def probs(self):
        p = np.zeros(self.K, np.float32)
        self.L.gpu_imm_get_probs(self.h, _ptr(p))
        return p

    def prob(self, k):
        return self.L.gpu_imm_get_prob(self.h, int(k))

    def uses_batch(self):
        return bool(self.L.gpu_imm_uses_batch(self.h))


def imm_build_grid(rhos, sigma_zs, mus, nu_state=0.0, nu_obs=0.0):
    """gpu_imm_build_grid as a list of (rho, sigma_z, mu, nu_state, nu_obs)."""
    L = load()
    r = np.ascontiguousarray(rhos, np.float32)
    s = np.ascontiguousarray(sigma_zs, np.float32)
    m = np.ascontiguousarray(mus, np.float32)
    n = ctypes.c_int(0)
    ptr = L.gpu_imm_build_grid(_ptr(r), len(r), _ptr(s), len(s), _ptr(m), len(m), float(nu_state),
                               float(nu_obs), ctypes.byref(n))
    if not ptr:
        raise GpuBpfError(last_error())
    arr = ctypes.cast(ptr, ctypes.POINTER(ImmModelParams * n.value)).contents
    out = [(a.rho, a.sigma_z, a.mu, a.nu_state, a.nu_obs) for a in arr]
    ctypes.CDLL(None).free(ctypes.c_void_p(ptr))
    return out


def run_rmse_apf(returns, true_h, n_particles, rho, sigma_z, mu, nu_state=0.0, nu_obs=0.0, seed=42):
    r = np.ascontiguousarray(returns, np.float64)
    h = np.ascontiguousarray(true_h, np.float64)
    return load().gpu_apf_run_rmse(_ptr(r), _ptr(h), len(r), int(n_particles), rho, sigma_z, mu,
                                   nu_state, nu_obs, int(seed))


def set_bands(n_particles, fracs, scales):
    """gpu_bpf_set_bands: process-wide sigma_z scale bands (adopted at each filter's next tick)."""
    f = np.ascontiguousarray(fracs, np.float32)
    sc = np.ascontiguousarray(scales, np.float32)
    load().gpu_bpf_set_bands(int(n_particles), len(sc), _ptr(f), _ptr(sc))


def set_adaptive_bands(n_particles, calm, alert, panic, thresh_alert=2.0, thresh_panic=4.0):
    """gpu_bpf_set_adaptive_bands; calm / alert / panic are (fracs, scales) pairs."""
    arrs = []
    args = [int(n_particles)]
    for fr, sc in (calm, alert, panic):
        f = np.ascontiguousarray(fr, np.float32)
        c = np.ascontiguousarray(sc, np.float32)
        arrs += [f, c]
        args += [_ptr(f), _ptr(c), len(c)]
    load().gpu_bpf_set_adaptive_bands(*args, float(thresh_alert), float(thresh_panic))


def reset_bands(n_particles=1):
    """One band of scale 1 (the default)."""
    set_bands(n_particles, [1.0], [1.0])


def run_rmse(returns, true_h, n_particles, rho, sigma_z, mu, nu_state=0.0, nu_obs=0.0, seed=42):
    r = np.ascontiguousarray(returns, np.float64)
    h = np.ascontiguousarray(true_h, np.float64)
    return load().gpu_bpf_run_rmse(_ptr(r), _ptr(h), len(r), int(n_particles), rho, sigma_z,
                                   mu, nu_state, nu_obs, int(seed))


MATH_ICDF, MATH_EXP, MATH_EXP_NONPOS, MATH_LOG = 0, 1, 2, 3


def eval_device_math(fn, x):
    """The device's per-particle arithmetic evaluated element-wise on the GPU
    (fn: MATH_ICDF, MATH_EXP, MATH_EXP_NONPOS, MATH_LOG)."""
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty_like(x)
    if load().gpu_bpf_eval_device_math(int(fn), _ptr(x), _ptr(out), x.size) != 0:
        raise GpuBpfError(last_error())
    return out
