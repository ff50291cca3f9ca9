"""Python mirror of the batched inner filters' C ABI
(include/gpu_bpf_batch.h) over ctypes.

A batch runs M SV bootstrap filters of uneven particle counts (at most
16384 each) in one launch per tick; filter m behaves as the plain filter
created with the same parameters, particle count and seed.  The compute path
is the HIP library; without it this module raises.
"""
import ctypes

import numpy as np

import gpu_bpf

BATCH_HEADER_PATH = gpu_bpf.os.path.join(gpu_bpf.REPO_DIR, "include", "gpu_bpf_batch.h")
MAX_N = 16384

_declared = False


def _lib():
    L = gpu_bpf.load()
    global _declared
    if not _declared:
        P, F, I = ctypes.c_void_p, ctypes.c_float, ctypes.c_int
        sig = {
            "gpu_bpf_batch_create": (P, [I, P, P, P, P, P, F, F]),
            "gpu_bpf_batch_destroy": (None, [P]),
            "gpu_bpf_batch_set_block": (I, [P, I]),
            "gpu_bpf_batch_get_block": (I, [P]),
            "gpu_bpf_batch_set_tile": (I, [P, I]),
            "gpu_bpf_batch_set_stats": (I, [P, I]),
            "gpu_bpf_batch_get_tile": (I, [P]),
            "gpu_bpf_batch_autotune": (I, [P, P, I]),
            "gpu_bpf_batch_run": (I, [P, P, I]),
            "gpu_bpf_batch_run_device": (I, [P, P, I]),
            "gpu_bpf_batch_synchronize": (I, [P]),
            "gpu_bpf_batch_get_timestep": (I, [P]),
            "gpu_bpf_batch_copy_stats": (I, [P, P, I]),
            "gpu_bpf_batch_get_loglik": (I, [P, P]),
            "gpu_bpf_batch_get_resample_counts": (I, [P, P]),
            "gpu_bpf_batch_copy_particles": (I, [P, I, P]),
            "gpu_bpf_batch_enable_score": (I, [P, I]),
            "gpu_bpf_batch_get_score": (I, [P, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _declared = True
    return L


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


class GpuBpfBatch:
    """M inner filters; per-filter arrays n, mu, phi, sigma_z, seeds."""

    def __init__(self, n, mu, phi, sigma_z, seeds, nu_obs=0.0, ess_threshold=0.0, block=None,
                 tile=None, stats=True):
        self.L = _lib()
        self.n = np.ascontiguousarray(n, np.int32)
        self.M = int(self.n.size)
        arrs = [np.ascontiguousarray(np.broadcast_to(np.asarray(v, np.float32), (self.M,)))
                for v in (mu, phi, sigma_z)]
        sd = np.ascontiguousarray(np.broadcast_to(np.asarray(seeds, np.int32), (self.M,)))
        self._keep = (self.n, *arrs, sd)
        self.h = self.L.gpu_bpf_batch_create(self.M, _ptr(self.n), _ptr(arrs[0]), _ptr(arrs[1]),
                                             _ptr(arrs[2]), _ptr(sd), float(nu_obs),
                                             float(ess_threshold))
        if not self.h:
            raise gpu_bpf.GpuBpfError(gpu_bpf.last_error())
        if block is not None:
            self._check(self.L.gpu_bpf_batch_set_block(self.h, int(block)))
        if tile is not None:
            self._check(self.L.gpu_bpf_batch_set_tile(self.h, int(tile)))
        if not stats:
            self.set_stats(False)

    def _check(self, rc):
        if rc < 0:
            raise gpu_bpf.GpuBpfError(gpu_bpf.last_error())
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.L.gpu_bpf_batch_destroy(self.h)
            self.h = None

    __del__ = close

    @property
    def block(self):
        return self.L.gpu_bpf_batch_get_block(self.h)

    @property
    def tile(self):
        """Particles per tile (256, 512 or 1024)."""
        return self.L.gpu_bpf_batch_get_tile(self.h)

    def set_tile(self, particles):
        self._check(self.L.gpu_bpf_batch_set_tile(self.h, int(particles)))

    def set_stats(self, on):
        """Per-tick statistics (mean, variance, ESS, the ring) on or off; off,
        the ticks form only the weight sums the filters need."""
        self._check(self.L.gpu_bpf_batch_set_stats(self.h, 1 if on else 0))

    def autotune(self, y, ticks=64):
        """Times every (block, tile) pair that fits; keeps the fastest and
        returns its block (its tile is self.tile)."""
        y = np.ascontiguousarray(y[:ticks], np.float32)
        return self._check(self.L.gpu_bpf_batch_autotune(self.h, _ptr(y), y.size))

    def run(self, y, stats=True):
        """Advance every filter over y; returns the per-tick statistics
        [T, M] (structured, as gpu_bpf.STATS_DTYPE) when stats is set."""
        y = np.ascontiguousarray(y, np.float32)
        self._check(self.L.gpu_bpf_batch_run(self.h, _ptr(y), y.size))
        if not stats:
            return None
        return self.stats(y.size)

    def run_device(self, y_dev_ptr, n_ticks):
        self._check(self.L.gpu_bpf_batch_run_device(self.h, ctypes.c_void_p(y_dev_ptr), n_ticks))

    def synchronize(self):
        self._check(self.L.gpu_bpf_batch_synchronize(self.h))

    def timestep(self):
        return self.L.gpu_bpf_batch_get_timestep(self.h)

    def stats(self, n_ticks):
        out = np.zeros((n_ticks, self.M), gpu_bpf.STATS_DTYPE)
        self._check(self.L.gpu_bpf_batch_copy_stats(self.h, _ptr(out), n_ticks))
        return out

    def loglik(self):
        ll = np.zeros(self.M, np.float64)
        self._check(self.L.gpu_bpf_batch_get_loglik(self.h, _ptr(ll)))
        return ll

    def resample_counts(self):
        c = np.zeros(self.M, np.int32)
        self._check(self.L.gpu_bpf_batch_get_resample_counts(self.h, _ptr(c)))
        return c

    def enable_score(self, estimator="pathwise"):
        """estimator: "pathwise" (ancestors fixed) or "fisher" (Fisher identity)."""
        self._check(self.L.gpu_bpf_batch_enable_score(self.h, 1 if estimator == "fisher" else 0))

    def score(self):
        sc = np.zeros((self.M, 3), np.float64)
        self._check(self.L.gpu_bpf_batch_get_score(self.h, _ptr(sc)))
        return sc

    def particles(self, m):
        x = np.empty(int(self.n[m]), np.float32)
        self._check(self.L.gpu_bpf_batch_copy_particles(self.h, int(m), _ptr(x)))
        return x


_smc2_declared = False


def _smc2_lib():
    L = _lib()
    global _smc2_declared
    if not _smc2_declared:
        P, F, I = ctypes.c_void_p, ctypes.c_float, ctypes.c_int
        sig = {
            "gpu_smc2_create": (P, [I, I, I, F, I]),
            "gpu_smc2_destroy": (None, [P]),
            "gpu_smc2_set_prior": (I, [P, P, P]),
            "gpu_smc2_set_rejuvenation": (I, [P, I, I]),
            "gpu_smc2_set_cpmmh_rho": (I, [P, F]),
            "gpu_smc2_set_proposal_scale": (I, [P, F, F]),
            "gpu_smc2_set_outer_threshold": (I, [P, F]),
            "gpu_smc2_set_capacity": (I, [P, I]),
            "gpu_smc2_init_from_prior": (I, [P]),
            "gpu_smc2_update": (F, [P, F]),
            "gpu_smc2_run": (I, [P, P, I]),
            "gpu_smc2_get_outer_ess": (F, [P]),
            "gpu_smc2_get_theta_mean": (I, [P, P]),
            "gpu_smc2_get_theta_std": (I, [P, P]),
            "gpu_smc2_copy_theta": (I, [P, P, P, P]),
            "gpu_smc2_get_loglik": (I, [P, P]),
            "gpu_smc2_get_counters": (I, [P, P]),
            "gpu_smc2_set_block": (I, [P, I]),
            "gpu_smc2_set_tile": (I, [P, I]),
            "gpu_smc2_set_mala_step": (I, [P, F]),
            "gpu_smc2_selftest_mcmc": (I, [I, I, I, I, P]),
            "gpu_smc2_get_history": (I, [P, P, P, I]),
            "gpu_smc2_get_score": (I, [P, P]),
            "gpu_smc2_systematic_ancestors": (I, [P, I, ctypes.c_double, P]),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _smc2_declared = True
    return L


COUNTER_NAMES = ("ticks", "outer_resamples", "proposed", "accepted", "replay_ticks",
                 "pt_forward", "pt_replay")


class GpuSmc2:
    """SMC^2 over batched inner SV filters (mirror of the gpu_smc2_* C ABI)."""

    def __init__(self, n_theta, n_min, n_max, nu_obs=0.0, seed=42, moves=5, cpmmh_rho=0.99,
                 outer_threshold=0.5, capacity=2048, prior=None, theta_scale=1.0, u_step=0.05,
                 block=None, kind="cpmmh", mala_step=0.5, tile=None):
        self.L = _smc2_lib()
        self.M = int(n_theta)
        self.h = self.L.gpu_smc2_create(self.M, int(n_min), int(n_max), float(nu_obs), int(seed))
        if not self.h:
            raise gpu_bpf.GpuBpfError(gpu_bpf.last_error())
        if prior is not None:
            lo = np.ascontiguousarray(prior[0], np.float32)
            hi = np.ascontiguousarray(prior[1], np.float32)
            self._check(self.L.gpu_smc2_set_prior(self.h, _ptr(lo), _ptr(hi)))
        self._check(self.L.gpu_smc2_set_rejuvenation(self.h, 1 if kind == "mala" else 0,
                                                     int(moves)))
        self._check(self.L.gpu_smc2_set_mala_step(self.h, float(mala_step)))
        self._check(self.L.gpu_smc2_set_cpmmh_rho(self.h, float(cpmmh_rho)))
        self._check(self.L.gpu_smc2_set_proposal_scale(self.h, float(theta_scale), float(u_step)))
        self._check(self.L.gpu_smc2_set_outer_threshold(self.h, float(outer_threshold)))
        self._check(self.L.gpu_smc2_set_capacity(self.h, int(capacity)))
        self._check(self.L.gpu_smc2_init_from_prior(self.h))
        if block is not None:
            self._check(self.L.gpu_smc2_set_block(self.h, int(block)))
        if tile is not None:
            self._check(self.L.gpu_smc2_set_tile(self.h, int(tile)))

    def _check(self, rc):
        if rc < 0:
            raise gpu_bpf.GpuBpfError(gpu_bpf.last_error())
        return rc

    def close(self):
        if getattr(self, "h", None):
            self.L.gpu_smc2_destroy(self.h)
            self.h = None

    __del__ = close

    def run(self, y):
        y = np.ascontiguousarray(y, np.float32)
        self._check(self.L.gpu_smc2_run(self.h, _ptr(y), y.size))

    def update(self, y):
        ess = self.L.gpu_smc2_update(self.h, float(y))
        if ess < 0:
            raise gpu_bpf.GpuBpfError(gpu_bpf.last_error())
        return ess

    def outer_ess(self):
        return self.L.gpu_smc2_get_outer_ess(self.h)

    def theta(self):
        th = np.zeros((self.M, 3), np.float32)
        lw = np.zeros(self.M, np.float64)
        n = np.zeros(self.M, np.int32)
        self._check(self.L.gpu_smc2_copy_theta(self.h, _ptr(th), _ptr(lw), _ptr(n)))
        return th, lw, n

    def theta_mean_std(self):
        m = np.zeros(3, np.float32)
        s = np.zeros(3, np.float32)
        self._check(self.L.gpu_smc2_get_theta_mean(self.h, _ptr(m)))
        self._check(self.L.gpu_smc2_get_theta_std(self.h, _ptr(s)))
        return m, s

    def loglik(self):
        ll = np.zeros(self.M, np.float64)
        self._check(self.L.gpu_smc2_get_loglik(self.h, _ptr(ll)))
        return ll

    def counters(self):
        c = np.zeros(7, np.int64)
        self._check(self.L.gpu_smc2_get_counters(self.h, _ptr(c)))
        return dict(zip(COUNTER_NAMES, c.tolist()))

    def score(self):
        """MALA runs: per theta-particle pathwise score [M, 3]."""
        sc = np.zeros((self.M, 3), np.float64)
        self._check(self.L.gpu_smc2_get_score(self.h, _ptr(sc)))
        return sc

    def history(self):
        """Outer resamples so far: (ticks done at each, total particle count of
        the resampled population)."""
        k = self._check(self.L.gpu_smc2_get_history(self.h, None, None, 0))
        t = np.zeros(max(k, 1), np.int32)
        s = np.zeros(max(k, 1), np.int64)
        self._check(self.L.gpu_smc2_get_history(self.h, _ptr(t), _ptr(s), k))
        return t[:k], s[:k]


def systematic_ancestors(log_w, u):
    """The outer layer's systematic resampling (host code of the library, no
    GPU): ancestors of len(log_w) slots for the uniform u."""
    lw = np.ascontiguousarray(log_w, np.float64)
    anc = np.zeros(lw.size, np.int32)
    rc = _smc2_lib().gpu_smc2_systematic_ancestors(_ptr(lw), lw.size, float(u), _ptr(anc))
    if rc < 0:
        raise gpu_bpf.GpuBpfError(gpu_bpf.last_error())
    return anc


def selftest_mcmc(kind, iters, seed=1, fault=0):
    """The rejuvenation moves on a closed-form 3-d Gaussian target (host code
    of the library, no GPU): returns (mean[3], cov[3, 3], acceptance)."""
    out = np.zeros(13, np.float64)
    rc = _smc2_lib().gpu_smc2_selftest_mcmc(1 if kind == "mala" else 0, int(iters), int(seed),
                                             int(fault), _ptr(out))
    if rc < 0:
        raise gpu_bpf.GpuBpfError(gpu_bpf.last_error())
    return out[:3], out[3:12].reshape(3, 3), out[12]
