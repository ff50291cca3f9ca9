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
            "gpu_bpf_batch_autotune": (I, [P, P, I]),
            "gpu_bpf_batch_run": (I, [P, P, I]),
            "gpu_bpf_batch_run_device": (I, [P, P, I]),
            "gpu_bpf_batch_synchronize": (I, [P]),
            "gpu_bpf_batch_get_timestep": (I, [P]),
            "gpu_bpf_batch_copy_stats": (I, [P, P, I]),
            "gpu_bpf_batch_get_loglik": (I, [P, P]),
            "gpu_bpf_batch_get_resample_counts": (I, [P, P]),
            "gpu_bpf_batch_copy_particles": (I, [P, I, P]),
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

    def __init__(self, n, mu, phi, sigma_z, seeds, nu_obs=0.0, ess_threshold=0.0, block=None):
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

    def autotune(self, y, ticks=64):
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

    def particles(self, m):
        x = np.empty(int(self.n[m]), np.float32)
        self._check(self.L.gpu_bpf_batch_copy_particles(self.h, int(m), _ptr(x)))
        return x
