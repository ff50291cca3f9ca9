"""Python mirror of include/smc2_rbpf.h over ctypes: SMC^2 over (rho, sigma_z,
mu_base, sigma_base) with the RBPF / OCSN inner filter.

The calls follow the reference host API one for one
(smc2/smc2_rbpf_batch.cuh:458-490: smc2_cuda_alloc / free / set_seed /
set_noise_capacity / set_fixed_lag / set_cpmmh_rho / set_proposal_std /
init_from_prior / update / get_theta_mean / get_theta_std / get_outer_ess);
the reference's struct-field access is replaced by the set_model / get_model /
set_rejuv / set_adaptive / get_counters / get_particles / get_inner calls, and
its compile-time SMC2_NOISE_FP16 by set_noise_precision (noise_bits=16).
The compute path is the HIP library _build/libgpubpf.so; there is no other.
"""
import ctypes
import math

import numpy as np

import gpu_bpf

PARAM_NAMES = ("rho", "sigma_z", "mu_base", "sigma_base")


class SVPrior(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in (
        "rho_mean", "rho_std", "sigma_z_mean", "sigma_z_std", "mu_base_mean", "mu_base_std",
        "sigma_base_mean", "sigma_base_std")]


class SVBounds(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in (
        "rho_min", "rho_max", "sigma_z_min", "sigma_z_max", "mu_base_min", "mu_base_max",
        "sigma_base_min", "sigma_base_max")]


class SVFixedCurves(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in ("mu_scale", "mu_rate", "sigma_scale", "sigma_rate")]


class SVCurve(ctypes.Structure):
    _fields_ = [(n, ctypes.c_float) for n in ("base", "scale", "rate")]


class SMC2Counters(ctypes.Structure):
    _fields_ = [(n, ctypes.c_int) for n in (
        "n_resamples", "n_rejuv_accepts", "n_rejuv_total", "t_current", "t_checkpoint",
        "noise_capacity")]


_bound = False


def load():
    """The HIP library with the smc2_cuda_* signatures set."""
    global _bound
    L = gpu_bpf.load()
    if _bound:
        return L
    P, F, I, D = ctypes.c_void_p, ctypes.c_float, ctypes.c_int, ctypes.c_double
    sig = {
        "smc2_cuda_alloc": (P, [I, I]),
        "smc2_cuda_free": (None, [P]),
        "smc2_cuda_set_seed": (None, [P, ctypes.c_uint64]),
        "smc2_cuda_set_noise_capacity": (None, [P, I]),
        "smc2_cuda_set_fixed_lag": (None, [P, I]),
        "smc2_cuda_set_cpmmh_rho": (None, [P, F]),
        "smc2_cuda_set_proposal_std": (None, [P, P]),
        "smc2_cuda_init_from_prior": (None, [P]),
        "smc2_cuda_update": (F, [P, F]),
        "smc2_cuda_get_theta_mean": (None, [P, P]),
        "smc2_cuda_get_theta_std": (None, [P, P]),
        "smc2_cuda_get_outer_ess": (F, [P]),
        "smc2_cuda_set_model": (I, [P, P, P, P, P]),
        "smc2_cuda_get_model": (I, [P, P, P, P, P]),
        "smc2_cuda_set_rejuv": (I, [P, I, F]),
        "smc2_cuda_set_adaptive": (I, [P, I]),
        "smc2_cuda_set_noise_precision": (I, [P, I]),
        "smc2_cuda_get_counters": (I, [P, ctypes.POINTER(SMC2Counters)]),
        "smc2_cuda_get_particles": (I, [P, P, P, P, P, P, P, P]),
        "smc2_cuda_get_inner": (I, [P, P, P, P, P]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _bound = True
    return L


def _ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p) if a is not None else None


def _err(what):
    msg = gpu_bpf.load().gpu_bpf_last_error()
    return gpu_bpf.GpuBpfError(f"{what}: {msg.decode() if msg else 'failed'}")


class Smc2Rbpf:
    """One SMC^2 learner (N_theta theta-particles x N_inner inner particles)."""

    def __init__(self, n_theta, n_inner, seed=12345, noise_capacity=None, fixed_lag=0,
                 cpmmh_rho=None, proposal_std=None, K_rejuv=None, ess_threshold=None,
                 adaptive=True, prior=None, bounds=None, theta_curve=None, fixed_curves=None,
                 noise_bits=32, init=True):
        self._L = load()
        self._s = self._L.smc2_cuda_alloc(n_theta, n_inner)
        if not self._s:
            raise _err("smc2_cuda_alloc")
        self.n_theta, self.n_inner = n_theta, n_inner
        L, s = self._L, self._s
        L.smc2_cuda_set_seed(s, seed)
        if noise_capacity is not None:
            L.smc2_cuda_set_noise_capacity(s, noise_capacity)
        L.smc2_cuda_set_fixed_lag(s, fixed_lag)
        if cpmmh_rho is not None:
            L.smc2_cuda_set_cpmmh_rho(s, cpmmh_rho)
        if proposal_std is not None:
            ps = np.ascontiguousarray(proposal_std, np.float32)
            L.smc2_cuda_set_proposal_std(s, _ptr(ps))
        if K_rejuv is not None or ess_threshold is not None:
            if L.smc2_cuda_set_rejuv(s, 5 if K_rejuv is None else K_rejuv,
                                     0.5 if ess_threshold is None else ess_threshold) != 0:
                raise _err("smc2_cuda_set_rejuv")
        if L.smc2_cuda_set_adaptive(s, 1 if adaptive else 0) != 0:
            raise _err("smc2_cuda_set_adaptive")
        if L.smc2_cuda_set_noise_precision(s, noise_bits) != 0:
            raise _err("smc2_cuda_set_noise_precision")
        if any(v is not None for v in (prior, bounds, theta_curve, fixed_curves)):
            args = [SVPrior(*prior) if prior is not None else None,
                    SVBounds(*bounds) if bounds is not None else None,
                    SVCurve(*theta_curve) if theta_curve is not None else None,
                    SVFixedCurves(*fixed_curves) if fixed_curves is not None else None]
            refs = [ctypes.byref(a) if a is not None else None for a in args]
            if L.smc2_cuda_set_model(s, *refs) != 0:
                raise _err("smc2_cuda_set_model")
        if init:
            self.init_from_prior()

    def init_from_prior(self):
        self._L.smc2_cuda_init_from_prior(self._s)
        msg = self._L.gpu_bpf_last_error()
        if msg:
            raise gpu_bpf.GpuBpfError(f"smc2_cuda_init_from_prior: {msg.decode()}")

    def update(self, y):
        ess = self._L.smc2_cuda_update(self._s, float(y))
        if math.isnan(ess):
            raise _err("smc2_cuda_update")
        return ess

    def run(self, y):
        return np.array([self.update(v) for v in np.asarray(y, np.float32)], np.float32)

    def theta_mean(self):
        out = np.zeros(4, np.float32)
        self._L.smc2_cuda_get_theta_mean(self._s, _ptr(out))
        return out

    def theta_std(self):
        out = np.zeros(4, np.float32)
        self._L.smc2_cuda_get_theta_std(self._s, _ptr(out))
        return out

    def outer_ess(self):
        return self._L.smc2_cuda_get_outer_ess(self._s)

    def counters(self):
        c = SMC2Counters()
        if self._L.smc2_cuda_get_counters(self._s, ctypes.byref(c)) != 0:
            raise _err("smc2_cuda_get_counters")
        return {n: getattr(c, n) for n, _ in SMC2Counters._fields_}

    def model(self):
        p, b, t, f = SVPrior(), SVBounds(), SVCurve(), SVFixedCurves()
        if self._L.smc2_cuda_get_model(self._s, ctypes.byref(p), ctypes.byref(b), ctypes.byref(t),
                                       ctypes.byref(f)) != 0:
            raise _err("smc2_cuda_get_model")
        return p, b, t, f

    def particles(self):
        n = self.n_theta
        th = np.zeros((4, n), np.float32)
        lw, ll = np.zeros(n), np.zeros(n)
        essi = np.zeros(n, np.float32)
        if self._L.smc2_cuda_get_particles(self._s, _ptr(th[0]), _ptr(th[1]), _ptr(th[2]), _ptr(th[3]),
                                           _ptr(lw), _ptr(ll), _ptr(essi)) != 0:
            raise _err("smc2_cuda_get_particles")
        return dict(theta=th, lw=lw, ll=ll, ess_inner=essi)

    def inner(self):
        a = np.zeros((4, self.n_theta * self.n_inner), np.float32)
        if self._L.smc2_cuda_get_inner(self._s, _ptr(a[0]), _ptr(a[1]), _ptr(a[2]), _ptr(a[3])) != 0:
            raise _err("smc2_cuda_get_inner")
        return a

    def close(self):
        if self._s:
            self._L.smc2_cuda_free(self._s)
            self._s = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
