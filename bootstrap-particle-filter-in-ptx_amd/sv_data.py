"""Synthetic stochastic-volatility observations: the reference's matched data
generator (tests/test_bpf_matched_dgp.cu:59-99 for its PCG32 + FP64
Box-Muller generator, :141-175 for the SV series), used to build the
benchmark workloads without the test oracle.

    x_0 ~ N(mu_0, sigma^2 / max(1 - phi^2, 1e-6)),
    x_t = mu_t + phi (x_{t-1} - mu_t) + sigma eps_t,   clamped to [-12, 4],
    y_t = exp(x_t / 2) eta_t,

eps and eta standard normal, or Student-t(nu) as z / sqrt(chi2_nu / nu) with
chi2 the sum of int(nu) squared normals (nu > 100 draws a normal).  The
arithmetic is plain IEEE double through the C library's log / cos / sqrt /
exp, so the series are bit-identical to any C restatement on the same libm.
"""
import math

import numpy as np

_MULT = 6364136223846793005
_MASK = (1 << 64) - 1


class _Pcg32:
    """PCG32 XSH-RR with the generator's seeding: inc = (seed * 12345) | 1,
    state advanced once, seed * 67890 added, advanced twice."""

    def __init__(self, seed):
        seed &= _MASK
        self.state = 0
        self.inc = ((seed * 12345) & _MASK) | 1
        self.next()
        self.state = (self.state + seed * 67890) & _MASK
        self.next()
        self.next()

    def next(self):
        old = self.state
        self.state = (old * _MULT + self.inc) & _MASK
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF

    def double(self):
        return (self.next() >> 11) * (1.0 / 2097152.0)

    def gauss(self):
        u1 = self.double()
        while u1 < 1e-10:
            u1 = self.double()
        u2 = self.double()
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * 3.14159265358979323846 * u2)

    def student_t(self, nu):
        if nu > 100.0:
            return self.gauss()
        z = self.gauss()
        chi = 0.0
        for _ in range(int(nu)):
            v = self.gauss()
            chi += v * v
        return z / math.sqrt(chi / nu)


def sv_series(n_ticks, phi, sigma, mu, nu_state=0.0, nu_obs=0.0, seed=42):
    """(true states, observations) as float64 arrays of n_ticks; mu is a
    scalar or a per-tick sequence (regime shifts)."""
    g = _Pcg32(int(seed))
    mus = [float(v) for v in mu] if np.ndim(mu) else [float(mu)] * n_ticks
    if len(mus) < n_ticks:
        raise ValueError("mu schedule shorter than the series")
    phi, sigma = float(phi), float(sigma)
    var_stat = (sigma * sigma) / max(1.0 - phi * phi, 1e-6)
    h = mus[0] + math.sqrt(var_stat) * g.gauss()
    xs = np.empty(n_ticks)
    ys = np.empty(n_ticks)
    for t in range(n_ticks):
        m = mus[t]
        if t > 0:
            eps = g.student_t(nu_state) if nu_state > 0.0 else g.gauss()
            h = m + phi * (h - m) + sigma * eps
        h = min(max(h, -12.0), 4.0)
        xs[t] = h
        eta = g.student_t(nu_obs) if nu_obs > 0.0 else g.gauss()
        ys[t] = math.exp(h / 2.0) * eta
    return xs, ys


def mu_schedule(n_ticks, block=1250, levels=(-1.5, -0.5, 0.5)):
    """Piecewise-constant mu: levels[(t // block) % 3] (the fourth block
    repeats the first level)."""
    return np.array([levels[(t // block) % 3] for t in range(n_ticks)], np.float32)


# theta(z) curve shapes of the regime-switching model (smc2/smc2_rbpf_batch.cu
# defaults): (base, scale, rate) of the mean-reversion speed, and (mu scale,
# mu rate, sigma scale, sigma rate) of the level and volatility curves
RB_TCURVE = (0.02, 0.08, 1.5)
RB_FIXED = (0.5, 1.0, 0.1, 1.0)


def rbpf_series(n_ticks, rho=0.95, sigma_z=0.10, mu_base=-1.0, sigma_base=0.15, fixed=RB_FIXED,
                tcurve=RB_TCURVE, seed=12345):
    """Regime-switching SV data of smc2/test_smc2_rbpf.cu:95-133 (FP64, NumPy
    normals): a latent AR(1) z~ drives z = 1.5 (1 + tanh z~), which sets the
    reversion theta(z), level mu(z) and volatility sigma(z) of
    h_t = (1 - theta) h_{t-1} + theta mu + sigma eps.  Returns
    (y = log r^2 floored at log 1e-30 as float32, h, z)."""
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
    y, hs, zs = np.zeros(n_ticks), np.zeros(n_ticks), np.zeros(n_ticks)
    for t in range(n_ticks):
        zt = rho * zt + sigma_z * rng.standard_normal()
        z = 1.5 * (1.0 + np.tanh(zt))
        th, mu, sh = curve(tb, ts, tr, z), curve(mu_base, ms, mr, z), curve(sigma_base, ss, sr, z)
        h = (1.0 - th) * h + th * mu + sh * rng.standard_normal()
        r2 = max((np.exp(h / 2.0) * rng.standard_normal()) ** 2, 1e-30)
        y[t], hs[t], zs[t] = np.log(r2), h, z
    return y.astype(np.float32), hs, zs
