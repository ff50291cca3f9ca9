"""Pins of the SV-SMC^2 oracle (oracle/smc2_oracle.c) on the CPU, before it
checks the GPU's gpu_smc2_* path (tests/test_gpu_smc2_oracle.py).

- Inner filters: one batched inner filter equals the single-filter oracle
  (bpf_oracle.c, 256-particle tiles, every tick resampled) bit for bit in its
  particles and to 1e-12 in its log-likelihood; its FP32 pathwise score
  agrees with the FP64 tangent oracle on the same genealogy (itself pinned by
  finite differences) to 1e-3 relative.
- Prior draws: n, theta of every theta-particle equal a pure-Python
  restatement of the documented outer-layer stream.
- Outer layer without resampling (threshold 0): every outer log-weight is its
  filter's log-likelihood exactly, and each filter is the single-filter
  oracle run of its (theta, n, seed).
- Identity moves (rho_cn = 1, zero random walks) replay the recorded noise bit
  for bit, so every move is accepted and the run equals the run without
  moves.
- Outer resampling against brute-force systematic resampling (non-finite
  weights weigh nothing); the moves' code leaves closed-form targets
  invariant; the Crank-Nicolson move keeps N(0, 1) invariant.
- Planted mistakes: each is caught by one of the checks above.
"""
import math

import numpy as np
import pytest

import sv_data

MASK = (1 << 64) - 1
MULT = 6364136223846793005


class HostStream:
    """The outer layer's PCG32 stream (XSH-RR, stream id -> increment)."""

    def __init__(self, seed, stream):
        self.s, self.inc = 0, ((stream << 1) | 1) & MASK
        self.next()
        self.s = (self.s + (seed & MASK)) & MASK
        self.next()

    def next(self):
        old = self.s
        self.s = (old * MULT + self.inc) & MASK
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((-rot) & 31))) & 0xFFFFFFFF

    def uniform(self):
        a = self.next() >> 5
        b = self.next() >> 6
        return ((a * 67108864 + b) + 0.5) / 9007199254740992.0


def prior_draws(M, n_min, n_max, seed, lo=(-2.0, 0.90, 0.05), hi=(0.0, 0.995, 0.40)):
    r = HostStream(seed, 0x5eed5c2)
    f32 = np.float32
    out = []
    for _ in range(M):
        lu = math.log(n_min) + r.uniform() * (math.log(n_max) - math.log(n_min))
        n = min(n_max, max(n_min, int(round_half_away(math.exp(lu)))))
        th = []
        for k in range(3):
            span = float(f32(hi[k]) - f32(lo[k]))          # float subtraction, as the library
            th.append(f32(float(f32(lo[k])) + r.uniform() * span))
        out.append((n, th, r.next() & 0x7FFFFFFF))
    return out


def round_half_away(v):
    return math.floor(v + 0.5) if v >= 0 else math.ceil(v - 0.5)


def _y(T, seed=5, nu=0.0):
    return sv_data.sv_series(T, 0.97, 0.15, -1.0, nu_obs=nu, seed=seed)[1].astype(np.float32)


@pytest.mark.parametrize("nu", [0.0, 5.0])
def test_inner_filter_is_the_single_filter(orc, nu):
    y = _y(80, nu=nu)
    for n, mu, phi, sig, seed in [(300, -1.1, 0.96, 0.2, 17), (4097, -0.4, 0.93, 0.31, 5)]:
        ll, sc, x = orc.sv_inner_run(y, n, mu, phi, sig, nu_obs=nu, seed=seed)
        ref, _, xo, _ = orc.run_filter(y, n, rho=phi, sigma_z=sig, mu=mu, nu_obs=nu, ess_threshold=0.0,
                                       seed=seed, tile=256, want_x=True)
        np.testing.assert_array_equal(x, xo)
        assert abs(ll - ref["ll_cum"][-1]) <= 1e-12 * abs(ref["ll_cum"][-1])


@pytest.mark.parametrize("nu", [0.0, 5.0])
def test_inner_score_matches_fp64_tangents(orc, nu):
    y = _y(100, nu=nu)
    n, mu, phi, sig, seed = 2048, -1.0, 0.97, 0.15, 9
    _, sc, _ = orc.sv_inner_run(y, n, mu, phi, sig, nu_obs=nu, seed=seed)
    _, ref, _, _ = orc.score(y, n, rho=phi, sigma_z=sig, mu=mu, nu_obs=nu, seed=seed, tile=256, x0_reparam=True)
    np.testing.assert_allclose(sc, ref, rtol=1e-3, atol=1e-3 * np.max(np.abs(ref)))


def test_prior_draws_follow_the_documented_stream(orc):
    y = _y(3)
    r = orc.run_sv_smc2(y, 40, 200, 5000, seed=123, moves=0, outer_threshold=0.0)
    want = prior_draws(40, 200, 5000, 123)
    np.testing.assert_array_equal(r["n"], [w[0] for w in want])
    np.testing.assert_array_equal(r["theta"], np.array([w[1] for w in want], np.float32))


@pytest.mark.parametrize("nu", [0.0, 5.0])
def test_without_outer_resampling_each_slot_is_the_single_filter(orc, nu):
    T, M = 60, 12
    y = _y(T, nu=nu)
    r = orc.run_sv_smc2(y, M, 100, 3000, seed=7, moves=2, nu_obs=nu, outer_threshold=0.0)
    assert r["counters"]["outer_resamples"] == 0 and r["counters"]["proposed"] == 0
    np.testing.assert_array_equal(r["logw"], r["ll"])
    for m, (n, th, seed) in enumerate(prior_draws(M, 100, 3000, 7)):
        ref = orc.run_filter(y, n, rho=float(th[1]), sigma_z=float(th[2]), mu=float(th[0]), nu_obs=nu,
                             ess_threshold=0.0, seed=seed, tile=256)[0]
        assert abs(r["ll"][m] - ref["ll_cum"][-1]) <= 1e-12 * abs(ref["ll_cum"][-1])
    assert r["counters"]["pt_forward"] == T * int(r["n"].sum())


@pytest.mark.parametrize("kind", ["cpmmh", "mala"])
def test_identity_moves_accept_everything_and_change_nothing(orc, kind):
    y = _y(120)
    kw = dict(seed=4, outer_threshold=0.9, kind=kind, theta_scale=0.0, cn_rho=1.0, u_step=0.0)
    if kind == "mala":
        kw["mala_step"] = 1e-30        # theta' = theta up to FP64 rounding of a 1e-15 step
    a = orc.run_sv_smc2(y, 24, 128, 1024, moves=3, **kw)
    b = orc.run_sv_smc2(y, 24, 128, 1024, moves=0, **kw)
    c = a["counters"]
    assert c["outer_resamples"] >= 2
    if kind == "cpmmh":
        assert c["accepted"] == c["proposed"] > 0
        for k in ("theta", "n", "logw", "ll"):
            np.testing.assert_array_equal(a[k], b[k])
        np.testing.assert_array_equal(a["history"][0], b["history"][0])
    else:
        # fresh noise: the replay is another filter, so only the bookkeeping is fixed
        assert c["proposed"] > 0 and c["accepted"] <= c["proposed"]


def test_outer_ancestors_brute_force(orc):
    rng = np.random.default_rng(0)
    for _ in range(20):
        M = int(rng.integers(1, 100))
        lw = rng.normal(0, 4, M)
        lw[rng.random(M) < 0.1] = np.nan
        lw[rng.random(M) < 0.1] = -np.inf
        if not np.isfinite(lw).any():
            lw[0] = 1.0
        u = float(rng.uniform(1e-9, 1 - 1e-9))
        ok = np.isfinite(lw)
        w = np.where(ok, np.exp(np.where(ok, lw - lw[ok].max(), 0)), 0.0)
        c = np.cumsum(w)
        want = [int(np.argmax(c >= (u + m) / M * c[-1])) for m in range(M)]
        np.testing.assert_array_equal(orc.smc2_ancestors(lw, u), want)
    with pytest.raises(ValueError):
        orc.smc2_ancestors([np.nan, -np.inf], 0.5)


M0 = np.array([-1.0, 0.95, 0.2])
C0 = np.array([0.04, 0.002, -0.003, 0.002, 0.0004, 0.0001, -0.003, 0.0001, 0.0025]).reshape(3, 3)


@pytest.mark.parametrize("kind", ["rw", "mala"])
def test_moves_leave_gaussian_invariant(orc, kind):
    mean, cov, acc = orc.smc2_selftest(kind, 400_000, seed=3)
    sd = np.sqrt(np.diag(C0))
    assert np.all(np.abs(mean - M0) < 0.05 * sd), mean
    np.testing.assert_allclose(np.diag(cov), np.diag(C0), rtol=0.05)


def test_crank_nicolson_keeps_the_noise_prior(orc):
    m, v, r = orc.smc2_selftest("cn", 400_000, seed=2)
    assert abs(m) < 0.02 and abs(v - 1.0) < 0.02 and abs(r - 0.6) < 0.01


def test_planted_mistakes_are_caught(orc):
    _, cov, _ = orc.smc2_selftest("mala", 400_000, seed=3, fault=1)
    assert np.max(np.abs(np.diag(cov) / np.diag(C0) - 1.0)) > 0.2
    _, v, _ = orc.smc2_selftest("cn", 200_000, seed=2, fault=2)
    assert abs(v - 1.0) > 0.2
    lw = np.log(np.array([0.1, 0.4, 0.2, 0.3]))
    ok = orc.smc2_ancestors(lw, 0.3)
    orc.lib().orc_smc2_set_fault(3)
    try:
        bad = orc.smc2_ancestors(lw, 0.3)
    finally:
        orc.lib().orc_smc2_set_fault(0)
    assert not np.array_equal(ok, bad)
