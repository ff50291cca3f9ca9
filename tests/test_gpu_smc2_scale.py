"""SV-SMC^2 at the configurations' full size, where the oracle is too slow:
properties that hold exactly whatever the size.

C3: M = 512 theta-particles, inner particle counts log-uniform in
[256, 8192] (the library's seeded prior draws, seed 42 as the benchmark),
T = 2000, CPMMH with K = 5 moves.  C4: M = 1024, counts log-uniform in
[512, 16384], T = 5000, Student-t (nu = 5) observations, MALA with K = 3.
Each runs on the benchmark's regime data (mu -1.5 / -0.5 / +0.5 every 1250
ticks) and on stationary data from a known theta.  Checked, in 32-tick
chunks:
- between outer resamples the outer log-weights grow by exactly the inner
  FP64 log-likelihood increments (to FP64 rounding);
- the outer ESS stays in [1, M]; every theta stays inside the prior box;
- counters: ticks = T, forward particle-ticks = sum over the segments
  between outer resamples of (total particles x segment length), replayed
  ticks and particle-ticks = K x the same over the resample ticks, proposals
  = K M per resample; acceptance in (0.05, 0.95);
- on the known-theta data the posterior concentrates on the truth.
Exact agreement with the oracle at M <= 64 is tests/test_gpu_smc2_oracle.py.
"""
import numpy as np
import pytest

import gpu_bpf_batch
import sv_data

pytestmark = pytest.mark.gpu

LO = np.array([-2.0, 0.90, 0.05])
HI = np.array([0.0, 0.995, 0.40])
TRUTH = np.array([-1.0, 0.97, 0.15])
CONFIGS = {
    "C3": dict(M=512, nmin=256, nmax=8192, T=2000, nu=0.0, kind="cpmmh", moves=5),
    "C4": dict(M=1024, nmin=512, nmax=16384, T=5000, nu=5.0, kind="mala", moves=3),
}


def say(*a):
    print("[case]", *a, flush=True)


def _data(cfg, regimes):
    T = cfg["T"]
    mu = sv_data.mu_schedule(T).astype(np.float64) if regimes else TRUTH[0]
    return sv_data.sv_series(T, TRUTH[1], TRUTH[2], mu, nu_obs=cfg["nu"], seed=42)[1].astype(np.float32)


@pytest.mark.parametrize("regimes", [True, False], ids=["regime_data", "known_theta"])
@pytest.mark.parametrize("name", list(CONFIGS))
def test_smc2_at_config_size(name, regimes):
    cfg = CONFIGS[name]
    M, T, K = cfg["M"], cfg["T"], cfg["moves"]
    y = _data(cfg, regimes)
    say(name, "regimes" if regimes else "known theta", "M", M, "T", T)
    s = gpu_bpf_batch.GpuSmc2(M, cfg["nmin"], cfg["nmax"], nu_obs=cfg["nu"], seed=42, kind=cfg["kind"],
                              moves=K, capacity=T, mala_step=0.5)
    th, lw, n0 = s.theta()
    assert np.all(n0 >= cfg["nmin"]) and np.all(n0 <= cfg["nmax"])
    ll = s.loglik()
    res_before = 0
    checked = 0
    for t0 in range(0, T, 32):
        s.run(y[t0:t0 + 32])
        th2, lw2, n2 = s.theta()
        ll2 = s.loglik()
        res = s.counters()["outer_resamples"]
        assert np.all(np.isfinite(ll2)) and np.all(np.isfinite(lw2))
        assert np.all(th2 >= LO.astype(np.float32)) and np.all(th2 <= HI.astype(np.float32))
        ess = s.outer_ess()
        assert 1.0 - 1e-6 <= ess <= M * (1 + 1e-6), ess
        if res == res_before:
            # no outer resample in the chunk: outer weights and filters continue
            np.testing.assert_array_equal(n2, n0)
            np.testing.assert_array_equal(th2, th)
            d_lw, d_ll = lw2 - lw, ll2 - ll
            assert np.all(np.abs(d_lw - d_ll) <= 1e-9 * np.maximum(1.0, np.abs(ll2))), \
                np.max(np.abs(d_lw - d_ll))
            checked += 1
        th, lw, n0, ll, res_before = th2, lw2, n2, ll2, res
    c = s.counters()
    ticks, sum_n = s.history()
    s_fin = s.theta_mean_std()
    sc = s.score() if cfg["kind"] == "mala" else None
    s.close()
    assert checked >= 10
    assert c["ticks"] == T and c["outer_resamples"] == len(ticks) >= 2
    assert np.all(np.diff(ticks) > 0) and ticks[-1] <= T
    # forward particle-ticks: the initial population until the first resample,
    # then each resampled population until the next
    s0 = gpu_bpf_batch.GpuSmc2(M, cfg["nmin"], cfg["nmax"], nu_obs=cfg["nu"], seed=42, kind=cfg["kind"],
                               moves=K, capacity=1)
    n_init = int(s0.theta()[2].sum())
    s0.close()
    seg = np.diff(np.concatenate([[0], ticks, [T]]))
    assert c["pt_forward"] == int(np.dot(np.concatenate([[n_init], sum_n]), seg))
    assert c["replay_ticks"] == K * int(ticks.sum())
    assert c["pt_replay"] == K * int(np.dot(sum_n, ticks))
    assert c["proposed"] == K * M * len(ticks)
    rate = c["accepted"] / c["proposed"]
    assert 0.05 < rate < 0.95, rate
    if sc is not None:
        assert np.all(np.isfinite(sc))
    mean, sd = s_fin
    say(name, "theta", mean, sd, "accept", rate, "resamples", len(ticks))
    prior_sd = (HI - LO) / np.sqrt(12.0)
    assert np.all(sd < 0.5 * prior_sd), (sd, prior_sd)
    if not regimes:
        # the posterior concentrates on the data's theta
        assert np.all(np.abs(mean - TRUTH) <= 4.0 * sd + 0.02 * (HI - LO)), (mean, sd)
