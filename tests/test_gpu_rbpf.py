"""GPU parity of the SMC^2 learner with the RBPF / OCSN inner filter
(include/smc2_rbpf.h, through its C ABI) against the CPU oracle.

Reference: smc2/smc2_rbpf_batch.cu, smc2/test_smc2_rbpf.cu.
Bar: theta-particles, inner states (z~, mu_h, var_h, log_w), resampling and
acceptance counts bit-exact; FP64 log-likelihoods and log-weights within
1e-12 relative (the device and host FP64 log differ by a few ulp); outer ESS
within 1e-6 relative.  Recovery: the reference's criterion (at least 3 of 4
parameters within 2 posterior standard deviations).
"""
import numpy as np
import pytest

import smc2_rbpf

pytestmark = pytest.mark.gpu

TRUTH = (0.95, 0.10, -1.0, 0.15)


def say(*a):
    print("[case]", *a, flush=True)


def gpu_run(nt, ni, y, seed=12345, K=5, ess_thr=0.5, lag=0, crho=None, adaptive=True, pstd=None,
            noise_bits=32):
    f = smc2_rbpf.Smc2Rbpf(nt, ni, seed=seed, fixed_lag=lag, cpmmh_rho=crho, proposal_std=pstd,
                           K_rejuv=K, ess_threshold=ess_thr, adaptive=adaptive, noise_bits=noise_bits)
    ess = f.run(y)
    out = f.particles()
    out["ess"] = ess
    out["inner"] = f.inner()
    out["counters"] = f.counters()
    out["mean"], out["std"] = f.theta_mean(), f.theta_std()
    f.close()
    return out


def orc_run(orc, nt, ni, y, seed=12345, K=5, ess_thr=0.5, lag=0, crho=0.99, adaptive=True,
            pstd=None, noise_bits=32):
    cfg = orc.rb_cfg(nt, ni, seed=seed, K_rejuv=K, ess_thr=ess_thr, lag=lag,
                     crho=0.99 if crho is None else crho, adaptive=adaptive,
                     pstd=orc.RB_PSTD if pstd is None else pstd, T=len(y), noise_bits=noise_bits)
    return orc.run_rb_smc2(cfg, y)


def assert_parity(g, o):
    c = g["counters"]
    assert (c["n_resamples"], c["n_rejuv_accepts"], c["n_rejuv_total"], c["t_checkpoint"]) == \
        (o["n_resamples"], o["n_acc"], o["n_tot"], o["t_cp"])
    assert np.array_equal(g["theta"], o["theta"])
    assert np.array_equal(g["inner"], o["inner"])
    np.testing.assert_allclose(g["ll"], o["ll"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(g["lw"], o["lw"], rtol=1e-12, atol=1e-12)
    np.testing.assert_allclose(g["ess"], o["ess"], rtol=1e-6)
    np.testing.assert_allclose(g["ess_inner"], o["ess_inner"], rtol=1e-6)


@pytest.mark.parametrize("nt, ni, T, K, ess_thr, lag", [
    (16, 64, 48, 2, 0.9, 0),
    (12, 128, 40, 2, 0.97, 8),
    (8, 256, 24, 1, 0.98, 0),
    (4, 512, 16, 1, 0.995, 4),
    (3, 1024, 10, 1, 0.95, 0),
])
def test_parity_with_oracle(orc, nt, ni, T, K, ess_thr, lag):
    say("rbpf parity", nt, ni, T, K, ess_thr, lag)
    y, _, _ = orc.rb_dgp(T, *TRUTH, seed=nt)
    g = gpu_run(nt, ni, y, K=K, ess_thr=ess_thr, lag=lag)
    o = orc_run(orc, nt, ni, y, K=K, ess_thr=ess_thr, lag=lag)
    assert o["n_resamples"] >= 2
    assert_parity(g, o)


def test_default_settings_parity(orc):
    """Defaults of smc2_cuda_alloc (K_rejuv 5, thresholds 0.5, rho_c 0.99)."""
    y, _, _ = orc.rb_dgp(300, *TRUTH, seed=4)
    say("rbpf defaults", 64, 64, 300)
    g = gpu_run(64, 64, y)
    o = orc_run(orc, 64, 64, y)
    assert o["n_resamples"] >= 1
    assert_parity(g, o)


@pytest.mark.parametrize("lag", [0, 5])
def test_identity_moves_change_nothing(orc, lag):
    """rho_c = 1 and a zero random walk: every move replays the current
    trajectory bit for bit, is accepted and changes no state."""
    y, _, _ = orc.rb_dgp(36, *TRUTH, seed=9)
    say("rbpf identity moves", lag)
    kw = dict(ess_thr=0.97, lag=lag, crho=1.0, adaptive=False, pstd=(0.0, 0.0, 0.0, 0.0))
    a = gpu_run(12, 128, y, K=0, **kw)
    b = gpu_run(12, 128, y, K=3, **kw)
    cb = b["counters"]
    assert cb["n_resamples"] >= 4
    assert cb["n_rejuv_accepts"] == cb["n_rejuv_total"] == 3 * 12 * cb["n_resamples"]
    assert np.array_equal(a["theta"], b["theta"]) and np.array_equal(a["inner"], b["inner"])
    np.testing.assert_allclose(b["ll"], a["ll"], rtol=1e-12)


def test_noise_history_growth(orc):
    """Past the default 2048-tick noise capacity the history is regrown row by
    row; resampling and fixed-lag moves after the growth stay exact."""
    T = 2100
    y, _, _ = orc.rb_dgp(T, *TRUTH, seed=21)
    say("rbpf growth", T)
    g = gpu_run(4, 64, y, K=1, ess_thr=1.0, lag=10)          # resample every tick
    assert g["counters"]["noise_capacity"] == 4096
    o = orc_run(orc, 4, 64, y, K=1, ess_thr=1.0, lag=10)
    assert o["resample_ticks"].max() > 2048
    assert_parity(g, o)


def test_errors_and_model_access():
    say("rbpf errors")
    with pytest.raises(Exception, match="N_inner"):
        smc2_rbpf.Smc2Rbpf(8, 100)
    with pytest.raises(Exception, match="N_theta"):
        smc2_rbpf.Smc2Rbpf(1, 64)
    f = smc2_rbpf.Smc2Rbpf(8, 64, init=False)
    with pytest.raises(Exception, match="init_from_prior"):
        f.update(0.0)
    p, b, t, fc = f.model()
    assert (p.rho_mean, b.rho_max, t.rate, fc.mu_scale) == (np.float32(0.95), np.float32(0.999),
                                                            np.float32(1.5), np.float32(0.5))
    f.close()
    g = smc2_rbpf.Smc2Rbpf(16, 64, prior=(0.9, 0.01, 0.1, 0.01, 0.5, 0.1, 0.2, 0.01))
    th = g.particles()["theta"]
    assert abs(th[0].mean() - 0.9) < 0.01 and abs(th[2].mean() - 0.5) < 0.1
    assert g.counters()["t_current"] == -1 and g.outer_ess() == pytest.approx(16.0)
    with pytest.raises(Exception, match="K_rejuv"):
        smc2_rbpf.Smc2Rbpf(8, 64, K_rejuv=-1)
    g.close()


def test_seed_zero_is_the_default_seed(orc):
    """smc2_cuda_set_seed(s, 0) selects the fixed seed 12345 (the reference
    seeds from the clock there)."""
    y, _, _ = orc.rb_dgp(20, *TRUTH, seed=3)
    say("rbpf seed 0")
    a = gpu_run(8, 64, y, seed=0, ess_thr=0.95, K=1)
    b = gpu_run(8, 64, y, seed=12345, ess_thr=0.95, K=1)
    assert np.array_equal(a["theta"], b["theta"]) and np.array_equal(a["inner"], b["inner"])
    c = gpu_run(8, 64, y, seed=7, ess_thr=0.95, K=1)
    assert not np.array_equal(a["inner"], c["inner"])


def test_parameter_recovery():
    """smc2/test_smc2_rbpf.cu:250-335: N_theta 1024, N_inner 512, T 1200,
    fixed lag 200; passes when at least 3 of the 4 learned parameters lie
    within 2 posterior standard deviations of the truth."""
    import oracle
    T = 1200
    y, _, _ = oracle.rb_dgp(T, *TRUTH, seed=12345)
    say("rbpf recovery", 1024, 512, T)
    f = smc2_rbpf.Smc2Rbpf(1024, 512, seed=12345, noise_capacity=T + 128, fixed_lag=200)
    f.run(y)
    m, s = f.theta_mean(), f.theta_std()
    c = f.counters()
    f.close()
    say("mean", m, "std", s, c)
    ok = sum(abs(m[q] - TRUTH[q]) <= 2 * s[q] for q in range(4))
    assert c["n_resamples"] >= 3 and c["n_rejuv_accepts"] > 0
    assert ok >= 3, (m, s)


def test_high_volatility_regime():
    """smc2/test_smc2_rbpf.cu:466-531: truth (0.98, 0.08, 2.0, 0.20), T 1000,
    widened prior (mu_base 0 +- 2, bounds [-5, 8]), fixed lag 200, seed 999;
    passes when at least 2 of 4 parameters lie within 2 posterior sd."""
    import oracle
    truth = (0.98, 0.08, 2.0, 0.20)
    y, _, _ = oracle.rb_dgp(1000, *truth, seed=999)
    say("rbpf high vol")
    prior = (0.96, 0.02, 0.10, 0.10, 0.0, 2.0, 0.15, 0.10)
    bounds = (0.8, 0.999, 0.01, 1.0, -5.0, 8.0, 0.01, 1.0)
    f = smc2_rbpf.Smc2Rbpf(1024, 512, seed=999, fixed_lag=200, prior=prior, bounds=bounds)
    f.run(y)
    m, s = f.theta_mean(), f.theta_std()
    f.close()
    say("mean", m, "std", s)
    assert sum(abs(m[q] - truth[q]) <= 2 * s[q] for q in range(4)) >= 2, (m, s)
    assert abs(m[2] - truth[2]) < 1.0



@pytest.mark.parametrize("nt, ni, T, lag", [(16, 64, 48, 0), (8, 256, 40, 6)])
def test_half_precision_noise_parity(orc, nt, ni, T, lag):
    """SMC2_NOISE_FP16 (smc2/smc2_noise_precision.cuh) as noise_bits = 16:
    histories stored as IEEE half, every noise used as stored; bit-exact with
    the oracle's half rounding, through resampling, moves and checkpoints."""
    say("rbpf fp16 noise", nt, ni, T, lag)
    y, _, _ = orc.rb_dgp(T, *TRUTH, seed=nt + 1)
    g = gpu_run(nt, ni, y, K=2, ess_thr=0.95, lag=lag, noise_bits=16)
    o = orc_run(orc, nt, ni, y, K=2, ess_thr=0.95, lag=lag, noise_bits=16)
    assert o["n_resamples"] >= 2
    assert_parity(g, o)
    f = gpu_run(nt, ni, y, K=2, ess_thr=0.95, lag=lag)        # FP32 histories differ
    assert not np.array_equal(f["inner"], g["inner"])
