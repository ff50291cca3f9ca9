"""SMC^2 over the batched inner filters, through the C ABI.

Pinned by reductions with exact answers: without outer resampling the
outer log-weights are the inner filters' FP64 log-likelihoods; a replay of
the stored noise with identity proposals (rho = 1, zero steps) reproduces
every filter bit for bit, so a run with such moves equals the run without
them; and on data from a known theta the posterior concentrates near it.
"""
import numpy as np
import pytest

import gpu_bpf_batch

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def ydata(orc):
    _, y = orc.dgp(600, 0.97, 0.15, -1.0, seed=7)
    return y.astype(np.float32)


def test_without_resampling_outer_weights_are_the_inner_logliks(ydata):
    say("no outer resampling")
    s = gpu_bpf_batch.GpuSmc2(64, 256, 2048, seed=3, outer_threshold=0.0, capacity=128)
    s.run(ydata[:100])
    th, lw, n = s.theta()
    ll = s.loglik()
    np.testing.assert_array_equal(lw, ll)
    c = s.counters()
    assert c["ticks"] == 100 and c["outer_resamples"] == 0
    assert c["pt_forward"] == 100 * int(n.sum())
    assert np.all((n >= 256) & (n <= 2048))
    s.close()


def test_identity_moves_replay_the_stored_noise_exactly(ydata):
    """rho = 1, zero theta and uniform steps: every replay reproduces the
    slot's filter, every move is accepted, and the run equals one without
    moves (the outer layer's random stream is separate from the moves')."""
    y = ydata[:90]
    say("identity moves")
    kw = dict(n_theta=48, n_min=256, n_max=4096, seed=11, outer_threshold=0.8, capacity=128)
    a = gpu_bpf_batch.GpuSmc2(moves=2, cpmmh_rho=1.0, theta_scale=0.0, u_step=0.0, **kw)
    b = gpu_bpf_batch.GpuSmc2(moves=0, **kw)
    a.run(y)
    b.run(y)
    ca, cb = a.counters(), b.counters()
    assert ca["outer_resamples"] >= 1 and ca["outer_resamples"] == cb["outer_resamples"]
    assert ca["accepted"] == ca["proposed"] > 0
    np.testing.assert_array_equal(a.loglik(), b.loglik())
    ta, la, na = a.theta()
    tb, lb, nb = b.theta()
    np.testing.assert_array_equal(ta, tb)
    np.testing.assert_array_equal(la, lb)
    np.testing.assert_array_equal(na, nb)
    a.close()
    b.close()


def test_cpmmh_posterior_concentrates(ydata):
    say("posterior")
    s = gpu_bpf_batch.GpuSmc2(256, 256, 1024, seed=5, capacity=512)
    s.run(ydata[:500])
    mean, sd = s.theta_mean_std()
    c = s.counters()
    assert c["outer_resamples"] >= 3
    assert 0.05 < c["accepted"] / c["proposed"] < 0.95
    assert abs(mean[0] - (-1.0)) < 3 * sd[0] + 0.3
    assert sd[0] < 0.4 and sd[2] < 0.12        # well inside the prior's spread
    s.close()


def test_update_returns_outer_ess(ydata):
    say("update")
    s = gpu_bpf_batch.GpuSmc2(32, 256, 512, seed=2, capacity=64)
    for t in range(20):
        ess = s.update(float(ydata[t]))
        assert 0.0 < ess <= 32.0 + 1e-6
    assert s.counters()["ticks"] == 20
    s.close()


def test_mala_rejuvenation_concentrates(ydata):
    """Score-driven (MALA) rejuvenation from the batch's per-filter pathwise
    scores, Student-t observations as the heavy-tailed configuration."""
    say("mala")
    s = gpu_bpf_batch.GpuSmc2(128, 512, 2048, nu_obs=5.0, seed=9, kind="mala", moves=3,
                              mala_step=0.5)
    s.run(ydata[:300])
    mean, sd = s.theta_mean_std()
    c = s.counters()
    assert c["outer_resamples"] >= 2
    assert 0.05 < c["accepted"] / c["proposed"] < 0.99
    assert np.all(np.isfinite(mean)) and sd[0] < 0.5
    th, lw, n = s.theta()
    assert np.all((n >= 512) & (n <= 2048))
    s.close()
