"""GPU parity of the online mu/rho learner (gpu_bpf_enable_mu_learning & co.)
against the CPU oracle.

Reference: gpu_bpf/gpu_bpf_ptx_full.cu:469-537 (accumulate every tick, update
every K ticks), :777-831 (enable / disable / set), kernel 14
(ptx_assembly/bpf_kernels_full.ptx:1116-1384).

Bar: the per-tick learner sums within 1e-10 relative of the oracle's (FP64
sums in another order); the learned (mu, rho) after every tick equal to the
oracle's bit for bit (the update is the same FP32 host arithmetic on the FP64
sums rounded to FP32), and then the particle states bit-exact as well.
"""
import numpy as np
import pytest

import gpu_bpf

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def _y(orc, T, nu=0.0, mu=-1.0):
    _, y = orc.dgp(T, 0.97, 0.15, mu, nu_obs=nu, seed=42)
    return y.astype(np.float32)


def _close(a, b, rtol):
    scale = np.maximum(np.abs(b), 1e-3)
    err = np.max(np.abs(a - b) / scale)
    assert err <= rtol, err


@pytest.mark.parametrize("n, nu, thr", [(5000, 0.0, 0.5), (8192, 5.0, 0.0), (3 * 2 ** 20 + 5, 0.0, 0.5)])
def test_learner_sums_per_tick(orc, n, nu, thr):
    """Learning on with no update due: the accumulated sums after each tick."""
    T = 12
    say("learn sums", n, nu, thr)
    y = _y(orc, T, nu)
    f = gpu_bpf.GpuBpf(n, 0.97, 0.15, -0.6, nu_obs=nu, seed=7, ess_threshold=thr)
    tile, _ = f.layout()
    f.enable_mu_learning(1, 10**9, 0.1, 10.0, 0.667)
    got = np.zeros((T, 4))
    for t in range(T):
        f.step(float(y[t]))
        got[t], k = f.learn_sums()
        assert k == t + 1
    x = f.particles()
    f.close()
    t_ref, ls, _, x_ref = orc.run_filter_learn(y, n, mu=-0.6, nu_obs=nu, ess_threshold=thr, seed=7,
                                               tile=tile, update_K=10**9, want_x=True, threads=8)
    assert np.array_equal(x, x_ref)
    _close(got, np.cumsum(ls, axis=0), 1e-10)


@pytest.mark.parametrize("learn_rho, nu, thr, K", [(False, 0.0, 0.5, 10), (True, 0.0, 0.0, 7),
                                                 (True, 5.0, 0.5, 1)])
def test_learned_parameters_bitexact(orc, learn_rho, nu, thr, K):
    T, n = 90, 20000
    say("learn update", learn_rho, nu, thr, K)
    y = _y(orc, T, nu)
    f = gpu_bpf.GpuBpf(n, 0.95, 0.15, 0.0, nu_obs=nu, seed=3, ess_threshold=thr)
    tile, _ = f.layout()
    f.enable_mu_learning(1, K, 0.1, 10.0, 0.667)
    if learn_rho:
        f.enable_rho_learning(True)
    par = np.zeros((T, 2), np.float32)
    for t in range(T):
        f.step(float(y[t]))
        par[t] = (f.get_mu(), f.get_rho())
    x = f.particles()
    f.close()
    _, _, par_ref, x_ref = orc.run_filter_learn(y, n, rho=0.95, mu=0.0, nu_obs=nu, ess_threshold=thr,
                                                seed=3, tile=tile, update_K=K, learn_rho=learn_rho,
                                                want_x=True, threads=8)
    assert par[-1, 0] != np.float32(0.0)               # the learner moved mu
    np.testing.assert_array_equal(par, par_ref)
    assert np.array_equal(x, x_ref)


def test_run_equals_stepping(orc):
    """gpu_bpf_run (observations in device memory) learns as tick-by-tick stepping does."""
    T, n = 60, 10000
    say("learn run vs step")
    y = _y(orc, T)
    a = gpu_bpf.GpuBpf(n, 0.97, 0.15, 0.0, seed=5, ess_threshold=0.5)
    a.enable_mu_learning(1, 8, 0.1, 10.0, 0.667)
    _, st_a = a.run(y)
    b = gpu_bpf.GpuBpf(n, 0.97, 0.15, 0.0, seed=5, ess_threshold=0.5)
    b.enable_mu_learning(1, 8, 0.1, 10.0, 0.667)
    for t in range(T):
        b.step(float(y[t]))
    assert a.get_mu() == b.get_mu()
    assert np.array_equal(a.particles(), b.particles())
    assert st_a["ll_cum"][-1] == b.stats()["ll_cum"]
    a.close()
    b.close()


def test_set_and_disable(orc):
    y = _y(orc, 20)
    say("learn set/disable")
    f = gpu_bpf.GpuBpf(4096, 0.97, 0.15, 0.0, seed=1)
    f.enable_mu_learning(1, 100, 0.1, 10.0, 0.667)
    for t in range(5):
        f.step(float(y[t]))
    s, k = f.learn_sums()
    assert k == 5 and np.all(np.isfinite(s)) and s[1] > 0
    f.set_mu(-1.0)                                     # restarts the accumulators
    s, k = f.learn_sums()
    assert k == 0 and np.all(s == 0.0)
    f.set_rho(1.5)
    assert f.get_rho() == np.float32(0.999)
    # the score pair refuses to run while learning is on
    g = gpu_bpf.GpuBpf(4096, 0.97, 0.15, -1.0, seed=1)
    g.enable_mu_learning(1, 10, 0.1, 10.0, 0.667)
    with pytest.raises(gpu_bpf.GpuBpfError):
        g.score_forward(y)
    g.disable_mu_learning()
    g.score_forward(y)
    g.close()
    f.disable_mu_learning()
    for t in range(5, 20):
        f.step(float(y[t]))
    assert f.get_mu() == np.float32(-1.0)
    f.close()
