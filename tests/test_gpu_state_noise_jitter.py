"""GPU parity of Student-t state noise (gpu_bpf_create with nu_state > 0) and
the Silverman jitter (gpu_bpf_set_silverman_shrink) against the CPU oracle.

References: gpu_bpf/gpu_bpf_ptx_full.cu:50-62,392-404,574-583 (chi2 state
noise), :708-739 (jitter); ptx_assembly/bpf_kernels_full.ptx:393-406,852-1113.
Bar: particle states bit-exact, statistics within 1e-12 relative.
"""
import numpy as np
import pytest

import gpu_bpf

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def _y(orc, T, nu=0.0):
    _, y = orc.dgp(T, 0.97, 0.15, -1.0, nu_obs=nu, seed=42)
    return y.astype(np.float32)


def _close(a, b, rtol=1e-12):
    scale = np.maximum(np.abs(b), 1.0)
    assert np.max(np.abs(a - b) / scale) <= rtol


@pytest.mark.parametrize("n, nu_state, nu_obs, thr", [
    (10000, 5.0, 0.0, 0.5),
    (8192, 3.0, 5.0, 0.0),
    (2 ** 21 + 77, 5.0, 0.0, 0.5),
])
def test_student_t_state_noise(orc, n, nu_state, nu_obs, thr):
    T = 20
    say("t state noise", n, nu_state, nu_obs, thr)
    y = _y(orc, T, nu_obs)
    f = gpu_bpf.GpuBpf(n, 0.97, 0.15, -1.0, nu_state=nu_state, nu_obs=nu_obs, seed=4, ess_threshold=thr)
    _, st = f.run(y)
    x = f.particles()
    tile, _ = f.layout()
    f.close()
    ref, _, xo = orc.run_filter_ext(y, n, nu_state=nu_state, nu_obs=nu_obs, ess_threshold=thr, seed=4,
                                    tile=tile, want_x=True, threads=8)
    assert np.array_equal(x, xo)
    np.testing.assert_array_equal(st["resampled"], ref["resampled"])
    _close(st["ll_cum"], ref["ll_cum"])


@pytest.mark.parametrize("n, shrink, thr, nu_state", [(8192, 0.5, 0.5, 0.0), (30000, 1.0, 0.0, 0.0),
                                                      (12000, 0.5, 0.5, 5.0)])
def test_silverman_jitter(orc, n, shrink, thr, nu_state):
    T = 40
    say("silverman", n, shrink, thr, nu_state)
    y = _y(orc, T)
    f = gpu_bpf.GpuBpf(n, 0.97, 0.15, -1.0, nu_state=nu_state, seed=6, ess_threshold=thr)
    f.set_silverman_shrink(shrink)
    _, st = f.run(y)
    x = f.particles()
    tile, _ = f.layout()
    f.close()
    ref, cnt, xo = orc.run_filter_ext(y, n, nu_state=nu_state, silverman=shrink, ess_threshold=thr, seed=6,
                                      tile=tile, want_x=True, threads=8)
    assert cnt > 0
    assert np.array_equal(x, xo)
    np.testing.assert_array_equal(st["resampled"], ref["resampled"])
    _close(st["ll_cum"], ref["ll_cum"])


def test_bad_nu_state_and_guards(orc):
    say("nu_state guards")
    with pytest.raises(gpu_bpf.GpuBpfError):
        gpu_bpf.GpuBpf(1024, 0.97, 0.15, -1.0, nu_state=0.3)
    with pytest.raises(gpu_bpf.GpuBpfError):
        gpu_bpf.GpuBpf(1024, 0.97, 0.15, -1.0, nu_state=100.0)
    y = _y(orc, 10)
    g = gpu_bpf.GpuBpf(4096, 0.97, 0.15, -1.0, nu_state=5.0)
    with pytest.raises(gpu_bpf.GpuBpfError):
        g.score_forward(y)
    g.close()
    # the learner runs with Student-t state noise (kernel variant with both)
    f = gpu_bpf.GpuBpf(4096, 0.97, 0.15, 0.0, nu_state=5.0, seed=2)
    f.enable_mu_learning(1, 5, 0.1, 10.0, 0.667)
    for t in range(10):
        f.step(float(y[t]))
    s, k = f.learn_sums()
    assert k == 0 and np.all(np.isfinite(s))
    assert np.isfinite(f.get_mu()) and f.get_mu() != 0.0
    f.close()
