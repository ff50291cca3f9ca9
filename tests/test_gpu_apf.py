"""GPU parity of the auxiliary particle filter (gpu_apf_*) against the CPU
oracle.  Reference: gpu_bpf/gpu_bpf_full.cuh:303-361,
gpu_bpf/gpu_bpf_ptx_full.cu:990-1190.  Bar: particle states bit-exact,
statistics within 1e-12 relative, BpfResult fields equal to the oracle's
float values within 1 ulp.
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
    fin = np.isfinite(b)
    assert np.array_equal(np.isfinite(a), fin)
    scale = np.maximum(np.abs(b[fin]), 1.0)
    assert np.max(np.abs(a[fin] - b[fin]) / scale) <= rtol


@pytest.mark.parametrize("n, nu_obs, nu_state", [(10000, 0.0, 0.0), (65536, 5.0, 0.0),
                                                 (2 ** 21 + 77, 0.0, 0.0), (8192, 0.0, 5.0),
                                                 (1, 0.0, 0.0), (257, 0.0, 0.0)])
def test_apf_parity(orc, n, nu_obs, nu_state):
    T = 30
    say("apf", n, nu_obs, nu_state)
    y = _y(orc, T, nu_obs)
    a = gpu_bpf.GpuApf(n, 0.97, 0.15, -1.0, nu_state=nu_state, nu_obs=nu_obs, seed=8)
    res, st = a.run(y)
    x = a.particles()
    tile, _ = a.layout()
    a.close()
    ref, xo = orc.run_apf(y, n, nu_obs=nu_obs, nu_state=nu_state, seed=8, tile=tile, want_x=True,
                          threads=8)
    assert np.array_equal(x, xo)
    for k in ("mean", "var", "ess", "ll_incr", "ll_cum", "ll_ref"):
        _close(st[k], ref[k])
    np.testing.assert_array_equal(res["h_mean"], st["h_mean"])
    ulp = np.spacing(np.abs(ref["log_lik"]).astype(np.float32))
    assert np.all(np.abs(res["log_lik"].astype(np.float64) - ref["log_lik"]) <= ulp)


def test_apf_step_equals_run(orc):
    T, n = 25, 20000
    say("apf step vs run")
    y = _y(orc, T)
    a = gpu_bpf.GpuApf(n, 0.97, 0.15, -1.0, seed=3)
    res, _ = a.run(y)
    b = gpu_bpf.GpuApf(n, 0.97, 0.15, -1.0, seed=3)
    for t in range(T):
        h, ll = b.step(float(y[t]))
        assert h == res["h_mean"][t] and ll == res["log_lik"][t]
    assert np.array_equal(a.particles(), b.particles())
    a.close()
    b.close()


def test_apf_run_rmse(orc):
    say("apf rmse")
    h, y = orc.dgp(400, 0.97, 0.15, -1.0, seed=42)
    r = gpu_bpf.run_rmse_apf(y, h, 20000, 0.97, 0.15, -1.0)
    a = gpu_bpf.GpuApf(20000, 0.97, 0.15, -1.0)
    res, _ = a.run(y.astype(np.float32))
    a.close()
    e = res["h_mean"][100:].astype(np.float64) - h[100:]
    assert abs(r - np.sqrt(np.mean(e * e))) < 1e-9
    assert 0.05 < r < 0.8


def test_apf_errors_and_empty_run():
    say("apf errors")
    with pytest.raises(gpu_bpf.GpuBpfError):
        gpu_bpf.GpuApf(0, 0.97, 0.15, -1.0)
    a = gpu_bpf.GpuApf(1000, 0.97, 0.15, -1.0)
    res, st = a.run(np.zeros(0, np.float32))
    assert len(res) == 0 and len(st) == 0
    a.close()
