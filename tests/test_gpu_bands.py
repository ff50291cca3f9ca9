"""GPU parity of the sigma_z scale bands (gpu_bpf_set_bands,
gpu_bpf_set_adaptive_bands) against the CPU oracle.

Reference: gpu_bpf/gpu_bpf_ptx_full.cu:257-334 and :585-588,767-769,
ptx_assembly/bpf_kernels_full.ptx:297-324.  Bar: particle states bit-exact,
statistics within 1e-12 relative, the adaptive regime of every tick equal.
The band configuration is process-wide, as in the reference; every test
restores the default (one band, scale 1) when it ends.
"""
import numpy as np
import pytest

import gpu_bpf

pytestmark = pytest.mark.gpu

CALM = ([0.99], [1.0, 4.0])
ALERT = ([0.90, 0.05, 0.03], [1.0, 2.0, 4.0, 8.0])
PANIC = ([0.70, 0.15, 0.10], [1.0, 2.0, 5.0, 12.0])


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(autouse=True)
def default_bands():
    gpu_bpf.reset_bands()
    yield
    gpu_bpf.reset_bands()


def _y(orc, T, seed=42):
    _, y = orc.dgp(T, 0.97, 0.15, -1.0, seed=seed)
    return y.astype(np.float32)


def _close(a, b, rtol=1e-12):
    scale = np.maximum(np.abs(b), 1.0)
    assert np.max(np.abs(a - b) / scale) <= rtol


@pytest.mark.parametrize("n, cfg, thr", [
    (10000, ([0.9], [1.0, 4.0]), 0.5),                         # boundary inside a tile
    (3 * 4096, ([2.0 / 3.0], [1.0, 3.0]), 0.0),               # boundary on a tile edge
    (2 ** 21 + 77, ([0.7, 0.15, 0.10], [1.0, 2.0, 5.0, 12.0]), 0.5),
    (5000, ([0.5], [2.5]), 0.0),                               # one band, scale 2.5
])
def test_static_bands(orc, n, cfg, thr):
    T = 25
    say("static bands", n, cfg, thr)
    y = _y(orc, T)
    f = gpu_bpf.GpuBpf(n, 0.97, 0.15, -1.0, seed=9, ess_threshold=thr)
    gpu_bpf.set_bands(n, *cfg)                 # after create: adopted at the next tick
    _, st = f.run(y)
    x = f.particles()
    tile, _ = f.layout()
    f.close()
    ref, _, xo = orc.run_filter_bands(y, n, cfg, ess_threshold=thr, seed=9, tile=tile, want_x=True,
                                      threads=8)
    assert np.array_equal(x, xo)
    np.testing.assert_array_equal(st["resampled"], ref["resampled"])
    _close(st["ll_cum"], ref["ll_cum"])


def test_adaptive_bands_stepping(orc):
    T, n = 200, 20000
    say("adaptive bands")
    y = _y(orc, T, seed=11)
    f = gpu_bpf.GpuBpf(n, 0.97, 0.15, -1.0, seed=5, ess_threshold=0.5)
    tile, _ = f.layout()
    gpu_bpf.set_adaptive_bands(n, CALM, ALERT, PANIC, 0.5, 1.0)
    reg = np.zeros(T, np.int32)
    ll = np.zeros(T)
    for t in range(T):
        f.step(float(y[t]))
        reg[t] = f.band_regime()
        ll[t] = f.stats()["ll_cum"]
    x = f.particles()
    f.close()
    ref, reg_o, xo = orc.run_filter_bands(y, n, (CALM, ALERT, PANIC), adaptive=True, thr_alert=0.5,
                                          thr_panic=1.0, ess_threshold=0.5, seed=5, tile=tile,
                                          want_x=True, threads=8)
    assert set(np.unique(reg_o)) == {0, 1, 2}
    np.testing.assert_array_equal(reg, reg_o)
    assert np.array_equal(x, xo)
    _close(ll, ref["ll_cum"])


def test_adaptive_run_equals_stepping(orc):
    T, n = 80, 8192
    say("adaptive run vs step")
    y = _y(orc, T, seed=11)
    gpu_bpf.set_adaptive_bands(n, CALM, ALERT, PANIC, 0.5, 1.0)
    a = gpu_bpf.GpuBpf(n, 0.97, 0.15, -1.0, seed=3, ess_threshold=0.5)
    _, st = a.run(y)
    b = gpu_bpf.GpuBpf(n, 0.97, 0.15, -1.0, seed=3, ess_threshold=0.5)
    for t in range(T):
        b.step(float(y[t]))
    assert a.band_regime() == b.band_regime()
    assert np.array_equal(a.particles(), b.particles())
    assert st["ll_cum"][-1] == b.stats()["ll_cum"]
    a.close()
    b.close()


def test_reset_and_score_guard(orc):
    y = _y(orc, 20)
    say("bands reset / score guard")
    ref = gpu_bpf.GpuBpf(4096, 0.97, 0.15, -1.0, seed=2)
    _, st_ref = ref.run(y)
    x_ref = ref.particles()
    ref.close()
    gpu_bpf.set_bands(4096, [0.5], [1.0, 3.0])
    g = gpu_bpf.GpuBpf(4096, 0.97, 0.15, -1.0, seed=2)
    with pytest.raises(gpu_bpf.GpuBpfError):
        g.score_forward(y)                     # the score needs one sigma for all particles
    g.close()
    gpu_bpf.reset_bands(4096)
    f = gpu_bpf.GpuBpf(4096, 0.97, 0.15, -1.0, seed=2)
    _, st = f.run(y)
    assert np.array_equal(f.particles(), x_ref)
    assert np.array_equal(st["ll_cum"], st_ref["ll_cum"])
    f.close()
