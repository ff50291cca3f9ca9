"""Every filter option at once against the oracle: online learner (mu and rho),
adaptive sigma_z bands, Student-t state noise and the Silverman jitter.  The
reference applies them in one tick in this order: regime switch before
propagating, learner update after the tick's sums, then resampling and the
jitter (whose clamp uses the updated mu) -- gpu_bpf/gpu_bpf_ptx_full.cu:
550-748.  Bar: particles bit-exact, learned parameters and regimes equal.
"""
import numpy as np
import pytest

import gpu_bpf

pytestmark = pytest.mark.gpu

CALM = ([0.99], [1.0, 4.0])
ALERT = ([0.90, 0.05, 0.03], [1.0, 2.0, 4.0, 8.0])
PANIC = ([0.70, 0.15, 0.10], [1.0, 2.0, 5.0, 12.0])


@pytest.fixture(autouse=True)
def default_bands():
    gpu_bpf.reset_bands()
    yield
    gpu_bpf.reset_bands()


@pytest.mark.parametrize("nu_state, shrink", [(5.0, 0.5), (0.0, 1.0)])
def test_all_options(orc, nu_state, shrink):
    T, n = 80, 12000
    print("[case] all options", nu_state, shrink, flush=True)
    _, y = orc.dgp(T, 0.97, 0.15, -1.0, seed=42)
    y = y.astype(np.float32)
    f = gpu_bpf.GpuBpf(n, 0.95, 0.15, 0.0, nu_state=nu_state, seed=13, ess_threshold=0.5)
    tile, _ = f.layout()
    gpu_bpf.set_adaptive_bands(n, CALM, ALERT, PANIC, 0.5, 1.0)
    f.enable_mu_learning(1, 10, 0.1, 10.0, 0.667)
    f.enable_rho_learning(True)
    f.set_silverman_shrink(shrink)
    par = np.zeros((T, 2), np.float32)
    reg = np.zeros(T, np.int32)
    for t in range(T):
        f.step(float(y[t]))
        par[t] = (f.get_mu(), f.get_rho())
        reg[t] = f.band_regime()
    x = f.particles()
    ll = f.stats()["ll_cum"]
    f.close()
    ref, par_o, reg_o, xo = orc.run_filter_all(
        y, n, rho=0.95, mu=0.0, ess_threshold=0.5, seed=13, tile=tile,
        learn=dict(update_K=10, learn_rho=True), bands=(CALM, ALERT, PANIC), adaptive=True,
        thr_alert=0.5, thr_panic=1.0, nu_state=nu_state, silverman=shrink, want_x=True, threads=8)
    assert ref["resampled"].sum() > 0 and len(set(reg_o.tolist())) > 1
    np.testing.assert_array_equal(reg, reg_o)
    np.testing.assert_array_equal(par, par_o)
    assert np.array_equal(x, xo)
    assert abs(ll - ref["ll_cum"][-1]) <= 1e-12 * max(1.0, abs(ref["ll_cum"][-1]))
