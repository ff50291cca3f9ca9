"""GPU parity of the SV-SMC^2 path (gpu_smc2_*) against the oracle's
restatement of the whole algorithm (oracle/smc2_oracle.c, pinned in
tests/test_smc2_oracle.py): inner filters, outer reweighting, the outer ESS
halt, systematic resampling of whole filters, and CPMMH / MALA moves.

Bar: theta of every theta-particle and its particle count bit for bit, the
counters and the outer-resample history exactly (so every ESS decision,
ancestor and MH decision agreed), outer log-weights and inner FP64
log-likelihoods within 1e-12 relative, MALA scores within 1e-9.  M <= 64,
so the oracle runs in seconds."""
import numpy as np
import pytest

import gpu_bpf_batch
import sv_data

pytestmark = pytest.mark.gpu

CASES = {
    # name: (M, n_min, n_max, T, kwargs)
    "cpmmh_gauss": (32, 128, 1024, 150, dict(kind="cpmmh", moves=3, outer_threshold=0.8, seed=3)),
    "cpmmh_settings": (17, 64, 2000, 120, dict(kind="cpmmh", moves=2, outer_threshold=0.9, seed=8,
                                               cpmmh_rho=0.6, u_step=0.2, theta_scale=1.5)),
    "mala_student_t": (32, 128, 1024, 150, dict(kind="mala", moves=3, outer_threshold=0.8, seed=5,
                                                nu_obs=5.0)),
    "mala_gauss_m64": (64, 256, 2048, 100, dict(kind="mala", moves=2, outer_threshold=0.7, seed=11,
                                                mala_step=0.3)),
    # inner tiles of 512 and 1024 particles (the oracle's tiled weights with the same tile)
    "cpmmh_tile512": (24, 128, 3000, 120, dict(kind="cpmmh", moves=3, outer_threshold=0.8, seed=13,
                                               tile=512)),
    "mala_student_t_tile1024": (16, 300, 4000, 100, dict(kind="mala", moves=2, outer_threshold=0.8,
                                                         seed=17, nu_obs=5.0, tile=1024)),
}


def say(*a):
    print("[case]", *a, flush=True)


@pytest.mark.parametrize("name", list(CASES))
def test_smc2_equals_oracle(orc, name):
    M, nmin, nmax, T, kw = CASES[name]
    nu = kw.get("nu_obs", 0.0)
    y = sv_data.sv_series(T, 0.97, 0.15, -1.0, nu_obs=nu, seed=21)[1].astype(np.float32)
    say("smc2 vs oracle", name)
    tile = kw.get("tile", 256)
    g = gpu_bpf_batch.GpuSmc2(M, nmin, nmax, capacity=T, block=256 if tile < 1024 else 128, **kw)
    g.run(y)
    th, lw, n = g.theta()
    ll = g.loglik()
    cnt = g.counters()
    hist = g.history()
    ess = g.outer_ess()
    sc = g.score() if kw["kind"] == "mala" else None
    g.close()
    okw = {k: v for k, v in kw.items() if k != "cpmmh_rho"}
    okw["cn_rho"] = kw.get("cpmmh_rho", 0.99)
    r = orc.run_sv_smc2(y, M, nmin, nmax, **okw)
    assert cnt["outer_resamples"] >= 2, cnt
    assert cnt == r["counters"]
    np.testing.assert_array_equal(hist[0], r["history"][0])
    np.testing.assert_array_equal(hist[1], r["history"][1])
    np.testing.assert_array_equal(n, r["n"])
    np.testing.assert_array_equal(th, r["theta"])
    for a, b in ((lw, r["logw"]), (ll, r["ll"])):
        assert np.max(np.abs(a - b) / np.maximum(np.abs(b), 1.0)) <= 1e-12
    want_ess = M if len(hist[0]) and hist[0][-1] == T else r["ess"][-1]
    assert abs(ess - want_ess) <= 1e-6 * want_ess
    if sc is not None:
        scale = np.max(np.abs(r["score"]), axis=0)
        assert np.all(np.abs(sc - r["score"]) <= 1e-9 * scale)
