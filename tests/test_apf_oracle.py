"""Oracle pins of the auxiliary particle filter (CPU).

Reference: gpu_bpf/gpu_bpf_ptx_full.cu:990-1047 (first stage, pair resampling,
propagate + correction) and :1102-1170 (the tick).  The reference draws its
noise with cuRAND, so its numbers are not reproducible here; the oracle's APF
runs on the bootstrap filter's PCG32 streams and is pinned by the mathematics:

* its likelihood estimate (first-stage mean weight times second-stage mean
  weight, per tick) is unbiased for p(y_1:T): the mean over seeds agrees with
  the exact point-mass (grid) likelihood within Monte-Carlo error;
* dropping the first-stage correction of the second-stage weight (a planted
  mistake) biases it far outside that error;
* tick 0 is the bootstrap filter's tick 0.
"""
import math

import numpy as np


def _y(orc, T, seed=42):
    _, y = orc.dgp(T, 0.97, 0.15, -1.0, seed=seed)
    return y.astype(np.float32)


def _apf_lls(orc, y, n=16384, seeds=range(1, 7)):
    return np.array([orc.run_apf(y, n, seed=s, threads=8)[0]["ll_cum"][-1] for s in seeds])


def test_apf_matches_exact_likelihood(orc):
    y = _y(orc, 100)
    g, _ = orc.grid_loglik(y, 0.97, 0.15, -1.0, n_grid=2048)
    v = _apf_lls(orc, y)
    se = v.std() / math.sqrt(len(v))
    assert abs(v.mean() - g) < 4 * se + 0.03, (v.mean(), g, se)


def test_planted_apf_mistake_is_caught(orc):
    y = _y(orc, 100)
    g, _ = orc.grid_loglik(y, 0.97, 0.15, -1.0, n_grid=2048)
    orc.lib().orc_set_fault(8)
    try:
        v = _apf_lls(orc, y, seeds=range(1, 3))
    finally:
        orc.lib().orc_set_fault(0)
    assert abs(v.mean() - g) > 1.0


def test_apf_tick0_is_the_bootstrap_tick0(orc):
    y = _y(orc, 1)
    a, xa = orc.run_apf(y, 5000, want_x=True)
    b, _, xb, _ = orc.run_filter(y, 5000, ess_threshold=0.0, want_x=True)
    assert np.array_equal(xa, xb)
    for k in ("mean", "var", "ess", "ll_incr", "ll_ref"):
        # same weights; the sums differ only in precision (80-bit vs FP64)
        np.testing.assert_allclose(a[k][0], b[k][0], rtol=1e-13, err_msg=k)


def test_apf_reference_loglik_is_the_second_stage(orc):
    """BpfResult.log_lik of an APF tick is log mean of the second-stage weights
    (gpu_bpf_ptx_full.cu:952-955); ll_incr adds the first stage."""
    y = _y(orc, 30)
    a, _ = orc.run_apf(y, 8192)
    assert np.all(np.abs(a["ll_incr"][1:] - a["ll_ref"][1:]) > 0)
    assert np.isfinite(a["ll_cum"]).all()
