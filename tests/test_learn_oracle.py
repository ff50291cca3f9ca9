"""Oracle pins of the online mu/rho learner (CPU).

Reference: kernel 14 `bpf_grad_alpha` (ptx_assembly/bpf_kernels_full.ptx:1116-1384)
accumulates sum w g, sum w g^2, sum w g d, sum w g^2 d^2 with g = d lw/dx of the
propagated particle and d = x_prev - mu; `bpf_grad_and_update`
(gpu_bpf/gpu_bpf_ptx_full.cu:469-537) turns them into a natural-gradient
Robbins-Monro step every K ticks.  The reference holds no vectors for either, so:

* the sums are checked against their definition evaluated in FP64 NumPy on the
  oracle's own particle states, and against an exact identity with the pathwise
  score (itself pinned by finite differences and by the torch restatement of
  tests/diff_bpf.py): with two ticks, x_0 held fixed and tick 0 always
  resampled, dLL/dmu = (1 - phi) sum w g and dLL/dphi = sum w g d at tick 1;
* the update is checked in FP64 against the reference's formula on chosen sums
  that exercise each branch (gate ramp, clip, clamps, rho);
* planted mistakes in the oracle (the 1/2 of eta^2, d from the wrong state)
  fail the identity.
"""
import math

import numpy as np
import pytest


def _y(orc, T, nu=0.0):
    _, y = orc.dgp(T, 0.97, 0.15, -1.0, nu_obs=nu, seed=42)
    return y.astype(np.float32)


@pytest.mark.parametrize("nu", [0.0, 5.0])
@pytest.mark.parametrize("theta", [(0.97, 0.15, -1.0), (0.9, 0.35, -1.7)])
def test_sums_equal_pathwise_score_at_tick_1(orc, nu, theta):
    phi, sig, mu = theta
    y = _y(orc, 2, nu)
    _, ls, _, _ = orc.run_filter_learn(y, 4096, rho=phi, sigma_z=sig, mu=mu, nu_obs=nu,
                                       update_K=10**9)
    _, sc, _, _ = orc.score(y, 4096, rho=phi, sigma_z=sig, mu=mu, nu_obs=nu, x0_reparam=False)
    assert ls[0, 2] == 0.0 and ls[0, 3] == 0.0          # tick 0 propagates nothing: d = 0
    np.testing.assert_allclose((1.0 - phi) * ls[1, 0], sc[0], rtol=2e-5)
    np.testing.assert_allclose(ls[1, 2], sc[1], rtol=2e-5)


@pytest.mark.parametrize("fault, nu, col", [(6, 0.0, 0), (7, 0.0, 2), (7, 5.0, 2)])
def test_planted_learner_mistakes_fail_the_identity(orc, fault, nu, col):
    phi, sig, mu = 0.9, 0.35, -1.7
    y = _y(orc, 2, nu)
    _, sc, _, _ = orc.score(y, 4096, rho=phi, sigma_z=sig, mu=mu, nu_obs=nu, x0_reparam=False)
    orc.lib().orc_set_fault(fault)
    try:
        _, ls, _, _ = orc.run_filter_learn(y, 4096, rho=phi, sigma_z=sig, mu=mu, nu_obs=nu,
                                           update_K=10**9)
    finally:
        orc.lib().orc_set_fault(0)
    got = (1.0 - phi) * ls[1, 0] if col == 0 else ls[1, 2]
    want = sc[0] if col == 0 else sc[1]
    assert abs(got - want) > 1e-3 * abs(want)


def _lw64(x, y, nu):
    eta2 = y * y * np.exp(-x)
    if nu > 0:
        c = math.lgamma((nu + 1) / 2) - math.lgamma(nu / 2) - 0.5 * math.log(nu * math.pi)
        return c - 0.5 * (nu + 1) * np.log1p(eta2 / nu) - 0.5 * x
    return -0.5 * math.log(2 * math.pi) - 0.5 * eta2 - 0.5 * x


def _g64(x, y, nu):
    eta2 = y * y * np.exp(-x)
    if nu > 0:
        return -0.5 + 0.5 * (nu + 1) * eta2 / (nu + eta2)
    return -0.5 + 0.5 * eta2


@pytest.mark.parametrize("nu", [0.0, 5.0])
def test_sums_match_their_definition(orc, nu):
    """Ticks 0 and 1 without resampling (carried weights): the four sums from
    the oracle's particle states, evaluated in FP64."""
    n, phi, sig, mu = 1000, 0.95, 0.25, -1.2
    y = _y(orc, 2, nu)
    thr = 1e-30       # positive and never reached: no resampling
    kw = dict(rho=phi, sigma_z=sig, mu=mu, nu_obs=nu, ess_threshold=thr, tile=256)
    _, _, _, x0 = orc.run_filter_learn(y[:1], n, update_K=10**9, want_x=True, **kw)
    t, ls, _, x1 = orc.run_filter_learn(y, n, update_K=10**9, want_x=True, **kw)
    assert t["resampled"].sum() == 0
    x0 = x0.astype(np.float64)
    x1 = x1.astype(np.float64)
    y0, y1 = float(y[0]), float(y[1])
    # tick 0: weights of x_0, d = 0
    l0 = _lw64(x0, y0, nu)
    w0 = np.exp(l0 - l0.max())
    w0 /= w0.sum()
    g0 = _g64(x0, y0, nu)
    np.testing.assert_allclose(ls[0, :2], [np.sum(w0 * g0), np.sum(w0 * g0 * g0)], rtol=1e-5)
    # tick 1: carried weights, d = x_0 - mu
    l1 = l0 + _lw64(x1, y1, nu)
    w1 = np.exp(l1 - l1.max())
    w1 /= w1.sum()
    g1 = _g64(x1, y1, nu)
    d = x0 - mu
    want = [np.sum(w1 * g1), np.sum(w1 * g1 * g1), np.sum(w1 * g1 * d), np.sum(w1 * g1 * g1 * d * d)]
    np.testing.assert_allclose(ls[1], want, rtol=1e-5, atol=1e-7)
    # weighted Cauchy-Schwarz
    assert ls[1, 1] >= ls[1, 0] ** 2 and ls[1, 3] >= ls[1, 2] ** 2


def _update64(sums, k, step, mu, rho, learn_rho=True, c=0.1, t0=10.0, gamma=0.667, clip=5.0):
    """The reference's update (gpu_bpf/gpu_bpf_ptx_full.cu:488-537) in FP64."""
    step += 1
    eta = c / (step + t0) ** gamma

    def direction(g, f):
        mg, mf = g / k, f / k
        dr = mg / max(mf, 1e-8)
        if clip > 0 and abs(dr) > clip:
            dr = math.copysign(clip, dr)
        gate = min(abs(mg) / max(math.sqrt(mf / k), 1e-10) / 2.0, 1.0)
        return gate * dr

    alpha = mu * (1 - rho) + eta * direction(sums[0], sums[1])
    mu_new = min(max(alpha / max(1 - rho, 1e-6), -5.0), 2.0)
    rho_new = rho
    if learn_rho:
        rho_new = min(max(rho + eta * direction(sums[2], sums[3]), 0.001), 0.999)
    return step, mu_new, rho_new


@pytest.mark.parametrize("case", [
    # (sums, k, step, mu, rho): gate 1 (snr 3), gate ramp (snr 0.5), clip, clamps
    ((15.0, 25.0, -6.0, 40.0), 50, 0, -1.0, 0.97),
    ((2.5, 25.0, 0.5, 4.0), 50, 0, -1.0, 0.97),
    ((500.0, 50.0, -500.0, 50.0), 50, 3, -0.5, 0.9),
    ((5000.0, 50.0, 5000.0, 50.0), 50, 0, 1.9, 0.998),
    ((-5000.0, 50.0, -5000.0, 50.0), 50, 0, -4.9, 0.002),
    ((0.0, 25.0, 0.0, 4.0), 50, 7, -1.3, 0.95),
    ((3.0, 2.0, 1.0, 1.0), 1, 0, -1.0, 0.97),
])
@pytest.mark.parametrize("learn_rho", [False, True])
def test_update_matches_reference_formula(orc, case, learn_rho):
    sums, k, step, mu, rho = case
    big_c = 50.0 if abs(sums[0]) >= 5000 else 0.1     # drive the clamps
    got = orc.learn_update(sums, k, step, mu, rho, learn_rho=learn_rho, c=big_c)
    want = _update64(sums, k, step, mu, rho, learn_rho=learn_rho, c=big_c)
    assert got[0] == want[0]
    np.testing.assert_allclose(got[1], want[1], rtol=2e-6, atol=2e-6)
    np.testing.assert_allclose(got[2], want[2] if learn_rho else np.float32(rho), rtol=2e-6, atol=2e-6)


def test_no_update_leaves_the_filter_unchanged(orc):
    y = _y(orc, 40)
    ref, _, x_ref, _ = orc.run_filter(y, 4096, mu=-0.4, ess_threshold=0.5, want_x=True)
    t, _, par, x = orc.run_filter_learn(y, 4096, mu=-0.4, ess_threshold=0.5, update_K=10**9,
                                        want_x=True)
    assert np.array_equal(t["ll_cum"], ref["ll_cum"]) and np.array_equal(x, x_ref)
    assert np.all(par[:, 0] == np.float32(-0.4))


def test_learner_moves_mu_towards_the_truth(orc):
    """Started at mu = 0 on data from mu = -1 (the setting of
    smc2/test_mu_learning.cu's phase 1), the learned mu ends much closer."""
    y = _y(orc, 600)
    _, _, par, _ = orc.run_filter_learn(y, 4096, mu=0.0, update_K=50, ess_threshold=0.5)
    assert abs(par[-1, 0] - (-1.0)) < 0.5 * abs(0.0 - (-1.0))
    assert np.all(par[:, 1] == np.float32(0.97))        # rho not learned
