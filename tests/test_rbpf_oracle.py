"""Oracle pins of the SMC^2 learner with the RBPF / OCSN inner filter (CPU).

The reference (smc2/smc2_rbpf_batch.cu) draws its noise with cuRAND, which is
not under the reference tree, and its tests check parameter recovery only
(smc2/test_smc2_rbpf.cu:250-335), so there are no golden vectors.  The
oracle (oracle/rbpf_oracle.c) is pinned by what the mathematics fixes:

* the OCSN table is the published mixture of log chi^2_1: weights sum to 1,
  mean psi(1/2) + log 2 = -1.2704, variance pi^2 / 2;
* the FP32 OCSN update equals the FP64 moment-matched mixture update;
* with sigma_z = 0 every inner particle is the same Kalman filter, so the
  inner log-likelihood is the FP64 assumed-density filter's, tick by tick;
* the tanh / erfc / Phi restatements match NumPy / SciPy;
* CPMMH with fully correlated noise and a zero random walk proposes the
  current theta with the current noise: every move is accepted and changes
  no state (the replay repeats the forward run bit for bit);
* without resampling, theta-particle 0 of a batched run is the single inner
  filter run of its theta;
* prior draws follow the bounds-truncated prior.
"""
import math

import numpy as np
import pytest
from scipy import special, stats

TRUTH = (0.95, 0.10, -1.0, 0.15)


def test_ocsn_table_is_log_chi2(orc):
    p, m, v = (a.astype(np.float64) for a in orc.rb_ocsn_table())
    assert abs(p.sum() - 1.0) < 1e-4
    mean = np.sum(p * m)
    var = np.sum(p * (v + m * m)) - mean ** 2
    assert abs(mean - (special.digamma(0.5) + math.log(2.0))) < 2e-3
    assert abs(var - math.pi ** 2 / 2) < 2e-2
    # and the density is close to that of log chi^2_1
    x = np.linspace(-12.0, 3.0, 301)
    mix = np.sum(p[:, None] * stats.norm.pdf(x[None], m[:, None], np.sqrt(v)[:, None]), axis=0)
    exact = stats.chi2.pdf(np.exp(x), 1) * np.exp(x)
    assert np.max(np.abs(mix - exact)) < 5e-3


def _ocsn_fp64(p, m, v, y, mp, vp):
    S = vp + v
    lc = np.log(p) + stats.norm.logpdf(y, mp + m, np.sqrt(S))
    ln = special.logsumexp(lc)
    w = np.exp(lc - ln)
    K = vp / S
    mk = mp + K * (y - mp - m)
    vk = (1.0 - K) * vp
    mu = np.sum(w * mk)
    return mu, np.sum(w * (vk + mk * mk)) - mu * mu, ln


def test_ocsn_update_matches_fp64(orc):
    p, m, v = (a.astype(np.float64) for a in orc.rb_ocsn_table())
    rng = np.random.default_rng(5)
    for _ in range(400):
        y = rng.uniform(-14.0, 3.0)
        mp = rng.uniform(-3.0, 1.5)
        vp = 10 ** rng.uniform(-3.0, 0.7)
        mu, var, ln = orc.rb_ocsn(y, mp, vp)
        mu_r, var_r, ln_r = _ocsn_fp64(p, m, v, np.float64(np.float32(y)), np.float64(np.float32(mp)),
                                       np.float64(np.float32(vp)))
        assert abs(ln - ln_r) < 2e-5 * (1.0 + abs(ln_r))
        assert abs(mu - mu_r) < 2e-5 * (1.0 + abs(mu_r))
        assert abs(var - var_r) < 2e-5 * (1.0 + mu_r * mu_r)


def _adf_fp64(y, th, cfg_t, fixed):
    """Kalman filter with the OCSN moment-matched update, FP64, one regime."""
    rho, sz, mub, sgb = th
    tb, ts, tr = cfg_t
    ms, mr, ss, sr = fixed
    z = 1.5

    def curve(b, s, r):
        return b + s * (1.0 - math.exp(-r * z))

    t, muz, sh = curve(tb, ts, tr), curve(mub, ms, mr), curve(sgb, ss, sr)
    phi = 1.0 - t
    mu, var = muz, sh * sh / max(1.0 - phi * phi, 1e-6)
    return mu, var, phi, t, muz, sh


def test_single_regime_is_the_assumed_density_filter(orc):
    p, m, v = (a.astype(np.float64) for a in orc.rb_ocsn_table())
    y, _, _ = orc.rb_dgp(300, seed=3)
    th = np.float32([0.95, 0.0, -1.0, 0.15])          # sigma_z = 0: z~ stays 0
    cfg = orc.rb_cfg(2, 128)
    lli, ess, st = orc.rb_inner_run(cfg, th, y)
    assert np.all(ess == 128.0)
    assert np.all(st[0] == 0.0) and np.all(st[1] == st[1][0]) and np.all(st[2] == st[2][0])
    mu, var, phi, t, muz, sh = _adf_fp64(y, th.astype(np.float64), orc.RB_TCURVE, orc.RB_FIXED)
    ref = np.zeros(len(y))
    for k, yk in enumerate(y.astype(np.float64)):
        mp = phi * mu + t * muz
        vp = max(phi * phi * var + sh * sh, 1e-8)
        mu, var, ref[k] = _ocsn_fp64(p, m, v, yk, mp, vp)
        var = max(var, 1e-6)
    np.testing.assert_allclose(lli, ref, rtol=0, atol=2e-5)
    assert abs(lli.sum() - ref.sum()) < 2e-3
    np.testing.assert_allclose(st[1][0], mu, rtol=0, atol=1e-4 * (1 + abs(mu)))


def test_special_function_restatements(orc):
    x = np.linspace(-9.0, 9.0, 4001, dtype=np.float32)
    np.testing.assert_allclose(orc.rb_math(0, x), np.tanh(x.astype(np.float64)), rtol=0, atol=2.5e-7)
    np.testing.assert_allclose(orc.rb_math(3, x), 1.5 * (1 + np.tanh(x.astype(np.float64))), rtol=0,
                               atol=1e-6)
    e = orc.rb_math(1, x[np.abs(x) < 4])
    r = special.erfc(x[np.abs(x) < 4].astype(np.float64))
    np.testing.assert_allclose(e, r, rtol=3e-7, atol=1e-7)
    u = orc.rb_math(2, x)
    ref = np.clip(stats.norm.cdf(x.astype(np.float64)), 1e-7, 1 - 1e-7)
    np.testing.assert_allclose(u, ref, rtol=0, atol=3e-7)
    assert u.min() >= np.float32(1e-7) and u.max() <= np.float32(1 - 1e-7)


@pytest.mark.parametrize("lag", [0, 6])
def test_identity_moves_change_nothing(orc, lag):
    """rho_c = 1 and a zero random walk: the move replays the current
    trajectory, so every move is accepted and leaves the state as it was."""
    y, _, _ = orc.rb_dgp(40, seed=9)
    common = dict(ess_thr=0.97, crho=1.0, adaptive=False, pstd=(0.0, 0.0, 0.0, 0.0), lag=lag)
    a = orc.run_rb_smc2(orc.rb_cfg(12, 64, K_rejuv=0, **common), y)
    b = orc.run_rb_smc2(orc.rb_cfg(12, 64, K_rejuv=3, **common), y)
    assert b["n_resamples"] >= 5 and b["n_resamples"] == a["n_resamples"]
    assert b["n_acc"] == b["n_tot"] == 3 * 12 * b["n_resamples"]
    assert np.array_equal(a["theta"], b["theta"]) and np.array_equal(a["inner"], b["inner"])
    if lag == 0:
        assert np.array_equal(a["ll"], b["ll"])
    else:
        np.testing.assert_allclose(b["ll"], a["ll"], rtol=1e-12)
    # a mistake in the replay (here: a different noise correlation) breaks it
    c = orc.run_rb_smc2(orc.rb_cfg(12, 64, K_rejuv=3, **dict(common, crho=0.9)), y)
    assert c["n_acc"] < c["n_tot"] or not np.array_equal(a["inner"], c["inner"])


def test_batched_run_is_the_inner_filter(orc):
    y, _, _ = orc.rb_dgp(60, seed=11)
    cfg = orc.rb_cfg(6, 128, ess_thr=0.0)
    r = orc.run_rb_smc2(cfg, y)
    assert r["n_resamples"] == 0
    lli, ess, st = orc.rb_inner_run(cfg, r["theta"][:, 0], y)
    acc = 0.0
    for v in lli:
        acc += v
    assert r["ll"][0] == acc and r["lw"][0] == acc
    assert np.array_equal(st, r["inner"][:, :128])


def test_prior_draws_are_truncated_prior(orc):
    n = 4096
    cfg = orc.rb_cfg(n, 64)
    th = orc.run_rb_smc2(cfg, np.zeros(0, np.float32))["theta"].astype(np.float64)
    pr, bd = orc.RB_PRIOR, orc.RB_BOUNDS
    for q in range(4):
        mu, sd, lo, hi = pr[2 * q], pr[2 * q + 1], bd[2 * q], bd[2 * q + 1]
        assert th[q].min() >= np.float32(lo) and th[q].max() <= np.float32(hi)
        tn = stats.truncnorm((lo - mu) / sd, (hi - mu) / sd, loc=mu, scale=sd)
        assert abs(th[q].mean() - tn.mean()) < 5 * tn.std() / math.sqrt(n)
        assert stats.kstest(th[q], tn.cdf).pvalue > 1e-3


def test_outer_resampling_and_moves_happen(orc):
    y, _, _ = orc.rb_dgp(80, *TRUTH, seed=2)
    r = orc.run_rb_smc2(orc.rb_cfg(32, 64, K_rejuv=2, lag=10), y)
    assert r["n_resamples"] >= 1 and 0 < r["n_acc"] < r["n_tot"]
    assert r["t_cp"] == 70
    assert np.all(np.isfinite(r["ll"])) and np.all(r["ess_inner"] > 0)


def test_half_rounding_is_ieee(orc):
    """The oracle's half-precision noise store rounds as NumPy's float16 (IEEE
    round to nearest even), including every halfway point and the overflow,
    subnormal and special cases."""
    rng = np.random.default_rng(0)
    h = np.arange(0, 0x7C00, dtype=np.uint16).view(np.float16).astype(np.float64)
    mid = ((h[:-1] + h[1:]) / 2).astype(np.float32)
    x = np.concatenate([rng.standard_normal(200_000).astype(np.float32) * 4, mid, -mid,
                        h.astype(np.float32), np.float32([65504, 65519.99, 65520, 1e6, -1e6, np.inf, -np.inf,
                                                          0.0, -0.0, 5.96e-8, 2.98e-8, 1e-9])])
    with np.errstate(over="ignore"):
        ref = x.astype(np.float16).astype(np.float32)
    assert np.array_equal(orc.rb_half_roundtrip(x), ref)
