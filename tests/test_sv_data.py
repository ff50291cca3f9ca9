"""The benchmark's data generator (sv_data.py, a Python restatement of the
reference's matched DGP, tests/test_bpf_matched_dgp.cu:59-175) against the
oracle's C restatement and the committed golden series: bit-identical
states and observations, with mu schedules, Student-t noise and the clamp."""
import os

import numpy as np
import pytest

import sv_data


@pytest.mark.parametrize("kw", [
    dict(phi=0.97, sigma=0.15, mu=-1.0, seed=42),
    dict(phi=0.99, sigma=0.4, mu=1.5, seed=7),             # reaches the upper clamp
    dict(phi=0.95, sigma=0.2, mu=-0.5, nu_obs=5.0, seed=11),
    dict(phi=0.9, sigma=0.3, mu=-1.0, nu_state=4.0, nu_obs=200.0, seed=-3),
])
def test_matches_oracle(orc, kw):
    T = 3000
    h, y = sv_data.sv_series(T, **kw)
    ho, yo = orc.dgp(T, kw["phi"], kw["sigma"], kw["mu"], kw.get("nu_state", 0.0),
                     kw.get("nu_obs", 0.0), seed=kw["seed"])
    np.testing.assert_array_equal(h, ho)
    np.testing.assert_array_equal(y, yo)


def test_regime_schedule_matches_oracle(orc):
    T = 5000
    mu_t = sv_data.mu_schedule(T)
    assert mu_t[0] == -1.5 and mu_t[1250] == -0.5 and mu_t[2500] == 0.5 and mu_t[3750] == -1.5
    for nu in (0.0, 5.0):
        h, y = sv_data.sv_series(T, 0.97, 0.15, mu_t.astype(np.float64), nu_obs=nu, seed=42)
        ho, yo = orc.dgp_regimes(mu_t.astype(np.float64), rho=0.97, sigma_z=0.15, nu_obs=nu, seed=42)
        np.testing.assert_array_equal(h, ho)
        np.testing.assert_array_equal(y, yo)


def test_matches_golden_series(golden_dir):
    g = np.load(os.path.join(golden_dir, "dgp_sv_seed42_T1000.npy"))   # rows: states, observations
    h, y = sv_data.sv_series(g.shape[1], 0.97, 0.15, -1.0, seed=42)
    np.testing.assert_array_equal(h, g[0])
    np.testing.assert_array_equal(y, g[1])


def test_rbpf_series_matches_oracle(orc):
    for args in [(1200, 0.95, 0.10, -1.0, 0.15), (500, 0.98, 0.08, 2.0, 0.20)]:
        a = sv_data.rbpf_series(*args, seed=12345)
        b = orc.rb_dgp(*args, seed=12345)
        for u, v in zip(a, b):
            np.testing.assert_array_equal(u, v)
