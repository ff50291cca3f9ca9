"""CPU tests of the SMC^2 rejuvenation moves' proposal and acceptance code
(the library's own host functions) on a closed-form target: a 3-d Gaussian
N(m, C) with its exact log-density and gradient.  A correct MH / MALA kernel
leaves it invariant, so long chains reproduce m and C; dropping the MALA
proposal-density correction (a planted fault) visibly shrinks the variance.
"""
import numpy as np
import pytest

import gpu_bpf_batch

M0 = np.array([-1.0, 0.95, 0.2])
C0 = np.array([0.04, 0.002, -0.003, 0.002, 0.0004, 0.0001, -0.003, 0.0001, 0.0025]).reshape(3, 3)


@pytest.mark.parametrize("kind", ["rw", "mala"])
def test_moves_leave_gaussian_target_invariant(kind):
    mean, cov, acc = gpu_bpf_batch.selftest_mcmc(kind, 400_000, seed=3)
    sd = np.sqrt(np.diag(C0))
    assert np.all(np.abs(mean - M0) < 0.05 * sd), mean
    np.testing.assert_allclose(np.diag(cov), np.diag(C0), rtol=0.05)
    corr = cov / np.sqrt(np.outer(np.diag(cov), np.diag(cov)))
    corr0 = C0 / np.outer(sd, sd)
    np.testing.assert_allclose(corr, corr0, atol=0.03)
    assert (0.15 < acc < 0.5) if kind == "rw" else (0.6 < acc < 0.99)


def test_dropped_mala_correction_is_detected():
    _, cov, _ = gpu_bpf_batch.selftest_mcmc("mala", 400_000, seed=3, fault=1)
    assert np.max(np.abs(np.diag(cov) / np.diag(C0) - 1.0)) > 0.2
