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


def _brute_ancestors(lw, u):
    """Systematic resampling by exhaustive search: slot m takes the smallest j
    whose exact running weight sum reaches (u + m) / n of the total; entries
    with non-finite log-weights weigh nothing."""
    lw = np.asarray(lw, np.float64)
    ok = np.isfinite(lw)
    w = np.where(ok, np.exp(np.where(ok, lw - lw[ok].max(), 0.0)), 0.0)
    c = np.cumsum(w)
    n = lw.size
    return np.array([int(np.argmax(c >= (u + m) / n * c[-1])) for m in range(n)])


@pytest.mark.parametrize("seed", range(6))
def test_outer_ancestors_match_brute_force(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 200))
    lw = rng.normal(0.0, 3.0, n) - 500.0
    if n > 3:
        lw[rng.integers(0, n, 3)] = [np.nan, -np.inf, np.nan]
    if not np.isfinite(lw).any():
        lw[0] = 0.0
    u = float(rng.uniform(1e-6, 1 - 1e-6))
    anc = gpu_bpf_batch.systematic_ancestors(lw, u)
    np.testing.assert_array_equal(anc, _brute_ancestors(lw, u))
    assert np.all(np.isfinite(lw[anc]))
    assert np.all(np.diff(anc) >= 0)


def test_outer_ancestors_nan_and_inf_never_chosen():
    lw = np.array([np.nan, 0.0, -np.inf, np.nan, 1.0, np.nan])
    for u in (1e-9, 0.3, 0.999999):
        anc = gpu_bpf_batch.systematic_ancestors(lw, u)
        assert set(anc.tolist()) <= {1, 4}
    # offspring counts are floor or ceil of n * w
    w = np.exp(np.array([0.0, 1.0]))
    w /= w.sum()
    cnt = np.bincount(gpu_bpf_batch.systematic_ancestors(lw, 0.5), minlength=6)[[1, 4]]
    assert np.all(np.abs(cnt - 6 * w) < 1.0)


def test_outer_ancestors_without_finite_weight_fail():
    import gpu_bpf

    with pytest.raises(gpu_bpf.GpuBpfError):
        gpu_bpf_batch.systematic_ancestors([np.nan, -np.inf, np.nan], 0.5)
