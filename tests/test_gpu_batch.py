"""Batched inner filters against the plain filter, through the C ABI.

Inner filter m of a batch must be the plain filter created with the same
(N_m, phi_m, sigma_m, mu_m, seed_m) and ESS threshold: particles bit-exact,
ancestors therefore equal, FP64 statistics and log-likelihood to 1e-12
relative (the same arithmetic in the same order; only the FP64 group sums
differ by additions of exact zeros).  M = 1 reduces to the plain filter.
"""
import numpy as np
import pytest

import gpu_bpf
import gpu_bpf_batch
import score_reference

pytestmark = pytest.mark.gpu

STAT_FIELDS = ("mean", "var", "ess", "ll_incr", "ll_cum", "ll_ref")


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def ydata(orc):
    _, y = orc.dgp(400, 0.97, 0.15, -1.0, seed=42)
    return y.astype(np.float32)


def filters(M, rng, nmin=256, nmax=16384):
    n = np.exp(rng.uniform(np.log(nmin), np.log(nmax), M)).round().astype(np.int32)
    mu = rng.uniform(-2.0, 0.0, M).astype(np.float32)
    phi = rng.uniform(0.90, 0.995, M).astype(np.float32)
    sig = rng.uniform(0.05, 0.40, M).astype(np.float32)
    seeds = rng.integers(1, 1 << 30, M).astype(np.int32)
    return n, mu, phi, sig, seeds


def plain(n, mu, phi, sig, seed, y, thr, nu=0.0):
    f = gpu_bpf.GpuBpf(int(n), float(phi), float(sig), float(mu), nu_obs=nu, seed=int(seed),
                       ess_threshold=thr)
    _, st = f.run(y)
    x = f.particles()
    cnt = f.resample_count()
    f.close()
    return st, x, cnt


def assert_same(st_b, st_p, rtol=1e-12):
    np.testing.assert_array_equal(st_b["resampled"], st_p["resampled"])
    for k in STAT_FIELDS:
        a, b = st_b[k], st_p[k]
        fin = np.isfinite(b)
        np.testing.assert_array_equal(np.isfinite(a), fin, err_msg=k)
        err = np.max(np.abs(a[fin] - b[fin]) / np.maximum(np.abs(b[fin]), 1.0)) if fin.any() else 0
        assert err <= rtol, (k, err)
    np.testing.assert_array_equal(st_b["h_mean"], st_p["h_mean"])


@pytest.mark.parametrize("thr", [0.0, 0.5])
def test_batch_equals_looped_plain_filters(ydata, thr):
    rng = np.random.default_rng(3)
    n, mu, phi, sig, seeds = filters(12, rng)
    n[:3] = [256, 257, 16384]          # one tile, a padded second tile, the largest filter
    T = 70                             # two captured 32-tick graphs and eager ticks
    y = ydata[:T]
    say("batch vs plain", thr, n.tolist())
    b = gpu_bpf_batch.GpuBpfBatch(n, mu, phi, sig, seeds, ess_threshold=thr, block=256)
    st = b.run(y)
    counts = b.resample_counts()
    ll = b.loglik()
    for m in range(len(n)):
        stp, xp, cp = plain(n[m], mu[m], phi[m], sig[m], seeds[m], y, thr)
        assert_same(st[:, m], stp)
        np.testing.assert_array_equal(b.particles(m), xp)
        assert counts[m] == cp
        assert ll[m] == st["ll_cum"][-1, m]
    b.close()


@pytest.mark.parametrize("block", [64, 128, 512, 1024])
def test_block_sizes_give_identical_results(ydata, block):
    rng = np.random.default_rng(5)
    n, mu, phi, sig, seeds = filters(9, rng)
    y = ydata[:40]
    say("block", block)
    ref = gpu_bpf_batch.GpuBpfBatch(n, mu, phi, sig, seeds, ess_threshold=0.5, block=256)
    sr = ref.run(y)
    b = gpu_bpf_batch.GpuBpfBatch(n, mu, phi, sig, seeds, ess_threshold=0.5, block=block)
    sb = b.run(y)
    for k in STAT_FIELDS + ("resampled",):
        np.testing.assert_array_equal(sb[k], sr[k], err_msg=k)
    for m in range(len(n)):
        np.testing.assert_array_equal(b.particles(m), ref.particles(m))
    b.close()
    ref.close()


@pytest.mark.parametrize("tile", [512, 1024])
@pytest.mark.parametrize("thr", [0.0, 0.5])
def test_tiles_equal_the_plain_filter_with_the_same_tile(ydata, tile, thr):
    """A batch with tiles of 256 K particles is, filter by filter, the plain
    filter at K particles per lane (the same tile)."""
    rng = np.random.default_rng(13)
    n, mu, phi, sig, seeds = filters(8, rng)
    n[:4] = [256, tile + 1, 3 * tile - 5, 16384]   # partial, padded, ragged and largest filters
    T = 45
    y = ydata[:T]
    say("tile", tile, thr, n.tolist())
    b = gpu_bpf_batch.GpuBpfBatch(n, mu, phi, sig, seeds, ess_threshold=thr,
                                  block=256 if tile < 1024 else 128, tile=tile)
    assert b.tile == tile
    st = b.run(y)
    counts = b.resample_counts()
    for m in range(len(n)):
        f = gpu_bpf.GpuBpf(int(n[m]), float(phi[m]), float(sig[m]), float(mu[m]), seed=int(seeds[m]),
                           ess_threshold=thr, items_per_thread=tile // 256)
        _, stp = f.run(y)
        assert f.layout()[0] == tile
        assert_same(st[:, m], stp)
        np.testing.assert_array_equal(b.particles(m), f.particles())
        assert counts[m] == f.resample_count()
        f.close()
    b.close()


@pytest.mark.parametrize("block,tile", [(64, 512), (1024, 512), (64, 1024), (512, 1024)])
def test_blocks_at_one_tile_give_identical_results(ydata, block, tile):
    rng = np.random.default_rng(15)
    n, mu, phi, sig, seeds = filters(7, rng)
    y = ydata[:40]
    say("block", block, "tile", tile)
    ref = gpu_bpf_batch.GpuBpfBatch(n, mu, phi, sig, seeds, ess_threshold=0.5, block=256 if tile < 1024 else 128,
                                    tile=tile)
    sr = ref.run(y)
    b = gpu_bpf_batch.GpuBpfBatch(n, mu, phi, sig, seeds, ess_threshold=0.5, block=block, tile=tile)
    sb = b.run(y)
    for k in STAT_FIELDS + ("resampled",):
        np.testing.assert_array_equal(sb[k], sr[k], err_msg=k)
    for m in range(len(n)):
        np.testing.assert_array_equal(b.particles(m), ref.particles(m))
    b.close()
    ref.close()


def test_student_t_batch(ydata):
    rng = np.random.default_rng(7)
    n, mu, phi, sig, seeds = filters(5, rng, 300, 5000)
    y = ydata[:50]
    say("student-t batch")
    b = gpu_bpf_batch.GpuBpfBatch(n, mu, phi, sig, seeds, nu_obs=5.0, ess_threshold=0.0)
    st = b.run(y)
    for m in range(len(n)):
        stp, xp, _ = plain(n[m], mu[m], phi[m], sig[m], seeds[m], y, 0.0, nu=5.0)
        assert_same(st[:, m], stp)
        np.testing.assert_array_equal(b.particles(m), xp)
    b.close()


def test_single_filter_batch_is_the_plain_filter(ydata):
    y = ydata[:100]
    say("M=1")
    b = gpu_bpf_batch.GpuBpfBatch([4096], [-1.0], [0.97], [0.15], [42], ess_threshold=0.5)
    st = b.run(y[:37])
    st2 = b.run(y[37:])
    stp, xp, cp = plain(4096, -1.0, 0.97, 0.15, 42, y, 0.5)
    assert_same(np.concatenate([st[:, 0], st2[:, 0]]), stp)
    np.testing.assert_array_equal(b.particles(0), xp)
    assert b.resample_counts()[0] == cp
    assert b.timestep() == 100
    b.close()


def test_autotune_picks_a_block(ydata):
    rng = np.random.default_rng(9)
    n, mu, phi, sig, seeds = filters(64, rng)
    say("autotune")
    b = gpu_bpf_batch.GpuBpfBatch(n, mu, phi, sig, seeds)
    blk = b.autotune(ydata, 40)
    assert blk in (64, 128, 256, 512, 1024) and b.block == blk
    assert b.tile in (256, 512, 1024)
    say("autotuned", blk, b.tile)
    b.run(ydata[:10], stats=False)
    assert b.timestep() == 10
    ll = b.loglik()
    # the tuned pair is the same filters: equal to a batch set to that tile
    ref = gpu_bpf_batch.GpuBpfBatch(n, mu, phi, sig, seeds, block=64 if blk != 64 else 128, tile=b.tile)
    ref.run(ydata[:10], stats=False)
    np.testing.assert_array_equal(ll, ref.loglik())
    ref.close()
    b.close()


@pytest.mark.parametrize("estimator", ["pathwise", "fisher"])
def test_autotune_with_score_keeps_the_score(ydata, estimator):
    """With the score enabled the autotuner times the score kernels; the
    tuned batch's score and log-likelihoods equal an untuned batch's."""
    rng = np.random.default_rng(11)
    n, mu, phi, sig, seeds = filters(24, rng, 256, 4000)
    say("autotune with score", estimator)
    b = gpu_bpf_batch.GpuBpfBatch(n, mu, phi, sig, seeds)
    b.enable_score(estimator)
    blk = b.autotune(ydata, 20)
    assert blk in (64, 128, 256, 512, 1024) and b.block == blk
    b.run(ydata[:30], stats=False)
    ref = gpu_bpf_batch.GpuBpfBatch(n, mu, phi, sig, seeds, block=64 if blk != 64 else 128, tile=b.tile)
    ref.enable_score(estimator)
    ref.run(ydata[:30], stats=False)
    np.testing.assert_array_equal(b.score(), ref.score())
    np.testing.assert_array_equal(b.loglik(), ref.loglik())
    b.close()
    ref.close()


@pytest.mark.parametrize("case", ["gauss", "ess_threshold", "student_t_score", "fisher_tile512"])
def test_statistics_off_leaves_the_filters_unchanged(ydata, case):
    """With the per-tick statistics off, the ticks skip the mean, variance and
    (without an ESS threshold) squared-weight sums; log-likelihoods, resample
    counts, particles and scores equal a batch with them on, bit for bit, over
    a captured 32-tick graph and single launches.  Statistics switched back on
    cover the ticks run since, and equal the other batch's."""
    rng = np.random.default_rng(21)
    n, mu, phi, sig, seeds = filters(12, rng, 256, 6000)
    nu = 5.0 if case == "student_t_score" else 0.0
    thr = 0.5 if case == "ess_threshold" else 0.0
    tile = 512 if case == "fisher_tile512" else None
    say("statistics off", case)
    on = gpu_bpf_batch.GpuBpfBatch(n, mu, phi, sig, seeds, nu_obs=nu, ess_threshold=thr, block=512, tile=tile)
    off = gpu_bpf_batch.GpuBpfBatch(n, mu, phi, sig, seeds, nu_obs=nu, ess_threshold=thr, block=512, tile=tile,
                                    stats=False)
    if case in ("student_t_score", "fisher_tile512"):
        est = "fisher" if case == "fisher_tile512" else "pathwise"
        on.enable_score(est)
        off.enable_score(est)
    on.run(ydata[:40], stats=False)
    off.run(ydata[:40], stats=False)
    np.testing.assert_array_equal(off.loglik(), on.loglik())
    np.testing.assert_array_equal(off.resample_counts(), on.resample_counts())
    for m in range(len(n)):
        np.testing.assert_array_equal(off.particles(m), on.particles(m))
    if case in ("student_t_score", "fisher_tile512"):
        np.testing.assert_array_equal(off.score(), on.score())
    with pytest.raises(gpu_bpf.GpuBpfError, match="statistics are off"):
        off.stats(1)
    off.set_stats(True)
    st_on = on.run(ydata[40:45])
    st_off = off.run(ydata[40:45])
    for k in st_on.dtype.names:
        np.testing.assert_array_equal(st_off[k], st_on[k], err_msg=k)
    with pytest.raises(gpu_bpf.GpuBpfError, match="available"):
        off.stats(6)
    on.close()
    off.close()


def test_batch_errors():
    with pytest.raises(gpu_bpf.GpuBpfError, match="particles"):
        gpu_bpf_batch.GpuBpfBatch([100, 20000], 0.0, 0.9, 0.1, [1, 2])
    with pytest.raises(gpu_bpf.GpuBpfError, match="set_block"):
        gpu_bpf_batch.GpuBpfBatch([100], 0.0, 0.9, 0.1, [1], block=96)
    with pytest.raises(gpu_bpf.GpuBpfError, match="set_tile"):
        gpu_bpf_batch.GpuBpfBatch([100], 0.0, 0.9, 0.1, [1], tile=300)
    # a (block, tile) pair whose LDS does not fit one CU is refused at the run
    b = gpu_bpf_batch.GpuBpfBatch([100], 0.0, 0.9, 0.1, [1], block=1024, tile=1024)
    with pytest.raises(gpu_bpf.GpuBpfError):
        b.run(np.zeros(3, np.float32), stats=False)
    # the statistics switch takes 0 or 1
    with pytest.raises(gpu_bpf.GpuBpfError, match="set_stats"):
        b._check(b.L.gpu_bpf_batch_set_stats(b.h, 2))
    b.close()


def test_batch_score_matches_adjoint_and_oracle(orc, ydata):
    """Forward-mode tangents of the batch give the single filter's reverse-mode
    pathwise score (x_0 differentiated) and the FP64 tangent oracle's, filter
    by filter, within 1e-4 relative (FP32 states and tangents)."""
    rng = np.random.default_rng(13)
    n, mu, phi, sig, seeds = filters(6, rng, 256, 6000)
    T = 120
    y = ydata[:T]
    say("batch score", n.tolist())
    b = gpu_bpf_batch.GpuBpfBatch(n, mu, phi, sig, seeds, ess_threshold=0.0)
    b.enable_score("pathwise")
    b.run(y, stats=False)
    sc = b.score()
    ll = b.loglik()
    for m in range(len(n)):
        f = gpu_bpf.GpuBpf(int(n[m]), float(phi[m]), float(sig[m]), float(mu[m]), seed=int(seeds[m]))
        llp = f.score_forward(y, 32)
        gp = f.score_backward(True)
        f.close()
        assert llp == ll[m]
        np.testing.assert_allclose(sc[m], gp, rtol=1e-4, atol=1e-4 * np.max(np.abs(gp)))
        if m < 2:
            og = orc.score(y, int(n[m]), rho=float(phi[m]), sigma_z=float(sig[m]), mu=float(mu[m]),
                           seed=int(seeds[m]), tile=256, x0_reparam=True)[1]
            np.testing.assert_allclose(sc[m], og, rtol=1e-4, atol=1e-4 * np.max(np.abs(og)))
    b.close()


def test_batch_score_student_t(orc, ydata):
    rng = np.random.default_rng(17)
    n, mu, phi, sig, seeds = filters(3, rng, 300, 3000)
    y = ydata[:80]
    say("batch score student-t")
    b = gpu_bpf_batch.GpuBpfBatch(n, mu, phi, sig, seeds, nu_obs=5.0)
    b.enable_score()
    b.run(y, stats=False)
    sc = b.score()
    for m in range(len(n)):
        og = orc.score(y, int(n[m]), rho=float(phi[m]), sigma_z=float(sig[m]), mu=float(mu[m]),
                       nu_obs=5.0, seed=int(seeds[m]), tile=256, x0_reparam=True)[1]
        np.testing.assert_allclose(sc[m], og, rtol=1e-4, atol=1e-4 * np.max(np.abs(og)))
    b.close()


def test_fisher_identity_score_matches_the_exact_score(orc, ydata):
    """Fisher-identity estimate sum_i w_i alpha_i (alpha = grad log p(x_0:t)
    along each ancestry) over 16 independent filters, against the exact score
    of the SV log-likelihood: central differences of the point-mass (grid)
    filter's log-likelihood, which is converged in the grid size here.  (The
    pathwise estimator -- ancestors held fixed, as tests/diff_bpf.py detaches
    them -- is biased for this quantity: on these data its mean is ~50% off
    in every component, so it is pinned against the oracle's tangent instead.)"""
    y = ydata[:60]
    th = np.array([-1.0, 0.97, 0.15])
    exact = score_reference.exact_score(orc, y, th)
    M = 16
    say("fisher vs exact", exact.tolist())
    b = gpu_bpf_batch.GpuBpfBatch(np.full(M, 16384), th[0], th[1], th[2], np.arange(1, M + 1))
    b.enable_score("fisher")
    b.run(y, stats=False)
    sc = b.score()
    b.close()
    mean, se = sc.mean(0), sc.std(0, ddof=1) / np.sqrt(M)
    say("fisher", mean.tolist(), se.tolist())
    assert np.all(np.abs(mean - exact) < 4 * se + 0.02 * np.abs(exact)), (mean, se, exact)
    assert np.all(se < 0.15 * np.abs(exact))        # the check has power
