"""GPU tests of the pathwise score (adjoint) through the C ABI.

Bar: the forward log-likelihood equals the plain filter's bit for bit; the
reverse sweep is bit-identical for every checkpoint spacing (the replay
reproduces the forward exactly); the score matches the FP64 tangent oracle on
the same genealogy within 1e-4 relative (FP32 particle states); a
torch.autograd.Function over the C ABI gives the gradient of the FP64 torch
restatement of tests/diff_bpf.py.
"""
import numpy as np
import pytest

import gpu_bpf
import gpu_bpf_batch
import score_reference

pytestmark = pytest.mark.gpu

RHO, SIG, MU = 0.97, 0.15, -1.0


def say(*a):
    print("[case]", *a, flush=True)


@pytest.fixture(scope="module")
def ydata(orc):
    _, y = orc.dgp(400, RHO, SIG, MU, seed=42)
    return y.astype(np.float32)


def gpu_score(n, y, every=64, x0=True, items=None, nu=0.0, theta=(MU, RHO, SIG)):
    mu, rho, sig = theta
    f = gpu_bpf.GpuBpf(n, rho, sig, mu, nu_obs=nu, seed=42, items_per_thread=items)
    ll = f.score_forward(y, every)
    g = f.score_backward(x0)
    tile, _ = f.layout()
    f.close()
    return ll, g, tile


def test_forward_equals_plain_filter(ydata):
    y = ydata[:200]
    n = 65536
    say("score-forward", n)
    ll, _, _ = gpu_score(n, y)
    f = gpu_bpf.GpuBpf(n, RHO, SIG, MU, seed=42, ess_threshold=0.0)
    _, st = f.run(y)
    f.close()
    assert ll == st["ll_cum"][-1]


def test_checkpoint_spacing_is_exact(ydata):
    y = ydata[:130]
    n = 20_000
    say("ckpt", n)
    ref = gpu_score(n, y, every=64)[1]
    for every in (1, 7, 130, 500):
        np.testing.assert_array_equal(gpu_score(n, y, every=every)[1], ref)


def test_overlapped_sweeps_are_exact(ydata):
    """At 2^20 particles segments replay on three streams at once, ahead of
    the reverse sweeps on a fourth; the score equals the one-segment run's bit
    for bit, a second backward call repeats it, and so does a second forward
    pass of the same shape after a restart (which keeps the buffers)."""
    y = ydata[:96]
    n = 1 << 20
    say("overlap", n)
    ref = gpu_score(n, y, every=96)[1]
    f = gpu_bpf.GpuBpf(n, RHO, SIG, MU, seed=42)
    ll1 = f.score_forward(y, 8)
    g1 = f.score_backward(True)
    g2 = f.score_backward(True)
    f.restart()
    ll2 = f.score_forward(y, 8)
    g3 = f.score_backward(True)
    f.close()
    np.testing.assert_array_equal(g1, ref)
    np.testing.assert_array_equal(g2, ref)
    np.testing.assert_array_equal(g3, ref)
    assert ll1 == ll2


def test_retiling_drops_the_score_pass(ydata):
    """A new tiling after a restart (gpu_bpf_set_items_per_thread) drops the
    forward pass laid out for the old tiles: the backward call then asks for
    a new forward pass, which gives the new tiling's score."""
    y = ydata[:40]
    n = 20_000
    say("retile", n)
    f = gpu_bpf.GpuBpf(n, RHO, SIG, MU, seed=42, items_per_thread=2)
    f.score_forward(y, 8)
    f.restart()
    f._check(f.L.gpu_bpf_set_items_per_thread(f.h, 4))
    with pytest.raises(gpu_bpf.GpuBpfError, match="score_forward first"):
        f.score_backward(True)
    f.score_forward(y, 8)
    g = f.score_backward(True)
    f.close()
    np.testing.assert_array_equal(g, gpu_score(n, y, every=8, items=4)[1])


@pytest.mark.parametrize("x0", [True, False])
@pytest.mark.parametrize("items", [1, 4, 16])
def test_score_matches_oracle(orc, ydata, x0, items):
    y = ydata[:150]
    n = 70_001
    say("score", n, x0, items)
    ll, g, tile = gpu_score(n, y, every=32, x0=x0, items=items)
    oll, og, oll32, _ = orc.score(y, n, tile=tile, x0_reparam=x0)
    assert abs(ll - oll32) <= 1e-9 * abs(oll32)
    np.testing.assert_allclose(g, og, rtol=1e-4, atol=1e-4 * np.max(np.abs(og)))


def test_student_t_score(orc, ydata):
    y = ydata[:100]
    n = 30_000
    say("score-t", n)
    _, g, tile = gpu_score(n, y, nu=5.0)
    og = orc.score(y, n, nu_obs=5.0, tile=tile)[1]
    np.testing.assert_allclose(g, og, rtol=1e-4, atol=1e-4 * np.max(np.abs(og)))


def test_differentiable_filter_config_prior_draw(orc):
    """theta from the uniform prior [-2,0]x[0.90,0.995]x[0.05,0.40] (host PCG32 seed 42)."""
    r = orc.host_stream(42, 3)
    u = [(v >> 9) / 8388608.0 for v in r]
    theta = (-2.0 + 2.0 * u[0], 0.90 + 0.095 * u[1], 0.05 + 0.35 * u[2])
    theta = tuple(float(np.float32(v)) for v in theta)
    _, y = orc.dgp(300, theta[1], theta[2], theta[0], seed=42)
    y = y.astype(np.float32)
    n = 1 << 18
    say("prior-theta", theta)
    _, g, tile = gpu_score(n, y, theta=theta)
    og = orc.score(y, n, mu=theta[0], rho=theta[1], sigma_z=theta[2], tile=tile)[1]
    np.testing.assert_allclose(g, og, rtol=1e-4, atol=1e-4 * np.max(np.abs(og)))


class FilterLogLik:
    """torch.autograd.Function over the C ABI (test harness only)."""

    @staticmethod
    def make(n, y, every=16, x0=True):
        import torch

        class F(torch.autograd.Function):
            @staticmethod
            def forward(ctx, theta):
                mu, phi, sig = [float(v) for v in theta.detach()]
                f = gpu_bpf.GpuBpf(n, phi, sig, mu, seed=42)
                ll = f.score_forward(y, every)
                ctx.f = f
                return torch.tensor(ll, dtype=torch.float64)

            @staticmethod
            def backward(ctx, gout):
                g = ctx.f.score_backward(x0)
                ctx.f.close()
                return gout * torch.as_tensor(g, dtype=torch.float64)

        return F


def test_autograd_function_matches_torch_restatement(orc, ydata):
    import torch

    y = ydata[:40]
    n = 1024
    say("autograd", n)
    theta0 = np.float32([MU, RHO, SIG]).astype(np.float64)
    F = FilterLogLik.make(n, y)
    th = torch.tensor(theta0, requires_grad=True)
    ll = F.apply(th)
    ll.backward()
    # same genealogy as the GPU run: the FP32 restatement's ancestors
    _, _, _, anc = orc.score(y, n, want_anc=True)
    z = np.array([[orc.normal(42, i, t) for i in range(n)] for t in range(len(y))])
    tl, tg = score_reference.grad(theta0, z, anc, y.astype(np.float64))
    assert abs(ll.item() - tl) <= 1e-5 * abs(tl)
    np.testing.assert_allclose(th.grad.numpy(), tg, rtol=1e-4, atol=1e-4 * np.max(np.abs(tg)))


def fisher_score(n, y, nu=0.0, theta=(MU, RHO, SIG), seed=42, items=None, want_x=False):
    mu, rho, sig = theta
    f = gpu_bpf.GpuBpf(n, rho, sig, mu, nu_obs=nu, seed=seed, items_per_thread=items)
    f.set_score_estimator("fisher")
    ll = f.score_forward(y)
    g = f.score_backward()
    x = f.particles() if want_x else None
    f.close()
    return (ll, g, x) if want_x else (ll, g)


@pytest.mark.parametrize("n,nu", [(3000, 0.0), (16384, 0.0), (5000, 5.0)])
def test_fisher_single_filter_equals_batch(ydata, n, nu):
    """The single filter's Fisher-identity estimate equals the batched
    filter's with one inner filter: the same particles, ancestors and FP32
    alpha vectors; only the association of the FP64 sums differs."""
    y = ydata[:150]
    theta = (-0.7, 0.95, 0.22)
    say("fisher single vs batch", n, nu)
    ll, g = fisher_score(n, y, nu=nu, theta=theta, seed=11)
    b = gpu_bpf_batch.GpuBpfBatch([n], theta[0], theta[1], theta[2], [11], nu_obs=nu, ess_threshold=0.0)
    b.enable_score("fisher")
    b.run(y, stats=False)
    gb, llb = b.score()[0], b.loglik()[0]
    b.close()
    assert abs(ll - llb) <= 1e-12 * max(abs(llb), 1.0)
    np.testing.assert_allclose(g, gb, rtol=1e-9, atol=1e-12 * np.max(np.abs(gb)))


@pytest.mark.parametrize("items", [None, 16])
def test_fisher_single_filter_matches_exact_score(orc, ydata, items):
    """Fisher-identity estimates of four 2^20-particle filters (independent
    seeds) against the exact score of the SV log-likelihood (central
    differences of the point-mass filter), with the default tiles (512
    particles) and with 4096-particle tiles."""
    y = ydata[:60]
    th = (MU, RHO, SIG)
    exact = score_reference.exact_score(orc, y, th)
    say("fisher single vs exact", items, exact.tolist())
    gs = np.array([fisher_score(1 << 20, y, seed=s, items=items)[1] for s in (1, 2, 3, 4)])
    mean, se = gs.mean(0), gs.std(0, ddof=1) / 2.0
    say("fisher single", mean.tolist(), se.tolist())
    assert np.all(np.abs(mean - exact) < 4 * se + 0.02 * np.abs(exact)), (mean, se, exact)
    assert np.all(se < 0.1 * np.abs(exact))        # the check has power


def test_fisher_estimator_argument_checked():
    f = gpu_bpf.GpuBpf(1000, RHO, SIG, MU, seed=1)
    with pytest.raises(KeyError):
        f.set_score_estimator("bogus")
    assert f.L.gpu_bpf_set_score_estimator(f.h, 7) == -1
    f.close()


@pytest.mark.parametrize("items", [1, 2, 4, 8, 16])
def test_fisher_every_tile_width(ydata, items):
    """The Fisher-identity kernels of every tile width (up to 4096 particles
    per tile, 16 per lane): the particles are the plain filter's bit for bit,
    and at tick 0 (initial vectors, no gather yet) the score is the same as
    with 256-particle tiles up to FP32 weight rounding.  Later ticks are not
    compared across widths: the tile maxima differ, so the FP32 weights, the
    CDF and hence some ancestors differ."""
    y = ydata[:60]
    n = 70_001
    say("fisher tile width", items)
    ll, g, x = fisher_score(n, y, items=items, want_x=True)
    f = gpu_bpf.GpuBpf(n, RHO, SIG, MU, seed=42, ess_threshold=0.0, items_per_thread=items)
    _, st = f.run(y)
    xp = f.particles()
    f.close()
    np.testing.assert_array_equal(x, xp)
    assert ll == st["ll_cum"][-1]
    _, g0 = fisher_score(n, y[:1], items=items)
    _, g1 = fisher_score(n, y[:1], items=1)
    np.testing.assert_allclose(g0, g1, rtol=1e-6, atol=1e-9 * np.max(np.abs(g1)))
