"""GPU parity of the IMM layer (gpu_imm_*) against the CPU oracle's filters
and an FP64 restatement of the reference's host mixing
(gpu_bpf/gpu_bpf.cu:943-1094), plus closed forms: one model is the plain
filter, and with an identity transition matrix the model probabilities are
the normalised products of each filter's per-tick likelihoods.
Bar: every model's BpfResult equal to the oracle's (float), the mixed outputs
equal to the restatement's within 1 ulp, probabilities within 1e-6.
"""
import math

import numpy as np
import pytest

import gpu_bpf

pytestmark = pytest.mark.gpu


def say(*a):
    print("[case]", *a, flush=True)


def _y(orc, T):
    _, y = orc.dgp(T, 0.97, 0.15, -1.0, seed=42)
    return y.astype(np.float32)


def _lse(x):
    mx = -1e30
    for v in x:
        if v > mx:
            mx = v
    return mx + math.log(sum(math.exp(v - mx) for v in x))


def _reference_imm(orc, models, n, y, seed, trans=None, tile=256):
    """The reference's IMM loop over the oracle's filters (FP64 host mixing)."""
    K, T = len(models), len(y)
    outs = []
    for k, (rho, sig, mu, nu_s, nu_o) in enumerate(models):
        t, _, _, _ = orc.run_filter(y, n, rho=rho, sigma_z=sig, mu=mu, nu_obs=nu_o, ess_threshold=0.5,
                                    seed=seed + 7919 * k, tile=tile, threads=8)
        outs.append(t)
    if trans is None:
        ps = 0.95
        pw = (1 - ps) / (K - 1) if K > 1 else 0.0
        logT = [[math.log(max(ps if i == j else pw, 1e-30)) for j in range(K)] for i in range(K)]
    else:
        logT = [[math.log(max(float(np.float32(trans[i][j])), 1e-30)) for j in range(K)] for i in range(K)]
    log_pi = [-math.log(K)] * K
    res, probs = [], []
    for t in range(T):
        pred = [_lse([logT[j][k] + log_pi[j] for j in range(K)]) for k in range(K)]
        ll = [float(outs[k]["log_lik"][t]) for k in range(K)]
        h = [float(outs[k]["h_mean"][t]) for k in range(K)]
        joint = [pred[k] + ll[k] for k in range(K)]
        lz = _lse(joint)
        log_pi = [joint[k] - lz for k in range(K)]
        hm = 0.0
        best, bl = 0, -1e30
        for k in range(K):
            hm += math.exp(log_pi[k]) * h[k]
            if log_pi[k] > bl:
                best, bl = k, log_pi[k]
        res.append((np.float32(hm), np.float32(lz), best, np.float32(math.exp(bl))))
        probs.append([math.exp(v) for v in log_pi])
    return res, np.array(probs), outs


def _ulp_close(a, b):
    return abs(float(a) - float(b)) <= float(np.spacing(np.abs(np.float32(b))))


def _run(models, n, y, seed, trans=None):
    imm = gpu_bpf.GpuImm(models, n, transition=trans, seed=seed)
    out, probs = [], []
    for t in range(len(y)):
        out.append(imm.step(float(y[t])))
        probs.append(imm.probs())
    batch = imm.uses_batch()
    imm.close()
    return out, np.array(probs), batch


@pytest.mark.parametrize("case", ["batch", "filters_nu", "filters_big"])
def test_imm_matches_reference_loop(orc, case):
    T = 60
    y = _y(orc, T)
    grid = gpu_bpf.imm_build_grid([0.95, 0.98], [0.15], [-1.5, -0.5])
    n = 4096
    if case == "filters_nu":
        grid = [g[:4] + ((5.0 if k % 2 else 0.0),) for k, g in enumerate(grid)]
    if case == "filters_big":
        n = 20000
    say("imm", case, len(grid), n)
    out, probs, batch = _run(grid, n, y, seed=11)
    assert batch == (case == "batch")
    ref, rprobs, _ = _reference_imm(orc, grid, n, y, seed=11)   # N <= 2^19: 256-particle tiles
    for t in range(T):
        h, vol, ll, best, bp = out[t]
        rh, rll, rbest, rbp = ref[t]
        assert _ulp_close(h, rh) and _ulp_close(ll, rll), (t, out[t], ref[t])
        assert best == rbest and _ulp_close(bp, rbp)
        assert abs(vol - math.exp(float(np.float32(h)) * 0.5)) <= 1e-6 * vol
    np.testing.assert_allclose(probs, rprobs, rtol=1e-6, atol=1e-7)


def test_one_model_is_the_plain_filter(orc):
    T, n = 40, 4096
    y = _y(orc, T)
    say("imm K=1")
    out, probs, _ = _run([(0.97, 0.15, -1.0, 0.0, 0.0)], n, y, seed=5, trans=[[1.0]])
    dflt, _, _ = _run([(0.97, 0.15, -1.0, 0.0, 0.0)], n, y, seed=5)
    f = gpu_bpf.GpuBpf(n, 0.97, 0.15, -1.0, seed=5)
    for t in range(T):
        h, ll = f.step(float(y[t]))
        assert out[t][0] == h and out[t][2] == ll and out[t][3] == 0 and out[t][4] == 1.0
        # the default matrix keeps 0.95 on the diagonal even for one model
        assert _ulp_close(dflt[t][2], np.float32(float(ll) + math.log(0.95)))
    f.close()
    assert np.all(probs == 1.0)


def test_identity_transition_is_bayes_factors(orc):
    T, n = 50, 4096
    y = _y(orc, T)
    models = [(0.97, 0.15, -1.0, 0.0, 0.0), (0.97, 0.15, -0.7, 0.0, 0.0), (0.95, 0.2, -1.0, 0.0, 0.0)]
    say("imm identity")
    out, probs, _ = _run(models, n, y, seed=2, trans=np.eye(3))
    _, _, outs = _reference_imm(orc, models, n, y, seed=2, trans=np.eye(3))
    cum = np.array([np.cumsum(o["log_lik"].astype(np.float64)) for o in outs])   # K x T
    want = np.exp(cum - cum.max(axis=0))
    want /= want.sum(axis=0)
    # the matrix is floored at 1e-30 off the diagonal, so the closed form
    # holds until a probability first falls near that floor
    ok = np.minimum.accumulate(want.min(axis=0)) > 1e-20
    assert ok.sum() >= 8
    np.testing.assert_allclose(probs[ok], want.T[ok], rtol=1e-5, atol=1e-7)


def test_too_many_models():
    with pytest.raises(gpu_bpf.GpuBpfError):
        gpu_bpf.GpuImm([(0.97, 0.15, -1.0, 0.0, 0.0)] * 65, 256)
