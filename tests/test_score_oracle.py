"""CPU tests pinning the score oracle (pathwise dLL/d(mu, phi, sigma) with
ancestors held fixed): central finite differences under common random
numbers with frozen ancestors, and the FP64 torch restatement of the
reference's autograd path (tests/diff_bpf.py)."""
import numpy as np
import pytest

import score_reference

RHO, SIG, MU = 0.97, 0.15, -1.0


@pytest.fixture(scope="module")
def setup(orc):
    _, y = orc.dgp(60, RHO, SIG, MU, seed=42)
    return y.astype(np.float32)


def normals(orc, n, T, seed=42):
    return np.array([[orc.normal(seed, i, t) for i in range(n)] for t in range(T)], np.float64)


def test_tangent_score_matches_central_differences(orc, setup):
    y = setup
    n = 1024
    ll, sc, _, anc = orc.score(y, n, want_anc=True)
    th = np.float32([MU, RHO, SIG]).astype(np.float64)
    h = 1e-6
    fd = np.zeros(3)
    for k in range(3):
        tp, tm = th.copy(), th.copy()
        tp[k] += h
        tm[k] -= h
        fd[k] = (orc.score(y, n, theta64=tp, anc_in=anc)[0] -
                 orc.score(y, n, theta64=tm, anc_in=anc)[0]) / (2 * h)
    np.testing.assert_allclose(sc, fd, rtol=2e-6, atol=1e-6)
    # replaying the recorded ancestors reproduces the log-likelihood exactly
    assert orc.score(y, n, anc_in=anc)[0] == ll


@pytest.mark.parametrize("x0", [True, False])
@pytest.mark.parametrize("nu", [0.0, 5.0])
def test_tangent_score_matches_torch_restatement(orc, setup, x0, nu):
    y = setup[:30]
    n = 512
    ll, sc, _, anc = orc.score(y, n, nu_obs=nu, x0_reparam=x0, want_anc=True)
    z = normals(orc, n, len(y))
    th = np.float32([MU, RHO, SIG]).astype(np.float64)
    tl, tg = score_reference.grad(th, z, anc, y.astype(np.float64), x0_reparam=x0, nu_obs=nu)
    assert abs(tl - ll) <= 1e-9 * abs(ll)
    np.testing.assert_allclose(sc, tg, rtol=1e-9, atol=1e-9)


def test_planted_fault_breaks_score_agreement(orc, setup):
    """Dropping the -x/2 term changes lw' and the check catches it."""
    y = setup[:30]
    n = 512
    _, _, _, anc = orc.score(y, n, want_anc=True)
    z = normals(orc, n, len(y))
    th = np.float32([MU, RHO, SIG]).astype(np.float64)
    _, tg = score_reference.grad(th, z, anc, y.astype(np.float64))
    with orc.planted_fault(1):
        _, sc_bad, _, _ = orc.score(y, n, anc_in=anc)
    assert np.max(np.abs(sc_bad - tg) / np.abs(tg)) > 1e-3
