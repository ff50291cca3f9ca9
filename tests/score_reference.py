"""FP64 torch restatement of the reference's gradient path (test harness only).

tests/diff_bpf.py:186-233 differentiates LL = sum_t [logsumexp(lw_t) - log N]
through the reparameterised predict step with the resampling indices
detached.  This restatement follows it with the conventions of the C filter
(tick 0 weights the stationary draw without propagating; no truncated BPTT:
chunk = T), and takes its normals and ancestors as inputs so that it runs on
the same random numbers and genealogy as the GPU filter.  Gaussian or
Student-t observation density as diff_bpf.py:169-184.
"""
import math

import torch


def loglik(theta, z, anc, y, x0_reparam=True, nu_obs=0.0):
    """theta: float64 tensor [mu, phi, sigma]; z: [T, N] normals; anc: [T, N]
    ancestors (tick t's resample); y: [T].  Returns the scalar LL."""
    mu, phi, sig = theta[0], theta[1], theta[2]
    T, N = z.shape
    sd = sig / torch.sqrt(torch.clamp(1.0 - phi * phi, min=1e-6))
    x = mu + sd * z[0]
    if not x0_reparam:
        x = x.detach()
    ll = torch.zeros((), dtype=torch.float64)
    for t in range(T):
        if t > 0:
            x = mu + phi * (x - mu) + sig * z[t]                    # diff_bpf.py:165-167
        yt = y[t]
        if nu_obs > 0:
            nu = torch.tensor(nu_obs, dtype=torch.float64)
            lw = (torch.lgamma((nu + 1) / 2) - torch.lgamma(nu / 2) - 0.5 * torch.log(nu * math.pi)
                  - 0.5 * x - (nu + 1) / 2 * torch.log1p(yt * yt * torch.exp(-x) / nu))
        else:
            lw = -0.5 * math.log(2 * math.pi) - 0.5 * yt * yt * torch.exp(-x) - 0.5 * x
        ll = ll + torch.logsumexp(lw, 0) - math.log(N)              # diff_bpf.py:222
        x = x[anc[t]]                                               # diff_bpf.py:231
    return ll


def grad(theta0, z, anc, y, x0_reparam=True, nu_obs=0.0):
    th = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    ll = loglik(th, torch.as_tensor(z, dtype=torch.float64), torch.as_tensor(anc, dtype=torch.long),
                torch.as_tensor(y, dtype=torch.float64), x0_reparam, nu_obs)
    ll.backward()
    return ll.item(), th.grad.numpy().copy()


def exact_score(orc, y, theta, h=(1e-4, 1e-5, 1e-5), n_grid=2048):
    """Score of the exact SV log-likelihood log p(y_0:T-1 | mu, phi, sigma)
    with Gaussian observations: central differences of the oracle's
    point-mass (grid) filter, which is converged in the grid size for these
    series.  theta = (mu, phi, sigma)."""
    import numpy as np

    th = np.asarray(theta, np.float64)
    out = np.zeros(3)
    for k in range(3):
        tp, tm = th.copy(), th.copy()
        tp[k] += h[k]
        tm[k] -= h[k]
        lp = orc.grid_loglik(y, rho=tp[1], sigma_z=tp[2], mu=tp[0], n_grid=n_grid)[0]
        lm = orc.grid_loglik(y, rho=tm[1], sigma_z=tm[2], mu=tm[0], n_grid=n_grid)[0]
        out[k] = (lp - lm) / (2 * h[k])
    return out
