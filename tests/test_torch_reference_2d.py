"""The dense two-factor restatement (oracle/torch_svpf2d.py) pinned piece by
piece against independent forms: the mixture-prior score and the likelihood
score / curvature against autograd of the log densities, the 2x2 Newton step
against numpy's matrix inverse, the transport's repulsion against autograd of
the 2-D IMQ kernel, the bandwidths against numpy sorts.  CPU only."""
import math

import numpy as np
import torch

from oracle import torch_svpf2d
from oracle.torch_svpf import log_w, student_t_const

dt = torch.float64


def theta(mu=-1.0, rf=0.8, sf=0.3, rs=0.995, ss=0.08, nu=5.0):
    return [torch.tensor(v, dtype=dt) for v in (mu, rf, sf, rs, ss, nu)]


def test_prior_and_likelihood_scores_and_newton_step():
    rng = np.random.default_rng(0)
    n = 9
    th = theta()
    mu, rf, sf, rs, ss, nu = th
    x = torch.tensor(rng.normal(size=(n, 2)) * [0.3, 0.5], dtype=dt)
    m = torch.tensor(rng.normal(size=(n, 2)) * [0.3, 0.5], dtype=dt)
    Dg = torch.stack([1 / sf ** 2, 1 / ss ** 2])
    # mixture prior score = grad of log sum_i exp(-(x-m_i)^T D (x-m_i)/2)
    xr = x.clone().requires_grad_(True)
    d = xr[:, None, :] - m[None, :, :]
    lse = torch.logsumexp(-0.5 * (d * d * Dg).sum(-1), dim=1).sum()
    (gref,) = torch.autograd.grad(lse, xr)
    d = x[:, None, :] - m[None, :, :]
    logr = -0.5 * (d * d * Dg).sum(-1)
    r = torch.exp(logr - logr.max(1, keepdim=True).values)
    gp = -(r[:, :, None] * d * Dg).sum(1) / (r.sum(1, keepdim=True) + 1e-8)
    assert torch.allclose(gp, gref, rtol=1e-6, atol=1e-6)
    # likelihood score and curvature in h, shared by both factors
    y = torch.tensor(0.7, dtype=dt)
    hr = (mu + x[:, 0] + x[:, 1]).clone().requires_grad_(True)
    lw = log_w(hr, y, nu, student_t_const(nu)).sum()
    (gl_ref,) = torch.autograd.grad(lw, hr, create_graph=True)
    (hl_ref,) = torch.autograd.grad(gl_ref.sum(), hr)
    A = y * y / (torch.exp(hr.detach()) + 1e-8) / nu
    gl = -0.5 + 0.5 * (nu + 1) * A / (1 + A)
    c = 0.5 * (nu + 1) * A / (1 + A) ** 2
    assert torch.allclose(gl, gl_ref.detach(), rtol=1e-6, atol=1e-8)
    assert torch.allclose(c, -hl_ref, rtol=1e-5, atol=1e-8)
    # closed-form 2x2 inverse of D + c 1 1^T
    for j in range(n):
        C = np.diag(Dg.numpy()) + c[j].item() * np.ones((2, 2))
        det = Dg[0] * Dg[1] + c[j] * (Dg[0] + Dg[1])
        Ci = np.array([[Dg[1] + c[j], -c[j]], [-c[j], Dg[0] + c[j]]]) / det.item()
        assert np.allclose(Ci, np.linalg.inv(C), rtol=1e-12)


def test_repulsion_is_half_the_imq_kernel_gradient():
    rng = np.random.default_rng(1)
    n = 7
    x = torch.tensor(rng.normal(size=(n, 2)), dtype=dt)
    ib = torch.tensor([1 / 0.3 ** 2, 1 / 0.7 ** 2], dtype=dt)
    Q = torch.tensor(rng.normal(size=(n, 2, 2)), dtype=dt)
    D = x[:, None, :] - x[None, :, :]
    K = 1.0 / (1.0 + (D * D * ib).sum(-1))
    rep = (Q[None] @ (D * ib)[..., None])[..., 0]
    got = ((K * K)[:, :, None] * rep).sum(1)
    want = torch.zeros(n, 2, dtype=dt)
    for i in range(n):
        for j in range(n):
            xj = x[j].clone().requires_grad_(True)
            dd = x[i] - xj
            k = 1.0 / (1.0 + (dd * dd * ib).sum())
            (gk,) = torch.autograd.grad(k, xj)
            want[i] += 0.5 * Q[j] @ gk
    assert torch.allclose(got, want, rtol=1e-10, atol=1e-12)


def test_tick_bandwidths_and_shapes():
    rng = np.random.default_rng(2)
    n, S = 33, 5
    cfg = dict(torch_svpf2d.DEFAULT_CFG, S=S)
    x = torch.tensor(rng.normal(size=(n, 2)) * [0.4, 0.6], dtype=dt)
    z = torch.tensor(rng.normal(size=(n, 2)), dtype=dt)
    xi = torch.tensor(rng.normal(size=(S, n, 2)), dtype=dt)
    th = theta()
    xn, v, pl = torch_svpf2d.tick(x, torch.zeros_like(x), torch.tensor(0.4, dtype=dt), z, xi, th, cfg)
    assert xn.shape == (n, 2) and torch.isfinite(xn).all() and torch.isfinite(pl)
    xp = x * torch.tensor([0.8, 0.995], dtype=dt) + torch.tensor([0.3, 0.08], dtype=dt) * z
    for c in range(2):
        a = xp[:, c].numpy()
        r2 = np.sort([(a[i] - a[j]) ** 2 for i in range(n) for j in range(i + 1, n)])
        assert torch_svpf2d.median_r2(xp[:, c]).item() == r2[(len(r2) - 1) // 2]
    # predictive loglik at the predicted particles
    h = 0.0 + (-1.0) + xp[:, 0] + xp[:, 1]
    ref = np.log(np.mean(np.exp(log_w(h, torch.tensor(0.4, dtype=dt), th[5], student_t_const(th[5])).numpy())))
    assert abs(pl.item() - ref) < 1e-12
