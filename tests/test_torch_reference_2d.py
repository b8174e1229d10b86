"""The dense two-factor restatement (oracle/torch_svpf2d.py) pinned piece by
piece against independent forms: the mixture-prior score and the likelihood
score / curvature against autograd of the log densities, the 2x2 Newton step
against numpy's matrix inverse, the transport's repulsion against autograd of
the 2-D IMQ kernel, the bandwidths against numpy sorts.  CPU only."""
import math

import numpy as np
import pytest
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


def test_single_particle_tick_is_a_hand_assembled_newton_step():
    """N = 1: K(0) = 1 and the repulsion vanishes, so each iteration moves the particle by
    eff phi / sqrt(v' + eps) with phi = 0.7 C^-1 g, g the AR(1) prior score (one mixture
    component: -D (x - m)) plus beta times the likelihood score (autograd of the Student-t
    log density), C = D + c 1 1^T with c the negative likelihood curvature (autograd),
    inverted by numpy."""
    th = theta()
    mu, rf, sf, rs, ss, nu = (t.item() for t in th)
    S = 3
    cfg = dict(torch_svpf2d.DEFAULT_CFG, S=S, temperature=0.0)
    x0 = torch.tensor([[0.2, -0.1]], dtype=dt)
    z = torch.tensor([[0.4, -1.1]], dtype=dt)
    y = torch.tensor(0.9, dtype=dt)
    xn, vn, _ = torch_svpf2d.tick(x0, torch.zeros_like(x0), y, z, torch.zeros(S, 1, 2, dtype=dt), th, cfg)
    m = np.array([rf * 0.2, rs * -0.1])
    x = m + np.array([sf, ss]) * np.array([0.4, -1.1])
    D = np.diag([1 / sf ** 2, 1 / ss ** 2])
    v = np.zeros(2)
    for k in range(S):
        beta = 0.3 + 0.7 * k / (S - 1)
        hr = torch.tensor(mu + x[0] + x[1], dtype=dt, requires_grad=True)
        lw = log_w(hr, y, th[5], student_t_const(th[5]))
        (gl,) = torch.autograd.grad(lw, hr, create_graph=True)
        (hl,) = torch.autograd.grad(gl, hr)
        g = np.clip(-D @ (x - m) + beta * gl.item(), -10, 10)
        C = D - hl.item() * np.ones((2, 2))
        phi = 0.7 * np.linalg.inv(C) @ g
        v = 0.9 * v + 0.1 * phi * phi
        x = x + 0.1 * math.sqrt(beta) * phi / np.sqrt(v + 1e-6)
    assert np.allclose(xn.numpy()[0], x, rtol=0, atol=1e-8)  # the 1e-8 guards of the restatement
    assert np.allclose(vn.numpy()[0], v, rtol=0, atol=1e-8)


def test_loss_gradient_matches_central_differences():
    """The summed predictive log-likelihood of a short run: autograd against FP64 central
    differences (step 1e-6) in each of the six parameters."""
    rng = np.random.default_rng(4)
    n, S, T = 6, 3, 3
    cfg = dict(torch_svpf2d.DEFAULT_CFG, S=S, temperature=0.5)
    x0 = torch.tensor(rng.normal(size=(n, 2)) * [0.4, 0.2], dtype=dt)
    z = torch.tensor(rng.normal(size=(T, n, 2)), dtype=dt)
    xi = torch.tensor(rng.normal(size=(T, S, n, 2)), dtype=dt)
    y = torch.tensor([0.3, -1.2, 0.5], dtype=dt)
    base = [-1.0, 0.8, 0.3, 0.97, 0.12, 6.0]
    th = [torch.tensor(v, dtype=dt, requires_grad=True) for v in base]
    pl, _, _ = torch_svpf2d.run(th, y, z, xi, x0, cfg)
    loss = -pl.sum()
    grads = torch.autograd.grad(loss, th)
    for k in range(6):
        e = 1e-6
        hi = [torch.tensor(v + (e if i == k else 0.0), dtype=dt) for i, v in enumerate(base)]
        lo = [torch.tensor(v - (e if i == k else 0.0), dtype=dt) for i, v in enumerate(base)]
        fd = (-torch_svpf2d.run(hi, y, z, xi, x0, cfg)[0].sum() + torch_svpf2d.run(lo, y, z, xi, x0, cfg)[0].sum()) / (2 * e)
        assert grads[k].item() == pytest.approx(fd.item(), rel=1e-5, abs=1e-7), k
