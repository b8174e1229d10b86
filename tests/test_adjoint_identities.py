"""The identities the hand-written adjoints rely on, checked in float64 with autograd.

The mixture-prior score of particle i with components m_j and per-coordinate scale sigma,
    g_i = -sigma^-2 sum_j r_ij (x_i - m_j) / sum_j r_ij,  r_ij = exp(-sum_c (x_ic - m_jc)^2 / (2 sigma_c^2)),
is homogeneous of degree -1 in coordinate c under (x_c, m_c, sigma_c) -> lambda (x_c, m_c, sigma_c) and invariant
under a common translation of x and m.  The adjoints (csrc/svpf_diff.hpp, csrc/svpf_diff2d.hpp) take the sigma
adjoint of the prior from the first identity instead of taping third responsibility moments:
    sigma_c d(gbar . g)/dsigma_c = -sum_i gbar_ic g_ic - sum_i x_ic xbar_ic - sum_j m_jc mbar_jc.
A dropped term, a wrong sign or a swapped coordinate fails these checks.
"""
import numpy as np
import pytest
import torch


def prior_score(x, m, sig):
    d = x[:, None, :] - m[None, :, :]
    logr = -0.5 * (d * d / (sig * sig)).sum(-1)
    r = torch.exp(logr - logr.max(dim=1, keepdim=True).values)
    return -(r[:, :, None] * d / (sig * sig)).sum(1) / r.sum(1, keepdim=True)


def grads(D, seed, centre=0.0):
    g = torch.Generator().manual_seed(seed)
    n = 40
    x = (torch.randn(n, D, generator=g, dtype=torch.float64) * 0.6 + centre).requires_grad_()
    m = (torch.randn(n, D, generator=g, dtype=torch.float64) * 0.6 + centre - 0.2).requires_grad_()
    sig = (torch.tensor([0.3, 0.08][:D], dtype=torch.float64)).requires_grad_()
    gbar = torch.randn(n, D, generator=g, dtype=torch.float64)
    gp = prior_score(x, m, sig)
    xb, mb, sb = torch.autograd.grad((gbar * gp).sum(), (x, m, sig))
    return x.detach(), m.detach(), sig.detach(), gbar, gp.detach(), xb, mb, sb


@pytest.mark.parametrize("D", [1, 2])
@pytest.mark.parametrize("centre", [0.0, 1.5])
def test_sigma_adjoint_by_homogeneity(D, centre):
    x, m, sig, gbar, gp, xb, mb, sb = grads(D, seed=3 + D, centre=centre)
    for c in range(D):
        euler = -(gbar[:, c] * gp[:, c]).sum() - (x[:, c] * xb[:, c]).sum() - (m[:, c] * mb[:, c]).sum()
        assert abs(float(sig[c] * sb[c] - euler)) <= 1e-10 * max(1.0, abs(float(euler)))


def test_homogeneity_is_per_coordinate():
    """In 2-D the other coordinate's terms do not enter: using both coordinates' position and mean
    adjoints for sigma_f (a plausible slip) gives a different value."""
    x, m, sig, gbar, gp, xb, mb, sb = grads(2, seed=9)
    wrong = -(gbar[:, 0] * gp[:, 0]).sum() - (x * xb).sum() - (m * mb).sum()
    assert abs(float(sig[0] * sb[0] - wrong)) > 1e-3 * abs(float(wrong))


@pytest.mark.parametrize("D", [1, 2])
def test_translation_invariance(D):
    x, m, sig, gbar, gp, xb, mb, sb = grads(D, seed=21)
    np.testing.assert_allclose((xb.sum(0) + mb.sum(0)).numpy(), 0.0, atol=1e-10)
