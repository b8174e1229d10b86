"""The dense PyTorch restatement (oracle/torch_svpf.py) pinned against the FP64
C oracle (same forward to rounding) and against FP64 central differences of
the oracle's loss (its autograd gradient).  CPU only."""
import numpy as np
import pytest
import torch

import oracle
from oracle import torch_svpf
from svpf_amd import config, data


def case(n=16, T=8, S=5, seed=0, gamma=0.0, mu=-1.0, rho=0.97, sz=0.15, nu=5.0, sign=0):
    cfg = config.plain_config(n=n, n_stein=S, mu=mu, rho=rho, sigma_z=sz, nu=nu, gamma=gamma, stein_sign=sign)
    y, _ = data.sv_series(T, mu=mu, rho=rho, sigma_z=sz, nu=nu, seed=seed)
    noise = data.predraw_noise(T, n, S, seed=seed + 1)
    h0 = data.initial_particles(n, mu, rho, sz, seed=seed + 2)
    return cfg, y, noise, h0


def oracle_loss(cfg, y, noise, h0):
    o = oracle.OracleFilter(cfg, h0)
    out = o.run(y, noise)
    return -float(np.sum(out["pred_loglik"])), out, o


def torch_run(cfg, y, noise, h0, theta=None):
    dt = torch.float64
    if theta is None:
        theta = [torch.tensor(cfg[k], dtype=dt, requires_grad=True) for k in ("mu", "rho", "sigma_z", "nu")]
    pl, x, v = torch_svpf.run(theta, torch.tensor(y, dtype=dt), torch.tensor(noise["z_pred"], dtype=dt),
                              torch.tensor(noise["xi"], dtype=dt), torch.tensor(h0, dtype=dt), cfg)
    return pl, x, v, theta


@pytest.mark.parametrize("gamma,sign", [(0.0, 0), (-0.3, 1)])
def test_forward_equals_oracle(gamma, sign):
    cfg, y, noise, h0 = case(n=24, T=10, gamma=gamma, sign=sign, seed=3)
    L, out, o = oracle_loss(cfg, y, noise, h0)
    pl, x, v, _ = torch_run(cfg, y, noise, h0)
    assert np.max(np.abs(pl.detach().numpy() - out["pred_loglik"])) < 1e-10
    assert np.max(np.abs(x.detach().numpy() - o.h)) < 1e-10
    assert np.max(np.abs(v.detach().numpy() - o.v)) < 1e-10


@pytest.mark.parametrize("seed", [0, 5])
def test_autograd_matches_oracle_central_differences(seed):
    cfg, y, noise, h0 = case(n=12, T=6, seed=seed, gamma=-0.2)
    pl, _, _, theta = torch_run(cfg, y, noise, h0)
    (-pl.sum()).backward()
    g = np.array([t.grad.item() for t in theta])
    fd = []
    for k in ("mu", "rho", "sigma_z", "nu"):
        # the loss is strongly curved in sigma_z: a 1e-4 step leaves an O(eps^2) error of 4e-4
        eps = 1e-6 * (10.0 if k == "nu" else 1.0)
        lp = oracle_loss(dict(cfg, **{k: cfg[k] + eps}), y, noise, h0)[0]
        lm = oracle_loss(dict(cfg, **{k: cfg[k] - eps}), y, noise, h0)[0]
        fd.append((lp - lm) / (2 * eps))
    fd = np.array(fd)
    assert np.all(np.abs(g - fd) <= 1e-6 * np.maximum(1.0, np.abs(fd))), (g, fd)


def test_median_matches_oracle():
    h = np.random.default_rng(1).standard_normal(31)
    assert torch_svpf.median_r2(torch.tensor(h)).item() == oracle.median_r2(h)
