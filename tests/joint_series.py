"""The synthetic series of the reference's joint-filter tests
(svpf_joint/test/test_svpf_joint.cu:28-60 and :190-226), regenerated with the
C library's own srand / rand through ctypes, so the draws are those of that test
binary on this C library; the Box-Muller arithmetic runs in float32 as there."""
from __future__ import annotations

import ctypes
import ctypes.util

import numpy as np

_libc = ctypes.CDLL(ctypes.util.find_library("c") or "libc.so.6")
_libc.rand.restype = ctypes.c_int
_libc.srand.argtypes = [ctypes.c_uint]
RAND_MAX = 2147483647
_F = np.float32


def _randn() -> float:
    u1 = (_F(_libc.rand()) + _F(1.0)) / (_F(RAND_MAX) + _F(2.0))
    u2 = (_F(_libc.rand()) + _F(1.0)) / (_F(RAND_MAX) + _F(2.0))
    return float(np.sqrt(_F(-2.0) * np.log(u1)) * np.cos(_F(2.0) * _F(np.pi) * u2))


def sv_series(T: int, mu: float, rho: float, sigma_z: float, seed: int):
    """generate_sv_data: h_0 = mu, then AR(1); y_t = exp(h_t / 2) * noise."""
    _libc.srand(seed)
    y = np.zeros(T, dtype=np.float32)
    h = np.zeros(T, dtype=np.float32)
    h[0] = mu
    for t in range(T):
        if t > 0:
            h[t] = _F(mu) + _F(rho) * (h[t - 1] - _F(mu)) + _F(sigma_z) * _F(_randn())
        y[t] = np.exp(h[t] * _F(0.5)) * _F(_randn())
    return y, h


def crash_series():
    """The crash-adaptation series: 200 calm steps (mu -3.5, rho 0.95, sigma 0.15, seed 42),
    then 100 crash steps (seed 123): 5 jumps around -1.0, then AR(1) with mu -1.0, rho 0.85,
    sigma 0.5."""
    T_calm, T_crash = 200, 100
    y = np.zeros(T_calm + T_crash, dtype=np.float32)
    h = np.zeros(T_calm + T_crash, dtype=np.float32)
    y[:T_calm], h[:T_calm] = sv_series(T_calm, -3.5, 0.95, 0.15, 42)
    _libc.srand(123)
    h_prev = h[T_calm - 1]
    for t in range(T_crash):
        if t < 5:
            ht = _F(-1.0) + _F(_randn()) * _F(0.5)
        else:
            ht = _F(-1.0) + _F(0.85) * (h_prev - _F(-1.0)) + _F(0.5) * _F(_randn())
        h[T_calm + t] = ht
        h_prev = ht
        y[T_calm + t] = np.exp(ht * _F(0.5)) * _F(_randn())
    return y, h, T_calm


def default_oracle_cfg(n: int, n_stein: int, **kw) -> dict:
    """The reference defaults (svpf_joint/src/svpf_joint.cu:21-72) as an oracle configuration."""
    from math import lgamma, log, pi

    nu = 30.0
    cfg = dict(n=n, n_stein_steps=n_stein, step_h=0.15, step_mu=0.005, step_rho=0.002, step_sigma=0.03,
               diffusion_mu=0.005, diffusion_rho=0.0005, diffusion_sigma=0.002, collapse_thresh_mu=0.05,
               collapse_thresh_rho=0.02, collapse_thresh_sigma=0.02, mu_prior_mean=-3.5, mu_prior_var=10.0,
               rho_prior_mean=2.0, rho_prior_var=5.0, sigma_prior_mean=-2.0, sigma_prior_var=5.0, nu=nu,
               student_t_const=lgamma((nu + 1) / 2) - lgamma(nu / 2) - 0.5 * log(pi * nu),
               student_t_implied_offset=0.0, lik_offset=0.10, prior_weight=0.05)
    cfg.update(kw)
    return cfg
