"""Synthetic stochastic-volatility series and pre-drawn filter noise.

The generating process is the reference's SV model (include/svpf.cuh:249-256,
py-torch/diff_svpf_torch.py:32-47):

    h_t = mu + rho (h_{t-1} - mu) + sigma_z eps_t,   y_t = exp(h_t / 2) eta_t,
    eta_t ~ Student-t(nu),  h_0 ~ N(mu, sigma_z^2 / (1 - rho^2)).

Everything is drawn from numpy's PCG64 generator with an explicit seed so that
the oracle, the CPU torch path and the device path consume identical inputs.
"""
from __future__ import annotations

import re

import numpy as np


def sv_series(T: int, mu=-1.0, rho=0.97, sigma_z=0.15, nu=5.0, seed=0):
    """Plain AR(1) log-volatility with Student-t returns.  Returns (y, h) float32."""
    rng = np.random.default_rng(seed)
    h = np.empty(T, dtype=np.float64)
    sd0 = sigma_z / np.sqrt(1.0 - rho * rho)
    h[0] = mu + sd0 * rng.standard_normal()
    for t in range(1, T):
        h[t] = mu + rho * (h[t - 1] - mu) + sigma_z * rng.standard_normal()
    eta = rng.standard_t(nu, size=T)
    y = np.exp(h / 2.0) * eta
    return y.astype(np.float32), h.astype(np.float32)


def regime_shift_series(T=2000, mu0=-1.0, mu1=1.5, t_shift=1000, jump_prob=0.02, jump=2.0,
                        rho_up=0.90, rho_down=0.98, sigma_z=0.2, nu=4.0, seed=0):
    """Regime-shift series: mu steps mu0 -> mu1 at t_shift, additive jumps of
    +jump with probability jump_prob, asymmetric persistence (rho_up when the
    log-vol rose on the previous step, rho_down otherwise)."""
    rng = np.random.default_rng(seed)
    h = np.empty(T, dtype=np.float64)
    h[0] = mu0 + sigma_z / np.sqrt(1.0 - rho_down ** 2) * rng.standard_normal()
    prev_up = False
    for t in range(1, T):
        mu = mu0 if t < t_shift else mu1
        rho = rho_up if prev_up else rho_down
        x = mu + rho * (h[t - 1] - mu) + sigma_z * rng.standard_normal()
        if rng.random() < jump_prob:
            x += jump
        prev_up = x > h[t - 1]
        h[t] = x
    eta = rng.standard_t(nu, size=T)
    y = np.exp(h / 2.0) * eta
    return y.astype(np.float32), h.astype(np.float32)


def two_factor_series(T=1000, mu=-1.0, rho_f=0.80, sigma_f=0.3, rho_s=0.995, sigma_s=0.08,
                      nu=5.0, seed=0):
    """Two-factor log-vol h = mu + h_f + h_s with independent AR(1) factors."""
    rng = np.random.default_rng(seed)
    hf = np.empty(T)
    hs = np.empty(T)
    hf[0] = sigma_f / np.sqrt(1 - rho_f ** 2) * rng.standard_normal()
    hs[0] = sigma_s / np.sqrt(1 - rho_s ** 2) * rng.standard_normal()
    for t in range(1, T):
        hf[t] = rho_f * hf[t - 1] + sigma_f * rng.standard_normal()
        hs[t] = rho_s * hs[t - 1] + sigma_s * rng.standard_normal()
    h = mu + hf + hs
    y = np.exp(h / 2.0) * rng.standard_t(nu, size=T)
    return y.astype(np.float32), h.astype(np.float32), hf.astype(np.float32), hs.astype(np.float32)


def initial_particles(n, mu, rho, sigma_z, seed):
    """h_0 particles from N(mu, sigma_z^2 / (1 - rho^2 + 1e-6)) clamped to [-15, 5]
    (src/svpf_optimized_graph.cu:84-97, :373-388), drawn on the host."""
    rng = np.random.default_rng(seed)
    sd = np.sqrt(sigma_z * sigma_z / (1.0 - rho * rho + 1e-6))
    h0 = mu + sd * rng.standard_normal(n)
    return np.clip(h0, -15.0, 5.0).astype(np.float32)


def predraw_noise(T, n, stein_rows, seed, mim=False, rejuv=False):
    """Pre-drawn per-tick noise: predict normals z[T, n], Langevin normals
    xi[T, stein_rows, n] and, when asked, MIM selector uniforms u[T, n] and
    rejuvenation uniforms / normals."""
    rng = np.random.default_rng(seed)
    out = {
        "z_pred": rng.standard_normal((T, n), dtype=np.float32),
        "xi": rng.standard_normal((T, stein_rows, n), dtype=np.float32),
    }
    if mim:
        out["u_pred"] = rng.random((T, n), dtype=np.float32)
    if rejuv:
        out["u_rej"] = rng.random((T, n), dtype=np.float32)
        out["z_rej"] = rng.standard_normal((T, n), dtype=np.float32)
    return out


def batched_thetas(B, seed=0, sizes=(128, 256, 512, 1024, 2048)):
    """Per-filter particle counts and parameters of the batched workload:
    N ~ U{sizes}, mu ~ U(-2, 0), rho ~ U(0.90, 0.995), sigma_z ~ U(0.05, 0.4),
    nu ~ U(3, 30)."""
    rng = np.random.default_rng(seed)
    n = rng.choice(np.asarray(sizes), size=B)
    mu = rng.uniform(-2.0, 0.0, B)
    rho = rng.uniform(0.90, 0.995, B)
    sz = rng.uniform(0.05, 0.4, B)
    nu = rng.uniform(3.0, 30.0, B)
    return n.astype(np.int32), mu, rho, sz, nu


# ---- market-data formats (include/market_data_loader.h of this repository,
# the reference's include/market_data_loader.h:37-237) ----

def load_returns_binary(path) -> np.ndarray:
    """<name>.bin: raw little-endian float32 returns."""
    return np.fromfile(path, dtype="<f4").astype(np.float32)


def load_returns_csv(path) -> np.ndarray:
    """<name>_returns.csv: one return per line (a blank line reads as 0, as atof does)."""
    with open(path) as f:
        return np.array([_atof(ln) for ln in f], dtype=np.float32)


def load_market_data_csv(path):
    """<name>.csv: header, then date,price,return,log_return.  Returns (log_returns, prices) float32."""
    with open(path) as f:
        rows = f.read().splitlines()[1:]
    cols = [r.split(",") for r in rows]
    ret = np.array([_atof(c[3]) if len(c) > 3 else 0.0 for c in cols], dtype=np.float32)
    px = np.array([_atof(c[1]) if len(c) > 1 else 0.0 for c in cols], dtype=np.float32)
    return ret, px


def return_stats(r):
    """(mean, std, min, max, n_3sigma, n_4sigma) as compute_return_stats: sums in
    double, population std, |r - mean| / std tested in float32."""
    r = np.asarray(r, dtype=np.float32)
    m = np.float32(np.sum(r.astype(np.float64)) / len(r))
    d = (r - m).astype(np.float64)
    sd = np.float32(np.sqrt(np.sum(d * d) / len(r)))
    z = np.abs((r - m) / sd)
    return float(m), float(sd), float(r.min()), float(r.max()), int((z > 3.0).sum()), int((z > 4.0).sum())


def _atof(s: str) -> float:
    """C atof: the longest leading decimal number, 0.0 if there is none."""
    m = re.match(r"\s*([+-]?(\d+\.?\d*|\.\d+)([eE][+-]?\d+)?)", s)
    return float(m.group(1)) if m else 0.0
