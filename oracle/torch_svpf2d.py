"""Dense PyTorch restatement of the two-factor (2-D) SVPF tick.

TEST INFRASTRUCTURE ONLY: the checker of the HIP two-factor engine and its
adjoint.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
import it.

Model (the reference's two-factor decomposition, include/svpf_two_factor.cuh
and src/svpf_two_factor.cu: log-vol h = mu + f + s with a fast and a slow AR(1)
factor).  The reference transports each coordinate with its own 1-D kernel;
this filter moves the particle x = (f, s) with one 2-D IMQ kernel and a 2x2
Newton preconditioner, the other parts following the 1-D step of
src/svpf_opt_kernels.cu coordinate-wise:

  predict     m_i = (rho_f f_i, rho_s s_i),  x'_i = m_i + (sigma_f z_f, sigma_s z_s)
  pred ll     logmeanexp_i log p(y_t | h_i = mu + f_i + s_i)       Student-t(nu)
  bandwidth   ell_c^2 = lower median of the pairwise (x_ic - x_jc)^2, bw_c = clamp(ell_c, 0.01, 2)
  metric      K_ij = 1 / (1 + (f_i-f_j)^2/bw_f^2 + (s_i-s_j)^2/bw_s^2)
  score       g_j = clamp(g_prior_j + beta (1, 1) g_lik(h_j), -10, 10) per coordinate,
              g_prior = -sum_i r_ij D (x_j - m_i) / sum_i r_ij, D = diag(1/sigma_f^2, 1/sigma_s^2),
              r_ij = exp(-(x_j - m_i)^T D (x_j - m_i) / 2)
  Newton      C_j = D + c_j 1 1^T with c_j = -d^2 log p(y|h)/dh^2 >= 0 (positive definite),
              s_j = 0.7 C_j^-1 g_j, repulsion Q_j = sgn2 C_j^-1
  transport   phi_i = (1/N) sum_j K_ij s_j + K_ij^2 Q_j M (x_i - x_j),  M = diag(1/bw_f^2, 1/bw_s^2)
  update      per coordinate RMSProp and SVLD noise as in 1-D; factors clamped to [-20, 20]
"""
from __future__ import annotations

import math

import torch

from .torch_svpf import log_w, logmeanexp, median_r2, student_t_const

FCLAMP = 20.0


def _ell2(xc, pair):
    """The median squared distance of one coordinate, or that of the given pair (a >= 0)."""
    if pair is not None and pair[0] >= 0:
        return (xc[pair[0]] - xc[pair[1]]) ** 2
    return median_r2(xc)


def tick(x, v, y_t, z, xi, theta, cfg, pair=None):
    """x [N, 2] positions (f, s); v [N, 2] RMSProp state; z [N, 2]; xi [S, N, 2];
    theta = (mu, rho_f, sigma_f, rho_s, sigma_s, nu).  pair: None, or per coordinate
    the indices (a, b) of the pair taken as that coordinate's median (a < 0: the
    median itself), to follow the pair choices of a float32 filter.  Returns
    (x, v, pred_ll)."""
    mu, rho_f, sig_f, rho_s, sig_s, nu = theta
    n = x.shape[0]
    S = cfg["S"]
    sgn2 = 2.0 if cfg.get("paper_sign", 0) else -2.0
    rho = torch.stack([rho_f, rho_s])
    sig = torch.stack([sig_f, sig_s])
    m = x * rho
    xp = torch.clamp(m + sig * z, -FCLAMP, FCLAMP)
    stc = student_t_const(nu)
    pred_ll = logmeanexp(log_w(mu + xp[:, 0] + xp[:, 1], y_t, nu, stc))
    bw = torch.stack([torch.clamp(torch.sqrt(_ell2(xp[:, c], None if pair is None else pair[c])), 0.01, 2.0)
                      for c in range(2)])
    ib = 1.0 / (bw * bw)
    Dg = 1.0 / (sig * sig)
    ysq = y_t * y_t
    temp = cfg.get("temperature", 1.0)
    h = xp
    for k in range(S):
        beta = 0.3 + 0.7 * k / (S - 1) if S > 1 else 1.0
        d = h[:, None, :] - m[None, :, :]                      # [j, i, c]
        logr = -0.5 * (d * d * Dg).sum(-1)
        lmax = logr.max(dim=1, keepdim=True).values
        r = torch.exp(logr - lmax)
        g_prior = -(r[:, :, None] * d * Dg).sum(1) / (r.sum(1, keepdim=True) + 1e-8)
        hh = mu + h[:, 0] + h[:, 1]
        A = ysq / (torch.exp(torch.clamp(hh, max=20.0)) + 1e-8) / nu
        g_lik = -0.5 + 0.5 * (nu + 1.0) * A / (1.0 + A) - cfg.get("lik_offset", 0.0)
        g = torch.clamp(g_prior + beta * g_lik[:, None], -10.0, 10.0)
        c = 0.5 * (nu + 1.0) * A / ((1.0 + A) ** 2)
        det = Dg[0] * Dg[1] + c * (Dg[0] + Dg[1])
        Cinv = torch.stack([torch.stack([Dg[1] + c, -c], -1), torch.stack([-c, Dg[0] + c], -1)], -2) / det[:, None, None]
        s = 0.7 * (Cinv @ g[:, :, None])[:, :, 0]
        Q = sgn2 * Cinv
        D = h[:, None, :] - h[None, :, :]                      # [i, j, c]
        K = 1.0 / (1.0 + (D * D * ib).sum(-1))
        rep = (Q[None, :, :, :] @ (D * ib)[:, :, :, None])[..., 0]   # [i, j, c]
        phi = (K[:, :, None] * s[None, :, :] + (K * K)[:, :, None] * rep).sum(1) / n
        v = cfg["rms_rho"] * v + (1.0 - cfg["rms_rho"]) * phi * phi
        eff = 0.1 * math.sqrt(beta)
        noise = math.sqrt(2.0 * eff * temp) * xi[k] if temp > 1e-6 else 0.0
        h = torch.clamp(h + eff * phi / torch.sqrt(v + cfg["rms_eps"]) + noise, -FCLAMP, FCLAMP)
    return h, v, pred_ll


def run(theta, y, z, xi, x0, cfg, v0=None, pairs=None):
    """T ticks: y [T], z [T, N, 2], xi [T, S, N, 2], x0 [N, 2]; pairs: None or [T][2][2]
    median pairs (see tick) -> (pred_ll [T], x_T, v_T)."""
    x = x0
    v = torch.zeros_like(x0) if v0 is None else v0
    out = []
    for t in range(y.shape[0]):
        x, v, pl = tick(x, v, y[t], z[t], xi[t], theta, cfg, None if pairs is None else pairs[t])
        out.append(pl)
    return torch.stack(out), x, v


DEFAULT_CFG = dict(S=5, paper_sign=0, temperature=1.0, rms_rho=0.9, rms_eps=1e-6, lik_offset=0.0)
