"""Dense PyTorch restatement of the differentiable SVPF tick.

TEST INFRASTRUCTURE ONLY: the autograd reference that the hand-written HIP
adjoint is checked against, and the CPU baseline timed beside it.  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import it.

It plays the role of the reference's PyTorch path (py-torch/diff_svpf_torch.py:
147-217: dense [N, N] kernels, predictive log-likelihood before the Stein
update, autograd through the whole filter) but computes the tick of the C++
step -- the same math as oracle/svpf_oracle.c svpfo_tick with the plain
feature set:

  predict        src/svpf_opt_kernels.cu:125-155   h = clamp(mu + rho (x - mu) + sigma z + lev)
  predictive ll  logmeanexp_i log p(y_t | h_i)     (the loss of py-torch/diff_svpf_torch.py:193-197)
  bandwidth      ell^2 = lower median of the pairwise r^2, bw = clamp(ell, 0.01, 2)
  Stein loop     src/svpf_opt_kernels.cu:354-521 (mixture prior, exact Student-t score, local Newton)
                 src/svpf_opt_kernels.cu:742-906 (IMQ transport with Newton), RMSProp, SVLD noise

Every clamp, the 1e-8 guards and the responsibility normaliser of the
reference are kept, so that float64 results agree with the C oracle to
rounding and autograd differentiates exactly the function the oracle
computes (piecewise: the median pair and the clamp pattern are those of the
evaluation point).
"""
from __future__ import annotations

import math

import torch


def student_t_const(nu):
    """src/svpf_optimized_graph.cu:123-125, differentiable in nu."""
    return torch.lgamma((nu + 1.0) / 2.0) - torch.lgamma(nu / 2.0) - 0.5 * torch.log(math.pi * nu)


def log_w(h, y, nu, st_const):
    """src/svpf_opt_kernels.cu:474-488."""
    vol = torch.exp(torch.clamp(h, max=20.0))
    A = (y * y) / (vol + 1e-8) / nu
    return st_const - 0.5 * h - (nu + 1.0) * 0.5 * torch.log1p(A)


def logmeanexp(x):
    m = x.max().detach()
    return m + torch.log(torch.clamp(torch.exp(x - m).mean(), min=1e-10))


def median_r2(h):
    """Lower median (rank floor((M-1)/2)) of the N(N-1)/2 pairwise squared distances."""
    n = h.shape[0]
    if n < 2:
        return h.sum() * 0.0  # no pair: 0, as the C oracle and the kernels
    iu = torch.triu_indices(n, n, offset=1, device=h.device)
    r2 = (h[iu[0]] - h[iu[1]]) ** 2
    k = (r2.numel() - 1) // 2
    return torch.kthvalue(r2, k + 1).values


def tick(x, v, y_t, y_prev, z, xi, theta, cfg, pair=None):
    """One tick.  x: positions after the previous tick [N]; v: RMSProp state [N];
    z: predict normals [N]; xi: Langevin normals [S, N]; theta = (mu, rho,
    sigma_z, nu) as 0-d tensors.  pair: None, or the indices (a, b) of the pair
    whose squared distance is taken as the median (a < 0: the median itself);
    a float32 filter and this float64 one can select different pairs of nearly
    equal distance, whose gradients differ.  Returns (x_new, v_new, pred_ll)."""
    mu, rho, sigma, nu = theta
    n = x.shape[0]
    S = cfg["stein_max_steps"]
    gamma = cfg["gamma"]
    sgn2 = 2.0 if cfg["stein_sign"] == 1 else -2.0
    # predict; the mixture means of the prior are the same m
    lev = gamma * y_prev / (torch.exp(x * 0.5) + 1e-8)
    m = mu + rho * (x - mu) + lev
    h = torch.clamp(m + sigma * z, -15.0, 5.0)
    stc = student_t_const(nu)
    pred_ll = logmeanexp(log_w(h, y_t, nu, stc))
    if pair is not None and pair[0] >= 0:
        ell2 = (h[pair[0]] - h[pair[1]]) ** 2
    else:
        ell2 = median_r2(h)
    bw = torch.clamp(torch.sqrt(ell2), 0.01, 2.0)
    ib = 1.0 / (bw * bw)
    inv_s2 = 1.0 / (sigma * sigma)
    ysq = y_t * y_t
    base_step = 0.1
    temp = cfg["temperature"] if cfg["use_svld"] else 0.0
    for k in range(S):
        beta = 0.3 + 0.7 * k / (S - 1) if S > 1 else 1.0
        # mixture prior (src/svpf_opt_kernels.cu:453-470): rows j = particles, columns i = components
        d = h[:, None] - m[None, :]
        logr = -d * d * (0.5 * inv_s2)
        lmax = logr.max(dim=1, keepdim=True).values
        r = torch.exp(logr - lmax)
        g_prior = -(r * d).sum(dim=1) * inv_s2 / (r.sum(dim=1) + 1e-8)
        vol = torch.exp(torch.clamp(h, max=20.0))
        A = ysq / (vol + 1e-8) / nu
        g_lik = -0.5 + 0.5 * (nu + 1.0) * A / (1.0 + A) - cfg["lik_offset"]
        g = torch.clamp(g_prior + beta * g_lik, -10.0, 10.0)
        h_lik = -0.5 * (nu + 1.0) * A / ((1.0 + A) ** 2)
        curv = torch.clamp(-(h_lik - inv_s2), 0.1, 100.0)
        s = 0.7 * g / curv
        c = sgn2 / curv
        # IMQ transport with local Newton (src/svpf_opt_kernels.cu:742-906)
        D = h[:, None] - h[None, :]
        K = 1.0 / (1.0 + D * D * ib)
        phi = (K * s[None, :] + c[None, :] * D * ib * K * K).sum(dim=1) / n
        v = cfg["rmsprop_rho"] * v + (1.0 - cfg["rmsprop_rho"]) * phi * phi
        eff = base_step * math.sqrt(beta)
        drift = eff * phi / torch.sqrt(v + cfg["rmsprop_eps"])
        noise = math.sqrt(2.0 * eff * temp) * xi[k] if temp > 1e-6 else 0.0
        h = torch.clamp(h + drift + noise, -15.0, 5.0)
    return h, v, pred_ll


def run(theta, y, z, xi, h0, cfg, v0=None, y_prev0=0.0, pairs=None):
    """T ticks.  y [T]; z [T, N]; xi [T, S, N]; h0 [N]; pairs: None or [T, 2]
    median pairs (see tick).  Returns (pred_ll [T], x_T, v_T).  The loss of the
    differentiable form is -pred_ll.sum()."""
    x = h0
    v = torch.zeros_like(h0) if v0 is None else v0
    out = []
    yp = torch.as_tensor(y_prev0, dtype=h0.dtype)
    for t in range(y.shape[0]):
        x, v, pl = tick(x, v, y[t], yp, z[t], xi[t], theta, cfg, None if pairs is None else pairs[t])
        out.append(pl)
        yp = y[t]
    return torch.stack(out), x, v
