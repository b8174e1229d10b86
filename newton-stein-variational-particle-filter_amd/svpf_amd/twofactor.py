"""Two-factor (2-D) SVPF on the GPU: B filters, particle x = (f, s), log-vol
h = mu + f + s, per-filter theta = (mu, rho_f, sigma_f, rho_s, sigma_s, nu).

``TwoFactorBatch`` owns the device buffers and drives ``svpf_diff2d_forward``
/ ``svpf_diff2d_backward`` of libsvpf_amd.so; the forward is also the batched
two-factor filter (per-tick records of the predictive log-likelihood, the
filtered h / f / s means, vol and the two bandwidths).  ``PredictiveLoglik2D``
is the ``torch.autograd.Function`` over it, and ``calibrate`` fits theta by
Adam on the summed negative predictive log-likelihood.

The reference decomposes the log-vol the same way (include/svpf_two_factor.cuh,
src/svpf_two_factor.cu) but moves each factor with its own 1-D kernel; this
filter uses one 2-D IMQ kernel with per-coordinate median bandwidths and a
2x2 Newton preconditioner (csrc/svpf_diff2d.hpp).  Record layouts mirror
csrc/svpf_types.h.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib

D2CONFIG_DTYPE = np.dtype([("n", "<i4"), ("S", "<i4"), ("T", "<i4"), ("C", "<i4"), ("paper_sign", "<i4"),
                           ("pad0", "<i4", (3,)), ("lik_offset", "<f4"), ("temperature", "<f4"), ("rms_rho", "<f4"),
                           ("rms_eps", "<f4"), ("pad1", "<f4", (8,))])
D2DESC_DTYPE = np.dtype([(f, "<u8") for f in ("theta", "y", "x0", "z", "xi", "out", "ell2", "ckpt", "tape", "gpl",
                                              "gtheta", "gx0", "xT", "seed")]
                        + [("filter_id", "<u4"), ("pad", "<i4"), ("drift", "<u8"), ("mpair", "<u8")])
assert D2CONFIG_DTYPE.itemsize == 80 and D2DESC_DTYPE.itemsize == 136

OUTW = 8
O_PRED_LL, O_HMEAN, O_FMEAN, O_SMEAN, O_VOL, O_BWF, O_BWS = range(7)
THETA_NAMES = ("mu", "rho_f", "sigma_f", "rho_s", "sigma_s", "nu")


def _bind(L):
    if getattr(L, "_svpf_d2_bound", False):
        return
    L.svpf_diff2d_tape_floats.restype = ctypes.c_size_t
    L.svpf_diff2d_tape_floats.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
    for name in ("svpf_diff2d_forward", "svpf_diff2d_backward"):
        f = getattr(L, name)
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    L._svpf_d2_bound = True


class TwoFactorBatch:
    """Device workspace for B two-factor filters of n particles over T ticks.

    y [B, T]; x0 [B, 2, n] initial factors; noise None (counter generator keyed
    by ``seeds``) or pre-drawn z [B, T, 2, n] and xi [B, T, S, 2, n].
    drift_budget_gb: the forward keeps every iteration's Stein drift for the
    backward (which then replays positions instead of recomputing the
    transport) when that buffer fits this many GB.
    record_pairs: keep, per tick and coordinate, the median pair (a, b) the
    adjoint differentiates through (``self.mpair`` [B, T, 2, 2], -1 where that
    bandwidth is clamped), filled by the backward.
    """

    def __init__(self, y, x0, n_stein=5, checkpoint_every=32, noise=None, seeds=None, lik_offset=0.0,
                 temperature=1.0, rmsprop_rho=0.9, rmsprop_eps=1e-6, paper_sign=False, device="cuda",
                 drift_budget_gb=32.0, record_pairs=False):
        self.L = _lib.load()
        _bind(self.L)
        dev = torch.device(device)
        f32 = dict(dtype=torch.float32, device=dev)
        self.y = torch.as_tensor(y, **f32).contiguous()
        self.x0 = torch.as_tensor(x0, **f32).contiguous()
        B, T = self.y.shape
        n = self.x0.shape[2]
        assert self.x0.shape[:2] == (B, 2) and 2 <= n <= 2048
        C = max(1, min(int(checkpoint_every), T))
        self.B, self.T, self.n, self.S, self.C = B, T, n, int(n_stein), C
        nseg = (T + C - 1) // C
        self.theta = torch.zeros((B, 6), **f32)
        self.out = torch.zeros((B, T, OUTW), **f32)
        self.ell2 = torch.zeros((B, T, 2), **f32)
        self.ckpt = torch.zeros((B, nseg, 4, n), **f32)
        tape_f = int(self.L.svpf_diff2d_tape_floats(n, self.S, C))
        self.tape = torch.empty((B, tape_f), **f32)
        self.gpl = torch.zeros((B, T), **f32)
        self.gtheta = torch.zeros((B, 6), **f32)
        self.gx0 = torch.zeros((B, 2, n), **f32)
        self.xT = torch.zeros((B, 2, n), **f32)
        # the Stein drift and the prior's shift of every forward iteration (the backward replays
        # positions from it and skips the prior's nearest-component pass), kept when it fits the
        # budget; otherwise the backward recomputes the transport
        drift_gb = B * T * self.S * 3 * n * 4 / 1e9
        self.drift = torch.empty((B, T, self.S, 3, n), **f32) if drift_gb <= drift_budget_gb else None
        self.mpair = torch.full((B, T, 2, 2), -1, dtype=torch.int32, device=dev) if record_pairs else None
        self.noise = None
        if noise is not None:
            self.noise = {k: torch.as_tensor(noise[k], **f32).contiguous() for k in ("z", "xi")}
            assert self.noise["z"].shape == (B, T, 2, n) and self.noise["xi"].shape == (B, T, self.S, 2, n)
        seeds = np.arange(B, dtype=np.uint64) if seeds is None else np.asarray(seeds, dtype=np.uint64)
        cfg = np.zeros(1, dtype=D2CONFIG_DTYPE)
        cfg["n"], cfg["S"], cfg["T"], cfg["C"] = n, self.S, T, C
        cfg["paper_sign"] = 1 if paper_sign else 0
        cfg["lik_offset"], cfg["temperature"] = lik_offset, temperature
        cfg["rms_rho"], cfg["rms_eps"] = rmsprop_rho, rmsprop_eps
        self.cfg = cfg
        desc = np.zeros(B, dtype=D2DESC_DTYPE)
        ptr = lambda t, b, stride: t.data_ptr() + 4 * b * stride
        for b in range(B):
            d = desc[b]
            d["theta"] = ptr(self.theta, b, 6)
            d["y"] = ptr(self.y, b, T)
            d["x0"] = ptr(self.x0, b, 2 * n)
            if self.noise is not None:
                d["z"] = ptr(self.noise["z"], b, T * 2 * n)
                d["xi"] = ptr(self.noise["xi"], b, T * self.S * 2 * n)
            d["out"] = ptr(self.out, b, T * OUTW)
            d["ell2"] = ptr(self.ell2, b, 2 * T)
            d["ckpt"] = ptr(self.ckpt, b, nseg * 4 * n)
            d["tape"] = ptr(self.tape, b, tape_f)
            d["gpl"] = ptr(self.gpl, b, T)
            d["gtheta"] = ptr(self.gtheta, b, 6)
            d["gx0"] = ptr(self.gx0, b, 2 * n)
            d["xT"] = ptr(self.xT, b, 2 * n)
            if self.drift is not None:
                d["drift"] = ptr(self.drift, b, T * self.S * 3 * n)
            if self.mpair is not None:
                d["mpair"] = ptr(self.mpair, b, 4 * T)
            d["seed"] = seeds[b]
            d["filter_id"] = b
        self.desc = torch.from_numpy(desc.view(np.uint8).reshape(-1).copy()).to(dev)
        self.forward_id = 0

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.y.device).cuda_stream)

    def forward(self, theta):
        """theta [B, 6] -> per-tick records [B, T, 8] (a view of the workspace); [..., 0] is pred_ll."""
        self.theta.copy_(theta.detach().to(self.theta.dtype))
        rc = self.L.svpf_diff2d_forward(self.cfg.ctypes.data, self.desc.data_ptr(), self.B, self._stream())
        if rc:
            raise RuntimeError(f"svpf_diff2d_forward failed ({rc})")
        self.forward_id += 1
        return self.out

    def backward(self, grad_pred_ll):
        """dL/d pred_ll [B, T] -> (dL/d theta [B, 6], dL/d x0 [B, 2, n]) for the theta of the last forward."""
        self.gpl.copy_(grad_pred_ll.detach().to(self.gpl.dtype))
        rc = self.L.svpf_diff2d_backward(self.cfg.ctypes.data, self.desc.data_ptr(), self.B, self._stream())
        if rc:
            raise RuntimeError(f"svpf_diff2d_backward failed ({rc})")
        return self.gtheta, self.gx0


class PredictiveLoglik2D(torch.autograd.Function):
    """pred_ll = PredictiveLoglik2D.apply(theta, batch): theta [B, 6] -> [B, T]."""

    @staticmethod
    def forward(ctx, theta, batch):
        ctx.batch = batch
        out = batch.forward(theta)[:, :, O_PRED_LL].clone()
        ctx.forward_id = batch.forward_id
        ctx.save_for_backward(theta.detach().clone())
        return out

    @staticmethod
    def backward(ctx, grad):
        batch = ctx.batch
        if batch.forward_id != ctx.forward_id:
            (theta,) = ctx.saved_tensors
            batch.forward(theta)
        gth, _ = batch.backward(grad.contiguous())
        return gth.clone(), None


def predictive_loglik_2d(theta, batch):
    return PredictiveLoglik2D.apply(theta, batch)


# unconstrained parameterisation: mu, atanh(rho_f), log(sigma_f), atanh(rho_s), log(sigma_s), log(nu - 2)
def to_unconstrained(theta):
    th = torch.as_tensor(theta, dtype=torch.float32)
    return torch.stack([th[..., 0], torch.atanh(th[..., 1]), torch.log(th[..., 2]), torch.atanh(th[..., 3]),
                        torch.log(th[..., 4]), torch.log(th[..., 5] - 2.0)], -1)


def to_constrained(u):
    return torch.stack([u[..., 0], torch.tanh(u[..., 1]), torch.exp(u[..., 2]), torch.tanh(u[..., 3]),
                        torch.exp(u[..., 4]), 2.0 + torch.exp(u[..., 5])], -1)


def calibrate(batch, theta0, epochs=200, lr=1e-2, fixed=None, shared=False, log=None):
    """Adam on the unconstrained parameters, loss = -sum_t pred_ll per filter.
    shared=False: every filter fits its own theta (theta0 [B, 6]); shared=True: one
    theta (theta0 [6]) for all filters, whose losses add up (series drawn from one
    generating theta).  fixed: names of parameters held at theta0.
    Returns (theta [B, 6] or [6], loss history [epochs, B])."""
    u = to_unconstrained(torch.as_tensor(theta0, device=batch.y.device)).clone().requires_grad_(True)
    mask = torch.ones(6, device=batch.y.device)
    for name in fixed or ():
        mask[THETA_NAMES.index(name)] = 0.0
    opt = torch.optim.Adam([u], lr=lr)
    hist = []
    for ep in range(epochs):
        opt.zero_grad()
        th = to_constrained(u)
        if shared:
            th = th.expand(batch.B, 6)
        pl = predictive_loglik_2d(th, batch)
        loss = -pl.sum(1)
        loss.sum().backward()
        u.grad *= mask
        opt.step()
        hist.append(loss.detach().cpu().numpy())
        if log and (ep % 20 == 0 or ep == epochs - 1):
            log(ep, float(loss.mean()))
    return to_constrained(u.detach()), np.stack(hist)
