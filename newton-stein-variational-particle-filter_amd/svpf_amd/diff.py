"""Differentiable SVPF on the GPU: B filters, per-filter theta = (mu, rho,
sigma_z, nu), loss terms = the predictive log-likelihood of every tick.

``DiffBatch`` owns the device buffers (observations, checkpoints, tape, ...)
and drives ``svpf_diff_forward`` / ``svpf_diff_backward`` of libsvpf_amd.so.
``PredictiveLoglik`` is the ``torch.autograd.Function`` over it: the forward
runs the filter (HIP kernels) and keeps a checkpoint every ``checkpoint_every``
ticks; the backward is the hand-written adjoint.  The reference's PyTorch path
(py-torch/diff_svpf_torch.py:147-217) gets the same gradient from autograd
over dense [N, N] tensors.

Feature set of the differentiable tick: predict (with leverage gamma), the
median-r^2 IMQ bandwidth, S Stein iterations with beta 0.3 -> 1.0, mixture
prior, exact Student-t score, local Newton, RMSProp, SVLD at ``temperature``.
Record layouts mirror csrc/svpf_types.h.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _lib

DCONFIG_DTYPE = np.dtype([("n", "<i4"), ("S", "<i4"), ("T", "<i4"), ("C", "<i4"), ("paper_sign", "<i4"),
                          ("pad0", "<i4", (3,)), ("gamma", "<f4"), ("lik_offset", "<f4"), ("temperature", "<f4"),
                          ("rms_rho", "<f4"), ("rms_eps", "<f4"), ("pad1", "<f4", (3,)), ("pad2", "<f4", (8,))])
DDESC_DTYPE = np.dtype([(f, "<u8") for f in ("theta", "y", "h0", "z", "xi", "pred_ll", "ell2", "ckpt", "tape", "gpl",
                                              "gtheta", "gh0", "hT", "seed")]
                       + [("filter_id", "<u4"), ("pad", "<i4"), ("drift", "<u8"), ("mpair", "<u8")])
assert DCONFIG_DTYPE.itemsize == 96 and DDESC_DTYPE.itemsize == 136


def _bind(L):
    if getattr(L, "_svpf_diff_bound", False):
        return
    L.svpf_diff_tape_floats.restype = ctypes.c_size_t
    L.svpf_diff_tape_floats.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
    for name in ("svpf_diff_forward", "svpf_diff_backward"):
        f = getattr(L, name)
        f.restype = ctypes.c_int
        f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]
    L._svpf_diff_bound = True


class DiffBatch:
    """Device workspace for B differentiable filters of n particles over T ticks.

    y [B, T] observations; h0 [B, n] initial particles; noise: None (counter
    generator keyed by ``seeds``) or a dict of pre-drawn tensors z [B, T, n]
    and xi [B, T, S, n].
    drift_budget_gb: the forward keeps every iteration's Stein drift for the
    backward (which then replays positions instead of recomputing the
    transport) when that buffer fits this many GB.
    record_pairs: keep, per tick, the median pair (a, b) the adjoint
    differentiates through (``self.mpair`` [B, T, 2], -1 where the bandwidth
    is clamped), filled by the backward.
    """

    def __init__(self, y, h0, n_stein=5, checkpoint_every=32, noise=None, seeds=None, gamma=0.0, lik_offset=0.0,
                 temperature=1.0, rmsprop_rho=0.9, rmsprop_eps=1e-6, paper_sign=False, device="cuda",
                 drift_budget_gb=32.0, record_pairs=False):
        self.L = _lib.load()
        _bind(self.L)
        dev = torch.device(device)
        f32 = dict(dtype=torch.float32, device=dev)
        self.y = torch.as_tensor(y, **f32).contiguous()
        self.h0 = torch.as_tensor(h0, **f32).contiguous()
        B, T = self.y.shape
        n = self.h0.shape[1]
        assert self.h0.shape[0] == B and 2 <= n <= 2048
        C = max(1, min(int(checkpoint_every), T))
        self.B, self.T, self.n, self.S, self.C = B, T, n, int(n_stein), C
        nseg = (T + C - 1) // C
        self.theta = torch.zeros((B, 4), **f32)
        self.pred_ll = torch.zeros((B, T), **f32)
        self.ell2 = torch.zeros((B, T), **f32)
        self.ckpt = torch.zeros((B, nseg, 2, n), **f32)
        tape_f = int(self.L.svpf_diff_tape_floats(n, self.S, C))
        self.tape = torch.empty((B, tape_f), **f32)
        self.gpl = torch.zeros((B, T), **f32)
        self.gtheta = torch.zeros((B, 4), **f32)
        self.gh0 = torch.zeros((B, n), **f32)
        self.hT = torch.zeros((B, n), **f32)
        # the Stein drift of every forward iteration (the backward replays positions from it),
        # kept when it fits the budget; otherwise the backward recomputes the transport
        drift_gb = B * T * self.S * n * 4 / 1e9
        self.drift = torch.empty((B, T, self.S, n), **f32) if drift_gb <= drift_budget_gb else None
        self.mpair = torch.full((B, T, 2), -1, dtype=torch.int32, device=dev) if record_pairs else None
        self.noise = None
        if noise is not None:
            self.noise = {k: torch.as_tensor(noise[k], **f32).contiguous() for k in ("z", "xi")}
            assert self.noise["z"].shape == (B, T, n) and self.noise["xi"].shape == (B, T, self.S, n)
        seeds = np.arange(B, dtype=np.uint64) if seeds is None else np.asarray(seeds, dtype=np.uint64)
        cfg = np.zeros(1, dtype=DCONFIG_DTYPE)
        cfg["n"], cfg["S"], cfg["T"], cfg["C"] = n, self.S, T, C
        cfg["paper_sign"] = 1 if paper_sign else 0
        cfg["gamma"], cfg["lik_offset"], cfg["temperature"] = gamma, lik_offset, temperature
        cfg["rms_rho"], cfg["rms_eps"] = rmsprop_rho, rmsprop_eps
        self.cfg = cfg
        desc = np.zeros(B, dtype=DDESC_DTYPE)
        ptr = lambda t, b, stride: t.data_ptr() + 4 * b * stride
        for b in range(B):
            d = desc[b]
            d["theta"] = ptr(self.theta, b, 4)
            d["y"] = ptr(self.y, b, T)
            d["h0"] = ptr(self.h0, b, n)
            if self.noise is not None:
                d["z"] = ptr(self.noise["z"], b, T * n)
                d["xi"] = ptr(self.noise["xi"], b, T * self.S * n)
            d["pred_ll"] = ptr(self.pred_ll, b, T)
            d["ell2"] = ptr(self.ell2, b, T)
            d["ckpt"] = ptr(self.ckpt, b, nseg * 2 * n)
            d["tape"] = ptr(self.tape, b, tape_f)
            d["gpl"] = ptr(self.gpl, b, T)
            d["gtheta"] = ptr(self.gtheta, b, 4)
            d["gh0"] = ptr(self.gh0, b, n)
            d["hT"] = ptr(self.hT, b, n)
            if self.drift is not None:
                d["drift"] = ptr(self.drift, b, T * self.S * n)
            if self.mpair is not None:
                d["mpair"] = ptr(self.mpair, b, 2 * T)
            d["seed"] = seeds[b]
            d["filter_id"] = b
        self.desc = torch.from_numpy(desc.view(np.uint8).reshape(-1).copy()).to(dev)
        self.forward_id = 0  # checkpoints on the device belong to this forward

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.y.device).cuda_stream)

    def forward(self, theta):
        """theta [B, 4] -> predictive log-likelihood [B, T] (a view of the workspace)."""
        self.theta.copy_(theta.detach().to(self.theta.dtype))
        rc = self.L.svpf_diff_forward(self.cfg.ctypes.data, self.desc.data_ptr(), self.B, self._stream())
        if rc:
            raise RuntimeError(f"svpf_diff_forward failed ({rc})")
        self.forward_id += 1
        return self.pred_ll

    def backward(self, grad_pred_ll):
        """dL/d pred_ll [B, T] -> (dL/d theta [B, 4], dL/d h0 [B, n]) for the theta of the last forward."""
        self.gpl.copy_(grad_pred_ll.detach().to(self.gpl.dtype))
        rc = self.L.svpf_diff_backward(self.cfg.ctypes.data, self.desc.data_ptr(), self.B, self._stream())
        if rc:
            raise RuntimeError(f"svpf_diff_backward failed ({rc})")
        return self.gtheta, self.gh0


class PredictiveLoglik(torch.autograd.Function):
    """pred_ll = PredictiveLoglik.apply(theta, batch): theta [B, 4] -> [B, T]."""

    @staticmethod
    def forward(ctx, theta, batch):
        ctx.batch = batch
        out = batch.forward(theta).clone()
        ctx.forward_id = batch.forward_id
        ctx.save_for_backward(theta.detach().clone())
        return out

    @staticmethod
    def backward(ctx, grad):
        batch = ctx.batch
        if batch.forward_id != ctx.forward_id:  # another forward reused the workspace: rebuild the checkpoints
            (theta,) = ctx.saved_tensors
            batch.forward(theta)
        gth, _ = batch.backward(grad.contiguous())
        return gth.clone(), None


def predictive_loglik(theta, batch):
    return PredictiveLoglik.apply(theta, batch)
