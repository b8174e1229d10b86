"""The single-block fused step svpf_fully_fused_step_v2 through the C ABI
(include/svpf_fully_fused_v2.cuh:24-116): one call advances one filter by one
tick, all state in caller-owned device arrays, on one workgroup.

V2_DEFAULTS are the arguments of the reference's own driver
(test/test_fully_fused_v2.cu:91-113) with the Student-t constants of nu = 5.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .config import implied_offset, student_t_const

V2_DEFAULTS = dict(
    n=512, timestep=0, y_t=0.0, y_prev=0.0, h_mean_prev=-4.5, vol_prev=float(np.exp(-2.25)), ksd_prev=0.1,
    rho=0.98, sigma_z=0.10, mu=-4.5, nu=5.0, lik_offset=0.345, student_t_const=student_t_const(5.0),
    student_t_implied_offset=implied_offset(5.0), gamma=0.0, delta_rho=0.0, delta_sigma=0.0,
    mim_jump_prob=0.25, mim_jump_scale=9.0, guide_strength_base=0.05, guide_strength_max=0.30,
    guide_innovation_threshold=1.0, guided_alpha_base=0.0, guided_alpha_shock=0.40,
    guided_innovation_thresh_predict=1.5, use_guide=1, use_guided_predict=1, use_guide_preserving=1,
    use_newton=1, use_full_newton=1, use_rejuvenation=1, rejuv_ksd_threshold=0.05, rejuv_prob=0.30,
    rejuv_blend=0.30, use_adaptive_mu=0, mu_ema_alpha=0.01, use_adaptive_sigma=1, sigma_boost_threshold=0.95,
    sigma_boost_max=3.2, step_size=0.25, temperature=0.45, rmsprop_rho=0.9, rmsprop_eps=1e-6, n_stein_steps=10,
    n_anneal_steps=5, stein_sign_mode=0,
)


def v2_args(**overrides) -> dict:
    a = dict(V2_DEFAULTS)
    for k, v in overrides.items():
        if k not in a:
            raise KeyError(f"unknown argument of the fused step: {k}")
        a[k] = v
    return a


class FusedV2:
    """Device buffers of one filter for svpf_fully_fused_step_v2 and a call
    wrapper.  h, h_prev, grad_log_p, log_weights, grad_v, inv_hessian: [n]
    float32 tensors on the GPU; the scalar slots (h_mean, vol, loglik,
    bandwidth, ksd, guide_mean, guide_var, mu_state) live in one device tensor."""

    SLOTS = ("h_mean", "vol", "loglik", "bandwidth", "ksd", "guide_mean", "guide_var", "mu_state")

    def __init__(self, h0, h_prev0=None, v0=None, guide_mean=0.0, guide_var=0.0, mu_state=0.0, device="cuda"):
        import torch

        self.torch = torch
        self.L = _lib.load()
        h0 = np.ascontiguousarray(h0, dtype=np.float32)
        self.n = int(h0.size)
        if not self.L.svpf_fully_fused_v2_supported(self.n):
            raise ValueError(f"svpf_fully_fused_step_v2 takes 1..1024 particles, got {self.n}")
        t = lambda a: torch.tensor(np.ascontiguousarray(a, dtype=np.float32), device=device)
        self.h = t(h0)
        self.h_prev = t(h0 if h_prev0 is None else h_prev0)
        self.grad = torch.zeros(self.n, device=device)
        self.logw = torch.zeros(self.n, device=device)
        self.v = torch.zeros(self.n, device=device) if v0 is None else t(v0)
        self.inv_hess = torch.zeros(self.n, device=device)
        self.scal = torch.zeros(len(self.SLOTS), device=device)
        self.scal[5] = guide_mean
        self.scal[6] = guide_var
        self.scal[7] = mu_state
        self._keep = []

    def _slot(self, name):
        return ctypes.c_void_p(self.scal.data_ptr() + 4 * self.SLOTS.index(name))

    def step(self, args: dict, noise=None, seed=0, stream_id=0, stream=None):
        """One call with the argument dict of v2_args (its n must equal this
        filter's).  noise: dict of device tensors z_state [n], u_jump [n],
        z_diffusion [rows, n], u_rejuv [n], z_rejuv [n], or None for the counter
        generator keyed by (seed, stream_id).  Returns the launch status."""
        if int(args["n"]) != self.n:
            raise ValueError("argument n differs from the filter's particle count")
        nz = _lib.SVPFV2Noise()
        nz.seed = seed
        nz.stream_id = stream_id
        self._keep = []
        if noise is not None:
            for k in ("z_state", "u_jump", "z_diffusion", "u_rejuv", "z_rejuv"):
                a = noise.get(k)
                if a is not None:
                    self._keep.append(a)
                    setattr(nz, k, ctypes.c_void_p(a.data_ptr()))
            if noise.get("z_diffusion") is not None:
                nz.diffusion_rows = int(noise["z_diffusion"].shape[0])
        ptrs = dict(h=self.h, h_prev=self.h_prev, grad_log_p=self.grad, log_weights=self.logw, d_grad_v=self.v,
                    d_inv_hessian=self.inv_hess)
        vals = []
        for name, kind in _lib.V2_ARGS:
            if name in ptrs:
                vals.append(ctypes.c_void_p(ptrs[name].data_ptr()))
            elif name == "noise":
                vals.append(ctypes.byref(nz))
            elif name == "stream":
                vals.append(None if stream is None else ctypes.c_void_p(stream))
            elif name.startswith("d_"):
                vals.append(self._slot(name[2:] if name != "d_bandwidth" else "bandwidth"))
            else:
                vals.append(args[name])
        return int(self.L.svpf_fully_fused_step_v2(*vals))

    def outputs(self) -> dict:
        s = self.scal.cpu().numpy()
        return {k: float(s[i]) for i, k in enumerate(self.SLOTS)}
