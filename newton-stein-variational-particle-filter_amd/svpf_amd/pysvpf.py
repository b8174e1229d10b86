"""Python surface of the reference's ``pysvpf`` extension module
(py-torch/svpf_python.cpp:483-654): class ``SVPF`` with the same constructor
arguments, ``set_defaults`` production configuration, setters, properties and
``initialize / step / run / reset``, driving libsvpf_amd.so through the C ABI.

Switching is an import change::

    from svpf_amd import pysvpf          # instead of: import pysvpf
    f = pysvpf.SVPF(n_particles=512, n_stein=8, nu=30.0)
    f.initialize(rho=0.97, sigma_z=0.15, mu=-3.5, gamma=-0.5, seed=42)
    vol, h_mean, loglik = f.run(returns)

Errors raise RuntimeError as the reference's module does
(py-torch/svpf_python.cpp:64-66, :353-355, :386-388).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib
from .config import student_t_const

__all__ = ["SVPF", "SVPFResult"]


@dataclass(frozen=True)
class SVPFResult:
    """py-torch/svpf_python.cpp:40-46."""
    log_lik: float
    vol_mean: float
    vol_std: float
    h_mean: float
    mu_estimate: float

    def __repr__(self):
        return f"<SVPFResult vol={self.vol_mean:.6f} h={self.h_mean:.6f} ll={self.log_lik:.6f}>"


class SVPF:
    """One filter on its own HIP stream (py-torch/svpf_python.cpp:52-477)."""

    def __init__(self, n_particles: int = 400, n_stein: int = 8, nu: float = 30.0):
        self._L = _lib.load()
        self._sp = self._L.svpf_create(int(n_particles), int(n_stein), float(nu), None)
        if not self._sp:
            raise RuntimeError("Failed to create SVPF filter")
        self._params = _lib.SVPFParams(0.97, 0.15, -3.5, -0.5)
        self._y_prev = 0.0
        self._initialized = False
        self.set_defaults()

    def __del__(self):
        sp = getattr(self, "_sp", None)
        try:
            if sp and not _lib.finalizing():
                self._L.svpf_destroy(sp)
                self._sp = None
        except Exception:  # the binding module is torn down: the process is exiting
            pass

    @property
    def _st(self):
        return self._sp.contents

    def _set(self, **fields):
        st = self._st
        for k, v in fields.items():
            setattr(st, k, v)
        self.invalidate_graph()

    # ---- configuration (py-torch/svpf_python.cpp:90-176) ----
    def set_defaults(self):
        """The reference module's production configuration."""
        self._set(
            nu=30.0, lik_offset=0.345,
            use_asymmetric_rho=1, rho_up=0.98, rho_down=0.93,
            use_svld=1, temperature=0.45, rmsprop_rho=0.9, rmsprop_eps=1e-6,
            use_annealing=1, n_anneal_steps=5,
            use_mim=1, mim_jump_prob=0.25, mim_jump_scale=9.0,
            use_adaptive_beta=1,
            use_rejuvenation=1, rejuv_ksd_threshold=0.05, rejuv_prob=0.30, rejuv_blend=0.30,
            use_newton=1, use_full_newton=1, use_exact_gradient=1,
            stein_min_steps=8, stein_max_steps=16, ksd_improvement_threshold=0.05,
            use_guided=1, guided_alpha_base=0.0, guided_alpha_shock=0.40, guided_innovation_threshold=1.5,
            use_guide=1, use_guide_preserving=1, guide_strength=0.05, guide_mean=0.0, guide_var=0.0,
            guide_K=0.0, guide_initialized=0,
            use_adaptive_guide=1, guide_strength_base=0.05, guide_strength_max=0.30,
            guide_innovation_threshold=1.0,
            use_adaptive_mu=1, mu_state=-3.5, mu_var=1.0, mu_process_var=0.001, mu_obs_var_scale=11.0,
            mu_min=-4.0, mu_max=-1.0,
            use_adaptive_sigma=1, sigma_boost_threshold=0.95, sigma_boost_max=3.2, sigma_z_effective=0.10,
            use_local_params=0, delta_rho=0.0, delta_sigma=0.0,
        )
        self._st.student_t_const = student_t_const(30.0)

    def set_adaptive_mu(self, enabled, process_var=0.001, obs_var_scale=11.0, mu_min=-4.0, mu_max=-1.0):
        self._set(use_adaptive_mu=int(bool(enabled)), mu_process_var=process_var, mu_obs_var_scale=obs_var_scale,
                  mu_min=mu_min, mu_max=mu_max)

    def set_adaptive_sigma(self, enabled, threshold=0.95, max_boost=3.2):
        self._set(use_adaptive_sigma=int(bool(enabled)), sigma_boost_threshold=threshold, sigma_boost_max=max_boost)

    def set_adaptive_guide(self, enabled, base=0.05, max=0.30, threshold=1.0):  # noqa: A002 (reference name)
        self._set(use_adaptive_guide=int(bool(enabled)), guide_strength_base=base, guide_strength_max=max,
                  guide_innovation_threshold=threshold)

    def set_asymmetric_rho(self, enabled, rho_up=0.98, rho_down=0.93):
        self._set(use_asymmetric_rho=int(bool(enabled)), rho_up=rho_up, rho_down=rho_down)

    def set_mim(self, enabled, jump_prob=0.25, jump_scale=9.0):
        self._set(use_mim=int(bool(enabled)), mim_jump_prob=jump_prob, mim_jump_scale=jump_scale)

    def set_svld(self, enabled, temperature=0.45):
        self._set(use_svld=int(bool(enabled)), temperature=temperature)

    def set_annealing(self, enabled, n_steps=5):
        self._set(use_annealing=int(bool(enabled)), n_anneal_steps=int(n_steps))

    def set_newton(self, enabled):
        self._set(use_newton=int(bool(enabled)))

    def set_guide(self, enabled, strength=0.05):
        self._set(use_guide=int(bool(enabled)), guide_strength=strength)

    def set_adaptive_beta(self, enabled):
        self._set(use_adaptive_beta=int(bool(enabled)))

    def set_full_newton(self, enabled):
        self._set(use_full_newton=int(bool(enabled)))

    def set_exact_gradient(self, enabled):
        self._set(use_exact_gradient=int(bool(enabled)))

    def set_stein_steps(self, min_steps=8, max_steps=16, ksd_threshold=0.05):
        self._set(stein_min_steps=int(min_steps), stein_max_steps=int(max_steps),
                  ksd_improvement_threshold=ksd_threshold)

    def set_guided_innovation(self, alpha_base=0.0, alpha_shock=0.40, threshold=1.5):
        self._set(guided_alpha_base=alpha_base, guided_alpha_shock=alpha_shock,
                  guided_innovation_threshold=threshold)

    def set_rejuvenation(self, enabled, ksd_threshold=0.05, prob=0.30, blend=0.30):
        self._set(use_rejuvenation=int(bool(enabled)), rejuv_ksd_threshold=ksd_threshold, rejuv_prob=prob,
                  rejuv_blend=blend)

    def set_params(self, rho, sigma_z, mu, gamma):
        self._params = _lib.SVPFParams(rho, sigma_z, mu, gamma)
        self.invalidate_graph()

    # ---- properties (py-torch/svpf_python.cpp:290-318) ----
    @property
    def nu(self):
        return self._st.nu

    @nu.setter
    def nu(self, v):
        self._set(nu=float(v))
        self._st.student_t_const = student_t_const(float(v))

    @property
    def lik_offset(self):
        return self._st.lik_offset

    @lik_offset.setter
    def lik_offset(self, v):
        self._set(lik_offset=float(v))

    @property
    def n_particles(self):
        return self._st.n_particles

    @property
    def n_stein_steps(self):
        return self._st.n_stein_steps

    # ---- core API (py-torch/svpf_python.cpp:320-415) ----
    def _check(self, what):
        rc = self._L.svpf_last_error(self._sp)
        if rc != _lib.SVPF_RC_OK:
            raise RuntimeError(f"{what} failed: {self._L.svpf_last_error_string(self._sp).decode()}")

    def initialize(self, rho=0.97, sigma_z=0.15, mu=-3.5, gamma=-0.5, seed=42):
        self._params = _lib.SVPFParams(rho, sigma_z, mu, gamma)
        self._L.svpf_initialize(self._sp, ctypes.byref(self._params), int(seed))
        self._check("initialize")
        if self._st.use_adaptive_mu:
            self._st.mu_state = mu
            self._st.mu_var = 0.01
        self._y_prev = 0.0
        self._initialized = True

    def step(self, y_t) -> SVPFResult:
        if not self._initialized:
            raise RuntimeError("Filter not initialized. Call initialize() first.")
        ll, vol, hm = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        self._L.svpf_step_graph(self._sp, float(y_t), self._y_prev, ctypes.byref(self._params), ll, vol, hm)
        self._check("step")
        self._y_prev = float(y_t)
        mu_est = self._st.mu_state if self._st.use_adaptive_mu else self._params.mu
        return SVPFResult(ll.value, vol.value, 0.0, hm.value, mu_est)

    def run(self, observations):
        """(vol[T], h_mean[T], loglik[T]) as float32 arrays."""
        if not self._initialized:
            raise RuntimeError("Filter not initialized. Call initialize() first.")
        obs = np.asarray(observations)
        if obs.ndim != 1:
            raise RuntimeError("Observations must be 1-dimensional")
        obs = np.ascontiguousarray(obs, dtype=np.float32)
        T = obs.size
        vol, hm, ll = (np.zeros(T, dtype=np.float32) for _ in range(3))
        for t in range(T):
            r = self.step(obs[t])
            vol[t], hm[t], ll[t] = r.vol_mean, r.h_mean, r.log_lik
        return vol, hm, ll

    def reset(self, seed=42):
        self._L.svpf_initialize(self._sp, ctypes.byref(self._params), int(seed))
        self._check("reset")
        if self._st.use_adaptive_mu:
            self._st.mu_state = self._params.mu
            self._st.mu_var = 0.01
        self._y_prev = 0.0

    def get_particles(self):
        h = np.zeros(self.n_particles, dtype=np.float32)
        self._L.svpf_get_particles(self._sp, h.ctypes.data)
        return h

    def get_ess(self):
        return float(self._L.svpf_get_ess(self._sp))

    def invalidate_graph(self):
        self._L.svpf_graph_invalidate(self._sp)

    def is_graph_captured(self):
        return bool(self._L.svpf_graph_is_captured(self._sp))
