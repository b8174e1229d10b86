"""Two-scale filter through the C ABI (include/svpf_multiscale.h; the
reference's include/svpf_multiscale.cuh): ctypes mirrors of MS_SVPF_Config,
MS_SVPF_Output and MS_SVPF, and ``TwoScale`` over ms_svpf_create / init /
step / destroy."""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_float, c_int, c_uint, c_void_p

from . import _lib

EXPORTED = (
    "ms_svpf_default_config", "ms_svpf_create", "ms_svpf_destroy", "ms_svpf_init", "ms_svpf_step",
    "ms_svpf_vol_for_execution", "ms_svpf_vol_for_position", "ms_svpf_vol_conservative", "ms_svpf_is_noise",
    "ms_svpf_regime_warning", "ms_svpf_print_state", "ms_svpf_print_config", "ms_svpf_set_reactive_dynamics",
    "ms_svpf_set_inertial_dynamics", "ms_svpf_set_weights", "ms_svpf_set_boost_config",
)


class MSConfig(ctypes.Structure):
    _fields_ = [("n_particles_reactive", c_int), ("n_particles_inertial", c_int), ("n_stein_reactive", c_int),
                ("n_stein_inertial", c_int), ("surprise_mild_threshold", c_float),
                ("surprise_high_threshold", c_float), ("boost_mild_factor", c_float),
                ("boost_high_factor", c_float), ("weight_reactive", c_float), ("weight_inertial", c_float)]


class MSOutput(ctypes.Structure):
    _fields_ = [(f, c_float) for f in ("vol_reactive", "vol_inertial", "h_reactive", "h_inertial",
                                        "loglik_reactive", "loglik_inertial", "vol_combined", "h_combined",
                                        "loglik_combined", "transient_magnitude", "transient_ratio",
                                        "ksd_reactive", "ksd_inertial")] + \
               [("stein_steps_reactive", c_int), ("stein_steps_inertial", c_int), ("surprise_reactive", c_float),
                ("cross_scale_boost", c_int), ("mim_prob_inertial_used", c_float), ("is_transient", c_int),
                ("is_regime_shift", c_int), ("is_calm", c_int)]


class MSState(ctypes.Structure):
    _fields_ = [("reactive", POINTER(_lib.SVPFState)), ("inertial", POINTER(_lib.SVPFState)), ("config", MSConfig),
                ("rho_reactive", c_float), ("sigma_z_reactive", c_float), ("rho_inertial", c_float),
                ("sigma_z_inertial", c_float), ("vol_prev", c_float), ("y_prev", c_float),
                ("surprise_ema", c_float), ("mim_jump_prob_inertial_base", c_float),
                ("stream_reactive", c_void_p), ("stream_inertial", c_void_p), ("initialized", c_int),
                ("tick_count", c_int)]


def bind(L=None):
    L = L or _lib.load()
    if getattr(L, "_svpf_ms_bound", False):
        return L
    MP, OP, CP = POINTER(MSState), POINTER(MSOutput), POINTER(MSConfig)
    PP = POINTER(_lib.SVPFParams)
    sig = {
        "ms_svpf_default_config": (MSConfig, []),
        "ms_svpf_create": (MP, [CP]),
        "ms_svpf_destroy": (None, [MP]),
        "ms_svpf_init": (None, [MP, c_float, PP, c_uint]),
        "ms_svpf_step": (None, [MP, c_float, PP, OP]),
        "ms_svpf_vol_for_execution": (c_float, [OP]),
        "ms_svpf_vol_for_position": (c_float, [OP]),
        "ms_svpf_vol_conservative": (c_float, [OP]),
        "ms_svpf_is_noise": (c_int, [OP]),
        "ms_svpf_regime_warning": (c_int, [OP]),
        "ms_svpf_print_state": (None, [MP, OP]),
        "ms_svpf_print_config": (None, [MP]),
        "ms_svpf_set_reactive_dynamics": (None, [MP, c_float, c_float, c_float]),
        "ms_svpf_set_inertial_dynamics": (None, [MP, c_float, c_float, c_float]),
        "ms_svpf_set_weights": (None, [MP, c_float, c_float]),
        "ms_svpf_set_boost_config": (None, [MP, c_float, c_float, c_float, c_float]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    L._svpf_ms_bound = True
    return L


class TwoScale:
    """ms = TwoScale(); ms.init(0.01, params, seed); out = ms.step(y_t, params)."""

    def __init__(self, config: MSConfig | None = None):
        self.L = bind()
        self.ms = self.L.ms_svpf_create(None if config is None else ctypes.byref(config))
        if not self.ms:
            raise RuntimeError("ms_svpf_create failed")

    @property
    def state(self) -> MSState:
        return self.ms.contents

    def init(self, initial_vol, params, seed):
        self.L.ms_svpf_init(self.ms, initial_vol, ctypes.byref(params), seed)

    def step(self, y_t, params) -> MSOutput:
        out = MSOutput()
        self.L.ms_svpf_step(self.ms, y_t, ctypes.byref(params), ctypes.byref(out))
        for sp in (self.state.reactive, self.state.inertial):
            _lib.check(sp, "ms_svpf_step")
        return out

    def close(self):
        if self.ms:
            self.L.ms_svpf_destroy(self.ms)
            self.ms = None

    def __del__(self):
        # nothing to do while the interpreter shuts down (the runtime may be gone, and the
        # binding module's globals with it)
        try:
            if _lib.finalizing():
                return
            self.close()
        except Exception:
            pass
