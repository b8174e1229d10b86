"""Diagnostics beside the step, through the C ABI of libsvpf_amd.so.

Health checks (include/svpf_diagnostics.h, the reference's
include/svpf_diagnostics.h and src/svpf_diagnostics.cu): ``SVPFDiagnostics``,
``STATUS`` (names -> codes) and the ``svpf_diag_*`` functions.  Score
diagnostics (include/svpf.h, the reference's include/svpf.cuh:895-1256 and
src/tools/svpf_gradient_diagnostic.cu): the posterior-weighted nu and sigma
scores, the shock state machine and the (placeholder) natural-gradient tuner.

``Health`` and ``Scores`` wrap them for a ``svpf_amd.filter.Filter``.
"""
from __future__ import annotations

import ctypes
from ctypes import c_bool, c_char, c_char_p, c_float, c_int, c_uint32, c_uint64, c_void_p

from . import _lib

STATUS = {
    "SVPF_OK": 0,
    "SVPF_ERR_CUDA_INIT": 100, "SVPF_ERR_CUDA_MALLOC": 101, "SVPF_ERR_CUDA_MEMCPY": 102,
    "SVPF_ERR_CUDA_KERNEL": 103, "SVPF_ERR_CUDA_GRAPH": 104, "SVPF_ERR_CUDA_SYNC": 105,
    "SVPF_ERR_CUDA_DEVICE": 106,
    "SVPF_ERR_NAN_PARTICLES": 200, "SVPF_ERR_NAN_WEIGHTS": 201, "SVPF_ERR_NAN_OUTPUT": 202,
    "SVPF_ERR_INF_PARTICLES": 203, "SVPF_ERR_INF_WEIGHTS": 204, "SVPF_ERR_PARTICLE_COLLAPSE": 210,
    "SVPF_ERR_WEIGHT_UNDERFLOW": 211, "SVPF_ERR_BANDWIDTH_ZERO": 212,
    "SVPF_ERR_INPUT_NAN": 300, "SVPF_ERR_INPUT_INF": 301, "SVPF_ERR_INPUT_EXTREME": 302,
    "SVPF_ERR_PARAMS_INVALID": 310,
    "SVPF_ERR_NOT_INITIALIZED": 400, "SVPF_ERR_ALREADY_DESTROYED": 401, "SVPF_ERR_STATE_CORRUPTED": 402,
    "SVPF_ERR_OUT_OF_MEMORY": 500, "SVPF_ERR_SHARED_MEM_EXCEEDED": 501,
    "SVPF_WARN_ESS_LOW": 1000, "SVPF_WARN_VOL_EXTREME": 1001, "SVPF_WARN_PARTICLES_CLUSTERED": 1002,
    "SVPF_WARN_GRAPH_RECAPTURE": 1003,
}


class SVPFDiagnostics(ctypes.Structure):
    _fields_ = [("status", c_int), ("error_count", c_uint32), ("warning_count", c_uint32), ("last_error", c_int),
                ("last_error_timestep", c_uint64), ("last_error_message", c_char * 256), ("ess", c_float),
                ("ess_ratio", c_float), ("particle_spread", c_float), ("weight_entropy", c_float),
                ("max_weight_ratio", c_float), ("vol_estimate", c_float), ("h_mean", c_float),
                ("bandwidth", c_float), ("loglik", c_float), ("steps_processed", c_uint64),
                ("graph_recaptures", c_uint64), ("avg_step_time_us", c_float), ("max_step_time_us", c_float),
                ("ess_warning_threshold", c_float), ("vol_max_threshold", c_float),
                ("vol_min_threshold", c_float), ("input_max_return", c_float)]


class SVPFGradientDiagnostics(ctypes.Structure):
    _fields_ = [("d_nu_grad", c_void_p), ("d_z_sq_mean", c_void_p), ("d_mu_grad", c_void_p),
                ("d_rho_grad", c_void_p), ("d_sigma_grad", c_void_p), ("d_fisher", c_void_p),
                ("d_fisher_inv", c_void_p), ("nu_gradient_ema", c_float), ("z_sq_ema", c_float),
                ("mu_gradient_ema", c_float), ("rho_gradient_ema", c_float), ("sigma_gradient_ema", c_float),
                ("shock_state", c_int), ("ticks_in_state", c_int), ("shock_threshold", c_float),
                ("shock_duration", c_int), ("recovery_duration", c_int), ("recovery_exit_threshold", c_float),
                ("enable_logging", c_bool), ("log_file", c_void_p), ("initialized", c_bool)]


class SVPFThetaUnconstrained(ctypes.Structure):
    _fields_ = [("mu", c_float), ("eta", c_float), ("kappa", c_float), ("kappa_nu", c_float)]


class SVPFNaturalGradientTuner(ctypes.Structure):
    _fields_ = [("theta", SVPFThetaUnconstrained), ("F", c_float * 16), ("F_ema_decay", c_float),
                ("F_reg", c_float), ("base_lr", c_float), ("lr_shock_mult", c_float), ("grad_clip", c_float),
                ("prior_weight", c_float), ("theta_prior", SVPFThetaUnconstrained), ("warmup_ticks", c_int),
                ("learning_enabled", c_bool)]


def bind(L=None):
    """Declare the argument types of the diagnostics entry points on the library."""
    L = L or _lib.load()
    if getattr(L, "_svpf_diag_bound", False):
        return L
    DP = ctypes.POINTER(SVPFDiagnostics)
    GP = ctypes.POINTER(SVPFGradientDiagnostics)
    TP = ctypes.POINTER(SVPFNaturalGradientTuner)
    SP = ctypes.POINTER(_lib.SVPFState)
    PP = ctypes.POINTER(_lib.SVPFParams)
    FP = ctypes.POINTER(c_float)
    sig = {
        "svpf_diag_init": (None, [DP]),
        "svpf_diag_check": (c_int, [c_void_p, DP]),
        "svpf_diag_check_input": (c_int, [c_float, DP]),
        "svpf_diag_message": (c_char_p, [DP]),
        "svpf_status_name": (c_char_p, [c_int]),
        "svpf_status_is_error": (c_bool, [c_int]),
        "svpf_status_is_warning": (c_bool, [c_int]),
        "svpf_diag_clear_error": (None, [DP]),
        "svpf_diag_print_summary": (None, [DP]),
        "svpf_diag_to_json": (c_int, [DP, c_char_p, c_int]),
        "svpf_gradient_diagnostic_create": (GP, [c_bool, c_char_p]),
        "svpf_gradient_diagnostic_destroy": (None, [GP]),
        "svpf_compute_nu_diagnostic": (None, [SP, GP, c_float, c_int, FP, FP]),
        "svpf_compute_nu_diagnostic_simple": (None, [SP, c_float, FP, FP]),
        "svpf_compute_sigma_diagnostic_simple": (None, [SP, PP, FP, FP]),
        "svpf_snapshot_particles": (None, [SP, c_void_p]),
        "svpf_update_shock_state": (None, [GP, c_float]),
        "svpf_get_shock_state": (c_int, [GP]),
        "svpf_should_learn": (c_bool, [GP]),
        "svpf_get_lr_multiplier": (c_float, [GP, c_float]),
        "svpf_tuner_create": (TP, [PP, c_float, c_float, c_float]),
        "svpf_tuner_destroy": (None, [TP]),
        "svpf_tuner_update": (None, [TP, GP, PP, FP]),
        "svpf_tuner_get_params": (None, [TP, PP, FP]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    L._svpf_diag_bound = True
    return L


class Health:
    """svpf_diag_* over one filter: ``check()`` after a step returns the status name."""

    def __init__(self, flt=None):
        self.L = bind()
        self.flt = flt
        self.d = SVPFDiagnostics()
        self.L.svpf_diag_init(ctypes.byref(self.d))

    def check(self) -> str:
        sp = self.flt.sp if self.flt is not None else None
        st = self.L.svpf_diag_check(ctypes.cast(sp, c_void_p) if sp else None, ctypes.byref(self.d))
        return self.L.svpf_status_name(st).decode()

    def check_input(self, y_t: float) -> str:
        return self.L.svpf_status_name(self.L.svpf_diag_check_input(y_t, ctypes.byref(self.d))).decode()

    def message(self) -> str:
        return self.L.svpf_diag_message(ctypes.byref(self.d)).decode()

    def to_json(self) -> str:
        buf = ctypes.create_string_buffer(1024)
        self.L.svpf_diag_to_json(ctypes.byref(self.d), buf, len(buf))
        return buf.value.decode()


class Scores:
    """Score diagnostics of one filter: ``nu(y_t)``, ``sigma()`` and the
    diagnostic state (EMAs, CSV log, shock state machine) of ``track(y_t, t)``."""

    def __init__(self, flt, log_path: str | None = None):
        self.L = bind()
        self.flt = flt
        self.g = self.L.svpf_gradient_diagnostic_create(log_path is not None,
                                                        None if log_path is None else log_path.encode())
        if not self.g:
            raise RuntimeError("svpf_gradient_diagnostic_create failed")

    def close(self):
        if self.g:
            self.L.svpf_gradient_diagnostic_destroy(self.g)
            self.g = None

    def __del__(self):
        # nothing to do while the interpreter shuts down (the runtime may be gone, and the
        # binding module's globals with it)
        try:
            if _lib.finalizing():
                return
            self.close()
        except Exception:
            pass

    def nu(self, y_t: float):
        g, z = c_float(float("nan")), c_float(float("nan"))
        self.L.svpf_compute_nu_diagnostic_simple(self.flt.sp, y_t, ctypes.byref(g), ctypes.byref(z))
        _lib.check(self.flt.sp, "svpf_compute_nu_diagnostic_simple")
        return g.value, z.value

    def sigma(self):
        g, e = c_float(float("nan")), c_float(float("nan"))
        self.L.svpf_compute_sigma_diagnostic_simple(self.flt.sp, ctypes.byref(self.flt.params), ctypes.byref(g),
                                                    ctypes.byref(e))
        _lib.check(self.flt.sp, "svpf_compute_sigma_diagnostic_simple")
        return g.value, e.value

    def track(self, y_t: float, timestep: int):
        g, z = c_float(float("nan")), c_float(float("nan"))
        self.L.svpf_compute_nu_diagnostic(self.flt.sp, self.g, y_t, timestep, ctypes.byref(g), ctypes.byref(z))
        _lib.check(self.flt.sp, "svpf_compute_nu_diagnostic")
        return g.value, z.value

    @property
    def state(self) -> SVPFGradientDiagnostics:
        return self.g.contents
