"""ctypes binding of libsvpf_amd.so (the C ABI declared in include/svpf.h and
include/svpf_fully_fused_v2.h).

The library is the product path: every call here runs the HIP kernels.  There
is no CPU fallback -- loading fails loudly when the library is missing.
"""
from __future__ import annotations

import atexit
import ctypes
import os
import sys

_PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_DIR, "libsvpf_amd.so")

SMOOTH_MAX_LAG = 8

c_int, c_float, c_void_p = ctypes.c_int, ctypes.c_float, ctypes.c_void_p
_FP = ctypes.POINTER(ctypes.c_float)


class SVPFParams(ctypes.Structure):
    _fields_ = [("rho", c_float), ("sigma_z", c_float), ("mu", c_float), ("gamma", c_float)]


class SVPFResult(ctypes.Structure):
    _fields_ = [("log_lik_increment", c_float), ("vol_mean", c_float), ("vol_std", c_float),
                ("h_mean", c_float), ("mu_estimate", c_float)]


# Field order of SVPFState in include/svpf.h.
_STATE_FIELDS = [
    ("h", _FP), ("h_prev", _FP), ("h_pred", _FP), ("grad_log_p", _FP), ("log_weights", _FP), ("d_grad_v", _FP),
    ("n_particles", c_int), ("n_stein_steps", c_int), ("nu", c_float), ("student_t_const", c_float),
    ("student_t_implied_offset", c_float), ("timestep", c_int), ("y_prev", c_float), ("stream", c_void_p),
    ("lik_offset", c_float), ("use_exact_gradient", c_int),
    ("use_svld", c_int), ("use_annealing", c_int), ("use_adaptive_beta", c_int), ("n_anneal_steps", c_int),
    ("temperature", c_float), ("rmsprop_rho", c_float), ("rmsprop_eps", c_float),
    ("use_mim", c_int), ("mim_jump_prob", c_float), ("mim_jump_scale", c_float),
    ("use_local_params", c_int), ("delta_rho", c_float), ("delta_sigma", c_float),
    ("use_asymmetric_rho", c_int), ("rho_up", c_float), ("rho_down", c_float),
    ("use_newton", c_int), ("use_full_newton", c_int),
    ("use_guided", c_int), ("guided_alpha_base", c_float), ("guided_alpha_shock", c_float),
    ("guided_innovation_threshold", c_float),
    ("use_rejuvenation", c_int), ("rejuv_ksd_threshold", c_float), ("rejuv_prob", c_float), ("rejuv_blend", c_float),
    ("use_guide", c_int), ("use_guide_preserving", c_int), ("guide_strength", c_float), ("guide_mean", c_float),
    ("guide_var", c_float), ("guide_K", c_float), ("guide_initialized", c_int),
    ("use_adaptive_guide", c_int), ("guide_strength_base", c_float), ("guide_strength_max", c_float),
    ("guide_innovation_threshold", c_float), ("vol_prev", c_float),
    ("use_adaptive_mu", c_int), ("mu_state", c_float), ("mu_var", c_float), ("mu_process_var", c_float),
    ("mu_obs_var_scale", c_float), ("mu_min", c_float), ("mu_max", c_float),
    ("use_adaptive_sigma", c_int), ("sigma_boost_threshold", c_float), ("sigma_boost_max", c_float),
    ("sigma_z_effective", c_float),
    ("stein_repulsive_sign", c_int), ("use_fan_mode", c_int),
    ("stein_min_steps", c_int), ("stein_max_steps", c_int), ("ksd_improvement_threshold", c_float),
    ("ksd_prev", c_float), ("stein_steps_used", c_int),
    ("use_student_t_state", c_int), ("nu_state", c_float),
    ("use_smoothing", c_int), ("smooth_lag", c_int), ("smooth_output_lag", c_int), ("smooth_head", c_int),
    ("smooth_h_mean", c_float * SMOOTH_MAX_LAG), ("smooth_h_var", c_float * SMOOTH_MAX_LAG),
    ("smooth_y", c_float * SMOOTH_MAX_LAG),
    ("use_persistent_kernel", c_int), ("persistent_kernel_supported", c_int), ("use_heun", c_int),
    ("use_two_factor", c_int), ("rho_fast", c_float), ("sigma_fast", c_float), ("rho_slow", c_float),
    ("sigma_slow", c_float), ("bw_floor_fast", c_float), ("bw_floor_slow", c_float),
    ("guide_mean_fast", c_float), ("guide_mean_slow", c_float), ("guide_var_fast", c_float),
    ("guide_var_slow", c_float), ("guide_2f_initialized", c_int),
    ("bandwidth_mode", c_int), ("use_linear_beta", c_int), ("use_ksd_early_stop", c_int),
    ("impl", c_void_p),
]


class SVPFState(ctypes.Structure):
    _fields_ = _STATE_FIELDS


# Status codes (include/svpf.h)
SVPF_RC_OK = 0
SVPF_RC_INVALID_ARGUMENT = 1
SVPF_RC_UNSUPPORTED = 2
SVPF_RC_HIP = 3
SVPF_RC_NOT_INITIALIZED = 4
SVPF_RC_OUT_OF_MEMORY = 5
SVPF_RC_TIMEOUT = 6
SVPF_RC_TOO_LARGE = 7
NOALLOC_MAX_TICKS = 32768



class SVPFV2Noise(ctypes.Structure):
    """svpf_v2_noise of include/svpf_fully_fused_v2.h."""
    _fields_ = [("seed", ctypes.c_ulonglong), ("stream_id", ctypes.c_uint), ("diffusion_rows", c_int),
                ("z_state", c_void_p), ("u_jump", c_void_p), ("z_diffusion", c_void_p), ("u_rejuv", c_void_p),
                ("z_rejuv", c_void_p)]


# Every symbol include/svpf.h, include/svpf_fully_fused_v2.h and
# include/svpf_diagnostics.h declare (non-inline).
EXPORTED = (
    "svpf_create", "svpf_destroy", "svpf_initialize", "svpf_step", "svpf_step_seeded", "svpf_run_sequence",
    "svpf_run_sequence_device", "svpf_step_adaptive", "svpf_step_graph", "svpf_graph_is_captured",
    "svpf_graph_invalidate", "svpf_get_particles", "svpf_get_stats", "svpf_get_ess", "svpf_get_ksd_stats",
    "svpf_last_error", "svpf_set_particles", "svpf_get_rmsprop_state", "svpf_set_predrawn_noise",
    "svpf_set_scalar_state", "svpf_get_last_outputs", "svpf_batched_run", "svpf_tile_capacity_query",
    "svpf_lds_bytes_query", "svpf_version", "svpf_set_coop_shape", "svpf_get_coop_shape",
    "svpf_last_error_string", "svpf_diff_tape_floats", "svpf_diff_forward", "svpf_diff_backward",
    "svpf_tile_occupancy_query", "svpf_diff2d_tape_floats", "svpf_diff2d_forward", "svpf_diff2d_backward",
    "svpf_set_grid_path", "svpf_get_path", "svpf_fully_fused_step_v2", "svpf_fully_fused_v2_supported",
    "svpf_coop_launch_is_cooperative", "svpf_set_coop_spin_limit", "svpf_digamma_f32",
    "svpf_get_two_factor_particles", "svpf_run_sequence_device_noalloc", "svpf_set_coop_fault",
    "svpf_set_coop_median_workgroup",
    # score diagnostics (include/svpf.h)
    "svpf_gradient_diagnostic_create", "svpf_gradient_diagnostic_destroy", "svpf_compute_nu_diagnostic",
    "svpf_compute_nu_diagnostic_simple", "svpf_compute_sigma_diagnostic_simple", "svpf_snapshot_particles",
    "svpf_update_shock_state", "svpf_get_shock_state", "svpf_should_learn", "svpf_get_lr_multiplier",
    "svpf_tuner_create", "svpf_tuner_destroy", "svpf_tuner_update", "svpf_tuner_get_params",
    # health checks (include/svpf_diagnostics.h)
    "svpf_diag_init", "svpf_diag_check", "svpf_diag_check_input", "svpf_diag_message", "svpf_status_name",
    "svpf_status_is_error", "svpf_status_is_warning", "svpf_diag_clear_error", "svpf_diag_print_summary",
    "svpf_diag_to_json",
)

# svpf_fully_fused_step_v2 argument kinds, in order (include/svpf_fully_fused_v2.h)
V2_ARGS = (
    [("h", "p"), ("h_prev", "p"), ("grad_log_p", "p"), ("log_weights", "p"), ("d_grad_v", "p"),
     ("d_inv_hessian", "p"), ("noise", "noise"), ("y_t", "f"), ("y_prev", "f"), ("h_mean_prev", "f"),
     ("vol_prev", "f"), ("ksd_prev", "f"), ("d_h_mean", "p"), ("d_vol", "p"), ("d_loglik", "p"),
     ("d_bandwidth", "p"), ("d_ksd", "p"), ("d_guide_mean", "p"), ("d_guide_var", "p"), ("rho", "f"),
     ("sigma_z", "f"), ("mu", "f"), ("nu", "f"), ("lik_offset", "f"), ("student_t_const", "f"),
     ("student_t_implied_offset", "f"), ("gamma", "f"), ("delta_rho", "f"), ("delta_sigma", "f"),
     ("mim_jump_prob", "f"), ("mim_jump_scale", "f"), ("guide_strength_base", "f"), ("guide_strength_max", "f"),
     ("guide_innovation_threshold", "f"), ("guided_alpha_base", "f"), ("guided_alpha_shock", "f"),
     ("guided_innovation_thresh_predict", "f"), ("use_guide", "i"), ("use_guided_predict", "i"),
     ("use_guide_preserving", "i"), ("use_newton", "i"), ("use_full_newton", "i"), ("use_rejuvenation", "i"),
     ("rejuv_ksd_threshold", "f"), ("rejuv_prob", "f"), ("rejuv_blend", "f"), ("use_adaptive_mu", "i"),
     ("d_mu_state", "p"), ("mu_ema_alpha", "f"), ("use_adaptive_sigma", "i"), ("sigma_boost_threshold", "f"),
     ("sigma_boost_max", "f"), ("step_size", "f"), ("temperature", "f"), ("rmsprop_rho", "f"),
     ("rmsprop_eps", "f"), ("n_stein_steps", "i"), ("n_anneal_steps", "i"), ("stein_sign_mode", "i"), ("n", "i"),
     ("timestep", "i"), ("stream", "p")]
)

_lib = None
_exiting = False


def _mark_exiting():
    global _exiting
    _exiting = True


atexit.register(_mark_exiting)


def finalizing() -> bool:
    """True once the interpreter is shutting down.  Destructors then skip the
    C-ABI destroy calls: the HIP runtime (and a profiler's tool library) may
    already be finalising, and the process's device memory goes with it.
    Explicit close() calls before that point free everything as usual."""
    return _exiting or sys.is_finalizing()


def load(path: str | None = None):
    """Load the library and declare signatures.  Raises if it is missing."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"libsvpf_amd.so not built ({p}); run __graft_entry__.build()")
    # PyTorch ships its own HIP runtime under the same soname (libamdhip64.so.7) as the
    # ROCm one this library links.  Whichever loads first serves the process; loading this
    # library first would leave torch with a second runtime that finds no GPU.  Load torch's
    # first, so torch tensors and this library share one runtime whatever the import order.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(p)
    SP = ctypes.POINTER(SVPFState)
    PP = ctypes.POINTER(SVPFParams)
    L.svpf_create.restype = SP
    L.svpf_create.argtypes = [c_int, c_int, c_float, c_void_p]
    L.svpf_destroy.restype = None
    L.svpf_destroy.argtypes = [SP]
    L.svpf_initialize.restype = None
    L.svpf_initialize.argtypes = [SP, PP, ctypes.c_ulonglong]
    L.svpf_step.restype = None
    L.svpf_step.argtypes = [SP, c_float, PP, ctypes.POINTER(SVPFResult)]
    L.svpf_step_seeded.restype = None
    L.svpf_step_seeded.argtypes = [SP, c_float, PP, ctypes.c_ulonglong, ctypes.POINTER(SVPFResult)]
    L.svpf_run_sequence.restype = None
    L.svpf_run_sequence.argtypes = [SP, c_void_p, c_int, PP, c_void_p, c_void_p]
    L.svpf_run_sequence_device.restype = None
    L.svpf_run_sequence_device.argtypes = [SP, c_void_p, c_int, PP, c_void_p, c_void_p]
    L.svpf_run_sequence_device_noalloc.restype = None
    L.svpf_run_sequence_device_noalloc.argtypes = [SP, c_void_p, c_int, PP, c_void_p, c_void_p]
    for name in ("svpf_step_adaptive", "svpf_step_graph"):
        f = getattr(L, name)
        f.restype = None
        f.argtypes = [SP, c_float, c_float, PP, _FP, _FP, _FP]
    L.svpf_graph_is_captured.restype = ctypes.c_bool
    L.svpf_graph_is_captured.argtypes = [SP]
    L.svpf_graph_invalidate.restype = None
    L.svpf_graph_invalidate.argtypes = [SP]
    L.svpf_get_particles.restype = None
    L.svpf_get_particles.argtypes = [SP, c_void_p]
    L.svpf_get_stats.restype = None
    L.svpf_get_stats.argtypes = [SP, _FP, _FP]
    L.svpf_get_ess.restype = c_float
    L.svpf_get_ess.argtypes = [SP]
    L.svpf_get_ksd_stats.restype = None
    L.svpf_get_ksd_stats.argtypes = [SP, _FP, ctypes.POINTER(c_int)]
    L.svpf_last_error.restype = c_int
    L.svpf_last_error.argtypes = [SP]
    L.svpf_last_error_string.restype = ctypes.c_char_p
    L.svpf_last_error_string.argtypes = [SP]
    L.svpf_set_particles.restype = c_int
    L.svpf_set_particles.argtypes = [SP, c_void_p, c_void_p, c_void_p]
    L.svpf_get_rmsprop_state.restype = c_int
    L.svpf_get_rmsprop_state.argtypes = [SP, c_void_p]
    L.svpf_set_predrawn_noise.restype = c_int
    L.svpf_set_predrawn_noise.argtypes = [SP, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_void_p, c_int]
    L.svpf_set_scalar_state.restype = c_int
    L.svpf_set_scalar_state.argtypes = [SP, c_float, c_float, c_float, c_float]
    L.svpf_get_last_outputs.restype = c_int
    L.svpf_get_last_outputs.argtypes = [SP, c_void_p]
    L.svpf_batched_run.restype = c_int
    L.svpf_batched_run.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                   c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_int, c_void_p, c_int, c_int,
                                   c_void_p]
    L.svpf_tile_capacity_query.restype = c_int
    L.svpf_tile_capacity_query.argtypes = [c_int, c_int]
    L.svpf_tile_occupancy_query.restype = c_int
    L.svpf_tile_occupancy_query.argtypes = [c_int, c_int, c_int]
    L.svpf_lds_bytes_query.restype = ctypes.c_size_t
    L.svpf_lds_bytes_query.argtypes = [c_int]
    L.svpf_set_coop_shape.restype = c_int
    L.svpf_set_coop_shape.argtypes = [SP, c_int, c_int]
    L.svpf_get_coop_shape.restype = c_int
    L.svpf_set_grid_path.restype = c_int
    L.svpf_set_grid_path.argtypes = [SP, c_int]
    L.svpf_get_path.restype = c_int
    L.svpf_get_path.argtypes = [SP]
    L.svpf_coop_launch_is_cooperative.restype = c_int
    L.svpf_coop_launch_is_cooperative.argtypes = [SP]
    L.svpf_set_coop_fault.restype = c_int
    L.svpf_set_coop_fault.argtypes = [SP, c_int]
    L.svpf_set_coop_spin_limit.restype = c_int
    L.svpf_set_coop_spin_limit.argtypes = [SP, c_int]
    L.svpf_set_coop_median_workgroup.restype = c_int
    L.svpf_set_coop_median_workgroup.argtypes = [SP, c_int]
    L.svpf_get_coop_shape.argtypes = [SP, ctypes.POINTER(c_int), ctypes.POINTER(c_int), ctypes.POINTER(c_int)]
    kinds = {"p": c_void_p, "f": c_float, "i": c_int, "noise": ctypes.POINTER(SVPFV2Noise)}
    L.svpf_fully_fused_step_v2.restype = c_int
    L.svpf_fully_fused_step_v2.argtypes = [kinds[k] for _, k in V2_ARGS]
    L.svpf_fully_fused_v2_supported.restype = ctypes.c_bool
    L.svpf_fully_fused_v2_supported.argtypes = [c_int]
    L.svpf_version.restype = ctypes.c_char_p
    L.svpf_version.argtypes = []
    L.svpf_get_two_factor_particles.restype = c_int
    L.svpf_get_two_factor_particles.argtypes = [SP, c_void_p, c_void_p]
    L.svpf_digamma_f32.restype = c_float
    L.svpf_digamma_f32.argtypes = [c_float]
    if path is None:
        _lib = L
    return L


def check(state_ptr, what: str):
    L = load()
    rc = L.svpf_last_error(state_ptr)
    if rc != SVPF_RC_OK:
        raise RuntimeError(f"{what} failed: {L.svpf_last_error_string(state_ptr).decode()} (status {rc})")
