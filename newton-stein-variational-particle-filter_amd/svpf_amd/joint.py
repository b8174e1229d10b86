"""Joint state-parameter filter through the C ABI (include/svpf_joint.h; the
reference's svpf_joint/include/svpf_joint.cuh): ctypes mirrors of
SVPFJointConfig, SVPFJointDiagnostics and SVPFJointState, and ``JointFilter``
over svpf_joint_create / initialize / step / get_vol / get_params / destroy."""
from __future__ import annotations

import ctypes
from ctypes import POINTER, c_float, c_int, c_uint64, c_void_p

import numpy as np

from . import _lib

EXPORTED = (
    "svpf_joint_default_config", "svpf_joint_create", "svpf_joint_initialize", "svpf_joint_step",
    "svpf_joint_get_vol", "svpf_joint_get_params", "svpf_joint_destroy", "svpf_joint_run_sequence_device",
    "svpf_joint_get_particles", "svpf_joint_set_particles", "svpf_joint_lds_bytes",
)
MAX_PARTICLES = 4096
RECW = 20

CONFIG_FIELDS = (
    "step_h", "step_mu", "step_rho", "step_sigma", "diffusion_mu", "diffusion_rho", "diffusion_sigma",
    "collapse_thresh_mu", "collapse_thresh_rho", "collapse_thresh_sigma", "mu_prior_mean", "mu_prior_var",
    "rho_prior_mean", "rho_prior_var", "sigma_prior_mean", "sigma_prior_var", "nu", "student_t_const",
    "student_t_implied_offset", "lik_offset", "prior_weight",
)
DIAG_FLOATS_A = ("h_mean", "vol_mean", "mu_mean", "rho_mean", "sigma_mean", "mu_std", "rho_std", "sigma_std",
                 "std_mu_tilde", "std_rho_tilde", "std_sigma_tilde")
DIAG_INTS = ("mu_collapsed", "rho_collapsed", "sigma_collapsed")
DIAG_FLOATS_B = ("ess", "log_likelihood")
RECORD_FIELDS = DIAG_FLOATS_A + DIAG_INTS + DIAG_FLOATS_B + ("bw_h", "bw_mu", "bw_rho", "bw_sigma")


class JointConfig(ctypes.Structure):
    _fields_ = [("n_particles", c_int), ("n_stein_steps", c_int)] + [(f, c_float) for f in CONFIG_FIELDS]


class JointDiagnostics(ctypes.Structure):
    _fields_ = [(f, c_float) for f in DIAG_FLOATS_A] + [(f, c_int) for f in DIAG_INTS] + \
               [(f, c_float) for f in DIAG_FLOATS_B]


_FP = POINTER(c_float)


class JointState(ctypes.Structure):
    _fields_ = [(f, _FP) for f in ("d_h", "d_h_prev", "d_mu_tilde", "d_rho_tilde", "d_sigma_tilde", "d_grad_h",
                                    "d_grad_mu", "d_grad_rho", "d_grad_sigma", "d_log_w")] + \
               [(f, c_float) for f in ("bw_h", "bw_mu", "bw_rho", "bw_sigma")] + \
               [("d_rng", c_void_p)] + [(f, _FP) for f in ("d_temp", "d_reduce_out", "d_param_mean", "d_param_std",
                                                            "d_std_unconstrained")] + \
               [("d_collapse_flags", POINTER(c_int)), ("config", JointConfig), ("timestep", c_int),
                ("y_prev", c_float), ("stream", c_void_p), ("seed", c_uint64), ("last_error", c_int),
                ("impl", c_void_p)]


def bind(L=None):
    L = L or _lib.load()
    if getattr(L, "_svpf_joint_bound", False):
        return L
    SP, CP, DP = POINTER(JointState), POINTER(JointConfig), POINTER(JointDiagnostics)
    sig = {
        "svpf_joint_default_config": (JointConfig, []),
        "svpf_joint_create": (SP, [CP, c_void_p]),
        "svpf_joint_initialize": (None, [SP, c_uint64]),
        "svpf_joint_step": (None, [SP, c_float, DP]),
        "svpf_joint_get_vol": (c_float, [SP]),
        "svpf_joint_get_params": (None, [SP, _FP, _FP, _FP]),
        "svpf_joint_destroy": (None, [SP]),
        "svpf_joint_run_sequence_device": (c_int, [SP, c_void_p, c_int, c_void_p]),
        "svpf_joint_get_particles": (c_int, [SP, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
        "svpf_joint_set_particles": (c_int, [SP, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]),
        "svpf_joint_lds_bytes": (c_int, [c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype, fn.argtypes = res, args
    L._svpf_joint_bound = True
    return L


def default_config() -> dict:
    c = bind().svpf_joint_default_config()
    d = {"n_particles": c.n_particles, "n_stein_steps": c.n_stein_steps}
    d.update({f: float(getattr(c, f)) for f in CONFIG_FIELDS})
    return d


def _ptr(a):
    return None if a is None else a.ctypes.data_as(c_void_p)


class JointFilter:
    """One joint filter on the GPU.  ``cfg`` overrides fields of the reference defaults."""

    def __init__(self, seed: int | None = None, stream: int = 0, **cfg):
        L = bind()
        self._L = L
        c = L.svpf_joint_default_config()
        for k, v in cfg.items():
            if not hasattr(c, k):
                raise KeyError(k)
            setattr(c, k, v)
        self.n = int(c.n_particles)
        self._s = L.svpf_joint_create(ctypes.byref(c), c_void_p(stream))
        if not self._s:
            raise ValueError("svpf_joint_create refused the configuration (n_particles in [1, 4096], "
                             "n_stein_steps >= 0) or could not allocate")
        if seed is not None:
            self.initialize(seed)

    @property
    def state(self) -> JointState:
        return self._s.contents

    def _check(self, what):
        e = self._s.contents.last_error
        if e:
            raise RuntimeError(f"{what} failed (hipError {e})")

    def initialize(self, seed: int):
        self._L.svpf_joint_initialize(self._s, seed)
        self._check("svpf_joint_initialize")

    def step(self, y_t: float, diag: bool = True) -> dict | None:
        d = JointDiagnostics()
        self._L.svpf_joint_step(self._s, float(y_t), ctypes.byref(d) if diag else None)
        self._check("svpf_joint_step")
        if not diag:
            return None
        return {f: getattr(d, f) for f, _ in JointDiagnostics._fields_}

    def vol(self) -> float:
        return float(self._L.svpf_joint_get_vol(self._s))

    def params(self) -> tuple[float, float, float]:
        mu, rho, sigma = c_float(), c_float(), c_float()
        self._L.svpf_joint_get_params(self._s, ctypes.byref(mu), ctypes.byref(rho), ctypes.byref(sigma))
        self._check("svpf_joint_get_params")
        return mu.value, rho.value, sigma.value

    def particles(self) -> np.ndarray:
        """[5, n] float32: h, h_prev, mu~, rho~, sigma~."""
        out = np.zeros((5, self.n), dtype=np.float32)
        rc = self._L.svpf_joint_get_particles(self._s, *[_ptr(out[k]) for k in range(5)])
        if rc:
            raise RuntimeError(f"svpf_joint_get_particles failed (hipError {rc})")
        return out

    def set_particles(self, h=None, h_prev=None, mu_tilde=None, rho_tilde=None, sigma_tilde=None):
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float32).reshape(self.n)
                for a in (h, h_prev, mu_tilde, rho_tilde, sigma_tilde)]
        rc = self._L.svpf_joint_set_particles(self._s, *[_ptr(a) for a in arrs])
        if rc:
            raise RuntimeError(f"svpf_joint_set_particles failed (hipError {rc})")

    def run_device(self, y_dev, records_dev=None):
        """T steps in one launch; y_dev (and records_dev, [T, 20]) are torch float32 tensors on the GPU."""
        T = int(y_dev.numel())
        rc = self._L.svpf_joint_run_sequence_device(self._s, c_void_p(y_dev.data_ptr()), T,
                                                    None if records_dev is None else c_void_p(records_dev.data_ptr()))
        if rc:
            raise RuntimeError(f"svpf_joint_run_sequence_device failed (hipError {rc})")

    def close(self):
        if getattr(self, "_s", None):
            self._L.svpf_joint_destroy(self._s)
            self._s = None

    def __del__(self):
        # nothing to do while the interpreter shuts down (the runtime may be gone, and the
        # binding module's globals with it)
        try:
            if _lib.finalizing():
                return
            self.close()
        except Exception:
            pass


def records_to_dicts(rec: np.ndarray) -> list[dict]:
    """Rows of SVPF_JOINT_RECW float32 words to dicts (flag words reinterpreted as int32)."""
    rec = np.ascontiguousarray(rec, dtype=np.float32).reshape(-1, RECW)
    ints = rec.view(np.int32)
    out = []
    for r, ri in zip(rec, ints):
        d = {}
        for k, f in enumerate(RECORD_FIELDS):
            d[f] = int(ri[k]) if f in DIAG_INTS else float(r[k])
        out.append(d)
    return out
