"""One filter on the C ABI: configuration from a field dict, per-tick stepping,
device runs and the parity hooks (explicit particles, pre-drawn noise)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib
from .batched import stationary_var
from .config import implied_offset, student_t_const

# config-dict field -> SVPFState field (identity unless listed)
_RENAMES = {"stein_sign": "stein_repulsive_sign"}
_PARAM_FIELDS = ("mu", "rho", "sigma_z", "gamma")
_SKIP = ("n",) + _PARAM_FIELDS


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Filter:
    """A single SVPF filter driven through libsvpf_amd.so."""

    def __init__(self, cfg: dict, stream=None, path: str = "auto", coop_shape=None, median_workgroup=True):
        """path: "auto" (the library's choice by n and features: one workgroup
        below n = 256, the cooperative multi-workgroup kernel for 256 <= n <
        1024 where it applies, else the grid path), "single" (one workgroup),
        "coop" (cooperative kernel when the features allow it, else one
        workgroup) or "grid" (per-tick sequence of grid-wide kernels, n >= 2).
        median_workgroup: with the median bandwidth the cooperative kernel
        selects the median in a workgroup of its own (False: in every row
        workgroup; results are identical)."""
        self.L = _lib.load()
        self.cfg = dict(cfg)
        n = int(cfg["n"])
        self.n = n
        self.sp = self.L.svpf_create(n, int(cfg.get("stein_max_steps", 5)), float(cfg["nu"]),
                                     ctypes.c_void_p(stream) if stream else None)
        if not self.sp:
            raise RuntimeError(f"svpf_create(n={n}) failed")
        self.params = _lib.SVPFParams(cfg["rho"], cfg["sigma_z"], cfg["mu"], cfg["gamma"])
        self._keep = []
        self.apply_config(cfg)
        if path in ("single", "coop"):
            rc = self.L.svpf_set_grid_path(self.sp, 0)
            if rc:
                raise RuntimeError(f"svpf_set_grid_path(n={n}, 0) rejected ({rc}): n is above the "
                                   "single-workgroup capacity")
            self.state.use_persistent_kernel = 1 if path == "coop" else 0
        elif path == "grid":
            rc = self.L.svpf_set_grid_path(self.sp, 1)
            if rc:
                raise RuntimeError(f"svpf_set_grid_path(n={n}) rejected ({rc})")
        if coop_shape is not None:
            rc = self.L.svpf_set_coop_shape(self.sp, int(coop_shape[0]), int(coop_shape[1]))
            if rc:
                raise RuntimeError(f"svpf_set_coop_shape{tuple(coop_shape)} rejected ({rc})")
        if not median_workgroup:
            rc = self.L.svpf_set_coop_median_workgroup(self.sp, 0)
            if rc:
                raise RuntimeError(f"svpf_set_coop_median_workgroup rejected ({rc})")

    def path(self):
        """"single", "coop" or "grid": the kernel path of the next step."""
        return {0: "single", 1: "coop", 2: "grid"}[self.L.svpf_get_path(self.sp)]

    def coop_shape(self):
        cs, g, st = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        self.L.svpf_get_coop_shape(self.sp, cs, g, st)
        return cs.value, g.value, st.value

    @property
    def state(self) -> _lib.SVPFState:
        return self.sp.contents

    def apply_config(self, cfg: dict):
        st = self.state
        for k, v in cfg.items():
            if k in _SKIP:
                continue
            name = _RENAMES.get(k, k)
            setattr(st, name, v)
        st.student_t_const = student_t_const(cfg["nu"])
        st.student_t_implied_offset = implied_offset(cfg["nu"])
        self.params = _lib.SVPFParams(cfg["rho"], cfg["sigma_z"], cfg["mu"], cfg["gamma"])
        self.cfg.update(cfg)

    def close(self):
        if self.sp:
            self.L.svpf_destroy(self.sp)
            self.sp = None

    def __del__(self):
        # nothing to do while the interpreter shuts down (the runtime may be gone, and the
        # binding module's globals with it)
        try:
            if _lib.finalizing():
                return
            self.close()
        except Exception:
            pass

    # ---- state ----
    def initialize(self, seed: int = 42):
        self.L.svpf_initialize(self.sp, ctypes.byref(self.params), seed)
        _lib.check(self.sp, "svpf_initialize")

    def set_particles(self, h, h_prev=None, v=None, h_mean_prev=-3.5, bw_sq_ema=0.0, return_ema=0.0,
                      return_var=0.0):
        """Identical starting state for parity runs (what svpf_initialize leaves
        in the scalar fields, with the given particles)."""
        h = _f32(h)
        hp = None if h_prev is None else _f32(h_prev)
        vv = None if v is None else _f32(v)
        rc = self.L.svpf_set_particles(self.sp, h.ctypes.data, None if hp is None else hp.ctypes.data,
                                       None if vv is None else vv.ctypes.data)
        if rc:
            raise RuntimeError(f"svpf_set_particles failed ({rc})")
        st = self.state
        c = self.cfg
        st.timestep = 0
        st.y_prev = 0.0
        st.guide_initialized = 0
        st.guide_mean = c["mu"]
        st.guide_var = stationary_var(c)
        st.nu_state = max(float(c.get("nu_state", 5.0)), 2.5)
        st.ksd_prev = 1e10
        st.vol_prev = 0.05
        st.mu_state = c["mu"]
        st.mu_var = 1.0
        st.sigma_z_effective = c["sigma_z"]
        self.L.svpf_set_scalar_state(self.sp, h_mean_prev, bw_sq_ema, return_ema, return_var)

    def set_noise(self, z_pred, xi, u_pred=None, u_rej=None, z_rej=None, t_base=0):
        """Pre-drawn noise as device tensors (torch) laid out [T, n] / [T, rows, n]."""
        if z_pred is None:
            rc = self.L.svpf_set_predrawn_noise(self.sp, None, None, None, 0, None, None, 0)
            self._keep = []
        else:
            rows = int(xi.shape[1])
            self._keep = [z_pred, xi, u_pred, u_rej, z_rej]
            ptr = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())
            rc = self.L.svpf_set_predrawn_noise(self.sp, ptr(z_pred), ptr(u_pred), ptr(xi), rows, ptr(u_rej),
                                                ptr(z_rej), t_base)
        if rc:
            raise RuntimeError(f"svpf_set_predrawn_noise failed ({rc})")

    # ---- stepping ----
    def step(self, y_t: float, y_prev: float):
        ll, vol, hm = ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
        self.L.svpf_step_graph(self.sp, y_t, y_prev, ctypes.byref(self.params), ll, vol, hm)
        _lib.check(self.sp, "svpf_step_graph")
        return ll.value, vol.value, hm.value

    def last_outputs(self):
        o = np.zeros(8, dtype=np.float32)
        self.L.svpf_get_last_outputs(self.sp, o.ctypes.data)
        return o

    def run_host(self, y):
        y = _f32(y)
        T = y.size
        ll = np.zeros(T, dtype=np.float32)
        vol = np.zeros(T, dtype=np.float32)
        self.L.svpf_run_sequence(self.sp, y.ctypes.data, T, ctypes.byref(self.params), ll.ctypes.data,
                                 vol.ctypes.data)
        _lib.check(self.sp, "svpf_run_sequence")
        return ll, vol

    def run_device(self, d_y, d_ll=None, d_vol=None):
        """d_y, d_ll, d_vol: device tensors (torch float32)."""
        T = int(d_y.numel())
        self.L.svpf_run_sequence_device(self.sp, ctypes.c_void_p(d_y.data_ptr()), T, ctypes.byref(self.params),
                                        None if d_ll is None else ctypes.c_void_p(d_ll.data_ptr()),
                                        None if d_vol is None else ctypes.c_void_p(d_vol.data_ptr()))
        _lib.check(self.sp, "svpf_run_sequence_device")

    def particles(self):
        h = np.zeros(self.n, dtype=np.float32)
        self.L.svpf_get_particles(self.sp, h.ctypes.data)
        return h

    def log_weights(self):
        """The log-weights of the last step (device array state.log_weights), read back synchronously."""
        lw = np.zeros(self.n, dtype=np.float32)
        hip = ctypes.CDLL("libamdhip64.so")
        hip.hipDeviceSynchronize()
        rc = hip.hipMemcpy(lw.ctypes.data_as(ctypes.c_void_p),
                           ctypes.c_void_p(ctypes.cast(self.state.log_weights, ctypes.c_void_p).value),
                           ctypes.c_size_t(4 * self.n), 2)  # device to host
        if rc != 0:
            raise RuntimeError(f"hipMemcpy of the log-weights failed ({rc})")
        return lw

    def two_factor_particles(self):
        """(h_fast, h_slow) of a two-factor filter, read back synchronously."""
        hf = np.zeros(self.n, dtype=np.float32)
        hs = np.zeros(self.n, dtype=np.float32)
        rc = self.L.svpf_get_two_factor_particles(self.sp, hf.ctypes.data, hs.ctypes.data)
        if rc:
            raise RuntimeError(f"svpf_get_two_factor_particles failed ({rc})")
        return hf, hs

    def rmsprop_state(self):
        v = np.zeros(self.n, dtype=np.float32)
        self.L.svpf_get_rmsprop_state(self.sp, v.ctypes.data)
        return v

    def ess(self):
        return float(self.L.svpf_get_ess(self.sp))
