"""ctypes wrapper of the FP64 oracle (oracle/svpf_oracle.c).

TEST INFRASTRUCTURE ONLY: the checker the HIP product path is compared against.
Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libsvpf_oracle.so")
_lib = None

CFG_INT_FIELDS = (
    "n", "bandwidth_mode", "use_exact_gradient", "use_newton", "use_full_newton",
    "stein_sign", "use_svld", "use_annealing", "n_anneal_steps", "use_adaptive_beta",
    "use_linear_beta", "stein_min_steps", "stein_max_steps", "use_mim",
    "use_asymmetric_rho", "use_local_params", "use_guided", "use_guide",
    "use_guide_preserving", "use_adaptive_guide", "use_rejuvenation", "use_fan_mode",
    "use_adaptive_mu", "use_adaptive_sigma", "use_ksd_early_stop", "use_heun", "use_student_t_state", "pad_",
)
CFG_DOUBLE_FIELDS = (
    "mu", "rho", "sigma_z", "gamma", "nu", "lik_offset", "temperature", "rmsprop_rho",
    "rmsprop_eps", "mim_jump_prob", "mim_jump_scale", "rho_up", "rho_down", "delta_rho",
    "delta_sigma", "guided_alpha_base", "guided_alpha_shock", "guided_innovation_threshold",
    "guide_strength_base", "guide_strength_max", "guide_innovation_threshold",
    "rejuv_ksd_threshold", "rejuv_prob", "rejuv_blend", "mu_process_var",
    "mu_obs_var_scale", "mu_min", "mu_max", "sigma_boost_threshold", "sigma_boost_max",
    "ksd_improvement_threshold", "nu_state",
)


class Cfg(ctypes.Structure):
    _fields_ = [(f, ctypes.c_int) for f in CFG_INT_FIELDS] + [(f, ctypes.c_double) for f in CFG_DOUBLE_FIELDS]


class State(ctypes.Structure):
    _fields_ = [
        ("timestep", ctypes.c_int), ("guide_initialized", ctypes.c_int),
        ("stein_steps_used", ctypes.c_int), ("pad_", ctypes.c_int),
        ("ksd_prev", ctypes.c_double), ("vol_prev", ctypes.c_double),
        ("h_mean_prev", ctypes.c_double), ("guide_mean", ctypes.c_double),
        ("guide_var", ctypes.c_double), ("bw_sq_ema", ctypes.c_double),
        ("return_ema", ctypes.c_double), ("return_var", ctypes.c_double),
        ("mu_state", ctypes.c_double), ("mu_var", ctypes.c_double),
        ("sigma_z_effective", ctypes.c_double),
    ]


class Noise(ctypes.Structure):
    _fields_ = [
        ("z_pred", ctypes.c_void_p), ("u_pred", ctypes.c_void_p), ("xi", ctypes.c_void_p),
        ("xi_rows", ctypes.c_int), ("pad_", ctypes.c_int),
        ("u_rej", ctypes.c_void_p), ("z_rej", ctypes.c_void_p),
        ("seed", ctypes.c_uint64), ("filter_id", ctypes.c_uint32), ("pad2_", ctypes.c_uint32),
    ]


V2_INT_FIELDS = ("n", "timestep", "use_guide", "use_guided_predict", "use_guide_preserving", "use_newton",
                 "use_full_newton", "use_rejuvenation", "use_adaptive_mu", "use_adaptive_sigma", "n_stein_steps",
                 "n_anneal_steps", "stein_sign_mode")
V2_DOUBLE_FIELDS = ("y_t", "y_prev", "h_mean_prev", "vol_prev", "ksd_prev", "rho", "sigma_z", "mu", "nu",
                    "lik_offset", "student_t_const", "student_t_implied_offset", "gamma", "delta_rho", "delta_sigma",
                    "mim_jump_prob", "mim_jump_scale", "guide_strength_base", "guide_strength_max",
                    "guide_innovation_threshold", "guided_alpha_base", "guided_alpha_shock",
                    "guided_innovation_thresh_predict", "rejuv_ksd_threshold", "rejuv_prob", "rejuv_blend",
                    "mu_ema_alpha", "sigma_boost_threshold", "sigma_boost_max", "step_size", "temperature",
                    "rmsprop_rho", "rmsprop_eps")


class V2Args(ctypes.Structure):
    _fields_ = ([(f, ctypes.c_int) for f in V2_INT_FIELDS] + [("pad_", ctypes.c_int)]
                + [(f, ctypes.c_double) for f in V2_DOUBLE_FIELDS])


class V2IO(ctypes.Structure):
    _fields_ = [(f, ctypes.c_double) for f in ("guide_mean", "guide_var", "mu_state", "h_mean", "vol", "loglik",
                                                "bandwidth", "ksd")]


class Out(ctypes.Structure):
    _fields_ = [
        ("loglik", ctypes.c_double), ("pred_loglik", ctypes.c_double), ("vol", ctypes.c_double),
        ("h_mean", ctypes.c_double), ("ksd", ctypes.c_double), ("bandwidth", ctypes.c_double),
        ("bw_sq", ctypes.c_double), ("steps", ctypes.c_int), ("pad_", ctypes.c_int),
    ]


def build(force=False):
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "-B" if force else "all"], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        D, I, P = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
        L.svpfo_student_t_const.restype = D
        L.svpfo_student_t_const.argtypes = [D]
        L.svpfo_implied_offset.restype = D
        L.svpfo_implied_offset.argtypes = [D]
        L.svpfo_log_w.restype = D
        L.svpfo_log_w.argtypes = [D, D, D, D]
        L.svpfo_median_r2.restype = D
        L.svpfo_median_r2.argtypes = [P, I]
        L.svpfo_median_r2_f32.restype = ctypes.c_float
        L.svpfo_median_r2_f32.argtypes = [P, I]
        L.svpfo_adaptive_beta.restype = D
        L.svpfo_adaptive_beta.argtypes = [D, I, I]
        L.svpfo_stein_budget.restype = I
        L.svpfo_stein_budget.argtypes = [P, P]
        L.svpfo_ekf_update.restype = None
        L.svpfo_ekf_update.argtypes = [P, P, D, D, D, D]
        L.svpfo_gradient.restype = None
        L.svpfo_gradient.argtypes = [P, P, P, D, D, D, D, D, P, P, P, P]
        L.svpfo_stein_phi.restype = None
        L.svpfo_stein_phi.argtypes = [P, P, P, P, D, P, P]
        L.svpfo_tick.restype = I
        L.svpfo_tick.argtypes = [P, P, P, P, P, D, D, P, P]
        L.svpfo_run.restype = I
        L.svpfo_run.argtypes = [P, P, P, P, P, I, P, ctypes.c_float, P, P, P, I, P, P,
                                P, P, P, P, P, P, P]
        L.svpfo_philox4x32_10.restype = None
        L.svpfo_philox4x32_10.argtypes = [P, P, P]
        L.svpfo_u01_from_bits.restype = D
        L.svpfo_u01_from_bits.argtypes = [ctypes.c_uint32]
        L.svpfo_normal_from_bits.restype = D
        L.svpfo_normal_from_bits.argtypes = [ctypes.c_uint32, ctypes.c_uint32]
        L.svpfo_v2_step.restype = I
        L.svpfo_v2_step.argtypes = [P, P, P, P, P, P, P, P, P]
        L.svpfo_student_t_draw.restype = D
        L.svpfo_student_t_draw.argtypes = [ctypes.c_uint64, ctypes.c_uint32, ctypes.c_uint32, ctypes.c_uint32, D]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def make_cfg(cfg: dict) -> Cfg:
    c = Cfg()
    for f in CFG_INT_FIELDS:
        if f != "pad_":
            setattr(c, f, int(cfg[f]))
    for f in CFG_DOUBLE_FIELDS:
        setattr(c, f, float(cfg[f]))
    return c


def make_state(cfg: dict, **kw) -> State:
    """Scalar state as svpf_initialize leaves it (src/svpf_optimized_graph.cu:349-451):
    ksd_prev = 1e10, EMA bandwidth unset, h_mean_prev = -3.5 (the value the
    optimized backend is created with, src/svpf_optimized_graph.cu:489-491)."""
    s = State()
    s.timestep = 0
    s.guide_initialized = 0
    s.ksd_prev = 1e10
    s.vol_prev = 0.05
    s.h_mean_prev = -3.5
    s.guide_mean = cfg["mu"]
    s.guide_var = cfg["sigma_z"] ** 2 / (1.0 - cfg["rho"] ** 2 + 1e-6)
    s.bw_sq_ema = 0.0
    s.return_ema = 0.0
    s.return_var = 0.0
    s.mu_state = cfg["mu"]
    s.mu_var = 1.0
    s.sigma_z_effective = cfg["sigma_z"]
    for k, v in kw.items():
        setattr(s, k, v)
    return s


def state_dict(s: State) -> dict:
    return {f: getattr(s, f) for f, _ in State._fields_ if f != "pad_"}


def philox(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    o = np.zeros(4, dtype=np.uint32)
    lib().svpfo_philox4x32_10(_ptr(c), _ptr(k), _ptr(o))
    return o


def median_r2(h):
    h = np.ascontiguousarray(h, dtype=np.float64)
    return lib().svpfo_median_r2(_ptr(h), h.size)


def median_r2_f32(h):
    h = np.ascontiguousarray(h, dtype=np.float32)
    return np.float32(lib().svpfo_median_r2_f32(_ptr(h), h.size))


def gradient(cfg, h, h_prev, y_t, y_prev, mu, sigma_z, beta):
    n = cfg["n"]
    c = make_cfg(cfg)
    h = np.ascontiguousarray(h, dtype=np.float64)
    hp = np.ascontiguousarray(h_prev, dtype=np.float64)
    g, lw, pc, cv = (np.zeros(n) for _ in range(4))
    lib().svpfo_gradient(ctypes.byref(c), _ptr(h), _ptr(hp), y_t, y_prev, mu, sigma_z, beta,
                         _ptr(g), _ptr(lw), _ptr(pc), _ptr(cv))
    return g, lw, pc, cv


def stein_phi(cfg, h, s, curv, inv_bw_sq):
    n = cfg["n"]
    c = make_cfg(cfg)
    h = np.ascontiguousarray(h, dtype=np.float64)
    s = np.ascontiguousarray(s, dtype=np.float64)
    curv = np.ascontiguousarray(curv, dtype=np.float64)
    phi = np.zeros(n)
    kp = np.zeros(n)
    lib().svpfo_stein_phi(ctypes.byref(c), _ptr(h), _ptr(s), _ptr(curv), inv_bw_sq, _ptr(phi), _ptr(kp))
    return phi, kp


class OracleFilter:
    """One filter run by the oracle: arrays in float64, scalar state in State."""

    def __init__(self, cfg: dict, h0, h_prev0=None, v0=None, **state_kw):
        self.cfg = dict(cfg)
        self.c = make_cfg(cfg)
        self.st = make_state(cfg, **state_kw)
        n = cfg["n"]
        self.h = np.array(h0, dtype=np.float64).reshape(n).copy()
        self.h_prev = self.h.copy() if h_prev0 is None else np.array(h_prev0, dtype=np.float64).copy()
        self.v = np.zeros(n) if v0 is None else np.array(v0, dtype=np.float64).copy()

    def tick(self, y_t, y_prev, z_pred=None, xi=None, u_pred=None, u_rej=None, z_rej=None,
             seed=0, filter_id=0):
        keep = []

        def f32(a):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float32)
            keep.append(a)
            return a

        nz = Noise()
        zp, xa = f32(z_pred), f32(xi)
        nz.z_pred = _ptr(zp)
        nz.xi = _ptr(xa)
        nz.xi_rows = 0 if xa is None else (xa.size // self.cfg["n"])
        nz.u_pred = _ptr(f32(u_pred))
        nz.u_rej = _ptr(f32(u_rej))
        nz.z_rej = _ptr(f32(z_rej))
        nz.seed = seed
        nz.filter_id = filter_id
        o = Out()
        rc = lib().svpfo_tick(ctypes.byref(self.c), ctypes.byref(self.st), _ptr(self.h), _ptr(self.h_prev),
                              _ptr(self.v), float(y_t), float(y_prev), ctypes.byref(nz), ctypes.byref(o))
        if rc:
            raise RuntimeError("oracle tick failed")
        return {f: getattr(o, f) for f, _ in Out._fields_ if f != "pad_"}

    def run(self, y, noise, y_prev0=0.0):
        T = len(y)
        n = self.cfg["n"]
        y = np.ascontiguousarray(y, dtype=np.float32)
        zp = np.ascontiguousarray(noise["z_pred"][:T], dtype=np.float32)
        xi = np.ascontiguousarray(noise["xi"][:T], dtype=np.float32)
        rows = xi.shape[1]
        up = noise.get("u_pred")
        ur = noise.get("u_rej")
        zr = noise.get("z_rej")
        up = None if up is None else np.ascontiguousarray(up[:T], dtype=np.float32)
        ur = None if ur is None else np.ascontiguousarray(ur[:T], dtype=np.float32)
        zr = None if zr is None else np.ascontiguousarray(zr[:T], dtype=np.float32)
        outs = {k: np.zeros(T) for k in ("loglik", "pred_loglik", "vol", "h_mean", "ksd", "bw")}
        steps = np.zeros(T, dtype=np.int32)
        rc = lib().svpfo_run(ctypes.byref(self.c), ctypes.byref(self.st), _ptr(self.h), _ptr(self.h_prev),
                             _ptr(self.v), T, _ptr(y), float(y_prev0), _ptr(zp), _ptr(up), _ptr(xi), rows,
                             _ptr(ur), _ptr(zr), _ptr(outs["loglik"]), _ptr(outs["pred_loglik"]),
                             _ptr(outs["vol"]), _ptr(outs["h_mean"]), _ptr(outs["ksd"]), _ptr(outs["bw"]),
                             _ptr(steps))
        if rc:
            raise RuntimeError("oracle run failed")
        outs["steps"] = steps
        return outs


def v2_step(args: dict, h, h_prev, v, io: dict, noise=None, seed=0, filter_id=0):
    """One call of the single-block fused step (svpfo_v2_step).  args: the
    scalar arguments (V2_INT_FIELDS, V2_DOUBLE_FIELDS); h, h_prev, v: float64
    arrays updated in place; io: guide_mean, guide_var, mu_state (read and
    updated).  noise: dict of float32 arrays z_pred [n], u_pred [n], xi
    [rows, n], u_rej [n], z_rej [n], or None for the counter generator keyed
    by (seed, filter_id, timestep).  Returns (grad, log_w, inv_hess, io)."""
    a = V2Args()
    for f in V2_INT_FIELDS:
        setattr(a, f, int(args[f]))
    for f in V2_DOUBLE_FIELDS:
        setattr(a, f, float(args[f]))
    n = a.n
    keep = []

    def f32(x):
        if x is None:
            return None
        x = np.ascontiguousarray(x, dtype=np.float32)
        keep.append(x)
        return x

    nz = Noise()
    noise = noise or {}
    xa = f32(noise.get("xi"))
    nz.z_pred = _ptr(f32(noise.get("z_pred")))
    nz.u_pred = _ptr(f32(noise.get("u_pred")))
    nz.xi = _ptr(xa)
    nz.xi_rows = 0 if xa is None else xa.size // n
    nz.u_rej = _ptr(f32(noise.get("u_rej")))
    nz.z_rej = _ptr(f32(noise.get("z_rej")))
    nz.seed = seed
    nz.filter_id = filter_id
    st = V2IO()
    for k in ("guide_mean", "guide_var", "mu_state"):
        setattr(st, k, float(io.get(k, 0.0)))
    grad, log_w, inv_hess = np.zeros(n), np.zeros(n), np.zeros(n)
    for arr in (h, h_prev, v):
        assert arr.dtype == np.float64 and arr.flags["C_CONTIGUOUS"] and arr.size == n
    rc = lib().svpfo_v2_step(ctypes.byref(a), _ptr(h), _ptr(h_prev), _ptr(grad), _ptr(log_w), _ptr(v),
                             _ptr(inv_hess), ctypes.byref(nz), ctypes.byref(st))
    if rc:
        raise RuntimeError("oracle v2 step failed")
    return grad, log_w, inv_hess, {f: getattr(st, f) for f, _ in V2IO._fields_}


# ---------------------------------------------------------------- initialisation, smoothing, two-factor
def init_particles(n, mu, rho, sigma_z, seed, student_t_state=0, nu_state=5.0):
    """svpfo_init_particles: svpf_initialize's particle draw (float64)."""
    L = lib()
    L.svpfo_init_particles.restype = None
    L.svpfo_init_particles.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_uint64, ctypes.c_void_p]
    h = np.zeros(n)
    L.svpfo_init_particles(n, mu, rho, sigma_z, int(student_t_state), nu_state, seed, _ptr(h))
    return h


class Smooth(ctypes.Structure):
    _fields_ = [("lag", ctypes.c_int), ("output_lag", ctypes.c_int), ("head", ctypes.c_int), ("pad_", ctypes.c_int),
                ("h_mean", ctypes.c_double * 8), ("h_var", ctypes.c_double * 8), ("y", ctypes.c_double * 8)]


class Smoother:
    """svpfo_smooth_step: the RTS-style correction of the reported h_mean over a
    window of `lag` ticks (buffers start as svpf_create leaves them)."""

    def __init__(self, lag=3, output_lag=1):
        self.s = Smooth()
        self.s.lag, self.s.output_lag, self.s.head = lag, output_lag, 0
        for i in range(8):
            self.s.h_var[i] = 1.0
        L = lib()
        L.svpfo_smooth_step.restype = ctypes.c_double
        L.svpfo_smooth_step.argtypes = [ctypes.c_void_p, ctypes.c_int] + [ctypes.c_double] * 6

    def step(self, timestep, h_mean, h_var, y_t, mu, rho, sigma_z):
        return lib().svpfo_smooth_step(ctypes.byref(self.s), int(timestep), h_mean, h_var, y_t, mu, rho, sigma_z)


TWOF_INT_FIELDS = ("n", "n_stein_steps", "n_anneal_steps", "use_annealing", "use_svld", "use_mim", "use_guide",
                   "use_guide_preserving", "use_exact_gradient", "stein_sign", "pad0_", "pad1_")
TWOF_DOUBLE_FIELDS = ("mu", "nu", "rho_fast", "sigma_fast", "rho_slow", "sigma_slow", "bw_floor_fast",
                      "bw_floor_slow", "guide_strength", "mim_jump_prob", "mim_jump_scale", "temperature",
                      "rmsprop_rho", "rmsprop_eps")


class Cfg2F(ctypes.Structure):
    _fields_ = [(f, ctypes.c_int) for f in TWOF_INT_FIELDS] + [(f, ctypes.c_double) for f in TWOF_DOUBLE_FIELDS]


class State2F(ctypes.Structure):
    _fields_ = [("timestep", ctypes.c_int), ("guide_2f_initialized", ctypes.c_int)] + [
        (f, ctypes.c_double) for f in ("guide_mean_fast", "guide_mean_slow", "guide_var_fast", "guide_var_slow",
                                       "mean_fast_prev", "mean_slow_prev", "vol_prev", "h_mean_prev")]


class Out2F(ctypes.Structure):
    _fields_ = [(f, ctypes.c_double) for f in ("loglik", "pred_loglik", "vol", "h_mean", "mean_fast", "mean_slow",
                                               "bw_fast", "bw_slow")] + [("steps", ctypes.c_int),
                                                                         ("pad_", ctypes.c_int)]


def _lib2f():
    L = lib()
    P = ctypes.c_void_p
    L.svpfo_tick_2f.restype = ctypes.c_int
    L.svpfo_tick_2f.argtypes = [P, P, P, P, P, P, P, P, ctypes.c_double, ctypes.c_uint64, ctypes.c_uint32, P]
    L.svpfo_ekf_2f_update.restype = None
    L.svpfo_ekf_2f_update.argtypes = [P, P, ctypes.c_double]
    L.svpfo_init_2f.restype = None
    L.svpfo_init_2f.argtypes = [ctypes.c_int] + [ctypes.c_double] * 4 + [ctypes.c_uint64, P, P]
    return L


def make_cfg_2f(cfg: dict) -> Cfg2F:
    c = Cfg2F()
    for f in TWOF_INT_FIELDS:
        if not f.startswith("pad"):
            setattr(c, f, int(cfg[f]))
    for f in TWOF_DOUBLE_FIELDS:
        setattr(c, f, float(cfg[f]))
    return c


def init_2f(n, rho_fast, sigma_fast, rho_slow, sigma_slow, seed):
    hf, hs = np.zeros(n), np.zeros(n)
    _lib2f().svpfo_init_2f(n, rho_fast, sigma_fast, rho_slow, sigma_slow, seed, _ptr(hf), _ptr(hs))
    return hf, hs


class TwoFactorOracle:
    """The coordinate-wise two-factor step run by the oracle (float64 arrays)."""

    def __init__(self, cfg: dict, hf0, hs0, hfp0=None, hsp0=None, vf0=None, vs0=None, **state_kw):
        n = int(cfg["n"])
        self.cfg = dict(cfg)
        self.c = make_cfg_2f(cfg)
        self.st = State2F()
        for k, v in state_kw.items():
            setattr(self.st, k, v)
        arr = lambda a, d: (np.array(d, dtype=np.float64) if a is None else np.array(a, dtype=np.float64)).reshape(n).copy()
        self.hf, self.hs = arr(hf0, hf0), arr(hs0, hs0)
        self.hfp, self.hsp = arr(hfp0, hf0), arr(hsp0, hs0)
        self.vf, self.vs = arr(vf0, np.zeros(n)), arr(vs0, np.zeros(n))

    def tick(self, y_t, seed=0, filter_id=0):
        o = Out2F()
        rc = _lib2f().svpfo_tick_2f(ctypes.byref(self.c), ctypes.byref(self.st), _ptr(self.hf), _ptr(self.hs),
                                    _ptr(self.hfp), _ptr(self.hsp), _ptr(self.vf), _ptr(self.vs), float(y_t), seed,
                                    filter_id, ctypes.byref(o))
        if rc:
            raise RuntimeError("oracle two-factor tick failed")
        return {f: getattr(o, f) for f, _ in Out2F._fields_ if f != "pad_"}

    def ekf(self, y_t):
        _lib2f().svpfo_ekf_2f_update(ctypes.byref(self.c), ctypes.byref(self.st), float(y_t))


# ---- the joint state-parameter filter (svpf_joint/) ----
JOINT_INT_FIELDS = ("n", "n_stein_steps")
JOINT_DOUBLE_FIELDS = (
    "step_h", "step_mu", "step_rho", "step_sigma", "diffusion_mu", "diffusion_rho", "diffusion_sigma",
    "collapse_thresh_mu", "collapse_thresh_rho", "collapse_thresh_sigma", "mu_prior_mean", "mu_prior_var",
    "rho_prior_mean", "rho_prior_var", "sigma_prior_mean", "sigma_prior_var", "nu", "student_t_const",
    "student_t_implied_offset", "lik_offset", "prior_weight",
)
JOINT_OUT_DOUBLE = ("h_mean", "vol_mean", "mu_mean", "rho_mean", "sigma_mean", "mu_std", "rho_std", "sigma_std",
                    "std_mu_tilde", "std_rho_tilde", "std_sigma_tilde", "ess", "log_likelihood")


class JointCfg(ctypes.Structure):
    _fields_ = [(f, ctypes.c_int) for f in JOINT_INT_FIELDS] + [(f, ctypes.c_double) for f in JOINT_DOUBLE_FIELDS]


class JointOut(ctypes.Structure):
    _fields_ = [(f, ctypes.c_double) for f in JOINT_OUT_DOUBLE] + [("bw", ctypes.c_double * 4)] + [
        (f, ctypes.c_int) for f in ("mu_collapsed", "rho_collapsed", "sigma_collapsed", "pad_")]


def _libj():
    L = lib()
    P = ctypes.c_void_p
    L.svpfo_joint_init.restype = None
    L.svpfo_joint_init.argtypes = [P, ctypes.c_uint64, P, P]
    L.svpfo_joint_step.restype = ctypes.c_int
    L.svpfo_joint_step.argtypes = [P, P, P, ctypes.c_double, ctypes.c_uint64, ctypes.c_uint32, P]
    L.svpfo_joint_grad.restype = None
    L.svpfo_joint_grad.argtypes = [P, P, ctypes.c_double, ctypes.c_double, P, P]
    return L


def make_cfg_joint(cfg: dict) -> JointCfg:
    c = JointCfg()
    for f in JOINT_INT_FIELDS:
        setattr(c, f, int(cfg[f]))
    for f in JOINT_DOUBLE_FIELDS:
        setattr(c, f, float(cfg[f]))
    return c


class JointOracle:
    """The joint state-parameter filter run by the oracle; x is [4, n] = h, mu~, rho~, sigma~ (float64)."""

    def __init__(self, cfg: dict, x0=None, hp0=None, seed=0):
        self.cfg = dict(cfg)
        self.c = make_cfg_joint(cfg)
        n = int(cfg["n"])
        if x0 is None:
            self.x, self.hp = np.zeros((4, n)), np.zeros(n)
            _libj().svpfo_joint_init(ctypes.byref(self.c), seed, _ptr(self.x), _ptr(self.hp))
        else:
            self.x = np.array(x0, dtype=np.float64).reshape(4, n).copy()
            self.hp = (self.x[0] if hp0 is None else np.asarray(hp0, dtype=np.float64)).reshape(n).copy()

    def step(self, y_t, seed, tick):
        o = JointOut()
        rc = _libj().svpfo_joint_step(ctypes.byref(self.c), _ptr(self.x), _ptr(self.hp), float(y_t), seed, tick,
                                      ctypes.byref(o))
        if rc:
            raise RuntimeError("oracle joint step failed")
        d = {f: getattr(o, f) for f in JOINT_OUT_DOUBLE}
        d["bw"] = list(o.bw)
        for f in ("mu_collapsed", "rho_collapsed", "sigma_collapsed"):
            d[f] = getattr(o, f)
        return d

    def grad(self, xj, h_prev, y_t):
        g, lw = np.zeros(4), np.zeros(1)
        xa = np.array(xj, dtype=np.float64).reshape(4)
        _libj().svpfo_joint_grad(ctypes.byref(self.c), _ptr(xa), float(h_prev), float(y_t), _ptr(g), _ptr(lw))
        return g, float(lw[0])
