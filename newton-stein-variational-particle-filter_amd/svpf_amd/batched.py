"""Batched form: B independent filters with their own particle counts and
parameters, advanced through ticks [t0, t1) by persistent launches (one per
tile group, the groups running concurrently) through svpf_batched_run.

Device memory comes from torch (plumbing only); every tick runs in the HIP
kernels.  Record layouts mirror csrc/svpf_types.h.
"""
from __future__ import annotations

import ctypes
import json
import os

import numpy as np

from . import _lib
from .config import implied_offset, student_t_const

# ---- flags (svpf_types.h) ----
F_EXACT_GRAD, F_NEWTON, F_FULL_NEWTON, F_SVLD, F_ANNEAL, F_ADAPT_BETA, F_LINEAR_BETA, F_MIM = (1 << k for k in range(8))
(F_ASYM_RHO, F_LOCAL_PARAMS, F_GUIDED, F_GUIDE, F_GUIDE_PRESERVING, F_ADAPT_GUIDE, F_REJUV,
 F_FAN) = (1 << k for k in range(8, 16))
F_ADAPT_MU, F_ADAPT_SIGMA, F_KSD_EARLY_STOP, F_PAPER_SIGN = (1 << k for k in range(16, 20))
F_HEUN, F_STUDENT_STATE = 1 << 20, 1 << 21
TILE_EXT = 0x100  # ORed into a tile id: the kernels with the Heun step and the Student-t state

_PARAM_FLOATS = ("mu", "rho", "sigma_z", "gamma", "nu", "lik_offset", "temperature", "rms_rho", "rms_eps", "mim_p",
                 "mim_scale", "rho_up", "rho_down", "delta_rho", "delta_sigma", "ga_base", "ga_shock", "ga_thr",
                 "gs_base", "gs_max", "gs_thr", "rej_thr", "rej_p", "rej_blend", "mu_q", "mu_rscale", "mu_min",
                 "mu_max", "sb_thr", "sb_max", "ksd_thr", "st_const", "implied_offset", "base_step")
FPARAMS_DTYPE = np.dtype([("n", "<i4"), ("flags", "<u4"), ("n_anneal", "<i4"), ("stein_min", "<i4"),
                          ("stein_max", "<i4"), ("bandwidth_mode", "<i4"), ("pad0", "<i4", (2,))]
                         + [(f, "<f4") for f in _PARAM_FLOATS] + [("nu_state", "<f4"), ("coop_spin_limit", "<i4")]
                         + [(f, "<f4") for f in ("rho_fast", "sigma_fast", "rho_slow", "sigma_slow", "bw_floor_fast",
                                                 "bw_floor_slow", "guide_strength")]
                         + [("n_stein", "<i4"), ("coop_fault_wg", "<i4"), ("pad1", "<f4", (11,))])
FSTATE_DTYPE = np.dtype([("timestep", "<i4"), ("guide_initialized", "<i4"), ("stein_steps_used", "<i4"),
                         ("status", "<i4")]
                        + [(f, "<f4") for f in ("ksd_prev", "vol_prev", "h_mean_prev", "guide_mean", "guide_var",
                                                "bw_sq_ema", "return_ema", "return_var", "mu_state", "mu_var",
                                                "sigma_z_eff", "y_prev", "last_loglik", "last_vol", "last_h_mean",
                                                "last_bw", "last_ksd", "last_pred_loglik", "last_guide_strength")]
                        + [("epoch", "<u4"), ("med_r2", "<f4")]
                        + [(f, "<f4") for f in ("g2f_mean_fast", "g2f_mean_slow", "g2f_var_fast", "g2f_var_slow",
                                                "mean_fast_prev", "mean_slow_prev")]
                        + [("g2f_initialized", "<i4")])
FDESC_DTYPE = np.dtype([("off", "<i8")] + [(f, "<u8") for f in ("y", "out", "z_pred", "u_pred", "xi", "u_rej", "z_rej",
                                                                  "log_w", "h_pred", "seed")]
                       + [("xi_rows", "<i4"), ("param_idx", "<i4"), ("filter_id", "<u4"), ("tile", "<i4")])
assert FPARAMS_DTYPE.itemsize == 256 and FSTATE_DTYPE.itemsize == 128 and FDESC_DTYPE.itemsize == 104

OUTW = 8
O_LOGLIK, O_VOL, O_HMEAN, O_KSD, O_BW, O_PRED_LL, O_STEPS, O_BWSQ = range(8)

# Tile table (csrc/svpf_kernels.h): id -> (rows per lane, lanes per row)
TILES = {0: (1, 4), 1: (1, 2), 2: (1, 1), 3: (2, 1), 4: (4, 1), 5: (8, 1), 6: (2, 2), 7: (4, 2)}
BLOCK_SIZES = (256, 512, 1024)


def valid_tiles(bs):
    """Tile ids instantiated at block size bs (1024 threads: rows <= 4, no split rows)."""
    return [t for t in TILES if bs in (256, 512) or t <= 4]

_TUNED_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_tiles.json")


def tile_capacity(tile, bs):
    r, cs = TILES[tile]
    return r * bs // cs


def default_tile(n):
    """Smallest-capacity tile of block 256 (then 512) holding n particles."""
    for bs in (256, 512):
        for t in (0, 1, 2, 3, 4, 5):
            if tile_capacity(t, bs) >= n:
                return bs, t
    raise ValueError(f"n={n} exceeds the single-workgroup engine")


def tuned_table():
    if os.path.exists(_TUNED_PATH):
        with open(_TUNED_PATH) as f:
            return {int(k): tuple(v) for k, v in json.load(f).items()}
    return {}


def choose_tile(n, table=None):
    table = tuned_table() if table is None else table
    cands = [k for k in table if k >= n]
    if cands:
        bs, t = table[min(cands)]
        if tile_capacity(t, bs) >= n:
            return bs, t
    return default_tile(n)


def pack_params(cfg: dict) -> np.ndarray:
    """svpf_fparams for one filter from a config dict (field names of SVPFState)."""
    p = np.zeros(1, dtype=FPARAMS_DTYPE)[0]
    f = 0
    bits = [("use_exact_gradient", F_EXACT_GRAD), ("use_newton", F_NEWTON), ("use_svld", F_SVLD),
            ("use_annealing", F_ANNEAL), ("use_adaptive_beta", F_ADAPT_BETA), ("use_linear_beta", F_LINEAR_BETA),
            ("use_mim", F_MIM), ("use_asymmetric_rho", F_ASYM_RHO), ("use_local_params", F_LOCAL_PARAMS),
            ("use_guided", F_GUIDED), ("use_guide", F_GUIDE), ("use_guide_preserving", F_GUIDE_PRESERVING),
            ("use_adaptive_guide", F_ADAPT_GUIDE), ("use_rejuvenation", F_REJUV), ("use_fan_mode", F_FAN),
            ("use_adaptive_mu", F_ADAPT_MU), ("use_adaptive_sigma", F_ADAPT_SIGMA),
            ("use_ksd_early_stop", F_KSD_EARLY_STOP), ("use_heun", F_HEUN),
            ("use_student_t_state", F_STUDENT_STATE)]
    for k, b in bits:
        if cfg[k]:
            f |= b
    if cfg["use_newton"] and cfg["use_full_newton"]:
        f |= F_FULL_NEWTON
    if cfg["stein_sign"] == 1:
        f |= F_PAPER_SIGN
    p["n"] = cfg["n"]
    p["flags"] = f
    p["coop_fault_wg"] = -1
    p["nu_state"] = max(float(cfg["nu_state"]), 2.5)
    p["n_anneal"] = cfg["n_anneal_steps"]
    p["stein_min"] = cfg["stein_min_steps"]
    p["stein_max"] = cfg["stein_max_steps"]
    p["bandwidth_mode"] = cfg["bandwidth_mode"]
    src = dict(mu="mu", rho="rho", sigma_z="sigma_z", gamma="gamma", nu="nu", lik_offset="lik_offset",
               temperature="temperature", rms_rho="rmsprop_rho", rms_eps="rmsprop_eps", mim_p="mim_jump_prob",
               mim_scale="mim_jump_scale", rho_up="rho_up", rho_down="rho_down", delta_rho="delta_rho",
               delta_sigma="delta_sigma", ga_base="guided_alpha_base", ga_shock="guided_alpha_shock",
               ga_thr="guided_innovation_threshold", gs_base="guide_strength_base", gs_max="guide_strength_max",
               gs_thr="guide_innovation_threshold", rej_thr="rejuv_ksd_threshold", rej_p="rejuv_prob",
               rej_blend="rejuv_blend", mu_q="mu_process_var", mu_rscale="mu_obs_var_scale", mu_min="mu_min",
               mu_max="mu_max", sb_thr="sigma_boost_threshold", sb_max="sigma_boost_max",
               ksd_thr="ksd_improvement_threshold")
    for k, v in src.items():
        p[k] = cfg[v]
    p["st_const"] = student_t_const(cfg["nu"])
    p["implied_offset"] = implied_offset(cfg["nu"])
    p["base_step"] = 0.1 * (0.5 if cfg["use_guide"] else 1.0)
    return p


def stationary_var(cfg: dict) -> float:
    """Stationary variance of the state, scaled by nu_s / (nu_s - 2) for the
    Student-t state with nu_s clamped to >= 2.5 (src/svpf_optimized_graph.cu:366-383)."""
    v = cfg["sigma_z"] ** 2 / (1.0 - cfg["rho"] ** 2 + 1e-6)
    if cfg.get("use_student_t_state"):
        nu_s = max(float(cfg["nu_state"]), 2.5)
        v *= nu_s / (nu_s - 2.0)
    return v


def initial_fstate(cfg: dict) -> np.ndarray:
    """svpf_fstate as svpf_create + svpf_initialize leave it (src/svpf_optimized_graph.cu:349-451)."""
    s = np.zeros(1, dtype=FSTATE_DTYPE)[0]
    s["ksd_prev"] = 1e10
    s["vol_prev"] = 0.05
    s["h_mean_prev"] = -3.5
    s["guide_mean"] = cfg["mu"]
    s["guide_var"] = stationary_var(cfg)
    s["mu_state"] = cfg["mu"]
    s["mu_var"] = 1.0
    s["sigma_z_eff"] = cfg["sigma_z"]
    s["stein_steps_used"] = cfg["stein_max_steps"]
    return s


class Batch:
    """B filters resident on one GPU.

    cfgs: list of config dicts (config.make_config); y: [B, T] observations;
    h0: list of initial particle arrays; seeds: per-filter counter-generator keys.
    noise (optional, for parity runs): dict of per-filter lists of host arrays
    z_pred [T, n], xi [T, rows, n], u_pred / u_rej / z_rej [T, n].
    """

    def __init__(self, cfgs, y, h0, seeds=None, noise=None, device="cuda", tile_table=None, persistent_grid=None):
        import torch

        self.torch = torch
        self.L = _lib.load()
        B = len(cfgs)
        self.B = B
        self.cfgs = cfgs
        self.n = np.array([c["n"] for c in cfgs], dtype=np.int64)
        y = np.ascontiguousarray(y, dtype=np.float32)
        assert y.shape[0] == B
        self.T = y.shape[1]
        self.device = device
        self.off = np.concatenate([[0], np.cumsum(self.n)[:-1]]).astype(np.int64)
        total = int(self.n.sum())
        self.pool = torch.zeros(5 * total, dtype=torch.float32, device=device)  # h | hp | v | log_w | h_pred
        self.total = total
        hcat = np.concatenate([np.asarray(h, dtype=np.float32) for h in h0])
        self.pool[:total] = torch.from_numpy(hcat).to(device)
        self.pool[total:2 * total] = torch.from_numpy(hcat).to(device)
        self.y = torch.from_numpy(y).to(device)
        self.out = torch.zeros((B, self.T, OUTW), dtype=torch.float32, device=device)
        params = np.stack([pack_params(c) for c in cfgs])
        fstate = np.stack([initial_fstate(c) for c in cfgs])
        self.params = torch.from_numpy(params.view(np.uint8).reshape(-1)).to(device)
        self.fstate = torch.from_numpy(fstate.view(np.uint8).reshape(-1).copy()).to(device)
        seeds = np.arange(B, dtype=np.uint64) if seeds is None else np.asarray(seeds, dtype=np.uint64)
        self._noise_t = []
        desc = np.zeros(B, dtype=FDESC_DTYPE)
        cost = self.n.astype(np.float64) ** 2
        table = tuned_table() if tile_table is None else tile_table
        # filters with the Heun step or the Student-t state run in the extended instantiation
        tiles = [(lambda bt, c: (bt[0], bt[1] | (TILE_EXT if c["use_heun"] or c["use_student_t_state"] else 0)))(
            choose_tile(int(nb), table), c) for nb, c in zip(self.n, cfgs)]
        base = self.pool.data_ptr()
        for b in range(B):
            d = desc[b]
            d["off"] = self.off[b]
            d["y"] = self.y.data_ptr() + 4 * b * self.T
            d["out"] = self.out.data_ptr() + 4 * b * self.T * OUTW
            d["log_w"] = base + 4 * (3 * total + self.off[b])
            d["h_pred"] = base + 4 * (4 * total + self.off[b])
            d["seed"] = seeds[b]
            d["param_idx"] = b
            d["filter_id"] = b
            d["tile"] = tiles[b][1]
            if noise is not None:
                for k in ("z_pred", "u_pred", "xi", "u_rej", "z_rej"):
                    arr = noise.get(k)
                    if arr is None or arr[b] is None:
                        continue
                    t = torch.from_numpy(np.ascontiguousarray(arr[b], dtype=np.float32)).to(device)
                    self._noise_t.append(t)
                    d[k] = t.data_ptr()
                    if k == "xi":
                        d["xi_rows"] = arr[b].shape[1]
        self.desc = torch.from_numpy(desc.view(np.uint8).reshape(-1).copy()).to(device)
        # groups: filters with the same (bs, tile), sorted by decreasing cost inside a group
        keys = sorted(set(tiles), key=lambda k: -max(cost[b] for b in range(B) if tiles[b] == k))
        order, gstart, gtile, gbs, gnmax = [], [0], [], [], []
        for k in keys:
            members = [b for b in range(B) if tiles[b] == k]
            members.sort(key=lambda b: -cost[b])
            order.extend(members)
            gstart.append(len(order))
            gbs.append(k[0])
            gtile.append(k[1])
            gnmax.append(int(max(self.n[b] for b in members)))
        cus = torch.cuda.get_device_properties(torch.device(device)).multi_processor_count
        ggrid = []
        for g in range(len(keys)):
            nf = gstart[g + 1] - gstart[g]
            per_cu = max(1, int(self.L.svpf_tile_occupancy_query(gtile[g], gbs[g], gnmax[g])))
            slots = cus * per_cu if persistent_grid is None else persistent_grid
            ggrid.append(max(1, min(nf, slots)))
        self.order = torch.tensor(order, dtype=torch.int32, device=device)
        # per group its work-queue counter, then per filter (in order) its progress word
        self.queues = torch.zeros(max(1, len(keys)) + len(order), dtype=torch.int32, device=device)
        self.groups = dict(start=np.array(gstart, np.int32), tile=np.array(gtile, np.int32),
                           bs=np.array(gbs, np.int32), grid=np.array(ggrid, np.int32),
                           nmax=np.array(gnmax, np.int32))
        self.n_groups = len(keys)
        self.tiles = tiles

    def run(self, t0=0, t1=None, stream=None):
        t1 = self.T if t1 is None else t1
        assert 0 <= t0 <= t1 <= self.T
        torch = self.torch
        s = torch.cuda.current_stream(self.device) if stream is None else stream
        g = self.groups
        P = lambda a: a.ctypes.data_as(ctypes.c_void_p)
        base = self.pool.data_ptr()
        rc = self.L.svpf_batched_run(
            ctypes.c_void_p(self.params.data_ptr()), ctypes.c_void_p(self.fstate.data_ptr()),
            ctypes.c_void_p(self.desc.data_ptr()), ctypes.c_void_p(base), ctypes.c_void_p(base + 4 * self.total),
            ctypes.c_void_p(base + 8 * self.total), ctypes.c_void_p(self.order.data_ptr()), P(g["start"]),
            P(g["tile"]), P(g["bs"]), P(g["grid"]), P(g["nmax"]), self.n_groups,
            ctypes.c_void_p(self.queues.data_ptr()), int(t0), int(t1), ctypes.c_void_p(s.cuda_stream))
        if rc != 0:
            raise RuntimeError(f"svpf_batched_run failed ({rc})")

    def outputs(self):
        return self.out.cpu().numpy()

    def particles(self, b):
        o, n = int(self.off[b]), int(self.n[b])
        return self.pool[o:o + n].cpu().numpy()

    def fstates(self):
        return self.fstate.cpu().numpy().view(FSTATE_DTYPE)
