"""Pins of the oracle's single-block fused step (svpfo_v2_step, following
src/svpf_fully_fused_v2.cu:90-712).

Where its semantics coincide with svpf_step_graph's, one call equals one tick of
the separately pinned graph-path oracle (Silverman bandwidth, linear beta over
two stages, MIM / local / guided predict, guide, rejuvenation, adaptive sigma).
The parts of its own are pinned by invariants: the last iteration's KSD takes
raw scores (equal with and without local Newton at one iteration), the EKF
re-initialisation, the N = 1 closed forms of the full-Newton quirk, and the
adaptive-mu EMA's end points."""
import ctypes
import math

import numpy as np
import pytest

import oracle
from svpf_amd import config
from svpf_amd.fused_v2 import v2_args


def _state(n, seed, mu=-1.0, spread=0.4):
    rng = np.random.default_rng(seed)
    h = mu + spread * rng.standard_normal(n)
    return h.astype(np.float32).astype(np.float64)


def _noise(n, rows, seed):
    rng = np.random.default_rng(seed)
    return {"z_pred": rng.standard_normal(n, dtype=np.float32), "u_pred": rng.random(n, dtype=np.float32),
            "xi": rng.standard_normal((rows, n), dtype=np.float32), "u_rej": rng.random(n, dtype=np.float32),
            "z_rej": rng.standard_normal(n, dtype=np.float32)}


@pytest.mark.parametrize("newton", ["plain", "full"])
@pytest.mark.parametrize("features", ["bare", "mim_local_guided", "guide_rejuv_sigma"])
def test_reduces_to_graph_tick(newton, features):
    n, S, mu, rho, sz, nu = 48, 5, -1.0, 0.95, 0.4, 6.0
    y_t, y_prev, hmp, volp, ksdp, t = 0.9, -0.3, -0.9, 0.6, 0.4, 12
    kw = dict(n=n, timestep=t, y_t=y_t, y_prev=y_prev, h_mean_prev=hmp, vol_prev=volp, ksd_prev=ksdp, rho=rho,
              sigma_z=sz, mu=mu, nu=nu, lik_offset=0.3, student_t_const=config.student_t_const(nu),
              student_t_implied_offset=config.implied_offset(nu), gamma=-0.2, n_stein_steps=S, n_anneal_steps=2,
              use_newton=0, use_full_newton=int(newton == "full"), stein_sign_mode=0, step_size=0.1,
              temperature=0.5, use_guide=0, use_guided_predict=0, use_rejuvenation=0, use_adaptive_sigma=0,
              delta_rho=0.0, delta_sigma=0.0, mim_jump_prob=0.0, guide_innovation_threshold=1e9)
    gkw = dict(n=n, bandwidth_mode=config.BW_SILVERMAN, use_exact_gradient=1, use_newton=int(newton == "full"),
               use_full_newton=int(newton == "full"), stein_sign=0, use_svld=1, temperature=0.5, use_annealing=1,
               n_anneal_steps=2, use_adaptive_beta=0, use_linear_beta=1, stein_min_steps=S, stein_max_steps=S,
               use_mim=0, use_asymmetric_rho=0, use_local_params=0, use_guided=0, use_guide=0, use_rejuvenation=0,
               use_adaptive_sigma=0, mu=mu, rho=rho, sigma_z=sz, gamma=-0.2, nu=nu, lik_offset=0.3)
    if features == "mim_local_guided":
        kw.update(mim_jump_prob=0.3, mim_jump_scale=1.5, delta_rho=0.02, delta_sigma=0.1, use_guided_predict=1,
                  guided_alpha_shock=0.4, guided_innovation_thresh_predict=0.2, y_t=2.5)
        gkw.update(use_mim=1, mim_jump_prob=0.3, mim_jump_scale=1.5, use_local_params=1, delta_rho=0.02,
                   delta_sigma=0.1, use_guided=1, guided_alpha_shock=0.4, guided_innovation_threshold=0.2)
    elif features == "guide_rejuv_sigma":
        kw.update(use_guide=1, use_guide_preserving=1, guide_strength_base=0.1, step_size=0.05, use_rejuvenation=1,
                  rejuv_ksd_threshold=0.05, rejuv_prob=0.4, rejuv_blend=0.3, use_adaptive_sigma=1,
                  sigma_boost_threshold=0.5, sigma_boost_max=2.0)
        gkw.update(use_guide=1, use_guide_preserving=1, guide_strength_base=0.1, use_rejuvenation=1,
                   rejuv_ksd_threshold=0.05, rejuv_prob=0.4, rejuv_blend=0.3, use_adaptive_sigma=1,
                   sigma_boost_threshold=0.5, sigma_boost_max=2.0)
    a = v2_args(**kw)
    cfg = config.make_config(**gkw)
    h0, hp0 = _state(n, 1), _state(n, 2)
    nz = _noise(n, S, 3)
    h, hp, v = h0.copy(), hp0.copy(), np.full(n, 0.05)
    _, _, _, io = oracle.v2_step(a, h, hp, v, {"guide_mean": 0.0, "guide_var": 0.0}, noise=nz)
    o = oracle.OracleFilter(cfg, h0, h_prev0=hp0, v0=np.full(n, 0.05), timestep=t, h_mean_prev=hmp,
                            vol_prev=volp, ksd_prev=ksdp)
    out = o.tick(y_t if features != "mim_local_guided" else 2.5, y_prev, **nz)
    assert np.max(np.abs(h - o.h)) <= 1e-12
    assert np.max(np.abs(v - o.v)) <= 1e-12
    assert np.array_equal(hp, o.h_prev)
    for k in ("h_mean", "vol", "loglik", "bandwidth", "ksd"):
        assert abs(io[k] - out[k]) <= 1e-12 * max(1.0, abs(out[k])), k
    if features == "guide_rejuv_sigma":
        assert abs(io["guide_mean"] - o.st.guide_mean) <= 1e-12
        assert abs(io["guide_var"] - o.st.guide_var) <= 1e-12


def test_ksd_takes_raw_scores():
    """One Stein iteration: the KSD is evaluated at the positions before the
    transport, with the raw (unpreconditioned) scores, so local Newton does not
    change it -- only the 1e-8 guard of the division differs.  (A wide cloud
    keeps the estimate above its floor at 0.)"""
    n = 40
    base = dict(n=n, n_stein_steps=1, n_anneal_steps=1, use_guide=0, use_guided_predict=0, use_rejuvenation=0,
                use_full_newton=0, mim_jump_prob=0.0, sigma_z=0.3, mu=-1.0, y_t=0.7)
    out = {}
    for newton in (0, 1):
        h, hp, v = _state(n, 4, spread=2.0), _state(n, 5, spread=2.0), np.zeros(n)
        _, _, _, io = oracle.v2_step(v2_args(use_newton=newton, **base), h, hp, v, {}, noise=_noise(n, 1, 6))
        out[newton] = io["ksd"]
    assert out[0] > 0.0
    assert abs(out[1] - out[0]) <= 1e-6 * out[0]


def test_full_newton_quirk_single_particle():
    """N = 1 (K = 1, no repulsion), use_newton and use_full_newton both set, as
    the reference's own driver does: the kernel-weighted curvature is
    curv + 2/bw^2, phi = 0.7 g / that, and the KSD's raw score of the column
    particle is divided by 0.7 curv (the stored term is the curvature, not its
    inverse): ksd^2 = g^2 / (0.7 curv + 1e-8) - 2/bw^2 (floored at 0)."""
    y_t, hv, mu, sz, nu = 1.2, -0.8, -1.0, 0.3, 5.0
    a = v2_args(n=1, n_stein_steps=1, n_anneal_steps=1, use_guide=0, use_guided_predict=0, use_rejuvenation=0,
                use_newton=1, use_full_newton=1, mim_jump_prob=0.0, temperature=0.0, y_t=y_t, mu=mu, sigma_z=sz,
                nu=nu, rho=0.9, gamma=0.0, use_adaptive_sigma=0, lik_offset=0.2, rmsprop_eps=1e-6)
    h, hp, v = np.array([hv]), np.array([hv]), np.array([1.0])
    nz = {"z_pred": np.array([0.3], np.float32)}
    g, lw, cvo, io = oracle.v2_step(a, h, hp, v, {}, noise=nz)
    hpred = mu + 0.9 * (hv - mu) + sz * float(np.float32(0.3))
    bw = 0.05  # Silverman with one particle: the variance floor 1e-6 gives 1.06e-3, clamped to 0.05
    ib = 1.0 / bw ** 2
    A = y_t ** 2 / (math.exp(hpred) + 1e-8) / nu
    gl = -0.5 + 0.5 * (nu + 1) * A / (1 + A) - 0.2
    # one mixture component, at the transition mean of the same particle (weight 1)
    m = mu + 0.9 * (hv - mu)
    gp = -(hpred - m) / sz ** 2 / (1 + 1e-8)
    gg = min(max(min(max(gp, -10), 10) + gl, -10), 10)
    cv = min(max(0.5 * (nu + 1) * A / (1 + A) ** 2 + 1 / sz ** 2, 0.1), 100)
    hw = min(max(cv + 2 * ib, 0.1), 100)
    phi = 0.7 * gg / hw
    vn = 0.9 * 1.0 + 0.1 * phi * phi
    assert abs(g[0] - gg) <= 1e-9 and abs(cvo[0] - cv) <= 1e-9
    assert abs(io["bandwidth"] - bw) <= 1e-15
    assert abs(h[0] - (hpred + 0.25 * phi / math.sqrt(vn + 1e-6))) <= 1e-9
    assert abs(io["ksd"] - math.sqrt(max(gg * gg / (0.7 * cv + 1e-8) - 2 * ib, 0.0))) <= 1e-9


def test_ekf_reinitialises_and_matches_the_graph_guide():
    """guide_var < 1e-6 on entry re-initialises the guide at (mu, stationary
    variance); the update then equals the graph path's EKF from the same start."""
    n = 16
    a = v2_args(n=n, y_t=0.8, mu=-1.2, rho=0.95, sigma_z=0.2, use_adaptive_mu=0, n_stein_steps=1, n_anneal_steps=1)
    h, hp, v = _state(n, 7, mu=-1.2), _state(n, 8, mu=-1.2), np.zeros(n)
    _, _, _, io = oracle.v2_step(a, h, hp, v, {"guide_mean": 5.0, "guide_var": 0.0}, noise=_noise(n, 1, 9))
    cfg = config.make_config(n=n, mu=-1.2, rho=0.95, sigma_z=0.2, nu=5.0)
    st = oracle.make_state(cfg)
    oracle.lib().svpfo_ekf_update(ctypes.byref(oracle.make_cfg(cfg)), ctypes.byref(st), 0.8, -1.2, 0.95, 0.2)
    assert abs(io["guide_mean"] - st.guide_mean) <= 1e-12
    assert abs(io["guide_var"] - st.guide_var) <= 1e-12


def test_adaptive_mu_ema_end_points():
    n = 16
    for alpha, want in ((0.0, lambda hm: -2.0), (1.0, lambda hm: min(max(hm, -10.0), 0.0))):
        a = v2_args(n=n, timestep=11, use_adaptive_mu=1, mu_ema_alpha=alpha, n_stein_steps=1, n_anneal_steps=1)
        h, hp, v = _state(n, 10, mu=-2.0), _state(n, 11, mu=-2.0), np.zeros(n)
        _, _, _, io = oracle.v2_step(a, h, hp, v, {"mu_state": -2.0}, noise=_noise(n, 1, 12))
        assert abs(io["mu_state"] - want(io["h_mean"])) <= 1e-12
    # at timestep <= 10 the estimate is left alone
    a = v2_args(n=n, timestep=10, use_adaptive_mu=1, mu_ema_alpha=1.0, n_stein_steps=1, n_anneal_steps=1)
    h, hp, v = _state(n, 10, mu=-2.0), _state(n, 11, mu=-2.0), np.zeros(n)
    _, _, _, io = oracle.v2_step(a, h, hp, v, {"mu_state": -2.0}, noise=_noise(n, 1, 12))
    assert io["mu_state"] == -2.0


@pytest.mark.parametrize("stages,middle", [(3, [0.55]), (4, [0.3 + 0.5 / 3, 0.3 + 1.0 / 3])])
def test_beta_ramp_recovered_from_a_single_particle(stages, middle):
    """The fused step's annealing ramp (src/svpf_fully_fused_v2.cu:441-451: 0.3 at the
    first stage, 1.0 at the last, 0.3 + 0.5 ai/(A-1) between), recovered from the
    motion of one particle.  At N = 1 the operator is the score itself (K(0) = 1, no
    repulsion); with no noise each stage moves the particle by
    0.1 sqrt(beta) phi / sqrt(v' + 1e-6), phi = the mixture-prior score plus beta times
    the likelihood score (central differences of scipy's Student-t).  The middle
    stages' beta are solved for from the final position and state (one iteration per
    stage); the graph path's linear rule (0.3 + 0.7 ai/(A-1)) would give other values."""
    import scipy.optimize
    import scipy.stats

    mu, rho, sig, nu, y, h0 = -1.0, 0.95, 0.3, 5.0, 0.8, -0.6
    a = v2_args(n=1, timestep=0, y_t=y, y_prev=0.0, gamma=0.0, n_stein_steps=stages, n_anneal_steps=stages,
                use_newton=0, use_full_newton=0, temperature=0.0, use_guide=0, use_guided_predict=0,
                use_rejuvenation=0, use_adaptive_mu=0, use_adaptive_sigma=0, mim_jump_prob=0.0, delta_rho=0.0,
                delta_sigma=0.0, step_size=0.1, sigma_z=sig, mu=mu, rho=rho, nu=nu, lik_offset=0.0,
                student_t_const=config.student_t_const(nu), student_t_implied_offset=config.implied_offset(nu))
    h, hp, v = np.array([h0]), np.array([h0]), np.zeros(1)
    nz = {"z_pred": np.zeros(1, np.float32), "u_pred": np.ones(1, np.float32),
          "xi": np.zeros((stages, 1), np.float32)}
    oracle.v2_step(a, h, hp, v, {"guide_mean": 0.0, "guide_var": 0.0}, noise=nz)
    m = mu + rho * (h0 - mu)

    def score(x):
        lp = lambda q: scipy.stats.t.logpdf(y * math.exp(-q / 2), nu) - q / 2
        return (lp(x + 1e-5) - lp(x - 1e-5)) / 2e-5

    def run(betas):
        x, vv = m, 0.0  # z = 0: the prediction is the component mean
        for b in betas:
            phi = float(np.clip(np.clip(-(x - m) / sig ** 2, -10, 10) + b * score(x), -10, 10))
            vv = 0.9 * vv + 0.1 * phi * phi
            x = x + 0.1 * math.sqrt(b) * phi / math.sqrt(vv + 1e-6)
        return x, vv

    if stages == 3:
        b = scipy.optimize.brentq(lambda q: run([0.3, q, 1.0])[0] - h[0], 0.31, 0.99, xtol=1e-12)
        got = [b]
    else:
        sol = scipy.optimize.fsolve(lambda q: [run([0.3, q[0], q[1], 1.0])[0] - h[0],
                                               run([0.3, q[0], q[1], 1.0])[1] - v[0]], [0.5, 0.7], xtol=1e-12)
        got = list(sol)
    assert np.allclose(got, middle, atol=1e-5), got
    x, vv = run([0.3] + list(middle) + [1.0])
    assert h[0] == pytest.approx(x, abs=1e-7) and v[0] == pytest.approx(vv, abs=1e-7)
