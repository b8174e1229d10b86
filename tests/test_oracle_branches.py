"""Pins the oracle's branchy stages (CPU only): MIM / asymmetric-rho / guided
predict, the EKF guide and its two application rules, adaptive guide strength,
adaptive sigma, adaptive mu, rejuvenation, the smoothing of the reported mean,
svpf_initialize's particle draw and the coordinate-wise two-factor step.

Each check reduces a branch to something fixed independently of the branch's
own code: another, simpler branch of the filter (a feature switched off), a
textbook identity (Gaussian conditioning, the exact Kalman update of a linear
model, precision-weighted averages), a library routine (scipy's Student-t,
numpy's moments) or a composition of separately pinned pieces (the Philox
known-answer vectors, the normal transform).  The configurations below make the
tick stop after the stage under test: a Stein budget of zero iterations
(stein_max_steps = 0 while timestep < 10) leaves the particles exactly where
predict, guide and rejuvenation put them.
"""
import math

import numpy as np
import pytest
import scipy.special
import scipy.stats

import oracle
from svpf_amd import config

L = oracle.lib()


def cfg0(n=16, **kw):
    """Predict only: no Stein iteration, no guide, no rejuvenation."""
    base = dict(n=n, stein_max_steps=0, stein_min_steps=0, use_guide=0, use_rejuvenation=0, use_mim=0,
                use_asymmetric_rho=0, use_adaptive_beta=0, gamma=0.0, mu=-1.0, rho=0.95, sigma_z=0.2)
    base.update(kw)
    return config.make_config(**base)


def tick(cfg, h0, y=0.3, y_prev=0.0, z=None, u=None, u_rej=None, z_rej=None, h_prev0=None, **state):
    n = cfg["n"]
    rng = np.random.default_rng(123)
    z = rng.standard_normal(n).astype(np.float32) if z is None else np.asarray(z, np.float32)
    f = oracle.OracleFilter(cfg, h0, h_prev0=h_prev0, **state)
    out = f.tick(y, y_prev, z_pred=z, xi=np.zeros((1, n), np.float32), u_pred=u, u_rej=u_rej, z_rej=z_rej)
    return f, out


def particles(n=16, seed=0, loc=-1.0, scale=0.4):
    return np.random.default_rng(seed).normal(loc, scale, n)


# ---------------------------------------------------------------- predict
def test_zero_stein_budget_leaves_the_plain_prediction():
    """The harness itself: with no Stein iteration the tick's particles are the AR(1) prediction."""
    n = 12
    h0 = particles(n)
    z = np.random.default_rng(5).standard_normal(n).astype(np.float32)
    f, out = tick(cfg0(n, gamma=-0.3), h0, y_prev=0.2, z=z)
    lev = -0.3 * 0.2 / (np.exp(h0 / 2) + 1e-8)
    want = np.clip(-1.0 + 0.95 * (h0 + 1.0) + 0.2 * z.astype(np.float64) + lev, -15, 5)
    assert np.allclose(f.h, want, atol=1e-12)
    assert out["steps"] == 0


def test_mim_without_jumps_equals_plain_predict():
    n = 16
    h0 = particles(n)
    u = np.random.default_rng(1).random(n).astype(np.float32)
    plain, _ = tick(cfg0(n), h0)
    mim, _ = tick(cfg0(n, use_mim=1, mim_jump_prob=0.0), h0, u=u)
    assert np.array_equal(plain.h, mim.h)


def test_mim_jump_scales_the_noise():
    n = 16
    h0 = particles(n)
    u = np.random.default_rng(2).random(n).astype(np.float32)  # all < 1
    mim, _ = tick(cfg0(n, use_mim=1, mim_jump_prob=1.0, mim_jump_scale=5.0), h0, u=u)
    wide, _ = tick(cfg0(n, sigma_z=1.0), h0)  # 0.2 * 5
    assert np.allclose(mim.h, wide.h, atol=1e-12)
    # a selector at or above the jump probability keeps the plain noise
    keep, _ = tick(cfg0(n, use_mim=1, mim_jump_prob=0.5, mim_jump_scale=5.0), h0, u=np.full(n, 0.5, np.float32))
    plain, _ = tick(cfg0(n), h0)
    assert np.allclose(keep.h, plain.h, atol=1e-12)


def test_asymmetric_rho_takes_rho_up_for_rising_particles():
    """rho = rho_up where h > h_prev, rho_down otherwise (src/svpf_opt_kernels.cu:193):
    a swap of the two fails this check."""
    n = 20
    h0 = particles(n)
    hp = h0 + np.where(np.arange(n) % 2 == 0, -0.3, 0.3)  # even particles rose
    u = np.ones(n, np.float32)
    c = cfg0(n, use_mim=1, mim_jump_prob=0.0, use_asymmetric_rho=1, rho_up=0.99, rho_down=0.5)
    asym, _ = tick(c, h0, u=u, h_prev0=hp)
    up, _ = tick(cfg0(n, rho=0.99), h0, h_prev0=hp)
    dn, _ = tick(cfg0(n, rho=0.5), h0, h_prev0=hp)
    rose = h0 > hp
    assert rose.sum() == n // 2
    assert np.allclose(asym.h[rose], up.h[rose], atol=1e-12)
    assert np.allclose(asym.h[~rose], dn.h[~rose], atol=1e-12)


@pytest.mark.parametrize("offset", [0.5, -0.5])
def test_local_params_follow_the_deviation_from_the_previous_mean(offset):
    """rho + delta_rho tanh(h - h_mean_prev), sigma (1 + delta_sigma |h - h_mean_prev|): with
    every particle at c the tick equals the plain predict with those two numbers."""
    n, c = 8, -1.2
    h0 = np.full(n, c)
    u = np.ones(n, np.float32)
    loc = cfg0(n, use_mim=1, mim_jump_prob=0.0, use_local_params=1, delta_rho=0.05, delta_sigma=0.4)
    f, _ = tick(loc, h0, u=u, h_mean_prev=c - offset)
    dev = offset
    ref, _ = tick(cfg0(n, rho=0.95 + 0.05 * math.tanh(dev), sigma_z=0.2 * (1 + 0.4 * abs(dev))), h0)
    assert np.allclose(f.h, ref.h, atol=1e-12)
    # no deviation: no effect
    f0, _ = tick(loc, h0, u=u, h_mean_prev=c)
    p0, _ = tick(cfg0(n), h0)
    assert np.allclose(f0.h, p0.h, atol=1e-12)


def test_guided_predict_with_zero_blend_equals_mim():
    n = 16
    h0 = particles(n)
    u = np.random.default_rng(3).random(n).astype(np.float32)
    kw = dict(use_mim=1, mim_jump_prob=0.1, mim_jump_scale=3.0, use_asymmetric_rho=1, rho_up=0.97, rho_down=0.9)
    mim, _ = tick(cfg0(n, **kw), h0, y=2.5, u=u, h_prev0=h0 - 0.1)
    gd, _ = tick(cfg0(n, use_guided=1, guided_alpha_base=0.0, guided_alpha_shock=0.0, **kw), h0, y=2.5, u=u,
                 h_prev0=h0 - 0.1)
    assert np.array_equal(mim.h, gd.h)
    # a shock blend that never activates (threshold out of reach) also equals MIM
    gq, _ = tick(cfg0(n, use_guided=1, guided_alpha_base=0.0, guided_alpha_shock=1.0,
                      guided_innovation_threshold=1e6, **kw), h0, y=2.5, u=u, h_prev0=h0 - 0.1)
    assert np.array_equal(mim.h, gq.h)


@pytest.mark.parametrize("y", [2.5, 1e-4])
def test_guided_predict_full_blend_lands_on_the_implied_mean(y):
    """alpha = 1: the proposal mean is log(y^2) + the Student-t implied offset, floored at -5."""
    n = 8
    h0 = particles(n)
    c = cfg0(n, use_guided=1, guided_alpha_base=1.0, guided_alpha_shock=1.0, nu=5.0)
    f, _ = tick(c, h0, y=y, u=np.ones(n, np.float32), z=np.zeros(n, np.float32))
    implied = -(math.log(5.0) + scipy.special.digamma(0.5) - scipy.special.digamma(2.5))
    want = max(math.log(y * y + 1e-10) + L.svpfo_implied_offset(5.0), -5.0)
    assert abs(L.svpfo_implied_offset(5.0) - implied) < 2e-3  # the 2-term psi of the reference
    assert np.allclose(f.h, want, atol=1e-12)


# ---------------------------------------------------------------- EKF guide and its application
def test_ekf_mean_is_the_precision_weighted_average():
    """The EKF update of the log y^2 observation (include/svpf_kernels.cuh:509-541) is the
    Gaussian posterior of a prior N(m_pred, P_pred) and an observation log y^2 + offset with
    variance R = 4.93 + 2: (m/P + o/R) / (1/P + 1/R)."""
    cfg = config.make_config(n=4, nu=5.0, mu=-1.0, rho=0.97, sigma_z=0.15)
    c = oracle.make_cfg(cfg)
    for y in (0.02, 0.4, 3.0):
        st = oracle.make_state(cfg, guide_initialized=1, guide_mean=-0.7, guide_var=0.3)
        L.svpfo_ekf_update(oracle.ctypes.byref(c), oracle.ctypes.byref(st), y, -1.0, 0.97, 0.15)
        m_pred = -1.0 + 0.97 * (-0.7 + 1.0)
        P_pred = 0.97 ** 2 * 0.3 + 0.15 ** 2
        o = math.log(y * y + 1e-8) + L.svpfo_implied_offset(5.0)
        R = 6.93
        assert st.guide_mean == pytest.approx((m_pred / P_pred + o / R) / (1 / P_pred + 1 / R), abs=1e-8)
        assert st.guide_var == pytest.approx(1.0 / (1 / P_pred + 1 / R), abs=1e-8)


def guide_cfg(n, **kw):
    return cfg0(n, use_guide=1, rho=1.0, **kw)


# rho = 1 makes the prediction the particles themselves (with z = 0); the EKF then starts
# from a given variance instead of the stationary one, which rho = 1 does not have
GUIDE_STATE = dict(guide_initialized=1, guide_mean=-0.8, guide_var=0.05)


def test_guide_preserving_strength_zero_is_identity_and_one_moves_the_mean():
    n = 16
    h0 = particles(n)
    z = np.zeros(n, np.float32)  # rho = 1, no noise: the prediction is h0
    m0 = float(h0.mean())
    off, _ = tick(cfg0(n, rho=1.0), h0, z=z, h_mean_prev=m0)
    g0, _ = tick(guide_cfg(n, use_guide_preserving=1, guide_strength_base=0.0), h0, z=z, h_mean_prev=m0,
                 **GUIDE_STATE)
    assert np.allclose(g0.h, off.h, atol=1e-12)
    g1, _ = tick(guide_cfg(n, use_guide_preserving=1, guide_strength_base=1.0), h0, z=z, h_mean_prev=m0,
                 **GUIDE_STATE)
    assert g1.h.mean() == pytest.approx(g1.st.guide_mean, abs=1e-12)
    assert np.allclose(g1.h - g1.h.mean(), h0 - m0, atol=1e-12)


def test_guide_contraction_blends_each_particle():
    n = 16
    h0 = particles(n)
    z = np.zeros(n, np.float32)
    g1, _ = tick(guide_cfg(n, use_guide_preserving=0, guide_strength_base=1.0), h0, z=z, **GUIDE_STATE)
    assert np.allclose(g1.h, g1.st.guide_mean, atol=1e-12)
    gh, _ = tick(guide_cfg(n, use_guide_preserving=0, guide_strength_base=0.5), h0, z=z, **GUIDE_STATE)
    assert np.allclose(gh.h, 0.5 * (h0 + gh.st.guide_mean), atol=1e-12)


def test_adaptive_guide_strength_below_and_far_above_its_threshold():
    """Below the innovation threshold the strength stays at its base; three units above
    it (severity 1) it reaches the maximum (src/svpf_optimized_graph.cu:1006-1039)."""
    n = 16
    h0 = particles(n)
    z = np.zeros(n, np.float32)
    kw = dict(use_guide_preserving=0, guide_strength_base=0.1, guide_strength_max=0.6, guide_innovation_threshold=1.0)
    vol = 0.5
    for y, want in ((0.3, 0.1), (0.5 * 4.5, 0.6)):  # |y| / vol = 0.6 and 4.5
        ad, _ = tick(guide_cfg(n, use_adaptive_guide=1, **kw), h0, y=y, z=z, timestep=1, vol_prev=vol,
                     **GUIDE_STATE)
        fx, _ = tick(guide_cfg(n, **dict(kw, guide_strength_base=want)), h0, y=y, z=z, timestep=1, vol_prev=vol,
                     **GUIDE_STATE)
        assert np.allclose(ad.h, fx.h, atol=1e-12), y


def test_adaptive_sigma_below_and_far_above_its_threshold():
    n = 16
    h0 = particles(n)
    vol = 0.5
    kw = dict(sigma_boost_threshold=1.0, sigma_boost_max=3.0)
    low, _ = tick(cfg0(n, use_adaptive_sigma=1, **kw), h0, y=0.3, timestep=1, vol_prev=vol)
    off, _ = tick(cfg0(n), h0, y=0.3, timestep=1, vol_prev=vol)
    assert np.array_equal(low.h, off.h)
    high, _ = tick(cfg0(n, use_adaptive_sigma=1, **kw), h0, y=0.5 * 4.5, timestep=1, vol_prev=vol)
    wide, _ = tick(cfg0(n, sigma_z=0.6), h0, y=0.5 * 4.5, timestep=1, vol_prev=vol)
    assert np.allclose(high.h, wide.h, atol=1e-12)
    assert high.st.sigma_z_effective == pytest.approx(0.6, abs=1e-12)


# ---------------------------------------------------------------- adaptive mu
def test_adaptive_mu_state_is_the_predict_mean():
    n = 16
    h0 = particles(n)
    ad, _ = tick(cfg0(n, use_adaptive_mu=1), h0, mu_state=-2.0)
    fx, _ = tick(cfg0(n, mu=-2.0), h0)
    assert np.allclose(ad.h, fx.h, atol=1e-12)


def test_adaptive_mu_update_is_a_precision_weighted_average():
    """After tick 10: mu <- (mu/P + h_mean/R) / (1/P + 1/R), P = mu_var + q, R = scale x bw^2,
    clamped to [mu_min, mu_max] (src/svpf_optimized_graph.cu:619-641)."""
    n = 16
    h0 = particles(n, loc=-2.0)
    kw = dict(use_adaptive_mu=1, mu_process_var=0.01, mu_obs_var_scale=2.0, mu_min=-6.0, mu_max=0.0,
              stein_min_steps=0, stein_max_steps=0)
    f, out = tick(cfg0(n, **kw), h0, timestep=11, mu_state=-1.5, mu_var=0.04)
    P, R = 0.04 + 0.01, 2.0 * out["bandwidth"] ** 2
    assert f.st.mu_state == pytest.approx((-1.5 / P + out["h_mean"] / R) / (1 / P + 1 / R), abs=1e-7)
    assert f.st.mu_var == pytest.approx(1.0 / (1 / P + 1 / R), abs=1e-7)
    # no uncertainty: mu stays; clamp to the bounds
    f0, _ = tick(cfg0(n, **dict(kw, mu_process_var=0.0)), h0, timestep=11, mu_state=-1.5, mu_var=0.0)
    assert f0.st.mu_state == -1.5
    fc, _ = tick(cfg0(n, **dict(kw, mu_obs_var_scale=0.0, mu_min=-1.8)), h0, timestep=11, mu_state=-1.5, mu_var=1.0)
    assert fc.st.mu_state == -1.8
    # at or before tick 10 the state is left alone
    fe, _ = tick(cfg0(n, **kw), h0, timestep=10, mu_state=-1.5, mu_var=0.04)
    assert fe.st.mu_state == -1.5


# ---------------------------------------------------------------- rejuvenation
def rejuv_cfg(n, **kw):
    return cfg0(n, use_guide=1, guide_strength_base=0.0, use_rejuvenation=1, rejuv_ksd_threshold=0.3,
                stein_min_steps=0, stein_max_steps=0, **kw)


def test_rejuvenation_probability_zero_is_identity():
    n = 16
    h0 = particles(n)
    ur = np.random.default_rng(4).random(n).astype(np.float32)
    zr = np.random.default_rng(5).standard_normal(n).astype(np.float32)
    a, _ = tick(rejuv_cfg(n, rejuv_prob=0.0, rejuv_blend=0.5), h0, u_rej=ur, z_rej=zr, timestep=11, ksd_prev=1.0)
    b, _ = tick(cfg0(n, use_guide=1, guide_strength_base=0.0, stein_min_steps=0, stein_max_steps=0), h0,
                timestep=11, ksd_prev=1.0)
    assert np.array_equal(a.h, b.h)


def test_rejuvenation_full_blend_redraws_from_the_guide():
    """p = 1, blend 1: h = m_EKF + sqrt(P_EKF) z for every particle (src/svpf_opt_kernels.cu:1248-1275)."""
    n = 16
    h0 = particles(n)
    ur = np.random.default_rng(4).random(n).astype(np.float32)
    zr = np.random.default_rng(5).standard_normal(n).astype(np.float32)
    f, _ = tick(rejuv_cfg(n, rejuv_prob=1.0, rejuv_blend=1.0), h0, u_rej=ur, z_rej=zr, timestep=11, ksd_prev=1.0)
    want = np.clip(f.st.guide_mean + math.sqrt(max(f.st.guide_var, 1e-6)) * zr.astype(np.float64), -15, 5)
    assert np.allclose(f.h, want, atol=1e-12)
    # half the particles selected (u < p), half kept
    ur2 = np.where(np.arange(n) % 2 == 0, 0.1, 0.9).astype(np.float32)
    g, _ = tick(rejuv_cfg(n, rejuv_prob=0.5, rejuv_blend=1.0), h0, u_rej=ur2, z_rej=zr, timestep=11, ksd_prev=1.0)
    sel = ur2 < 0.5
    assert np.allclose(g.h[sel], want[sel], atol=1e-12)
    assert np.allclose(g.h[~sel], np.clip(-1.0 + 0.95 * (h0[~sel] + 1.0) + 0.2 * np.random.default_rng(123)
                                          .standard_normal(n).astype(np.float32)[~sel].astype(np.float64), -15, 5),
                       atol=1e-12)


@pytest.mark.parametrize("state", [dict(timestep=11, ksd_prev=0.1), dict(timestep=10, ksd_prev=1.0)])
def test_rejuvenation_gates(state):
    """No rejuvenation while the previous KSD is below its threshold or before tick 11."""
    n = 16
    h0 = particles(n)
    ur = np.zeros(n, np.float32)
    zr = np.ones(n, np.float32)
    a, _ = tick(rejuv_cfg(n, rejuv_prob=1.0, rejuv_blend=1.0), h0, u_rej=ur, z_rej=zr, **state)
    b, _ = tick(rejuv_cfg(n, rejuv_prob=0.0, rejuv_blend=1.0), h0, u_rej=ur, z_rej=zr, **state)
    assert np.array_equal(a.h, b.h)


# ---------------------------------------------------------------- adaptive beta and the Stein budget
def test_adaptive_beta_properties():
    """src/svpf_optimized_graph.cu:1066-1092: the last stage is 1; otherwise beta falls
    (weakly) as the previous KSD rises, is continuous in it, linear between the KSD
    breakpoints, grows with the stage, and stays in [0.1, 1]."""
    ksd = np.concatenate([np.linspace(0.0, 1.0, 401), [10.0, 1e10]])
    for na in (2, 3, 5):
        assert all(L.svpfo_adaptive_beta(k, na - 1, na) == 1.0 for k in ksd)
        for ai in range(na - 1):
            b = np.array([L.svpfo_adaptive_beta(k, ai, na) for k in ksd])
            assert (np.diff(b) <= 1e-15).all() and (b >= 0.1).all() and (b <= 1.0).all()
            assert np.abs(np.diff(b[:401])).max() < 0.01  # continuous over the grid step 0.0025
            lo, hi = L.svpfo_adaptive_beta(0.05, ai, na), L.svpfo_adaptive_beta(0.5, ai, na)
            assert L.svpfo_adaptive_beta(0.275, ai, na) == pytest.approx(0.5 * (lo + hi), abs=1e-12)
            assert b[0] > b[-1]
            if ai > 0:
                assert L.svpfo_adaptive_beta(0.2, ai, na) > L.svpfo_adaptive_beta(0.2, ai - 1, na)


def test_stein_budget_properties():
    """src/svpf_optimized_graph.cu:1097-1112: the maximum before tick 10; afterwards a
    non-decreasing integer in the previous KSD from the minimum (KSD <= 0.05) to the
    maximum (KSD >= 0.5)."""
    import ctypes

    c = oracle.make_cfg(config.make_config(stein_min_steps=4, stein_max_steps=12))
    st = oracle.State()

    def budget(ts, k):
        st.timestep, st.ksd_prev = ts, k
        return L.svpfo_stein_budget(ctypes.byref(c), ctypes.byref(st))

    assert {budget(ts, k) for ts in range(10) for k in (0.0, 0.3, 9.0)} == {12}
    ks = np.linspace(0.0, 0.7, 281)
    b = np.array([budget(20, k) for k in ks])
    assert (np.diff(b) >= 0).all() and b[0] == 4 and b[-1] == 12
    assert (b[ks <= 0.05] == 4).all() and (b[ks >= 0.5] == 12).all()
    assert set(b.tolist()) == set(range(4, 13))


# ---------------------------------------------------------------- smoothing
def test_smoothing_window_of_one_is_identity():
    s = oracle.Smoother(lag=1, output_lag=1)
    for t in range(10):
        assert s.step(t, -1.0 + 0.1 * t, 0.05, 0.2, -1.0, 0.97, 0.15) == -1.0 + 0.1 * t


def test_smoothing_output_lag_zero_is_raw():
    s = oracle.Smoother(lag=4, output_lag=0)
    for t in range(10):
        assert s.step(t, math.sin(t), 0.05, 0.2, -1.0, 0.97, 0.15) == math.sin(t)


def test_smoothing_two_tick_window_is_gaussian_conditioning():
    """Window 2, output lag 1: once the window is full each output is the previous raw
    estimate conditioned on the newest one under the AR(1) link -- the Gaussian
    conditional mean of h_{t-1} ~ N(m, v) given h_t = mu + rho (h_{t-1} - mu) + sigma eps."""
    mu, rho, sig = -1.0, 0.95, 0.2
    s = oracle.Smoother(lag=2, output_lag=1)
    rng = np.random.default_rng(0)
    m = rng.normal(-1.0, 0.3, 12)
    v = rng.uniform(0.01, 0.2, 12)
    for t in range(12):
        out = s.step(t, m[t], v[t], 0.1, mu, rho, sig)
        if t < 2:
            assert out == m[t]
            continue
        C = np.array([[v[t - 1], rho * v[t - 1]], [rho * v[t - 1], rho * rho * v[t - 1] + sig * sig]])
        mean = np.array([m[t - 1], mu + rho * (m[t - 1] - mu)])
        cond_mean = mean[0] + C[0, 1] / C[1, 1] * (m[t] - mean[1])
        cond_var = C[0, 0] - C[0, 1] ** 2 / C[1, 1]
        assert out == pytest.approx(cond_mean, abs=1e-7)
        i = (s.s.head - 2) % 2
        assert s.s.h_var[i] == pytest.approx(cond_var, abs=1e-7)


# ---------------------------------------------------------------- svpf_initialize
def test_initial_particles_compose_the_generator_and_the_normal_transform():
    n, mu, rho, sig, seed = 64, -1.0, 0.97, 0.15, 0x1234_5678_9ABC
    h = oracle.init_particles(n, mu, rho, sig, seed)
    sd = math.sqrt(sig * sig / (1 - rho * rho + 1e-6))
    for i in range(n):
        o = oracle.philox([i, 0, 255, 0], [seed & 0xFFFFFFFF, seed >> 32])
        z = L.svpfo_normal_from_bits(int(o[0]), int(o[1]))
        assert h[i] == pytest.approx(min(max(mu + sd * z, -15.0), 5.0), abs=1e-12)


def test_initial_particles_follow_the_stationary_law_and_clamp():
    n = 200000
    h = oracle.init_particles(n, -1.0, 0.9, 0.3, 7)
    sd = math.sqrt(0.09 / (1 - 0.81 + 1e-6))
    assert scipy.stats.kstest((h + 1.0) / sd, "norm").pvalue > 1e-3
    ht = oracle.init_particles(n, -1.0, 0.9, 0.3, 7, student_t_state=1, nu_state=5.0)
    assert np.var(ht) / np.var(h) == pytest.approx(5.0 / 3.0, rel=1e-9)  # same normals, scaled
    hc = oracle.init_particles(5000, 4.9, 0.5, 2.0, 3)
    assert hc.max() == 5.0 and (hc <= 5.0).all() and (hc == 5.0).sum() > 100


# ---------------------------------------------------------------- two-factor step
def cfg2f(n, **kw):
    c = dict(n=n, n_stein_steps=1, n_anneal_steps=1, use_annealing=1, use_svld=0, use_mim=0, use_guide=0,
             use_guide_preserving=1, use_exact_gradient=1, stein_sign=0, mu=-1.0, nu=5.0, rho_fast=0.8,
             sigma_fast=0.3, rho_slow=0.995, sigma_slow=0.08, bw_floor_fast=0.01, bw_floor_slow=0.05,
             guide_strength=0.12, mim_jump_prob=0.25, mim_jump_scale=9.0, temperature=0.45, rmsprop_rho=0.9,
             rmsprop_eps=1e-6)
    c.update(kw)
    return c


def normal_at(seed, filt, tick_, purpose, i, words=(0, 1)):
    o = oracle.philox([i, tick_, purpose, filt], [seed & 0xFFFFFFFF, seed >> 32])
    return L.svpfo_normal_from_bits(int(o[words[0]]), int(o[words[1]])), o


def test_two_factor_ekf_is_the_exact_kalman_update_of_the_linear_model():
    """The innovation split (include/svpf_two_factor.cuh:245-300) equals the Kalman update
    of x = (f, s) with prior N(m, diag(Pf, Ps)) and observation log y^2 + offset =
    mu + f + s + noise (variance 4.93 + 2), in means and marginal variances."""
    c = cfg2f(4)
    o = oracle.TwoFactorOracle(c, np.zeros(4), np.zeros(4), guide_2f_initialized=1, guide_mean_fast=0.2,
                               guide_mean_slow=-0.1, guide_var_fast=0.05, guide_var_slow=0.02)
    y = 0.7
    o.ekf(y)
    m = np.array([0.8 * 0.2, 0.995 * -0.1])
    P = np.diag([0.64 * 0.05 + 0.09, 0.995 ** 2 * 0.02 + 0.0064])
    H = np.array([[1.0, 1.0]])
    z = math.log(y * y + 1e-8) + L.svpfo_implied_offset(5.0) - (-1.0)
    S = H @ P @ H.T + 6.93
    K = P @ H.T / S
    post_m = m + (K * (z - H @ m)).ravel()
    post_P = P - K @ H @ P
    assert o.st.guide_mean_fast == pytest.approx(post_m[0], abs=1e-9)
    assert o.st.guide_mean_slow == pytest.approx(post_m[1], abs=1e-9)
    assert o.st.guide_var_fast == pytest.approx(post_P[0, 0], abs=1e-9)
    assert o.st.guide_var_slow == pytest.approx(post_P[1, 1], abs=1e-9)
    # the first call starts from the stationary laws of the mean-zero factors
    o2 = oracle.TwoFactorOracle(c, np.zeros(4), np.zeros(4))
    o2.ekf(y)
    P0 = np.diag([0.09 / (1 - 0.64) * 0.64 + 0.09, 0.0064 / (1 - 0.995 ** 2) * 0.995 ** 2 + 0.0064])
    K0 = P0 @ H.T / (H @ P0 @ H.T + 6.93)
    assert o2.st.guide_mean_fast == pytest.approx(K0[0, 0] * z, abs=1e-9)


def test_two_factor_single_particle_tick_assembled_by_hand():
    """N = 1: K(0) = 1 and no repulsion, so one Stein iteration moves each factor by
    eff g / sqrt(v' + eps) with g the factor's AR(1) prior score plus its share of the
    likelihood score (the derivative of scipy's Student-t log density, by central
    differences), shares sigma_f^2 : sigma_s^2."""
    seed, y = 99, 0.6
    c = cfg2f(1, n_anneal_steps=2, n_stein_steps=1)
    hf0, hs0 = np.array([0.3]), np.array([-0.2])
    o = oracle.TwoFactorOracle(c, hf0, hs0)
    out = o.tick(y, seed=seed)
    zf, _ = normal_at(seed, 0, 0, 0, 0)
    zs, _ = normal_at(seed, 0, 0, 2, 0)
    f, s = 0.8 * 0.3 + 0.3 * zf, 0.995 * -0.2 + 0.08 * zs
    fp, sp = 0.3, -0.2

    def loglik(h):
        return scipy.stats.t.logpdf(y * math.exp(-h / 2), 5.0) - h / 2

    s2f, s2s = 0.09 + 1e-8, 0.0064 + 1e-8
    wf = s2f / (s2f + s2s)
    vf = vs = 0.0
    for a in range(2):
        beta = (a + 1) / 2
        h = -1.0 + f + s
        e = 1e-5
        gl = (loglik(h + e) - loglik(h - e)) / (2 * e)
        gf = -(f - 0.8 * fp) / s2f + beta * wf * gl
        gs = -(s - 0.995 * sp) / s2s + beta * (1 - wf) * gl
        eff = 0.5 / 1 * beta
        vf = 0.9 * vf + 0.1 * gf * gf
        vs = 0.9 * vs + 0.1 * gs * gs
        f = f + eff * gf / math.sqrt(vf + 1e-6)
        s = s + eff * gs / math.sqrt(vs + 1e-6)
    assert o.hf[0] == pytest.approx(f, abs=1e-6)
    assert o.hs[0] == pytest.approx(s, abs=1e-6)
    assert out["steps"] == 2
    assert out["h_mean"] == pytest.approx(-1.0 + f + s, abs=1e-6)


def test_two_factor_predict_mim_and_outputs():
    """Predict: AR(1) of each mean-zero factor from the counter generator's normals, the MIM
    jump on the fast factor only; outputs: h = mu + f + s, vol = mean exp(h/2), the
    Silverman bandwidths with their floors, and the log-likelihood at the positions of
    the last gradient evaluation (here the predicted ones) against scipy."""
    n, seed, y = 48, 5, -0.4
    rng = np.random.default_rng(8)
    hf0, hs0 = rng.normal(0, 0.4, n), rng.normal(0, 0.3, n)
    c = cfg2f(n, use_mim=1, mim_jump_prob=0.5, mim_jump_scale=9.0, n_stein_steps=1, n_anneal_steps=1,
              bw_floor_slow=0.5)
    o = oracle.TwoFactorOracle(c, hf0, hs0)
    out = o.tick(y, seed=seed)
    f, s = np.zeros(n), np.zeros(n)
    jumps = 0
    for i in range(n):
        zf, of = normal_at(seed, 0, 0, 0, i)
        u = L.svpfo_u01_from_bits(int(of[2]))
        if u < 0.5:
            zf *= 9.0
            jumps += 1
        zs, _ = normal_at(seed, 0, 0, 2, i)
        f[i], s[i] = min(max(0.8 * hf0[i] + 0.3 * zf, -15.0), 5.0), 0.995 * hs0[i] + 0.08 * zs  # clamped
    assert 10 < jumps < 38
    assert np.array_equal(o.hfp, hf0) and np.array_equal(o.hsp, hs0)  # the previous values, for the prior
    h = -1.0 + f + s
    ll = scipy.special.logsumexp(scipy.stats.t.logpdf(y * np.exp(-h / 2), 5.0) - h / 2) - math.log(n)
    assert out["pred_loglik"] == pytest.approx(ll, abs=1e-7)
    assert out["loglik"] == pytest.approx(ll, abs=1e-7)
    assert out["bw_fast"] == pytest.approx(max(1.06 * np.std(f) * n ** -0.2, 0.01), rel=1e-9)
    assert out["bw_slow"] == pytest.approx(max(1.06 * np.std(s) * n ** -0.2, 0.5), rel=1e-9)
    assert out["h_mean"] == pytest.approx(-1.0 + o.hf.mean() + o.hs.mean(), abs=1e-12)
    assert out["vol"] == pytest.approx(np.mean(np.exp(0.5 * (-1.0 + o.hf + o.hs))), abs=1e-12)


def hand_2f(c, hf0, hs0, y, seed, shift=(0.0, 0.0)):
    """Brute-force two-factor tick for tiny N from separately pinned pieces: the
    counter generator's normals, the likelihood score by central differences of
    scipy's Student-t, numpy's standard deviation, and the IMQ operator written out
    pair by pair (no noise, no MIM; `shift` moves each predicted cloud, as a
    variance-preserving guide of strength 1 does)."""
    n = len(hf0)
    f = np.array([0.8 * hf0[i] + 0.3 * normal_at(seed, 0, 0, 0, i)[0] for i in range(n)]) + shift[0]
    s = np.array([0.995 * hs0[i] + 0.08 * normal_at(seed, 0, 0, 2, i)[0] for i in range(n)]) + shift[1]
    nu, mu = c["nu"], c["mu"]

    def score(h):
        lp = lambda x: scipy.stats.t.logpdf(y * math.exp(-x / 2), nu) - x / 2
        return (lp(h + 1e-5) - lp(h - 1e-5)) / 2e-5

    bwf = max(1.06 * np.std(f) * n ** -0.2, c["bw_floor_fast"])
    bws = max(1.06 * np.std(s) * n ** -0.2, c["bw_floor_slow"])
    s2f, s2s = c["sigma_fast"] ** 2 + 1e-8, c["sigma_slow"] ** 2 + 1e-8
    wf = s2f / (s2f + s2s)
    sgn = 1.0 if c["stein_sign"] == 1 else -1.0
    vf, vs = np.zeros(n), np.zeros(n)
    A = c["n_anneal_steps"]
    for a in range(A):
        beta = (a + 1) / A
        for _ in range(c["n_stein_steps"]):
            gl = np.array([score(mu + f[i] + s[i]) for i in range(n)])
            gf = -(f - 0.8 * hf0) / s2f + beta * wf * gl
            gs = -(s - 0.995 * hs0) / s2s + beta * (1 - wf) * gl
            eff = 0.5 / c["n_stein_steps"] * beta
            new = []
            for x, g, v, bw in ((f, gf, vf, bwf), (s, gs, vs, bws)):
                phi = np.zeros(n)
                for i in range(n):
                    for j in range(n):
                        K = 1.0 / (1.0 + (x[i] - x[j]) ** 2 / bw ** 2)
                        phi[i] += K * g[j] + sgn * 2 * (x[i] - x[j]) / bw ** 2 * K * K
                phi /= n
                v[:] = 0.9 * v + 0.1 * phi * phi
                new.append(x + eff * phi / np.sqrt(v + 1e-6))
            f, s = new
    return f, s


@pytest.mark.parametrize("sign", [0, 1])
@pytest.mark.parametrize("n", [2, 3])
def test_two_factor_small_tick_assembled_by_hand(n, sign):
    seed, y = 21, -0.5
    c = cfg2f(n, stein_sign=sign, n_anneal_steps=2, n_stein_steps=2, bw_floor_fast=0.05, bw_floor_slow=0.02)
    rng = np.random.default_rng(n + 10 * sign)
    hf0, hs0 = rng.normal(0, 0.4, n), rng.normal(0, 0.1, n)
    o = oracle.TwoFactorOracle(c, hf0, hs0)
    o.tick(y, seed=seed)
    f, s = hand_2f(c, hf0, hs0, y, seed)
    assert np.allclose(o.hf, f, atol=1e-6)
    assert np.allclose(o.hs, s, atol=1e-6)


def test_two_factor_guide_uses_the_previous_factor_means():
    """Guide-preserving with strength 1 moves each factor's cloud by (EKF mean - the factor
    mean after the previous tick) and keeps its deviations; strength 0 leaves it."""
    n, seed = 16, 4
    rng = np.random.default_rng(9)
    hf0, hs0 = rng.normal(0.1, 0.4, n), rng.normal(-0.05, 0.1, n)
    base = dict(n_stein_steps=1, use_guide=1, use_guide_preserving=1)
    prev = dict(mean_fast_prev=0.25, mean_slow_prev=-0.1)
    o0 = oracle.TwoFactorOracle(cfg2f(n, guide_strength=0.0, **base), hf0, hs0, **prev)
    o1 = oracle.TwoFactorOracle(cfg2f(n, guide_strength=1.0, **base), hf0, hs0, **prev)
    r = oracle.TwoFactorOracle(cfg2f(n, **dict(base, use_guide=0)), hf0, hs0, **prev)
    out0 = o0.tick(0.3, seed=seed)
    o1.tick(0.3, seed=seed)
    outr = r.tick(0.3, seed=seed)
    assert out0["pred_loglik"] == outr["pred_loglik"]
    assert np.allclose(o0.hf, r.hf, atol=1e-12) and np.allclose(o0.hs, r.hs, atol=1e-12)
    # strength 1: each predicted cloud moved by (EKF mean - the factor's previous mean)
    shift = (o1.st.guide_mean_fast - 0.25, o1.st.guide_mean_slow + 0.1)
    assert abs(shift[0]) > 1e-3 and abs(shift[1]) > 1e-4
    f, s_ = hand_2f(cfg2f(n, **base), hf0, hs0, 0.3, seed, shift=shift)
    assert np.allclose(o1.hf, f, atol=1e-6) and np.allclose(o1.hs, s_, atol=1e-6)
