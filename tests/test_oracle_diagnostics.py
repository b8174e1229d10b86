"""Pins of the diagnostics oracle (oracle/diagnostics.py) by what the
mathematics fixes: scipy's digamma and log-densities, ESS closed forms and a
hand-traced run of the shock state machine."""
import numpy as np
import pytest
from scipy import special, stats

from oracle import diagnostics as od


def test_digamma_matches_scipy():
    x = np.concatenate([np.linspace(0.05, 6.5, 400), np.linspace(6.5, 60.0, 200)])
    np.testing.assert_allclose(od.digamma(x), special.digamma(x), rtol=0, atol=5e-9)


@pytest.mark.parametrize("nu", [2.5, 4.0, 5.0, 12.0, 30.0])
def test_nu_score_is_the_derivative_of_the_student_t_log_density(nu):
    rng = np.random.default_rng(1)
    h = rng.normal(-1.0, 1.0, 64)
    y = 0.07
    g, z2 = od.nu_score_particles(h, y, nu)
    z = y / (np.exp(0.5 * h) + 1e-8)
    # log p(y | h, nu) = log t_nu(z) - h/2 (the Jacobian of y = exp(h/2) z does not depend on nu)
    f = lambda v: stats.t.logpdf(z, v) - 0.5 * h
    d = 1e-5
    fd = (f(nu + d) - f(nu - d)) / (2 * d)
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-8)
    np.testing.assert_allclose(z2, z * z, rtol=1e-15)


def test_sigma_score_is_the_derivative_of_the_transition_log_density():
    rng = np.random.default_rng(2)
    hp = rng.normal(-1.0, 0.5, 64)
    mu, rho, sigma = -1.0, 0.97, 0.15
    h = mu + rho * (hp - mu) + rng.normal(0.0, 0.2, 64)
    g, e2 = od.sigma_score_particles(h, hp, mu, rho, sigma)
    eps = h - mu - rho * (hp - mu)
    f = lambda s: stats.norm.logpdf(eps, 0.0, s)
    d = 1e-6
    np.testing.assert_allclose(g, (f(sigma + d) - f(sigma - d)) / (2 * d), rtol=1e-6)
    np.testing.assert_allclose(e2, (eps / sigma) ** 2, rtol=1e-14)


def test_weighted_means_and_ess_closed_forms():
    rng = np.random.default_rng(3)
    n = 300
    h = rng.normal(-1.0, 0.3, n)
    # equal weights: plain means, ESS = N
    lw = np.full(n, -7.0)
    g, z2 = od.nu_score(h, lw, 0.05, 5.0)
    gp, zp = od.nu_score_particles(h, 0.05, 5.0)
    assert abs(g - gp.mean()) < 1e-9 and abs(z2 - zp.mean()) < 1e-9
    st, m = od.health_check(h, lw)
    assert st == "OK" and abs(m["ess"] - n) < 1e-9 and abs(m["h_mean"] - h.mean()) < 1e-12
    # one live weight: ESS = 1, and the score is that particle's
    lw = np.full(n, -np.inf)
    lw[17] = 3.0
    st, m = od.health_check(h, lw)
    assert st == "ESS_LOW" and abs(m["ess"] - 1.0) < 1e-12
    g, _ = od.nu_score(h, lw, 0.05, 5.0)
    assert abs(g - gp[17] / (1 + 1e-8)) < 1e-12
    # two log-weights log 1 and log 3: ESS = 16 / 10
    st, m = od.health_check(h[:2], np.log([1.0, 3.0]))
    assert abs(m["ess"] - 1.6) < 1e-12


def test_health_decisions():
    h = np.full(8, -1.0)
    assert od.health_check(np.r_[h, np.nan], np.zeros(9))[0] == "NAN_IN_PARTICLES"
    assert od.health_check(h, np.r_[np.zeros(7), np.nan])[0] == "NAN_IN_PARTICLES"
    assert od.health_check(np.r_[h, 2e10], np.zeros(9))[0] == "INF_IN_PARTICLES"
    assert od.health_check(h, np.full(8, -np.inf))[0] == "WEIGHT_UNDERFLOW"
    assert od.health_check(np.full(8, 5.0), np.zeros(8))[0] == "VOL_EXTREME"  # exp(2.5) = 12.2 > 10
    assert od.health_check(np.full(8, -19.0), np.zeros(8))[0] == "VOL_EXTREME"  # exp(-9.5) < 1e-4
    assert od.check_input(float("nan")) == "INPUT_IS_NAN"
    assert od.check_input(float("inf")) == "INPUT_IS_INF"
    assert od.check_input(-0.6) == "INPUT_EXTREME_VALUE"
    assert od.check_input(0.5) == "OK"


def test_shock_machine_hand_traced():
    s = od.ShockState(threshold=9.0, shock_duration=3, recovery_duration=4, exit_threshold=4.0)
    seq = [1.0, 10.0, 0.0, 0.0, 0.0, 5.0, 5.0, 5.0, 5.0, 1.0, 12.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0]
    want = [0, 1, 1, 1, 2, 2, 2, 2, 2, 0, 1, 1, 1, 2, 2, 2, 2, 0]
    assert [s.update(z) for z in seq] == want
    assert s.lr_multiplier(2.0) == 1.0
