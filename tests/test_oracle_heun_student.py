"""Pins of the oracle's Heun step and Student-t state option against checks
that do not restate its formulas: finite differences of the Student-t mixture
log-density, the second derivative of a single Student-t component, a
single-particle Heun step assembled from the oracle's (separately pinned)
gradient, and the distribution of the Student-t draws."""
import math

import numpy as np
import pytest
from scipy import stats

import oracle
from svpf_amd import config


def _mix_logpdf(h, m, nu_s, s2):
    l = -0.5 * (nu_s + 1.0) * np.log1p((h - m) ** 2 / (nu_s * s2))
    mx = l.max()
    return mx + math.log(np.exp(l - mx).sum())


@pytest.mark.parametrize("nu_s", [3.0, 5.0, 12.0])
def test_student_t_prior_gradient_is_the_mixture_score(nu_s):
    """g_prior(h_j) = d/dh log sum_i p_t(h | m_i) (beta = 0 leaves the prior alone)."""
    n = 7
    rng = np.random.default_rng(3)
    cfg = config.plain_config(n=n, mu=-1.0, rho=0.95, sigma_z=0.3, gamma=0.0,
                              use_student_t_state=1, nu_state=nu_s)
    hp = rng.standard_normal(n) * 0.4 - 1.0
    h = rng.standard_normal(n) * 0.4 - 1.0
    g, _, _, _ = oracle.gradient(cfg, h, hp, 0.3, 0.1, -1.0, 0.3, 0.0)
    m = -1.0 + 0.95 * (hp + 1.0)
    eps = 1e-6
    for j in range(n):
        fd = (_mix_logpdf(h[j] + eps, m, nu_s, 0.09) - _mix_logpdf(h[j] - eps, m, nu_s, 0.09)) / (2 * eps)
        assert abs(g[j] - fd) <= 1e-6 * max(1.0, abs(fd)), (j, g[j], fd)


def test_student_t_prior_hessian_of_one_component():
    """With one component the prior Hessian is d^2/dh^2 log p_t(h | m); y = 0
    removes the likelihood's curvature, so curv = -d^2 log p_t."""
    nu_s, s2 = 5.0, 0.3 ** 2
    cfg = config.plain_config(n=1, mu=-1.0, rho=0.95, sigma_z=0.3, gamma=0.0, use_newton=1,
                              use_exact_gradient=1, use_student_t_state=1, nu_state=nu_s)
    for hv, hpv in [(-0.8, -1.1), (-1.0, -1.0), (-1.9, -0.7)]:
        _, _, _, cv = oracle.gradient(cfg, np.array([hv]), np.array([hpv]), 0.0, 0.0, -1.0, 0.3, 1.0)
        m = -1.0 + 0.95 * (hpv + 1.0)
        f = lambda x: -0.5 * (nu_s + 1.0) * math.log1p((x - m) ** 2 / (nu_s * s2))
        eps = 1e-4
        d2 = (f(hv + eps) - 2 * f(hv) + f(hv - eps)) / eps ** 2
        want = min(max(-d2, 0.1), 100.0)
        assert abs(cv[0] - want) <= 1e-5 * max(1.0, abs(want)), (hv, cv[0], want)


def test_student_t_draws_follow_the_t_distribution():
    lib = oracle.lib()
    nu = 5.0
    draws = np.array([lib.svpfo_student_t_draw(7, 0, 3, i, nu) for i in range(200000)])
    assert abs(draws.mean()) < 0.02
    tail = np.mean(np.abs(draws) > 3.0)
    assert abs(tail - 2 * stats.t.sf(3.0, nu)) < 0.003
    assert abs(np.mean(np.abs(draws) > 1.0) - 2 * stats.t.sf(1.0, nu)) < 0.005
    # above 30 degrees of freedom the draw is the predict normal itself
    for i in range(5):
        o = oracle.philox([i, 3, 0, 0], [7, 0])
        assert lib.svpfo_student_t_draw(7, 0, 3, i, 40.0) == lib.svpfo_normal_from_bits(int(o[0]), int(o[1]))


def test_heun_single_particle_step():
    """n = 1: the Stein operator is the score (K = 1, no repulsion), so a Heun
    iteration is h~ = h + eps g(h)/sqrt(v + e), v' = rho v + (1 - rho) p^2,
    h' = h + eps p / sqrt(v' + e) with p = (g(h) + g(h~)) / 2; the KSD of the
    last iteration is sqrt(max(g(h~)^2 - 2/bw^2, 0))."""
    sz = 0.8  # keeps the scores inside the +-10 clamp, so every iteration moves the particle
    cfg = config.plain_config(n=1, n_stein=3, mu=-1.0, rho=0.97, sigma_z=sz, nu=5.0,
                              use_heun=1, temperature=0.0, use_svld=0)
    h0, y_t = np.array([-0.9]), 0.9
    z = float(np.float32(0.2))
    # RMSProp state 1 (from v = 0 the first predictor step is scaled by 1/sqrt(1e-6) and clamps)
    f = oracle.OracleFilter(cfg, h0, v0=np.array([1.0]))
    out = f.tick(y_t, 0.0, z_pred=np.array([z], np.float32), xi=np.zeros((3, 1), np.float32))
    h = float(np.clip(-1.0 + 0.97 * (h0[0] + 1.0) + sz * z, -15, 5))
    hp = h0.copy()
    v = 1.0
    S = cfg["stein_max_steps"]
    for k in range(S):
        beta = 0.3 + 0.7 * k / (S - 1)
        eff = 0.1 * math.sqrt(beta)
        g1 = oracle.gradient(cfg, np.array([h]), hp, y_t, 0.0, -1.0, sz, beta)[0][0]
        ht = min(max(h + eff * g1 / math.sqrt(v + 1e-6), -15.0), 5.0)
        g2 = oracle.gradient(cfg, np.array([ht]), hp, y_t, 0.0, -1.0, sz, beta)[0][0]
        assert abs(g1) < 10.0 and abs(g2) < 10.0 and g1 != g2
        p = 0.5 * (g1 + g2)
        v = 0.9 * v + 0.1 * p * p
        h = min(max(h + eff * p / math.sqrt(v + 1e-6), -15.0), 5.0)
    assert abs(f.h[0] - h) <= 1e-12
    assert abs(f.v[0] - v) <= 1e-12
    assert out["steps"] == S
    assert abs(out["ksd"] - math.sqrt(max(g2 * g2 - 2.0 / out["bandwidth"] ** 2, 0.0))) <= 1e-12
