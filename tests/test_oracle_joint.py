"""Pins of the oracle's joint state-parameter filter (oracle/svpf_oracle.c,
svpfo_joint_*; the reference's svpf_joint/src/svpf_joint_kernels.cu and
svpf_joint/src/svpf_joint.cu) by checks that do not restate its code:

- the score in the calm regime is the gradient of an explicit log density
  (central differences), which fixes every sign, chain-rule factor and constant of
  the transition, likelihood-surrogate, prior and mean-reversion terms;
- the log weight is scipy's Student-t log density with the log-volatility Jacobian;
- the surprise branches: continuity at their thresholds and saturation points;
- one particle: the kernel mass floor of 32 and the boosted step sizes;
- many particles: the repulsion from torch autograd of the 4-D IMQ kernel;
- the diagnostics from numpy and scipy on the returned particles;
- the initial draw and the predict law by their moments;
- the reference tests' own behaviour checks (test_svpf_joint.cu) on their series.
"""
import math

import numpy as np
import pytest
import torch
from scipy import special, stats

import oracle
from joint_series import crash_series, default_oracle_cfg, sv_series


def cfg(n=64, S=3, **kw):
    return default_oracle_cfg(n, S, **kw)


def grad_of(c, x, hp, y):
    return oracle.JointOracle(c, x0=np.zeros((4, c["n"]))).grad(x, hp, y)


def calm_log_density(c, x, hp, y):
    """Log density whose gradient is the calm-regime score (z^2 in (1.5, 4), sigma >= 0.03)."""
    h, mt, rt, st = x
    rho = 1.0 / (1.0 + math.exp(-rt))
    var = math.exp(2.0 * st)
    diff = h - (mt + rho * (hp - mt))
    lik = -((math.log(y * y + 1e-10) - h + c["lik_offset"]) ** 2) / (2.0 * 1.4)
    trans = -diff * diff / (2.0 * var) - st
    mr = -0.1 * (st - c["sigma_prior_mean"]) ** 2
    prior = -c["prior_weight"] * ((mt - c["mu_prior_mean"]) ** 2 / (2 * c["mu_prior_var"]) +
                                  (rt - c["rho_prior_mean"]) ** 2 / (2 * c["rho_prior_var"]) +
                                  (st - c["sigma_prior_mean"]) ** 2 / (2 * c["sigma_prior_var"]))
    return lik + trans + mr + prior


@pytest.mark.parametrize("seed", range(6))
def test_calm_score_is_gradient_of_log_density(seed):
    rng = np.random.default_rng(seed)
    c = cfg()
    hp = rng.uniform(-5, -2)
    y = math.sqrt(rng.uniform(1.6, 3.9) * math.exp(hp))  # z^2 against h_prev in (1.5, 4)
    x = np.array([hp + rng.normal(0, 0.1), rng.uniform(-4.5, -2.5), rng.uniform(0.5, 3.0), rng.uniform(-1.5, -0.8)])
    g, _ = grad_of(c, x, hp, y)
    assert abs(g[0]) < 10 and np.all(np.abs(g[1:]) < 5), "clips must be inactive for this check"
    eps = 1e-6
    for q in range(4):
        xp, xm = x.copy(), x.copy()
        xp[q] += eps
        xm[q] -= eps
        fd = (calm_log_density(c, xp, hp, y) - calm_log_density(c, xm, hp, y)) / (2 * eps)
        assert g[q] == pytest.approx(fd, rel=1e-6, abs=1e-6), q


def test_calm_check_fails_a_planted_sign_error():
    """The density check above would catch a flipped rho chain-rule factor."""
    c = cfg()
    hp, y = -3.0, math.sqrt(2.5 * math.exp(-3.0))
    x = np.array([-2.7, -3.4, 1.5, -1.2])
    g, _ = grad_of(c, x, hp, y)
    eps = 1e-6
    xp, xm = x.copy(), x.copy()
    xp[2] += eps
    xm[2] -= eps
    fd = (calm_log_density(c, xp, hp, y) - calm_log_density(c, xm, hp, y)) / (2 * eps)
    planted = -(g[2] - c["prior_weight"] * -(x[2] - c["rho_prior_mean"]) / c["rho_prior_var"]) + \
        c["prior_weight"] * -(x[2] - c["rho_prior_mean"]) / c["rho_prior_var"]
    assert g[2] == pytest.approx(fd, rel=1e-6)
    assert planted != pytest.approx(fd, rel=1e-3)


@pytest.mark.parametrize("nu", [3.0, 30.0])
def test_log_weight_is_student_t_density(nu):
    c = cfg(nu=nu, student_t_const=math.lgamma((nu + 1) / 2) - math.lgamma(nu / 2) - 0.5 * math.log(math.pi * nu))
    rng = np.random.default_rng(1)
    for _ in range(20):
        h, y = rng.uniform(-8, 2), rng.normal(0, 0.2)
        _, lw = grad_of(c, np.array([h, -3.5, 2.0, -2.0]), h, y)
        # density of y = exp(h / 2) * t_nu: t.logpdf(y exp(-h/2)) - h/2, with the reference's
        # guard exp(h) + 1e-8 in the scale (svpf_joint_kernels.cu:187)
        ref = stats.t.logpdf(y / math.sqrt(math.exp(h) + 1e-8), nu) - h / 2
        assert lw == pytest.approx(ref, rel=1e-6, abs=1e-6)


def _score_at(c, z2, sigma_t=-2.0, h_off=0.3):
    hp = -3.0
    y = math.sqrt(z2 * (math.exp(hp) + 1e-8))  # z^2 = y^2 / (exp(h_prev) + 1e-8) exactly
    x = np.array([hp + h_off, -3.2, 2.0, sigma_t])
    return grad_of(c, x, hp, y)[0]


def test_surprise_branches_are_continuous_and_saturate():
    c = cfg()
    e = 1e-9
    # observation force on sigma: zero on both sides of 1.5 and 9; capped at 5 from z^2 = 19 on
    g15a, g15b = _score_at(c, 1.5 - e), _score_at(c, 1.5 + e)
    assert g15a[3] == pytest.approx(g15b[3], abs=1e-6)
    # the h-spring inflation starts at z^2 = 4 (continuous there) and saturates at z^2 = 24
    g4a, g4b = _score_at(c, 4 - e), _score_at(c, 4 + e)
    assert g4a[0] == pytest.approx(g4b[0], abs=1e-6)
    # above 9 the h spring uses max(var * inflate, 1) and weight 0.3: with sigma = 1.5 the
    # inflated variance exceeds 1, so the spring is 0.3 * diff / (var * inflate)
    def spring(z):  # the h score minus its likelihood-surrogate term
        lik = (math.log(z * (math.exp(-3.0) + 1e-8) + 1e-10) - (-2.7) + c["lik_offset"]) / 1.4
        return _score_at(c, z, sigma_t=math.log(1.5))[0] - lik

    s24, s30, s100, s20 = spring(24.0), spring(30.0), spring(100.0), spring(20.0)
    assert s24 == pytest.approx(s30, rel=1e-6) and s30 == pytest.approx(s100, rel=1e-6)
    assert s20 != pytest.approx(s24, rel=1e-3)
    # the sigma force grows with slope 0.5 above 9 until it reaches 5 at 19
    f = [_score_at(c, z, sigma_t=math.log(1.5))[3] for z in (10.0, 12.0, 20.0, 25.0)]
    assert f[1] - f[0] == pytest.approx(1.0, rel=1e-9)
    assert f[2] - f[0] == pytest.approx(4.5, rel=1e-9)  # 0.5 at z^2 = 10, the cap 5 at 20
    assert f[3] == pytest.approx(f[2], abs=1e-9)
    # the panic switch at 9 is a jump of the h spring: weight 0.3 and variance floor 1
    lo, hi = _score_at(c, 9 - e, h_off=0.3), _score_at(c, 9 + e, h_off=0.3)
    lik = (math.log(9 * (math.exp(-3.0) + 1e-8) + 1e-10) - (-2.7) + c["lik_offset"]) / 1.4
    var = math.exp(-4.0)
    diff_lo = lo[0] - lik  # -diff / (var * inflate(9))
    diff_hi = hi[0] - lik  # -0.3 * diff / 1
    assert diff_lo * var * (1 + 0.3 * 5) == pytest.approx(diff_hi / 0.3, rel=1e-6)


def test_sigma_clip_is_wider_in_panic():
    c = cfg()
    hp = -3.0
    x = np.array([hp + 3.0, -3.5, 2.0, -4.0])  # huge transition error for a tiny sigma
    calm = grad_of(c, x, hp, math.sqrt(2.0 * math.exp(hp)))[0]
    panic = grad_of(c, x, hp, math.sqrt(50.0 * math.exp(hp)))[0]
    assert calm[3] == 5.0 and panic[3] == 10.0


def _predicted(c, x0, y, seed, tick):
    """The particles after predict (a zero-iteration step from the same state and counters)."""
    o = oracle.JointOracle(dict(c, n_stein_steps=0), x0=x0)
    o.step(y, seed, tick)
    return o.x.copy(), o.hp.copy()


@pytest.mark.parametrize("z2", [2.0, 20.0])
def test_single_particle_uses_mass_floor_and_boosts(z2):
    c = cfg(n=1, S=3, diffusion_mu=0.0, diffusion_rho=0.0, diffusion_sigma=0.0)
    x0 = np.array([[-3.0], [-3.4], [2.0], [-1.5]])
    y = math.sqrt(z2 * math.exp(-3.0))
    xp, hp = _predicted(c, x0, y, 7, 3)
    boost = np.array([5.0, 2.0, 1.0, 4.0]) if z2 > 9 else np.ones(4)
    step = np.array([c["step_h"], c["step_mu"], c["step_rho"], c["step_sigma"]]) * boost
    x = xp[:, 0].copy()
    lo, hi = np.array([-15, -10, -5, -5.0]), np.array([5, 5, 5, 2.0])
    for _ in range(3):
        g, _ = grad_of(c, x, hp[0], y)
        x = np.clip(x + step * g / 32.0, lo, hi)  # the kernel of a particle with itself is 1 < 32
    o = oracle.JointOracle(c, x0=x0)
    o.step(y, 7, 3)
    np.testing.assert_allclose(o.x[:, 0], x, rtol=1e-12, atol=1e-12)


def test_repulsion_matches_autograd_of_imq_kernel():
    n = 40
    c = cfg(n=n, S=1, diffusion_mu=0.05, diffusion_rho=0.05, diffusion_sigma=0.05)
    rng = np.random.default_rng(3)
    x0 = np.stack([rng.normal(-3.3, 0.4, n), rng.normal(-3.5, 0.2, n), rng.normal(2.0, 0.2, n),
                   rng.normal(-2.0, 0.2, n)])
    y = 0.02
    xp, hp = _predicted(c, x0, y, 11, 5)
    o = oracle.JointOracle(c, x0=x0)
    out = o.step(y, 11, 5)
    bw = np.array(out["bw"])
    ib = 1.0 / (bw ** 2 + 1e-8)
    G = np.stack([grad_of(c, xp[:, j], hp[j], y)[0] for j in range(n)])  # [n, 4]
    X = torch.tensor(xp.T, dtype=torch.float64, requires_grad=True)  # [n, 4]
    Xj = X.detach()
    # K[i, j] = 1 / (1 + sum_q (x_iq - x_jq)^2 ib_q); grad wrt x_i of sum_j K[i, j] e_j
    D = X[:, None, :] - Xj[None, :, :]
    K = 1.0 / (1.0 + (D * D * torch.tensor(ib)).sum(-1))
    phi = torch.zeros(n, 4, dtype=torch.float64)
    for j in range(n):
        (gk,) = torch.autograd.grad(K[:, j].sum(), X, retain_graph=True)
        phi[j] = gk.sum(0)
    Kd = K.detach().numpy()
    phi = phi.numpy() + Kd.T @ G
    norm = np.clip(Kd.sum(0), 32.0, n)
    z2 = y * y / (np.exp(hp) + 1e-8)
    boost = np.where(z2[:, None] > 9, np.array([5.0, 2.0, 1.0, 4.0]), 1.0)
    step = np.array([c["step_h"], c["step_mu"], c["step_rho"], c["step_sigma"]]) * boost
    want = np.clip(xp.T + step * phi / norm[:, None], [-15, -10, -5, -5], [5, 5, 5, 2]).T
    np.testing.assert_allclose(o.x, want, rtol=1e-10, atol=1e-12)


def test_bandwidths_scale_and_clamp():
    n = 500
    c = cfg(n=n, S=0, diffusion_mu=0.0, diffusion_rho=0.0, diffusion_sigma=0.0)
    rng = np.random.default_rng(4)
    base = np.stack([np.clip(rng.normal(-5.0, 6.0, n), -15, 5), rng.normal(-3.5, 0.1, n), np.full(n, 2.0),
                     rng.normal(-2.0, 0.2, n)])
    out = oracle.JointOracle(c, x0=base).step(0.03, 1, 0)
    bw = np.array(out["bw"])
    assert bw[2] == pytest.approx(0.01)  # identical rho~: variance 0, floor of the clamp
    assert bw[0] == pytest.approx(2.0)  # wide h spread: ceiling of the clamp
    # mu~ is untouched by predict without diffusion: bandwidth ~ spread, doubling with it
    wide = base.copy()
    wide[1] = -3.5 + 2.0 * (base[1] + 3.5)
    bw2 = np.array(oracle.JointOracle(c, x0=wide).step(0.03, 1, 0)["bw"])
    assert bw2[1] == pytest.approx(2.0 * bw[1], rel=1e-12)
    assert bw[1] == pytest.approx(math.sqrt(2.0 * np.var(base[1]) / math.log(n + 1)), rel=1e-12)


def test_diagnostics_from_numpy():
    n = 300
    c = cfg(n=n, S=1)
    o0 = oracle.JointOracle(c, seed=5)
    x0 = o0.x.copy()
    y = 0.05
    xp, hp = _predicted(c, x0, y, 5, 0)
    o = oracle.JointOracle(c, x0=x0)
    d = o.step(y, 5, 0)
    x = o.x
    rho, sig = special.expit(x[2]), np.exp(x[3])
    assert d["h_mean"] == pytest.approx(x[0].mean(), rel=1e-12)
    assert d["vol_mean"] == pytest.approx(math.exp(0.5 * x[0].mean()), rel=1e-12)
    assert (d["mu_mean"], d["rho_mean"], d["sigma_mean"]) == pytest.approx((x[1].mean(), rho.mean(), sig.mean()))
    assert (d["mu_std"], d["rho_std"], d["sigma_std"]) == pytest.approx((x[1].std(), rho.std(), sig.std()), rel=1e-9)
    assert (d["std_mu_tilde"], d["std_rho_tilde"], d["std_sigma_tilde"]) == pytest.approx(
        (x[1].std(), x[2].std(), x[3].std()), rel=1e-9)
    # the weights are those of the last score evaluation: before the single iteration
    lw = np.array([grad_of(c, xp[:, j], hp[j], y)[1] for j in range(n)])
    w = np.exp(lw - lw.max())
    assert d["ess"] == pytest.approx(1.0 / np.sum((w / w.sum()) ** 2), rel=1e-9)
    assert d["log_likelihood"] == pytest.approx(special.logsumexp(lw) - math.log(n), rel=1e-12)
    flags = oracle.JointOracle(dict(c, collapse_thresh_mu=10.0, collapse_thresh_rho=0.0,
                                    collapse_thresh_sigma=10.0), x0=x0).step(y, 5, 0)
    assert (flags["mu_collapsed"], flags["rho_collapsed"], flags["sigma_collapsed"]) == (1, 0, 1)


def test_initial_draw_moments():
    n = 40000
    o = oracle.JointOracle(cfg(n=n), seed=12345)
    m, s = o.x.mean(1), o.x.std(1)
    np.testing.assert_allclose(m, [-3.5, -3.5, 2.0, -2.0], atol=0.02)
    np.testing.assert_allclose(s, [0.5, 0.2, 0.2, 0.2], rtol=0.02)
    assert np.abs(np.corrcoef(o.x) - np.eye(4)).max() < 0.03
    np.testing.assert_array_equal(o.hp, o.x[0])


def test_predict_uses_the_diffused_parameters():
    n = 40000
    c = cfg(n=n, S=0, diffusion_mu=1.0, diffusion_rho=0.3, diffusion_sigma=0.1)
    x0 = np.stack([np.full(n, -2.0), np.full(n, -3.5), np.full(n, 2.0), np.full(n, -1.0)])
    o = oracle.JointOracle(c, x0=x0)
    o.step(0.03, 9, 4)
    np.testing.assert_allclose(o.x[1:].std(1), [1.0, 0.3, 0.1], rtol=0.02)
    np.testing.assert_array_equal(o.hp, x0[0])
    # h = mu~ + rho (h - mu~) + sigma z with the diffused mu~: its variance carries (1 - rho)^2 var(mu~)
    rho = 1 / (1 + math.exp(-2.0))
    sig2 = np.mean(np.exp(2 * o.x[3]))
    assert o.x[0].var() == pytest.approx((1 - rho) ** 2 * 1.0 + sig2, rel=0.05)
    assert o.x[0].mean() == pytest.approx(-3.5 + rho * (-2.0 + 3.5), abs=0.02)


def test_zero_iterations_weights_the_predicted_particles():
    n = 64
    c = cfg(n=n, S=0)
    o = oracle.JointOracle(c, seed=2)
    d = o.step(0.01, 2, 0)
    lw = np.array([grad_of(c, o.x[:, j], o.hp[j], 0.01)[1] for j in range(n)])
    assert d["log_likelihood"] == pytest.approx(special.logsumexp(lw) - math.log(n), rel=1e-12)


def test_reference_basic():
    """test_svpf_joint.cu:66-105: 256 particles, 3 iterations, ten observations of 0.01."""
    o = oracle.JointOracle(cfg(n=256, S=3), seed=12345)
    for t in range(10):
        d = o.step(0.01, 12345, t)
    assert all(np.isfinite(v) for k, v in d.items() if k != "bw")
    assert 1.0 <= d["ess"] <= 256.0


def test_reference_calm_tracking():
    """test_svpf_joint.cu:112-177: 512 particles, 5 iterations, 500 calm observations; the
    reference warns when |rho_hat - 0.95| > 0.15."""
    y, h = sv_series(500, -3.5, 0.95, 0.15, 42)
    o = oracle.JointOracle(cfg(n=512, S=5), seed=12345)
    for t in range(500):
        d = o.step(float(y[t]), 12345, t)
    assert abs(d["rho_mean"] - 0.95) <= 0.15
    assert abs(d["h_mean"] - h[-1]) < 1.5


def test_reference_crash_adaptation():
    """test_svpf_joint.cu:184-289: sigma's peak in the first 20 crash steps against its value
    before the crash; the reference reports SUCCESS above 1.5x."""
    y, h, T_calm = crash_series()
    o = oracle.JointOracle(cfg(n=512, S=5), seed=12345)
    peak, before = 0.0, None
    for t in range(len(y)):
        d = o.step(float(y[t]), 12345, t)
        if t == T_calm - 1:
            before = d["sigma_mean"]
        if T_calm <= t < T_calm + 20:
            peak = max(peak, d["sigma_mean"])
    assert peak > 1.5 * before
