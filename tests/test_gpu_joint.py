"""The joint state-parameter filter's HIP kernel (csrc/svpf_joint.hip) through
the C ABI (include/svpf_joint.h) against the FP64 oracle (svpfo_joint_*).

Tolerances: one step from the same float32 state is compared particle by particle
at 5e-5 absolute (float32 pair sums of up to 4096 terms, five iterations), or at
20x the oracle's own change under a 4-ulp float32 perturbation of the state where
that is larger (many iterations amplify rounding); the diagnostics at 1e-5
relative for means, 1e-4 for spreads, ESS and bandwidths."""
import numpy as np
import pytest

import oracle
from joint_series import crash_series, sv_series
from svpf_amd import joint

pytestmark = pytest.mark.gpu

TILE_SIZES = [1, 17, 128, 200, 256, 512, 1000, 2048, 3000, 4096]


def oracle_cfg(f: joint.JointFilter) -> dict:
    """The filter's own float32 configuration as an oracle configuration."""
    c = f.state.config
    d = {"n": c.n_particles, "n_stein_steps": c.n_stein_steps}
    d.update({k: float(getattr(c, k)) for k in joint.CONFIG_FIELDS})
    return d


def random_state(n, seed):
    rng = np.random.default_rng(seed)
    x = np.stack([rng.normal(-3.4, 0.5, n), rng.normal(-3.5, 0.25, n), rng.normal(2.0, 0.3, n),
                  rng.normal(-2.0, 0.3, n)]).astype(np.float32)
    hp = (x[0] + rng.normal(0, 0.2, n)).astype(np.float32)
    return x, hp


def compare_diag(g: dict, o: dict, what=""):
    for k in ("h_mean", "vol_mean", "mu_mean", "rho_mean", "sigma_mean"):
        assert g[k] == pytest.approx(o[k], rel=1e-5, abs=1e-6), (what, k)
    for k in ("mu_std", "rho_std", "sigma_std", "std_mu_tilde", "std_rho_tilde", "std_sigma_tilde", "ess"):
        assert g[k] == pytest.approx(o[k], rel=1e-4, abs=1e-6), (what, k)
    assert g["log_likelihood"] == pytest.approx(o["log_likelihood"], abs=1e-4), what
    for k in ("mu_collapsed", "rho_collapsed", "sigma_collapsed"):
        assert g[k] == o[k], (what, k)


def test_initialize_matches_oracle(torch_cuda):
    for n in (1, 100, 512, 4096):
        f = joint.JointFilter(seed=12345, n_particles=n)
        o = oracle.JointOracle(oracle_cfg(f), seed=12345)
        p = f.particles()
        np.testing.assert_allclose(p[[0, 2, 3, 4]], o.x, rtol=0, atol=3e-6, err_msg=f"n={n}")
        np.testing.assert_array_equal(p[1], p[0])
        mu, rho, sigma = f.params()
        assert mu == pytest.approx(o.x[1].mean(), rel=1e-5)
        assert rho == pytest.approx(np.mean(1 / (1 + np.exp(-o.x[2]))), rel=1e-5)
        assert sigma == pytest.approx(np.mean(np.exp(o.x[3])), rel=1e-5)
        assert f.vol() == pytest.approx(np.exp(0.5 * o.x[0].mean()), rel=1e-5)
        f.close()


REGIMES = {
    "calm": dict(y=0.6 * np.exp(-3.4 / 2)),
    "surprise": dict(y=6.0 * np.exp(-3.4 / 2)),
    "quiet": dict(y=1e-4),
    "zero_iterations": dict(y=0.2 * np.exp(-3.4 / 2), n_stein_steps=0),
    "ten_iterations": dict(y=1.5 * np.exp(-3.4 / 2), n_stein_steps=10),
}


@pytest.mark.parametrize("regime", list(REGIMES))
@pytest.mark.parametrize("n", TILE_SIZES)
def test_one_step_matches_oracle(torch_cuda, n, regime):
    spec = dict(REGIMES[regime])
    y = float(np.float32(spec.pop("y")))
    f = joint.JointFilter(seed=777, n_particles=n, **spec)
    x, hp = random_state(n, n)
    f.set_particles(x[0], hp, x[1], x[2], x[3])
    tick = 4321
    f.state.timestep = tick
    g = f.step(y)
    o = oracle.JointOracle(oracle_cfg(f), x0=x.astype(np.float64), hp0=hp.astype(np.float64))
    od = o.step(y, 777, tick)
    p = f.particles()
    err = np.abs(p[[0, 2, 3, 4]] - o.x).max()
    # the step's own sensitivity to float32 rounding: the oracle from the state perturbed by
    # 4 float32 ulps per coordinate, the size of the rounding a float32 sum of a few thousand
    # pair terms carries
    rng = np.random.default_rng(1)
    xq = x.astype(np.float64) * (1.0 + rng.choice([-1.0, 1.0], x.shape) * 2.0 ** -22)
    oq = oracle.JointOracle(oracle_cfg(f), x0=xq, hp0=hp.astype(np.float64))
    oq.step(y, 777, tick)
    sens = np.abs(oq.x - o.x).max()
    assert err < max(5e-5, 20.0 * sens), f"max particle error {err}, oracle sensitivity {sens}"
    np.testing.assert_array_equal(p[1], x[0])  # h_prev is the state before predict
    compare_diag(g, od, f"{regime} n={n}")
    st = f.state
    np.testing.assert_allclose([st.bw_h, st.bw_mu, st.bw_rho, st.bw_sigma], od["bw"], rtol=1e-4)
    assert st.timestep == tick + 1 and st.y_prev == np.float32(y)
    f.close()


@pytest.mark.parametrize("n", [512, 3000])
def test_sequence_launch_equals_single_steps(torch_cuda, n):
    torch = torch_cuda
    T = 25
    y, _ = sv_series(T, -3.5, 0.95, 0.15, 42)
    a = joint.JointFilter(seed=99, n_particles=n)
    b = joint.JointFilter(seed=99, n_particles=n)
    diags = [a.step(float(v)) for v in y]
    yd = torch.tensor(y, device="cuda")
    rec = torch.zeros((T, joint.RECW), dtype=torch.float32, device="cuda")
    b.run_device(yd, rec)
    np.testing.assert_array_equal(a.particles(), b.particles())
    rows = joint.records_to_dicts(rec.cpu().numpy())
    for t in range(T):
        for k, v in diags[t].items():
            assert rows[t][k] == v, (t, k)
    assert b.state.timestep == T
    st = a.state
    assert [rows[-1]["bw_h"], rows[-1]["bw_mu"], rows[-1]["bw_rho"], rows[-1]["bw_sigma"]] == \
        [st.bw_h, st.bw_mu, st.bw_rho, st.bw_sigma]
    a.close()
    b.close()


def test_lockstep_with_oracle_on_the_reference_series(torch_cuda):
    """40 steps on the calm series without re-synchronising: the float32 kernel stays on
    the FP64 oracle's path."""
    n, T = 256, 40
    y, _ = sv_series(T, -3.5, 0.95, 0.15, 42)
    f = joint.JointFilter(seed=12345, n_particles=n)
    o = oracle.JointOracle(oracle_cfg(f), seed=12345)
    worst = 0.0
    for t in range(T):
        g = f.step(float(y[t]))
        od = o.step(float(y[t]), 12345, t)
        for k in ("h_mean", "mu_mean", "rho_mean", "sigma_mean"):
            worst = max(worst, abs(g[k] - od[k]) / max(abs(od[k]), 1e-3))
    assert worst < 1e-3, worst
    err = np.abs(f.particles()[[0, 2, 3, 4]] - o.x).max()
    assert err < 2e-3, err
    f.close()


def test_reference_basic(torch_cuda):
    """test_svpf_joint.cu:66-105."""
    f = joint.JointFilter(n_particles=256, n_stein_steps=3)
    f.initialize(12345)
    mu, rho, sigma = f.params()
    assert np.isfinite([mu, rho, sigma]).all()
    for _ in range(10):
        d = f.step(0.01)
    assert all(np.isfinite(v) for v in d.values())
    assert 1.0 <= d["ess"] <= 256.0
    f.close()


def test_reference_calm_tracking(torch_cuda):
    """test_svpf_joint.cu:112-177: |rho_hat - 0.95| <= 0.15 after 500 calm steps; the
    last-step h estimate is the particle mean the reference test recomputes."""
    y, h = sv_series(500, -3.5, 0.95, 0.15, 42)
    f = joint.JointFilter(seed=12345, n_particles=512, n_stein_steps=5)
    for t in range(500):
        d = f.step(float(y[t]), diag=(t == 499))
    assert abs(d["rho_mean"] - 0.95) <= 0.15
    assert d["h_mean"] == pytest.approx(float(f.particles()[0].mean()), abs=1e-5)
    assert abs(d["h_mean"] - h[-1]) < 1.5
    f.close()


def test_reference_crash_adaptation_and_oracle_agreement(torch_cuda):
    """test_svpf_joint.cu:184-289: sigma's peak over the first 20 crash steps exceeds 1.5x its
    pre-crash value; the GPU run follows the oracle's sigma path."""
    y, _, T_calm = crash_series()
    f = joint.JointFilter(seed=12345, n_particles=512, n_stein_steps=5)
    o = oracle.JointOracle(oracle_cfg(f), seed=12345)
    gs, os_ = [], []
    for t in range(len(y)):
        gs.append(f.step(float(y[t]))["sigma_mean"])
        os_.append(o.step(float(y[t]), 12345, t)["sigma_mean"])
    before, peak = gs[T_calm - 1], max(gs[T_calm:T_calm + 20])
    assert peak > 1.5 * before
    np.testing.assert_allclose(gs, os_, rtol=2e-2)
    f.close()


def test_reference_diversity_run(torch_cuda):
    """test_svpf_joint.cu:296-344: 1000 steps at 256 particles; every diagnostic finite, and the
    collapse flags agree with the unconstrained spreads."""
    torch = torch_cuda
    y, _ = sv_series(1000, -3.5, 0.95, 0.15, 42)
    f = joint.JointFilter(seed=12345, n_particles=256, n_stein_steps=5)
    rec = torch.zeros((1000, joint.RECW), dtype=torch.float32, device="cuda")
    f.run_device(torch.tensor(y, device="cuda"), rec)
    rows = joint.records_to_dicts(rec.cpu().numpy())
    c = f.state.config
    for r in rows:
        assert all(np.isfinite(v) for v in r.values())
        assert r["mu_collapsed"] == int(r["std_mu_tilde"] < c.collapse_thresh_mu)
        assert r["rho_collapsed"] == int(r["std_rho_tilde"] < c.collapse_thresh_rho)
        assert r["sigma_collapsed"] == int(r["std_sigma_tilde"] < c.collapse_thresh_sigma)
    f.close()


def test_collapse_warning_is_printed(torch_cuda, capfd):
    f = joint.JointFilter(seed=5, n_particles=128, collapse_thresh_mu=10.0, collapse_thresh_rho=10.0,
                          collapse_thresh_sigma=10.0)
    d = f.step(0.01)
    assert (d["mu_collapsed"], d["rho_collapsed"], d["sigma_collapsed"]) == (1, 1, 1)
    f.close()
    out = capfd.readouterr().out
    assert "WARNING [t=0]: Diversity collapsed - mu:1 rho:1 sigma:1" in out


def test_getters_agree_with_the_last_diagnostics(torch_cuda):
    f = joint.JointFilter(seed=8, n_particles=300)
    for v in (0.01, -0.02, 0.03):
        d = f.step(v)
    mu, rho, sigma = f.params()
    assert (mu, rho, sigma) == (d["mu_mean"], d["rho_mean"], d["sigma_mean"])
    assert f.vol() == pytest.approx(np.exp(0.5 * d["h_mean"]), rel=1e-6)
    f.close()
