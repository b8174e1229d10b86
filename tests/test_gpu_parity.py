"""Parity of the HIP path (through the C ABI) against the FP64 oracle.

Tolerances: per-particle |dh| <= 1e-4 after one tick from identical state;
filtered log-vol (h_mean) RMSE <= 1e-4 over whole trajectories; every column of
the per-tick record is compared: log-likelihood and predictive log-likelihood
to 1e-3 (absolute), vol to 1e-4, the KSD of the last Stein iteration to 2e-3
relative (absolute below 0.05; the reference's sign of the mixed kernel
derivative changes it by far more), the bandwidth to 1e-3 relative (a pair
distance inherits the particles' absolute error over the whole trajectory) and the
Stein step count exactly; the median bandwidth selection is bit-exact against
a float32 sort of the same particles.
"""
import ctypes
import os

import numpy as np
import pytest

import oracle
from svpf_amd import config, data
from svpf_amd.filter import Filter

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))


def _hip():
    return ctypes.CDLL("libamdhip64.so")


def d2h(ptr, n):
    out = np.zeros(n, dtype=np.float32)
    rc = _hip().hipMemcpy(out.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(ctypes.cast(ptr, ctypes.c_void_p).value),
                          ctypes.c_size_t(4 * n), 2)
    assert rc == 0
    return out


def dev(torch, a):
    return None if a is None else torch.tensor(np.ascontiguousarray(a, dtype=np.float32), device="cuda")


def run_gpu(torch, cfg, y, noise, h0, path="auto", coop_shape=None):
    f = Filter(cfg, path=path, coop_shape=coop_shape)
    f.set_particles(h0)
    tz = {k: dev(torch, noise.get(k)) for k in ("z_pred", "xi", "u_pred", "u_rej", "z_rej")}
    f.set_noise(tz["z_pred"], tz["xi"], tz["u_pred"], tz["u_rej"], tz["z_rej"])
    rec = np.zeros((len(y), 8), dtype=np.float64)
    yp = 0.0
    for t in range(len(y)):
        f.step(float(y[t]), yp)
        rec[t] = f.last_outputs()
        yp = float(y[t])
    return f, rec


def run_oracle(cfg, y, noise, h0):
    o = oracle.OracleFilter(cfg, h0)
    out = o.run(y, noise)
    return o, out


def compare(rec, out, tol_hmean=1e-4, tol_ll=1e-3, tol_vol=1e-4, tol_ksd=2e-3, tol_bw=1e-3):
    """Per-tick device records [T, 8] against the oracle's outputs (every column the oracle reports)."""
    dh = rec[:, 2] - out["h_mean"]
    assert np.sqrt(np.mean(dh ** 2)) <= tol_hmean, np.abs(dh).max()
    assert np.max(np.abs(rec[:, 1] - out["vol"])) <= tol_vol * max(1.0, np.abs(out["vol"]).max())
    assert np.max(np.abs(rec[:, 0] - out["loglik"])) <= tol_ll
    assert np.array_equal(rec[:, 6].astype(int), np.asarray(out["steps"]).astype(int))
    if "ksd" in out:
        k = np.asarray(out["ksd"], dtype=np.float64)
        err = np.abs(rec[:, 3] - k) / np.maximum(0.05, np.abs(k))
        assert err.max() <= tol_ksd, (err.max(), k[np.argmax(err)], rec[np.argmax(err), 3])
    bw = out.get("bw", out.get("bandwidth"))
    if bw is not None:
        bw = np.asarray(bw, dtype=np.float64)
        assert np.max(np.abs(rec[:, 4] - bw) / bw) <= tol_bw
    if "pred_loglik" in out:
        assert np.max(np.abs(rec[:, 5] - out["pred_loglik"])) <= tol_ll


@pytest.mark.parametrize("path", ["single", "coop"])
def test_one_tick_from_identical_state(torch_cuda, path):
    n = 512
    cfg = config.plain_config(n=n)
    rng = np.random.default_rng(11)
    h0 = data.initial_particles(n, -1.0, 0.97, 0.15, seed=5)
    noise = data.predraw_noise(1, n, 5, seed=6)
    y = np.array([0.42], dtype=np.float32)
    f, rec = run_gpu(torch_cuda, cfg, y, noise, h0, path=path)
    o, out = run_oracle(cfg, y, noise, h0)
    h_gpu = f.particles()
    assert np.max(np.abs(h_gpu - o.h)) <= 1e-4
    assert np.max(np.abs(f.rmsprop_state() - o.v)) <= 1e-4
    # median selection: bit-exact against a float32 sort of the device's predicted particles
    hp = d2h(f.state.h_pred, n)
    assert np.float32(rec[0, 7]) == oracle.median_r2_f32(hp)
    del rng


@pytest.mark.parametrize("path", ["single", "coop"])
def test_plain_filter_trajectory_matches_golden(torch_cuda, path):
    from golden.make_golden import plain_inputs

    g = np.load(os.path.join(HERE, "golden", "plain_n512_t1000.npz"))
    cfg, y, h_true, noise, h0 = plain_inputs()
    f, rec = run_gpu(torch_cuda, cfg, y, noise, h0, path=path)
    out = {k: g[k] for k in ("h_mean", "vol", "loglik", "steps", "ksd", "bw", "pred_loglik")}
    compare(rec, out)
    assert np.sqrt(np.mean((f.particles() - g["h_final"]) ** 2)) <= 1e-4


FEATURE_CONFIGS = {
    "reference_defaults": dict(),
    "reference_defaults_exact_newton": dict(use_exact_gradient=1, lik_offset=0.25, use_newton=1),
    "full_newton": dict(use_newton=1, use_full_newton=1, use_exact_gradient=1, stein_sign=1),
    "plain_svgd_legacy": dict(use_newton=0, use_mim=0, use_guide=0, use_rejuvenation=0, stein_sign=0),
    "silverman_paper": dict(bandwidth_mode=2, stein_sign=1, use_newton=1),
    "median_guided": dict(bandwidth_mode=1, use_guided=1, use_adaptive_guide=1, use_newton=1),
    "production": dict(nu=30.0, lik_offset=0.345, temperature=0.45, n_anneal_steps=5, mim_jump_prob=0.25,
                       mim_jump_scale=9.0, rejuv_ksd_threshold=0.05, use_newton=1, use_full_newton=1,
                       use_exact_gradient=1, stein_min_steps=8, stein_max_steps=16, use_guided=1,
                       guided_alpha_shock=0.40, use_adaptive_guide=1, use_adaptive_mu=1,
                       mu_obs_var_scale=11.0, mu_min=-4.0, use_adaptive_sigma=1, sigma_boost_threshold=0.95,
                       sigma_boost_max=3.2),
    "ksd_early_stop": dict(use_newton=1, use_ksd_early_stop=1, stein_min_steps=2, stein_max_steps=8,
                           bandwidth_mode=1),
    "fan_mode": dict(use_fan_mode=1, stein_sign=1),
}


@pytest.mark.parametrize("path", ["single", "coop"])
@pytest.mark.parametrize("name", sorted(FEATURE_CONFIGS))
def test_feature_configs_short_run(torch_cuda, name, path):
    n, T = 300, 25
    cfg = config.make_config(n=n, mu=-1.0, rho=0.97, sigma_z=0.15, gamma=-0.2, **FEATURE_CONFIGS[name])
    y, _ = data.sv_series(T, mu=-1.0, rho=0.97, sigma_z=0.15, nu=5.0, seed=3)
    y[12] *= 6.0  # a shock to trigger the adaptive branches
    rows = cfg["stein_max_steps"]
    noise = data.predraw_noise(T, n, rows, seed=4, mim=True, rejuv=True)
    h0 = data.initial_particles(n, -1.0, 0.97, 0.15, seed=5)
    f, rec = run_gpu(torch_cuda, cfg, y, noise, h0, path=path)
    o, out = run_oracle(cfg, y, noise, h0)
    compare(rec, out, tol_hmean=1e-4, tol_ll=2e-3, tol_vol=2e-4)


@pytest.mark.parametrize("path", ["single", "coop"])
@pytest.mark.parametrize("n", [1, 2, 3, 65, 400, 2048, 4096])
def test_particle_counts(torch_cuda, n, path):
    T = 6
    cfg = config.plain_config(n=n)
    y, _ = data.sv_series(T, seed=8)
    noise = data.predraw_noise(T, n, 5, seed=9)
    h0 = data.initial_particles(n, -1.0, 0.97, 0.15, seed=10)
    f, rec = run_gpu(torch_cuda, cfg, y, noise, h0, path=path)
    o, out = run_oracle(cfg, y, noise, h0)
    compare(rec, out, tol_ll=1e-3)
    assert np.max(np.abs(f.particles() - o.h)) <= 2e-4


def test_counter_generator_matches_oracle(torch_cuda):
    n, T = 256, 20
    cfg = config.plain_config(n=n)
    y, _ = data.sv_series(T, seed=12)
    h0 = data.initial_particles(n, -1.0, 0.97, 0.15, seed=13)
    f = Filter(cfg)
    f.set_particles(h0)  # no pre-drawn noise: the counter generator, keyed by the default seed 42
    o = oracle.OracleFilter(cfg, h0)
    recs, outs = [], []
    yp = 0.0
    seed = 42
    for t in range(T):
        f.step(float(y[t]), yp)
        recs.append(f.last_outputs())
        outs.append(o.tick(float(y[t]), yp, seed=seed))
        yp = float(y[t])
    hm = np.array([r[2] for r in recs]) - np.array([q["h_mean"] for q in outs])
    assert np.sqrt(np.mean(hm ** 2)) <= 1e-4
    assert np.max(np.abs(f.particles() - o.h)) <= 5e-4


def test_device_run_equals_per_tick_steps(torch_cuda):
    torch = torch_cuda
    n, T = 512, 40
    cfg = config.plain_config(n=n)
    y, _ = data.sv_series(T, seed=14)
    noise = data.predraw_noise(T, n, 5, seed=15)
    h0 = data.initial_particles(n, -1.0, 0.97, 0.15, seed=16)
    f1, rec = run_gpu(torch, cfg, y, noise, h0)
    f2 = Filter(cfg)
    f2.set_particles(h0)
    f2.set_noise(dev(torch, noise["z_pred"]), dev(torch, noise["xi"]))
    dy = dev(torch, y)
    ll = torch.zeros(T, device="cuda")
    vol = torch.zeros(T, device="cuda")
    f2.run_device(dy, ll, vol)
    # the device loop starts with y_prev = 0 like svpf_run_sequence (src/svpf_optimized_graph.cu:1414-1424)
    assert np.array_equal(ll.cpu().numpy(), rec[:, 0].astype(np.float32))
    assert np.array_equal(vol.cpu().numpy(), rec[:, 1].astype(np.float32))
    assert np.array_equal(f1.particles(), f2.particles())
    f3 = Filter(cfg)
    f3.set_particles(h0)
    f3.set_noise(dev(torch, noise["z_pred"]), dev(torch, noise["xi"]))
    ll3, vol3 = f3.run_host(y)
    assert np.array_equal(ll3, rec[:, 0].astype(np.float32))


def test_unsupported_and_uninitialized_fail_loudly(torch_cuda):
    from svpf_amd import _lib

    cfg = config.plain_config(n=64)
    f = Filter(cfg)
    with pytest.raises(RuntimeError, match="not initialized"):
        f.step(0.1, 0.0)
    f.initialize(7)
    f.state.use_two_factor = 1
    f.state.n_stein_steps = 0  # the two-factor step needs at least one Stein iteration
    with pytest.raises(RuntimeError, match="invalid"):
        f.step(0.1, 0.0)
    f.state.use_two_factor = 0
    f.step(0.1, 0.0)
    assert f.L.svpf_last_error(f.sp) == _lib.SVPF_RC_OK
    assert not f.L.svpf_create(0, 5, 5.0, None)


def test_graph_capture_replays(torch_cuda):
    cfg = config.plain_config(n=128)
    f = Filter(cfg)
    f.initialize(3)
    assert not f.L.svpf_graph_is_captured(f.sp)
    f.step(0.1, 0.0)
    assert f.L.svpf_graph_is_captured(f.sp)
    a = [f.step(0.2, 0.1) for _ in range(3)]
    f.L.svpf_graph_invalidate(f.sp)
    assert not f.L.svpf_graph_is_captured(f.sp)
    assert all(np.isfinite(x).all() for x in a)


@pytest.mark.parametrize("shape", [(4, 8), (8, 8), (16, 8), (32, 8), (8, 1), (16, 1)])
def test_coop_shapes(torch_cuda, shape):
    """The cooperative kernel at several workgroup counts / placements (N=512)."""
    n, T = 512, 8
    cfg = config.plain_config(n=n)
    y, _ = data.sv_series(T, seed=21)
    noise = data.predraw_noise(T, n, 5, seed=22)
    h0 = data.initial_particles(n, -1.0, 0.97, 0.15, seed=23)
    f, rec = run_gpu(torch_cuda, cfg, y, noise, h0, path="coop", coop_shape=shape)
    assert f.coop_shape()[0] == shape[0] and f.coop_shape()[2] == shape[1]
    o, out = run_oracle(cfg, y, noise, h0)
    compare(rec, out)
    assert np.max(np.abs(f.particles() - o.h)) <= 1e-4


@pytest.mark.parametrize("name,n,shape", [("plain", 512, None), ("plain", 256, None), ("plain", 768, (8, 8)),
                                          ("plain", 512, (8, 1)), ("median_guided", 512, None),
                                          ("production_median", 384, None)])
def test_coop_median_workgroup_is_bitwise_neutral(torch_cuda, name, n, shape):
    """With the median bandwidth the cooperative kernel selects the median in a
    workgroup of its own while the row workgroups compute the first gradient and
    exchange its records unscaled.  Per-tick records, particles and RMSProp state
    equal, bit for bit, those of the kernel whose every row workgroup selects the
    median itself (per-tick steps and the device loop), and match the oracle."""
    torch = torch_cuda
    T = 12
    if name == "plain":
        cfg = config.plain_config(n=n)
    else:
        kw = dict(FEATURE_CONFIGS["median_guided" if name == "median_guided" else "production"])
        kw["bandwidth_mode"] = 1
        cfg = config.make_config(n=n, mu=-1.0, rho=0.97, sigma_z=0.15, gamma=-0.2, **kw)
    y, _ = data.sv_series(T, seed=61)
    rows = cfg.get("stein_max_steps", 5)
    noise = data.predraw_noise(T, n, rows, seed=62, mim=True, rejuv=True)
    h0 = data.initial_particles(n, -1.0, 0.97, 0.15, seed=63)
    recs, parts, rms = [], [], []
    for mw in (True, False):
        print(f"coop median workgroup {mw}: {name} n={n} shape={shape}", flush=True)
        f = Filter(cfg, path="coop", coop_shape=shape, median_workgroup=mw)
        assert f.path() == "coop"
        f.set_particles(h0)
        tz = {k: dev(torch, noise.get(k)) for k in ("z_pred", "xi", "u_pred", "u_rej", "z_rej")}
        f.set_noise(tz["z_pred"], tz["xi"], tz["u_pred"], tz["u_rej"], tz["z_rej"])
        rec = np.zeros((T, 8), dtype=np.float32)
        yp = 0.0
        for t in range(T - 4):
            f.step(float(y[t]), yp)
            rec[t] = f.last_outputs()
            yp = float(y[t])
        ll = torch.zeros(4, device="cuda")
        f.run_device(dev(torch, y[T - 4:]), ll)
        rec[T - 4:, 0] = ll.cpu().numpy()
        recs.append(rec)
        parts.append(f.particles())
        rms.append(f.rmsprop_state())
        f.close()
    assert np.array_equal(recs[0], recs[1])
    assert np.array_equal(parts[0], parts[1])
    assert np.array_equal(rms[0], rms[1])
    o, out = run_oracle(cfg, y[: T - 4], {k: (v[: T - 4] if v is not None else None) for k, v in noise.items()}, h0)
    assert np.sqrt(np.mean((recs[0][: T - 4, 2] - np.asarray(out["h_mean"])) ** 2)) <= 1e-4


def test_coop_device_run_equals_per_tick(torch_cuda):
    torch = torch_cuda
    n, T = 512, 30
    cfg = config.plain_config(n=n)
    y, _ = data.sv_series(T, seed=24)
    noise = data.predraw_noise(T, n, 5, seed=25)
    h0 = data.initial_particles(n, -1.0, 0.97, 0.15, seed=26)
    f1, rec = run_gpu(torch, cfg, y, noise, h0, path="coop")
    f2 = Filter(cfg, path="coop")
    f2.set_particles(h0)
    f2.set_noise(dev(torch, noise["z_pred"]), dev(torch, noise["xi"]))
    ll = torch.zeros(T, device="cuda")
    f2.run_device(dev(torch, y), ll)
    assert np.array_equal(ll.cpu().numpy(), rec[:, 0].astype(np.float32))
    assert np.array_equal(f1.particles(), f2.particles())


def test_coop_exchange_timeout_repeats_with_a_cooperative_launch(torch_cuda):
    """The cooperative kernel runs as a plain launch.  A forced time-out (the
    exchanges wait for a tag nobody publishes, 4 polls) writes nothing back; the
    step is repeated with a cooperative launch (forced to time out as well, so it
    reports SVPF_RC_TIMEOUT) and the filter keeps cooperative launches.  With
    the default bound its next tick equals that of a filter that never timed out."""
    from svpf_amd import _lib

    n = 512
    cfg = config.plain_config(n=n)
    h0 = data.initial_particles(n, -1.0, 0.97, 0.15, seed=31)
    noise = data.predraw_noise(1, n, 5, seed=32)
    y0 = 0.37

    def make():
        f = Filter(cfg, path="coop")
        f.set_particles(h0)
        f.set_noise(dev(torch_cuda, noise["z_pred"]), dev(torch_cuda, noise["xi"]))
        assert f.path() == "coop"
        return f

    f = make()
    L = f.L
    assert L.svpf_coop_launch_is_cooperative(f.sp) == 0
    assert L.svpf_set_coop_spin_limit(f.sp, -4) == _lib.SVPF_RC_OK
    print("coop exchange time-out", flush=True)
    with pytest.raises(RuntimeError, match="time"):
        f.step(y0, 0.0)
    assert L.svpf_last_error(f.sp) == _lib.SVPF_RC_TIMEOUT
    assert L.svpf_coop_launch_is_cooperative(f.sp) == 1
    assert np.array_equal(f.particles(), h0.astype(np.float32))
    assert f.state.timestep == 0
    assert L.svpf_set_coop_spin_limit(f.sp, 0) == _lib.SVPF_RC_OK
    f.step(y0, 0.0)
    g = make()
    g.step(y0, 0.0)
    assert L.svpf_coop_launch_is_cooperative(g.sp) == 0
    assert np.array_equal(f.particles(), g.particles())
    assert np.array_equal(f.last_outputs(), g.last_outputs())
    assert L.svpf_set_coop_spin_limit(f.sp, -(1 << 21)) == _lib.SVPF_RC_INVALID_ARGUMENT


def test_initialize_draws_the_oracle_particles(torch_cuda):
    """svpf_initialize (src/svpf_optimized_graph.cu:84-97, 349-451): particles from the counter
    generator keyed by the seed, against the oracle's draw; the Student-t state widens them."""
    for n, kw in ((1000, {}), (333, dict(use_student_t_state=1, nu_state=4.0))):
        cfg = config.make_config(n=n, mu=-1.2, rho=0.95, sigma_z=0.25, **kw)
        f = Filter(cfg)
        f.initialize(seed=987654321)
        want = oracle.init_particles(n, -1.2, 0.95, 0.25, 987654321, student_t_state=cfg["use_student_t_state"],
                                     nu_state=cfg["nu_state"])
        got = f.particles()
        assert np.max(np.abs(got - want)) <= 1e-5 * max(1.0, np.abs(want).max())
        assert np.array_equal(d2h(f.state.h_prev, n), got)
        assert np.array_equal(f.rmsprop_state(), np.zeros(n, np.float32))
        assert f.state.timestep == 0 and f.state.ksd_prev == pytest.approx(1e10)
        f.close()


@pytest.mark.parametrize("lag,out_lag,adaptive_mu", [(3, 1, 0), (5, 2, 0), (4, 1, 1)])
def test_smoothing_matches_oracle(torch_cuda, lag, out_lag, adaptive_mu):
    """use_smoothing: the step's reported h_mean is the RTS-corrected estimate output_lag
    ticks back (src/svpf_optimized_graph.cu:619-717), fed by the tick's h_mean and
    bandwidth^2, with the adaptive mean as it was before the tick's update."""
    n, T = 256, 30
    extra = dict(use_adaptive_mu=1, mu_obs_var_scale=11.0, mu_min=-4.0, mu_max=0.0) if adaptive_mu else {}
    cfg = config.plain_config(n=n, **extra)
    y, _ = data.sv_series(T, seed=41)
    noise = data.predraw_noise(T, n, 5, seed=42)
    h0 = data.initial_particles(n, -1.0, 0.97, 0.15, seed=43)
    f = Filter(cfg)
    f.set_particles(h0)
    f.set_noise(dev(torch_cuda, noise["z_pred"]), dev(torch_cuda, noise["xi"]))
    f.state.use_smoothing = 1
    f.state.smooth_lag = lag
    f.state.smooth_output_lag = out_lag
    o = oracle.OracleFilter(cfg, h0)
    sm = oracle.Smoother(lag, out_lag)
    yp = 0.0
    for t in range(T):
        hm = f.step(float(y[t]), yp)[2]
        mu_before = o.st.mu_state if adaptive_mu else cfg["mu"]
        r = o.tick(float(y[t]), yp, z_pred=noise["z_pred"][t], xi=noise["xi"][t])
        want = sm.step(t, r["h_mean"], r["bandwidth"] ** 2, float(y[t]), mu_before, cfg["rho"], cfg["sigma_z"])
        assert abs(hm - want) <= 1e-4, (t, hm, want)
        yp = float(y[t])
    f.close()


@pytest.mark.parametrize("wg", [0, 7, 15])
def test_coop_partial_timeout_commits_nothing(torch_cuda, wg):
    """One workgroup's waits at the final exchange give up while every other workgroup
    completes (host-set fault, every kernel exits normally).  The launch must commit
    nothing -- particles, previous particles, RMSProp state and scalars stay as they
    were -- and report the time-out; with the fault cleared, the filter's next tick
    equals that of a filter that never failed.  Also for a 6-tick device loop whose
    last tick fails."""
    from svpf_amd import _lib

    torch = torch_cuda
    n = 512
    cfg = config.plain_config(n=n)
    h0 = data.initial_particles(n, -1.0, 0.97, 0.15, seed=51)
    noise = data.predraw_noise(6, n, 5, seed=52)
    y = data.sv_series(6, seed=53)[0]

    def make():
        f = Filter(cfg, path="coop")
        f.set_particles(h0)
        f.set_noise(dev(torch, noise["z_pred"]), dev(torch, noise["xi"]))
        assert f.path() == "coop" and f.coop_shape()[1] == 16
        return f

    f = make()
    L = f.L
    assert L.svpf_set_coop_fault(f.sp, wg) == _lib.SVPF_RC_OK
    print(f"coop fault in workgroup {wg}", flush=True)
    with pytest.raises(RuntimeError, match="time"):
        f.step(float(y[0]), 0.0)
    assert L.svpf_last_error(f.sp) == _lib.SVPF_RC_TIMEOUT
    assert np.array_equal(f.particles(), h0.astype(np.float32))
    assert np.array_equal(d2h(f.state.h_prev, n), h0.astype(np.float32))
    assert np.array_equal(f.rmsprop_state(), np.zeros(n, np.float32))
    assert f.state.timestep == 0
    # the device loop: the last tick's final exchange fails, nothing of the 6 ticks is committed
    with pytest.raises(RuntimeError, match="time"):
        f.run_device(dev(torch, y))
    assert np.array_equal(f.particles(), h0.astype(np.float32)) and f.state.timestep == 0
    assert L.svpf_set_coop_fault(f.sp, -1) == _lib.SVPF_RC_OK
    f.step(float(y[0]), 0.0)
    g = make()
    g.step(float(y[0]), 0.0)
    assert np.array_equal(f.particles(), g.particles())
    assert np.array_equal(f.rmsprop_state(), g.rmsprop_state())
    assert np.array_equal(f.last_outputs(), g.last_outputs())
    assert L.svpf_set_coop_fault(f.sp, 16) == _lib.SVPF_RC_INVALID_ARGUMENT
    f.close()
    g.close()
