"""Parity of the Heun predictor-corrector Stein step and the Student-t state
(noise draw and mixture prior) against the FP64 oracle, through the C ABI, on
the single-workgroup kernel (its extended instantiation, tile ids with the
extended bit), the batched launch and the grid path.

The cooperative kernel does not take these features: the automatic path choice
runs such filters on one workgroup below 256 particles and on the grid path from
256 up.  Tolerances are those of the other tests: per-particle |dh| <= 1e-4
after one tick from identical state, h_mean RMSE <= 1e-4 over short
trajectories and equal Stein iteration counts.
"""
import numpy as np
import pytest

import oracle
from svpf_amd import _lib, config, data
from svpf_amd.batched import TILE_EXT, Batch
from svpf_amd.filter import Filter
from test_gpu_parity import compare, d2h, dev, run_gpu, run_oracle

pytestmark = pytest.mark.gpu

HEUN_CONFIGS = {
    "heun_plain": dict(use_heun=1, use_newton=0),
    "heun_local_newton": dict(use_heun=1, use_newton=1, use_exact_gradient=1, lik_offset=0.25),
    "heun_full_newton": dict(use_heun=1, use_newton=1, use_full_newton=1, use_exact_gradient=1, stein_sign=1),
    "heun_ksd_early_stop": dict(use_heun=1, use_newton=1, use_ksd_early_stop=1, stein_min_steps=2,
                                stein_max_steps=8, bandwidth_mode=1),
    "heun_annealed_svld": dict(use_heun=1, use_newton=1, use_full_newton=1, n_anneal_steps=3, temperature=0.45,
                               stein_max_steps=9),
}

STUDENT_CONFIGS = {
    "student_plain": dict(use_student_t_state=1, nu_state=5.0, use_newton=0),
    "student_newton": dict(use_student_t_state=1, nu_state=6.0, use_newton=1, use_exact_gradient=1),
    "student_full_newton": dict(use_student_t_state=1, nu_state=8.0, use_newton=1, use_full_newton=1,
                                use_exact_gradient=1, stein_sign=1),
    "student_heun_median": dict(use_student_t_state=1, nu_state=3.0, use_heun=1, use_newton=1,
                                bandwidth_mode=1),
}


def _series(T, seed):
    y, _ = data.sv_series(T, mu=-1.0, rho=0.97, sigma_z=0.15, nu=5.0, seed=seed)
    y[T // 2] *= 6.0  # a shock: large steps and the adaptive branches
    return y


@pytest.mark.parametrize("path", ["single", "grid"])
@pytest.mark.parametrize("name", ["heun_plain", "heun_full_newton", "student_newton", "student_heun_median"])
def test_one_tick_from_identical_state(torch_cuda, name, path):
    """From a warm RMSProp state (v = 0.05): at v = 0 the first predictor step is
    eff / sqrt(1e-6) = 100 sqrt(beta) times the operator, which scales float32
    rounding of the operator into the 1e-4 range and tests conditioning, not parity."""
    n = 400
    feats = {**HEUN_CONFIGS, **STUDENT_CONFIGS}[name]
    cfg = config.make_config(n=n, mu=-1.0, rho=0.97, sigma_z=0.15, gamma=-0.2, **feats)
    h0 = data.initial_particles(n, -1.0, 0.97, 0.15, seed=5)
    v0 = np.full(n, 0.05, dtype=np.float32)
    noise = data.predraw_noise(1, n, cfg["stein_max_steps"], seed=6, mim=True)
    y = np.array([0.42], dtype=np.float32)
    print(f"one tick {name} {path}", flush=True)
    f = Filter(cfg, path=path)
    f.set_particles(h0, v=v0)
    f.set_noise(dev(torch_cuda, noise["z_pred"]), dev(torch_cuda, noise["xi"]), dev(torch_cuda, noise["u_pred"]))
    f.step(float(y[0]), 0.0)
    rec = f.last_outputs()[None].astype(np.float64)
    assert f.path() == path
    o = oracle.OracleFilter(cfg, h0, v0=v0.astype(np.float64))
    out = o.run(y, noise)
    assert np.max(np.abs(f.particles() - o.h)) <= 1e-4
    assert np.max(np.abs(f.rmsprop_state() - o.v)) <= 1e-4
    compare(rec, out)


@pytest.mark.parametrize("path", ["single", "grid"])
@pytest.mark.parametrize("name", sorted(HEUN_CONFIGS) + sorted(STUDENT_CONFIGS))
def test_feature_configs_short_run(torch_cuda, name, path):
    """25 ticks with a 6-sigma shock.  The configurations are well conditioned
    here: started from particles perturbed by 1e-7 relative, the FP64 oracle's
    h_mean moves by at most 5e-6.  (With nu_s = 4 and local Newton it moves by
    1.7e-3 to 0.12, which no float32 run can follow.)"""
    n, T = 300, 25
    feats = {**HEUN_CONFIGS, **STUDENT_CONFIGS}[name]
    cfg = config.make_config(n=n, mu=-1.0, rho=0.97, sigma_z=0.15, gamma=-0.2, **feats)
    y = _series(T, 3)
    noise = data.predraw_noise(T, n, cfg["stein_max_steps"], seed=4, mim=True, rejuv=True)
    h0 = data.initial_particles(n, -1.0, 0.97, 0.15, seed=5)
    print(f"short run {name} {path}", flush=True)
    f, rec = run_gpu(torch_cuda, cfg, y, noise, h0, path=path)
    assert f.path() == path
    o, out = run_oracle(cfg, y, noise, h0)
    compare(rec, out, tol_hmean=1e-4, tol_ll=2e-3, tol_vol=2e-4)


@pytest.mark.parametrize("path", ["single", "grid"])
@pytest.mark.parametrize("nu_state", [3.0, 5.0, 40.0])
def test_student_t_counter_generator_matches_oracle(torch_cuda, nu_state, path):
    """No pre-drawn noise: the Student-t state draw (z sqrt(k / chi2_k) of the
    counter generator's normals, a plain normal above 30 degrees of freedom)
    and the Student-t prior, against the oracle keyed by the same seed.

    Every tick's predicted particles are checked draw by draw against the
    transition applied to the device's own previous particles.  The trajectory
    is compared over the first 3 ticks: with heavy-tailed state noise the
    filter soon becomes sensitive to rounding -- on this series the FP64 oracle
    itself, started from particles perturbed by 1e-7 relative, moves by < 4e-7
    in h_mean over ticks 0-2 but by up to 5e-4 (nu_s = 5) and 6e-5 (nu_s = 3) by
    tick 7, against 1e-8 for the Gaussian state."""
    n, T, T_cmp = 256, 8, 3
    cfg = config.plain_config(n=n, use_student_t_state=1, nu_state=nu_state, use_newton=1)
    y, _ = data.sv_series(T, mu=-1.0, rho=0.97, sigma_z=0.15, nu=5.0, seed=12)
    h0 = data.initial_particles(n, -1.0, 0.97, 0.15, seed=13)
    f = Filter(cfg, path=path)
    f.set_particles(h0)
    assert f.path() == path
    o = oracle.OracleFilter(cfg, h0)
    lib = oracle.lib()
    mu, rho, sz, gam = cfg["mu"], cfg["rho"], cfg["sigma_z"], cfg["gamma"]
    print(f"student-t counter generator nu_state={nu_state} {path}", flush=True)
    hm = []
    yp = 0.0
    for t in range(T):
        hprev = f.particles().astype(np.float64)
        f.step(float(y[t]), yp)
        draws = np.array([lib.svpfo_student_t_draw(42, 0, t, i, nu_state) for i in range(n)])
        lev = gam * yp / (np.exp(hprev * 0.5) + 1e-8)
        want = np.clip(mu + rho * (hprev - mu) + sz * draws + lev, -15.0, 5.0)
        assert np.max(np.abs(d2h(f.state.h_pred, n) - want)) <= 5e-6, t
        if t < T_cmp:
            hm.append(f.last_outputs()[2] - o.tick(float(y[t]), yp, seed=42)["h_mean"])
            if t == T_cmp - 1:
                assert np.max(np.abs(f.particles() - o.h)) <= 1e-4
        yp = float(y[t])
    assert np.sqrt(np.mean(np.square(hm))) <= 1e-4


@pytest.mark.parametrize("n", [2, 65, 1000, 4096, 8192])
def test_heun_student_particle_counts(torch_cuda, n):
    """Warm RMSProp state (v = 0.05): from v = 0 the n = 2 filter is
    ill-conditioned under Heun (a 1e-7 relative perturbation of the FP64
    oracle's particles moves its h_mean by RMSE 1e-4 over these 5 ticks; 1e-7
    from the warm state)."""
    T = 5
    cfg = config.plain_config(n=n, use_heun=1, use_student_t_state=1, nu_state=5.0, use_newton=1,
                              use_full_newton=1)
    y = _series(T, 8)
    noise = data.predraw_noise(T, n, 5, seed=9)
    h0 = data.initial_particles(n, -1.0, 0.97, 0.15, seed=10)
    v0 = np.full(n, 0.05, dtype=np.float32)
    print(f"heun + student-t n={n}", flush=True)
    f = Filter(cfg)
    f.set_particles(h0, v=v0)
    f.set_noise(dev(torch_cuda, noise["z_pred"]), dev(torch_cuda, noise["xi"]))
    rec = np.zeros((T, 8))
    yp = 0.0
    for t in range(T):
        f.step(float(y[t]), yp)
        rec[t] = f.last_outputs()
        yp = float(y[t])
    assert f.path() == ("single" if n < 256 else "grid")
    o = oracle.OracleFilter(cfg, h0, v0=v0.astype(np.float64))
    out = o.run(y, noise)
    compare(rec, out, tol_ll=1e-3)
    assert np.max(np.abs(f.particles() - o.h)) <= 1e-4


def test_batched_mixed_features(torch_cuda):
    """One batched run: plain, Heun and Student-t filters of different N; the
    extended filters form their own tile groups; each filter matches the oracle."""
    T = 6
    specs = [(128, {}), (200, dict(use_heun=1)), (160, dict(use_student_t_state=1, nu_state=4.0, use_newton=1)),
             (300, dict(use_heun=1, use_student_t_state=1, nu_state=6.0, use_newton=1, use_full_newton=1))]
    cfgs, ys, h0s, noise = [], [], [], {"z_pred": [], "xi": []}
    for b, (n, feats) in enumerate(specs):
        cfgs.append(config.plain_config(n=n, **feats))
        ys.append(_series(T, 60 + b))
        h0s.append(data.initial_particles(n, -1.0, 0.97, 0.15, seed=70 + b))
        nz = data.predraw_noise(T, n, 5, seed=80 + b)
        noise["z_pred"].append(nz["z_pred"])
        noise["xi"].append(nz["xi"])
    print("batched mixed features", flush=True)
    bt = Batch(cfgs, np.stack(ys), h0s, noise=noise)
    assert [bool(t[1] & TILE_EXT) for t in bt.tiles] == [False, True, True, True]
    bt.run(0, T)
    torch_cuda.cuda.synchronize()
    assert (bt.fstates()["status"] == 0).all()
    out = bt.outputs()
    for b in range(len(specs)):
        o = oracle.OracleFilter(cfgs[b], h0s[b])
        ref = o.run(ys[b], {"z_pred": noise["z_pred"][b], "xi": noise["xi"][b]})
        assert np.sqrt(np.mean((out[b, :, 2] - ref["h_mean"]) ** 2)) <= 1e-4, b
        assert np.array_equal(out[b, :, 6].astype(int), ref["steps"])
        assert np.max(np.abs(bt.particles(b) - o.h)) <= 2e-4, b


def test_paths(torch_cuda):
    """Heun / Student-t filters: one workgroup below 256 particles, the grid path
    from 256 up (the cooperative kernel does not take them), one workgroup on
    request up to its capacity, the grid path alone above it."""
    assert Filter(config.plain_config(n=128, use_heun=1)).path() == "single"
    for n in (256, 512, 1024, 4096, 16384):
        assert Filter(config.plain_config(n=n, use_heun=1, use_student_t_state=1)).path() == "grid", n
    assert Filter(config.plain_config(n=512, use_student_t_state=1), path="single").path() == "single"
    big = Filter(config.plain_config(n=8192, use_heun=1))
    assert big.L.svpf_set_grid_path(big.sp, 0) == _lib.SVPF_RC_UNSUPPORTED
    big.initialize(3)
    big.step(0.1, 0.0)
    assert big.L.svpf_last_error(big.sp) == _lib.SVPF_RC_OK


def test_grid_heun_student_n16384_one_tick(torch_cuda):
    """The largest filter with both features, one tick from identical state with
    pre-drawn noise (2 Stein iterations keep the FP64 oracle within a minute)."""
    n = 16384
    cfg = config.full_feature_config(n=n, stein_min_steps=2, stein_max_steps=2, use_heun=1,
                                     use_student_t_state=1, nu_state=6.0)
    h0 = data.initial_particles(n, -1.0, 0.98, 0.2, seed=41)
    v0 = np.full(n, 0.05, dtype=np.float32)
    noise = data.predraw_noise(1, n, 2, seed=42, mim=True)
    y = np.array([0.9], dtype=np.float32)
    print("grid heun + student-t n=16384", flush=True)
    f = Filter(cfg)
    f.set_particles(h0, v=v0)
    f.set_noise(dev(torch_cuda, noise["z_pred"]), dev(torch_cuda, noise["xi"]), dev(torch_cuda, noise["u_pred"]))
    f.step(float(y[0]), 0.0)
    assert f.path() == "grid"
    o = oracle.OracleFilter(cfg, h0, v0=v0.astype(np.float64))
    out = o.run(y, noise)
    assert np.max(np.abs(f.particles() - o.h)) <= 1e-4
    compare(f.last_outputs()[None].astype(np.float64), out)
