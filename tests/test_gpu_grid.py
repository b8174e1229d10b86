"""Parity of the grid path (one large filter as a per-tick sequence of grid-wide
kernels) against the FP64 oracle and against the single-workgroup kernel.

The grid path serves filters above the single-workgroup LDS capacity (up to
16384 particles, the full-feature regime-shift workload at N=16384); it can be
forced on at any n, which is how it is checked against the oracle at sizes the
oracle runs in seconds.  Tolerances are those of the single-workgroup tests:
per-particle |dh| <= 1e-4 after one tick from identical state, h_mean RMSE <=
1e-4 over short trajectories, the Stein iteration counts equal, and the median
bandwidth bit-exact against a float32 sort of the device's predicted particles.
"""
import numpy as np
import pytest

import oracle
from svpf_amd import config, data
from svpf_amd.filter import Filter
from test_gpu_parity import FEATURE_CONFIGS, compare, d2h, dev, run_gpu, run_oracle

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [512, 5000])
def test_grid_one_tick_from_identical_state(torch_cuda, n):
    cfg = config.plain_config(n=n)
    h0 = data.initial_particles(n, -1.0, 0.97, 0.15, seed=5)
    noise = data.predraw_noise(1, n, 5, seed=6)
    y = np.array([0.42], dtype=np.float32)
    print(f"grid one tick n={n}", flush=True)
    f, rec = run_gpu(torch_cuda, cfg, y, noise, h0, path="grid")
    assert f.path() == "grid"
    o, out = run_oracle(cfg, y, noise, h0)
    assert np.max(np.abs(f.particles() - o.h)) <= 1e-4
    assert np.max(np.abs(f.rmsprop_state() - o.v)) <= 1e-4
    hp = d2h(f.state.h_pred, n)
    assert np.float32(rec[0, 7]) == oracle.median_r2_f32(hp)
    compare(rec, out)


def test_grid_median_heavy_ties(torch_cuda):
    """Predicted particles with heavy ties (initial values rounded to 0.1, zero
    predict and Langevin noise): the selection needs more rounds than the grid
    launches at this size, so the final workgroup finishes it; the bandwidth
    stays bit-exact and the tick matches the oracle."""
    n = 3000
    cfg = config.plain_config(n=n)
    rng = np.random.default_rng(41)
    h0 = np.round(rng.standard_normal(n) * 0.6 - 1.0, 1).astype(np.float32)
    noise = {"z_pred": np.zeros((1, n), np.float32), "xi": np.zeros((1, 5, n), np.float32)}
    y = np.array([0.3], dtype=np.float32)
    print("grid median heavy ties", flush=True)
    f, rec = run_gpu(torch_cuda, cfg, y, noise, h0, path="grid")
    hp = d2h(f.state.h_pred, n)
    assert np.unique(hp).size < n // 10
    assert np.float32(rec[0, 7]) == oracle.median_r2_f32(hp)
    o, out = run_oracle(cfg, y, noise, h0)
    assert np.max(np.abs(f.particles() - o.h)) <= 1e-4
    compare(rec, out)


@pytest.mark.parametrize("name", sorted(FEATURE_CONFIGS))
def test_grid_feature_configs_short_run(torch_cuda, name):
    n, T = 300, 25
    cfg = config.make_config(n=n, mu=-1.0, rho=0.97, sigma_z=0.15, gamma=-0.2, **FEATURE_CONFIGS[name])
    y, _ = data.sv_series(T, mu=-1.0, rho=0.97, sigma_z=0.15, nu=5.0, seed=3)
    y[12] *= 6.0
    rows = cfg["stein_max_steps"]
    noise = data.predraw_noise(T, n, rows, seed=4, mim=True, rejuv=True)
    h0 = data.initial_particles(n, -1.0, 0.97, 0.15, seed=5)
    print(f"grid feature config {name}", flush=True)
    f, rec = run_gpu(torch_cuda, cfg, y, noise, h0, path="grid")
    o, out = run_oracle(cfg, y, noise, h0)
    compare(rec, out, tol_hmean=1e-4, tol_ll=2e-3, tol_vol=2e-4)


@pytest.mark.parametrize("n", [2, 3, 65, 1000, 4097, 6000])
def test_grid_particle_counts(torch_cuda, n):
    T = 6
    cfg = config.plain_config(n=n)
    y, _ = data.sv_series(T, seed=8)
    noise = data.predraw_noise(T, n, 5, seed=9)
    h0 = data.initial_particles(n, -1.0, 0.97, 0.15, seed=10)
    print(f"grid particle count n={n}", flush=True)
    f, rec = run_gpu(torch_cuda, cfg, y, noise, h0, path="grid")
    o, out = run_oracle(cfg, y, noise, h0)
    compare(rec, out, tol_ll=1e-3)
    assert np.max(np.abs(f.particles() - o.h)) <= 2e-4


def test_grid_equals_single_workgroup_full_features(torch_cuda):
    """Full feature set (MIM, asymmetric rho, EKF guide, SVLD, KSD-adaptive beta,
    early stop) at N=4096 on a regime-shift series, counter generator: the grid
    path and the single-workgroup kernel agree."""
    n, T = 4096, 30
    cfg = config.full_feature_config(n=n)
    y, _ = data.regime_shift_series(T=T, t_shift=15, seed=2)
    h0 = data.initial_particles(n, -1.0, 0.98, 0.2, seed=3)
    recs = {}
    for path in ("single", "grid"):
        print(f"grid vs single, path {path}", flush=True)
        f = Filter(cfg, path=path)
        f.set_particles(h0)
        assert f.path() == path
        r = np.zeros((T, 8))
        yp = 0.0
        for t in range(T):
            f.step(float(y[t]), yp)
            r[t] = f.last_outputs()
            yp = float(y[t])
        recs[path] = r
    a, b = recs["single"], recs["grid"]
    assert np.sqrt(np.mean((a[:, 2] - b[:, 2]) ** 2)) <= 1e-4
    assert np.max(np.abs(a[:, 0] - b[:, 0])) <= 1e-3
    assert np.array_equal(a[:, 6], b[:, 6])


def test_grid_n16384_one_tick_matches_oracle(torch_cuda):
    """The full-feature workload at its real size, one tick from identical state
    with pre-drawn noise (3 Stein iterations keep the FP64 oracle within seconds)."""
    n = 16384
    cfg = config.full_feature_config(n=n, stein_min_steps=2, stein_max_steps=3)
    h0 = data.initial_particles(n, -1.0, 0.98, 0.2, seed=31)
    noise = data.predraw_noise(1, n, 3, seed=32, mim=True)
    y = np.array([0.9], dtype=np.float32)
    print("grid n=16384 one tick", flush=True)
    f, rec = run_gpu(torch_cuda, cfg, y, noise, h0)
    assert f.path() == "grid"
    o, out = run_oracle(cfg, y, noise, h0)
    assert np.max(np.abs(f.particles() - o.h)) <= 1e-4
    hp = d2h(f.state.h_pred, n)
    assert np.float32(rec[0, 7]) == oracle.median_r2_f32(hp)
    compare(rec, out)


def test_grid_n16384_deterministic_device_run(torch_cuda):
    """N=16384 over 20 ticks of the regime-shift series with the counter
    generator: sane records, bitwise identical per-tick (graph) and device-loop
    runs (the grid reductions use no atomics)."""
    torch = torch_cuda
    n, T = 16384, 20
    cfg = config.full_feature_config(n=n)
    y, _ = data.regime_shift_series(T=T, t_shift=10, seed=4)
    h0 = data.initial_particles(n, -1.0, 0.98, 0.2, seed=5)
    print("grid n=16384 per-tick run", flush=True)
    f1 = Filter(cfg)
    f1.set_particles(h0)
    rec = np.zeros((T, 8), dtype=np.float32)
    yp = 0.0
    for t in range(T):
        f1.step(float(y[t]), yp)
        rec[t] = f1.last_outputs()
        yp = float(y[t])
    assert np.isfinite(rec).all()
    assert ((rec[:, 6] >= 1) & (rec[:, 6] <= cfg["stein_max_steps"])).all()
    assert (rec[:, 4] > 0).all()
    print("grid n=16384 device run", flush=True)
    f2 = Filter(cfg)
    f2.set_particles(h0)
    ll = torch.zeros(T, device="cuda")
    vol = torch.zeros(T, device="cuda")
    f2.run_device(dev(torch, y), ll, vol)
    assert np.array_equal(ll.cpu().numpy(), rec[:, 0])
    assert np.array_equal(vol.cpu().numpy(), rec[:, 1])
    assert np.array_equal(f1.particles(), f2.particles())


def test_grid_path_switching(torch_cuda):
    from svpf_amd import _lib

    f = Filter(config.plain_config(n=256))
    assert f.path() == "coop"  # automatic choice at 256 <= n < 1024 with the cooperative kernel's features
    assert f.L.svpf_set_grid_path(f.sp, 1) == _lib.SVPF_RC_OK and f.path() == "grid"
    assert f.L.svpf_set_grid_path(f.sp, 0) == _lib.SVPF_RC_OK and f.path() == "coop"
    f.state.use_persistent_kernel = 0
    assert f.path() == "single"
    assert f.L.svpf_set_grid_path(f.sp, -1) == _lib.SVPF_RC_OK and f.path() == "grid"  # no cooperative kernel
    assert f.L.svpf_set_grid_path(f.sp, 2) == _lib.SVPF_RC_INVALID_ARGUMENT
    e = Filter(config.full_feature_config(n=512))  # KSD early stop: no cooperative kernel
    assert e.path() == "grid"
    assert Filter(config.plain_config(n=128)).path() == "single"
    assert Filter(config.plain_config(n=1024)).path() == "grid"
    g = Filter(config.plain_config(n=16384))
    assert g.path() == "grid"
    assert g.L.svpf_set_grid_path(g.sp, 0) == _lib.SVPF_RC_UNSUPPORTED
    assert not g.L.svpf_create(16385, 5, 5.0, None)
