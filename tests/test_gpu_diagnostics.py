"""Diagnostics beside the step on the GPU, through the C ABI: svpf_diag_check
and the nu / sigma scores against the float64 oracle (oracle/diagnostics.py)
on the device's own arrays after real steps of each kernel path, and on
planted arrays for the error cases (NaN, Inf, weight underflow, one live
weight).  Tolerances: ESS and the scores to 1e-4 relative (float32 sums over
up to 16384 particles), h_mean to 1e-5; statuses exact."""
import ctypes

import numpy as np
import pytest

from oracle import diagnostics as od
from svpf_amd import config, data
from svpf_amd import diagnostics as dg
from svpf_amd.filter import Filter
from test_gpu_parity import d2h

pytestmark = pytest.mark.gpu


def h2d(ptr, a):
    a = np.ascontiguousarray(a, dtype=np.float32)
    rc = ctypes.CDLL("libamdhip64.so").hipMemcpy(ctypes.c_void_p(ctypes.cast(ptr, ctypes.c_void_p).value),
                                                 a.ctypes.data_as(ctypes.c_void_p), ctypes.c_size_t(a.nbytes), 1)
    assert rc == 0


def _close(a, b, rel=1e-4):
    assert abs(a - b) <= rel * max(1.0, abs(b)), (a, b)


def _stepped(n, T=12, seed=0):
    cfg = config.plain_config(n=n)
    f = Filter(cfg)
    f.initialize(seed)
    y, _ = data.sv_series(T, seed=seed)
    yp = 0.0
    for t in range(T):
        f.step(float(y[t]), yp)
        yp = float(y[t])
    return f, cfg, y


@pytest.mark.parametrize("n,path", [(100, "single"), (512, "coop"), (4096, "grid")])
def test_checks_and_scores_after_steps(torch_cuda, n, path):
    print(f"diagnostics after steps n={n}", flush=True)
    f, cfg, y = _stepped(n)
    assert f.path() == path
    h = d2h(f.state.h, n)
    lw = d2h(f.state.log_weights, n)
    hd = dg.Health(f)
    st = hd.check()
    want, m = od.health_check(h, lw)
    assert st == want, (st, hd.message())
    _close(hd.d.ess, m["ess"])
    _close(hd.d.ess_ratio, m["ess_ratio"])
    assert abs(hd.d.h_mean - m["h_mean"]) <= 1e-5
    _close(hd.d.vol_estimate, m["vol_estimate"])
    assert hd.d.steps_processed == 1
    sc = dg.Scores(f)
    g, z2 = sc.nu(float(y[-1]))
    og, oz = od.nu_score(h, lw, float(y[-1]), cfg["nu"])
    _close(g, og)
    _close(z2, oz)
    # sigma score over h_pred given h_prev, at the state's effective sigma
    hp, hv = d2h(f.state.h_pred, n), d2h(f.state.h_prev, n)
    sig = f.state.sigma_z_effective if f.state.sigma_z_effective >= 1e-6 else cfg["sigma_z"]
    g, e2 = sc.sigma()
    og, oe = od.sigma_score(hp, hv, lw, cfg["mu"], cfg["rho"], sig)
    _close(g, og)
    _close(e2, oe)
    # plain predict: h_pred = mu + rho (h_prev - mu) + sigma z, so eps / sigma is the N(0, 1) draw
    eps = (hp.astype(np.float64) - cfg["mu"] - cfg["rho"] * (hv - cfg["mu"])) / cfg["sigma_z"]
    assert 0.6 < np.mean(eps * eps) < 1.5 and abs(np.mean(eps)) < 0.3
    sc.close()


@pytest.mark.parametrize("n", [2, 1000, 16384])
def test_planted_arrays(torch_cuda, n):
    print(f"diagnostics planted n={n}", flush=True)
    f = Filter(config.plain_config(n=n))
    f.initialize(3)
    rng = np.random.default_rng(n)
    h = rng.normal(-1.0, 0.4, n).astype(np.float32)
    cases = {
        "uniform": (h, np.full(n, -3.0, np.float32)),
        "random": (h, rng.normal(0.0, 2.0, n).astype(np.float32)),
        "one_live": (h, np.where(np.arange(n) == n // 2, 1.5, -np.inf).astype(np.float32)),
        "nan_h": (np.where(np.arange(n) == n - 1, np.nan, h).astype(np.float32), np.zeros(n, np.float32)),
        "nan_w": (h, np.where(np.arange(n) == 0, np.nan, 0.0).astype(np.float32)),
        "inf_h": (np.where(np.arange(n) == n // 3, np.inf, h).astype(np.float32), np.zeros(n, np.float32)),
        "underflow": (h, np.full(n, -np.inf, np.float32)),
        "vol_high": (np.full(n, 5.0, np.float32), np.zeros(n, np.float32)),
    }
    for name, (hh, ll) in cases.items():
        h2d(f.state.h, hh)
        h2d(f.state.log_weights, ll)
        hd = dg.Health(f)
        st = hd.check()
        want, m = od.health_check(hh, ll)
        assert st == want, (name, st, hd.message())
        if m:
            _close(hd.d.ess, m["ess"])
            assert abs(hd.d.h_mean - m["h_mean"]) <= 1e-5, name
        if name in ("uniform", "random", "one_live"):
            sc = dg.Scores(f)
            g, z2 = sc.nu(0.04)
            og, oz = od.nu_score(hh, ll, 0.04, 5.0)
            _close(g, og)
            _close(z2, oz)
            sc.close()
    assert od.health_check(*cases["one_live"])[1]["ess"] == 1.0


def test_tracking_state_and_log(torch_cuda, tmp_path):
    print("diagnostics tracking", flush=True)
    f, cfg, y = _stepped(256, T=4)
    log = tmp_path / "nu.csv"
    sc = dg.Scores(f, log_path=str(log))
    ema_g, ema_z = 0.0, 1.0
    yp = float(y[-1])
    ys = data.sv_series(6, seed=9)[0]
    for t in range(6):
        f.step(float(ys[t]), yp)
        yp = float(ys[t])
        g, z2 = sc.track(float(ys[t]), t)
        og, oz = od.nu_score(d2h(f.state.h, 256), d2h(f.state.log_weights, 256), float(ys[t]), cfg["nu"])
        _close(g, og)
        _close(z2, oz)
        ema_g, ema_z = 0.9 * ema_g + 0.1 * g, 0.9 * ema_z + 0.1 * z2
        assert abs(sc.state.nu_gradient_ema - ema_g) <= 1e-6 and abs(sc.state.z_sq_ema - ema_z) <= 1e-6
        # the diagnostic's own device buffers hold the last result
        assert d2h(sc.state.d_nu_grad, 2).tolist() == [np.float32(g), np.float32(z2)]
    sc.close()
    rows = log.read_text().strip().split("\n")
    assert rows[0] == "t,y_t,vol,z_sq,nu_grad,nu_grad_ema,z_sq_ema,shock_state" and len(rows) == 7
    assert rows[-1].split(",")[0] == "5"


def test_snapshot_particles(torch_cuda):
    torch = torch_cuda
    print("diagnostics snapshot", flush=True)
    f, _, _ = _stepped(700, T=3)
    buf = torch.zeros(700, device="cuda")
    dg.bind().svpf_snapshot_particles(f.sp, ctypes.c_void_p(buf.data_ptr()))
    torch.cuda.synchronize()
    assert np.array_equal(buf.cpu().numpy(), d2h(f.state.h, 700))
