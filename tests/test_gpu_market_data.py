"""The filter on real return series: the reference's SPY 2008-crisis and
2020-covid series (fixtures in tests/golden/market_data/, read with the
market-data loaders), through each kernel path, against the FP64 oracle on the
same pre-drawn noise.  mu is set from the data the way the reference's own
real-data tests do (2 log of the realized std, test/test_svpf_gradient_sigma.cu:64-74).
Bar: h_mean RMSE <= 1e-4 over the whole series (these series are well
conditioned: the oracle started from particles perturbed by 1e-7 relative
moves h_mean by < 1e-6)."""
import os

import numpy as np
import pytest

from svpf_amd import config, data
from test_gpu_parity import compare, run_gpu, run_oracle

pytestmark = pytest.mark.gpu

FIX = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "market_data")


@pytest.mark.parametrize("name,n,path", [("spy_2008_crisis", 128, "single"), ("spy_2008_crisis", 512, "coop"),
                                         ("spy_2020_covid", 1024, "grid")])
def test_real_series_matches_oracle(torch_cuda, name, n, path):
    y = data.load_returns_binary(os.path.join(FIX, name + ".bin"))
    mu = float(2 * np.log(np.std(y.astype(np.float64))))
    cfg = config.plain_config(n=n, mu=mu, rho=0.98, sigma_z=0.15, nu=5.0)
    noise = data.predraw_noise(len(y), n, 5, seed=21)
    h0 = data.initial_particles(n, mu, 0.98, 0.15, seed=22)
    print(f"market data {name} n={n} path={path} T={len(y)}", flush=True)
    f, rec = run_gpu(torch_cuda, cfg, y, noise, h0, path=path)
    assert f.path() == path
    o, out = run_oracle(cfg, y, noise, h0)
    compare(rec, out, tol_ll=1e-3)


@pytest.mark.parametrize("name", ["spy_2008_crisis", "spy_2020_covid"])
@pytest.mark.parametrize("n,path", [(200, "single"), (512, "coop"), (1024, "grid")])
def test_real_series_full_feature_set_matches_oracle(torch_cuda, name, n, path):
    """The production feature set (guided MIM predict with asymmetric rho, EKF guide with
    adaptive strength, adaptive mu and sigma, full Newton, rejuvenation; the configuration
    the reference's pysvpf set_defaults and two-scale filter use) on the real series, every
    path against the FP64 oracle.  Same bars as the plain case; the FP64 oracle moves h_mean
    by < 4e-7 when its start is perturbed by 1e-7 relative on these series and settings."""
    from test_gpu_parity import FEATURE_CONFIGS

    y = data.load_returns_binary(os.path.join(FIX, name + ".bin"))
    mu = float(2 * np.log(np.std(y.astype(np.float64))))
    cfg = config.make_config(n=n, mu=mu, rho=0.98, sigma_z=0.15, gamma=-0.2, **FEATURE_CONFIGS["production"])
    noise = data.predraw_noise(len(y), n, cfg["stein_max_steps"], seed=21, mim=True, rejuv=True)
    h0 = data.initial_particles(n, mu, 0.98, 0.15, seed=22)
    print(f"market data {name} production n={n} path={path} T={len(y)}", flush=True)
    f, rec = run_gpu(torch_cuda, cfg, y, noise, h0, path=path)
    assert f.path() == path
    o, out = run_oracle(cfg, y, noise, h0)
    compare(rec, out, tol_ll=2e-3, tol_vol=2e-4)
    f.close()
