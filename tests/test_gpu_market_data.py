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
