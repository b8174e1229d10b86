"""The reference-shaped Python module (svpf_amd.pysvpf) on the GPU: the
production defaults of py-torch/svpf_python.cpp:90-176 run through the C ABI
and match the FP64 oracle tick by tick on the counter generator."""
import numpy as np
import pytest

import oracle
from svpf_amd import config, data, pysvpf

pytestmark = pytest.mark.gpu


def cfg_from_state(f):
    st, p = f._st, f._params
    cfg = {}
    for k in config.REFERENCE_DEFAULTS:
        if k == "n":
            cfg[k] = st.n_particles
        elif k == "stein_sign":
            cfg[k] = st.stein_repulsive_sign
        elif k in ("mu", "rho", "sigma_z", "gamma"):
            cfg[k] = getattr(p, k)
        else:
            cfg[k] = getattr(st, k)
    return config.make_config(**cfg)


def test_production_defaults_match_oracle(torch_cuda):
    n, T = 256, 30
    f = pysvpf.SVPF(n_particles=n, n_stein=8, nu=30.0)
    f.initialize(rho=0.97, sigma_z=0.15, mu=-3.5, gamma=-0.5, seed=42)
    h0 = f.get_particles()
    cfg = cfg_from_state(f)
    assert cfg["use_full_newton"] == 1 and cfg["use_guided"] == 1 and cfg["stein_max_steps"] == 16
    o = oracle.OracleFilter(cfg, h0, mu_state=-3.5, mu_var=0.01)
    y, _ = data.sv_series(T, mu=-3.5, rho=0.97, sigma_z=0.15, nu=5.0, seed=31)
    y[15] *= 5.0
    d = []
    yp = 0.0
    for t in range(T):
        r = f.step(float(y[t]))
        q = o.tick(float(y[t]), yp, seed=42)
        d.append(r.h_mean - q["h_mean"])
        assert abs(r.log_lik - q["loglik"]) <= 2e-3
        yp = float(y[t])
    assert np.sqrt(np.mean(np.square(d))) <= 1e-4
    assert np.max(np.abs(f.get_particles() - o.h)) <= 5e-4


def test_run_and_errors(torch_cuda):
    f = pysvpf.SVPF(n_particles=128)
    with pytest.raises(RuntimeError, match="not initialized"):
        f.step(0.01)
    f.initialize(seed=7)
    p0 = f.get_particles()
    y, _ = data.sv_series(20, mu=-3.5, seed=32)
    vol, hm, ll = f.run(y)
    assert vol.shape == hm.shape == ll.shape == (20,)
    assert np.isfinite(vol).all() and np.isfinite(hm).all() and np.isfinite(ll).all()
    assert f.is_graph_captured()
    with pytest.raises(RuntimeError, match="1-dimensional"):
        f.run(np.zeros((2, 2), np.float32))
    # reset re-draws the particles as initialize did; like the reference's svpf_initialize it
    # keeps the device scalars of the step (src/svpf_optimized_graph.cu:481-494)
    f.reset(seed=7)
    assert np.array_equal(f.get_particles(), p0)
    vol2, _, _ = f.run(y)
    assert np.isfinite(vol2).all()
