"""Parity of the single-workgroup fused step (svpf_fully_fused_step_v2, the C
ABI of include/svpf_fully_fused_v2.h) against the FP64 oracle's restatement of
src/svpf_fully_fused_v2.cu:90-712.

Tolerances are those of the graph-path tests: per-particle |dh| <= 1e-4 after
one call from identical state, outputs to 1e-4 (loglik 1e-3), h_mean RMSE <=
1e-4 over short trajectories driven the way the reference's own driver does
(test/test_fully_fused_v2.cu:148-186: h_mean, vol and ksd of one call feed the
next)."""
import numpy as np
import pytest

import oracle
from svpf_amd.fused_v2 import FusedV2, v2_args

pytestmark = pytest.mark.gpu

CONFIGS = {
    # the reference driver's arguments: use_newton and use_full_newton both set
    "driver_defaults": dict(),
    "plain": dict(use_newton=0, use_full_newton=0, use_guide=0, use_guided_predict=0, use_rejuvenation=0),
    "local_newton": dict(use_newton=1, use_full_newton=0),
    "full_newton_only": dict(use_newton=0, use_full_newton=1, stein_sign_mode=1),
    "local_params_adaptive_mu": dict(delta_rho=0.02, delta_sigma=0.1, use_adaptive_mu=1, mu_ema_alpha=0.05,
                                     use_guide_preserving=0, gamma=-0.3),
}


def _h0(n, seed, mu=-4.5, sd=0.5):
    return (mu + sd * np.random.default_rng(seed).standard_normal(n)).astype(np.float32)


def _noise(n, rows, seed):
    rng = np.random.default_rng(seed)
    return {"z_pred": rng.standard_normal(n, dtype=np.float32), "u_pred": rng.random(n, dtype=np.float32),
            "xi": rng.standard_normal((rows, n), dtype=np.float32), "u_rej": rng.random(n, dtype=np.float32),
            "z_rej": rng.standard_normal(n, dtype=np.float32)}


def _dev(torch, nz):
    m = {"z_pred": "z_state", "u_pred": "u_jump", "xi": "z_diffusion", "u_rej": "u_rejuv", "z_rej": "z_rejuv"}
    return {m[k]: torch.tensor(v, device="cuda") for k, v in nz.items()}


@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_one_call_from_identical_state(torch_cuda, name):
    n = 512
    a = v2_args(n=n, timestep=12, y_t=0.3, y_prev=-0.1, ksd_prev=0.4, **CONFIGS[name])
    h0, hp0 = _h0(n, 1), _h0(n, 2)
    v0 = np.full(n, 0.05, np.float32)
    io0 = {"guide_mean": -4.4, "guide_var": 0.02, "mu_state": -4.3}
    nz = _noise(n, a["n_stein_steps"], 3)
    print(f"fused v2 one call {name}", flush=True)
    f = FusedV2(h0, hp0, v0, **io0)
    assert f.step(a, _dev(torch_cuda, nz)) == 0
    h, hp, v = h0.astype(np.float64), hp0.astype(np.float64), v0.astype(np.float64)
    g, lw, cv, io = oracle.v2_step(a, h, hp, v, io0, noise=nz)
    out = f.outputs()
    assert np.max(np.abs(f.h.cpu().numpy() - h)) <= 1e-4
    assert np.max(np.abs(f.v.cpu().numpy() - v)) <= 1e-4
    assert np.array_equal(f.h_prev.cpu().numpy(), hp.astype(np.float32))
    assert np.max(np.abs(f.grad.cpu().numpy() - g)) <= 1e-3
    assert np.max(np.abs(f.logw.cpu().numpy() - lw)) <= 1e-4
    if a["use_newton"] or a["use_full_newton"]:
        assert np.max(np.abs(f.inv_hess.cpu().numpy() - cv) / np.maximum(1.0, cv)) <= 1e-4
    for k in ("h_mean", "vol", "bandwidth", "ksd", "guide_mean", "guide_var", "mu_state"):
        assert abs(out[k] - io[k]) <= 1e-4 * max(1.0, abs(io[k])), k
    assert abs(out["loglik"] - io["loglik"]) <= 1e-3


@pytest.mark.parametrize("name", ["driver_defaults", "local_params_adaptive_mu"])
def test_trajectory_matches_oracle(torch_cuda, name):
    """20 calls on the reference driver's synthetic data shape (mu = -4.5,
    rho = 0.98, sigma_z = 0.1), each call's h_mean / vol / ksd feeding the next,
    from a spread cloud with a warm RMSProp state.  These configurations are
    well conditioned there: the FP64 oracle started from particles perturbed by
    1e-7 relative moves h_mean by < 2e-7.  Others are not (vanilla SVGD: 2.5e-4;
    full Newton alone with the paper sign: 0.17; the driver's start with every
    particle at -4.5 and v = 0: 6e-5), so they are checked call by call above."""
    torch = torch_cuda
    n, T = 256, 20
    base = v2_args(n=n, **CONFIGS[name])
    rng = np.random.default_rng(5)
    ht = np.zeros(T)
    ht[0] = -4.5
    for t in range(1, T):
        ht[t] = -4.5 + 0.98 * (ht[t - 1] + 4.5) + 0.1 * rng.standard_normal()
    y = (np.exp(ht / 2) * rng.standard_normal(T)).astype(np.float32)
    y[T // 2] *= 5.0
    h0 = (-4.5 + 0.3 * np.random.default_rng(3).standard_normal(n)).astype(np.float32)
    v0 = np.full(n, 0.05, np.float32)
    f = FusedV2(h0, v0=v0, mu_state=-4.5)
    h, hp, v = h0.astype(np.float64), h0.astype(np.float64), v0.astype(np.float64)
    io = {"guide_mean": 0.0, "guide_var": 0.0, "mu_state": -4.5}
    gpu_prev = dict(y_prev=0.0, h_mean_prev=-4.5, vol_prev=float(np.exp(-2.25)), ksd_prev=0.1)
    orc_prev = dict(gpu_prev)
    dh = []
    print(f"fused v2 trajectory {name}", flush=True)
    for t in range(T):
        nz = _noise(n, base["n_stein_steps"], 100 + t)
        assert f.step(dict(base, timestep=t, y_t=float(y[t]), **gpu_prev), _dev(torch, nz)) == 0
        out = f.outputs()
        _, _, _, io = oracle.v2_step(dict(base, timestep=t, y_t=float(y[t]), **orc_prev), h, hp, v, io, noise=nz)
        dh.append(out["h_mean"] - io["h_mean"])
        gpu_prev = dict(y_prev=float(y[t]), h_mean_prev=out["h_mean"], vol_prev=out["vol"], ksd_prev=out["ksd"])
        orc_prev = dict(y_prev=float(y[t]), h_mean_prev=io["h_mean"], vol_prev=io["vol"], ksd_prev=io["ksd"])
    assert np.sqrt(np.mean(np.square(dh))) <= 1e-4
    assert np.max(np.abs(f.h.cpu().numpy() - h)) <= 2e-4


@pytest.mark.parametrize("n", [1, 2, 64, 100, 256, 300, 1000, 1024])
def test_particle_counts(torch_cuda, n):
    a = v2_args(n=n, timestep=3, y_t=0.2, n_stein_steps=4, n_anneal_steps=2)
    h0, hp0 = _h0(n, 7), _h0(n, 8)
    v0 = np.full(n, 0.05, np.float32)
    nz = _noise(n, 4, 9)
    print(f"fused v2 n={n}", flush=True)
    f = FusedV2(h0, hp0, v0)
    assert f.step(a, _dev(torch_cuda, nz)) == 0
    h, hp, v = h0.astype(np.float64), hp0.astype(np.float64), v0.astype(np.float64)
    _, _, _, io = oracle.v2_step(a, h, hp, v, {}, noise=nz)
    out = f.outputs()
    assert np.max(np.abs(f.h.cpu().numpy() - h)) <= 1e-4
    assert abs(out["h_mean"] - io["h_mean"]) <= 1e-4
    assert abs(out["ksd"] - io["ksd"]) <= 1e-4 * max(1.0, io["ksd"])


def test_counter_generator_matches_oracle(torch_cuda):
    n = 300
    a = v2_args(n=n, timestep=14, y_t=0.25, ksd_prev=0.4)
    h0, hp0 = _h0(n, 10), _h0(n, 11)
    v0 = np.full(n, 0.05, np.float32)
    print("fused v2 counter generator", flush=True)
    f = FusedV2(h0, hp0, v0)
    assert f.step(a, None, seed=1234, stream_id=7) == 0
    h, hp, v = h0.astype(np.float64), hp0.astype(np.float64), v0.astype(np.float64)
    _, _, _, io = oracle.v2_step(a, h, hp, v, {}, noise=None, seed=1234, filter_id=7)
    assert np.max(np.abs(f.h.cpu().numpy() - h)) <= 1e-4
    assert abs(f.outputs()["h_mean"] - io["h_mean"]) <= 1e-4


def test_rejects_bad_calls(torch_cuda):
    from svpf_amd import _lib

    L = _lib.load()
    f = FusedV2(_h0(64, 1))
    assert f.step(v2_args(n=64)) == 0  # counter generator, reference driver arguments
    assert not L.svpf_fully_fused_v2_supported(1025)
    with pytest.raises(ValueError):
        FusedV2(np.zeros(1025, np.float32))
    # n = 0 and NULL buffers are refused before any launch
    vals = []
    for name, kind in _lib.V2_ARGS:
        vals.append(None if kind in ("p", "noise") else (0 if kind == "i" else 0.0))
    assert L.svpf_fully_fused_step_v2(*vals) != 0
