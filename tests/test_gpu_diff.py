"""The differentiable form on the GPU (svpf_diff_forward / svpf_diff_backward
through the C ABI): the forward against the FP64 oracle, the hand-written
adjoint against the float64 autograd of the dense PyTorch restatement
(oracle/torch_svpf.py, itself pinned to FP64 central differences) and
against FP64 central differences of the oracle, all to 1e-3
relative, and the gradient's independence of the checkpoint interval."""
import numpy as np
import pytest
import torch

import oracle
from oracle import torch_svpf
from svpf_amd import config, data
from svpf_amd.diff import DiffBatch, predictive_loglik

pytestmark = pytest.mark.gpu

THETAS = [(-1.0, 0.97, 0.15, 5.0), (-0.5, 0.93, 0.30, 8.0), (-1.6, 0.99, 0.08, 20.0)]
# Gradient checks use well-conditioned filters: on some series a tight cloud (e.g. sigma_z = 0.08,
# rho = 0.99) leaves RMSProp's v near 0, where eps phi / sqrt(v + 1e-6) amplifies rounding so much
# that the gradient of the FP32 filter differs from the FP64 one by a factor (a numpy emulation of
# this adjoint in float32 against float64 shows the same), whatever the adjoint.
GRAD_THETAS = [(-1.0, 0.97, 0.15, 5.0), (-0.5, 0.93, 0.30, 8.0), (-1.6, 0.98, 0.12, 20.0)]


def inputs(n, T, S, thetas, seed=0):
    ys, h0s, zs, xis = [], [], [], []
    for b, (mu, rho, sz, nu) in enumerate(thetas):
        y, _ = data.sv_series(T, mu=mu, rho=rho, sigma_z=sz, nu=nu, seed=seed + 10 * b)
        nz = data.predraw_noise(T, n, S, seed=seed + 10 * b + 1)
        ys.append(y)
        h0s.append(data.initial_particles(n, mu, rho, sz, seed=seed + 10 * b + 2))
        zs.append(nz["z_pred"])
        xis.append(nz["xi"])
    return np.stack(ys), np.stack(h0s), np.stack(zs), np.stack(xis)


def make_batch(n, T, S, thetas, C, gamma=0.0, sign=False, seed=0, drift_budget_gb=32.0):
    y, h0, z, xi = inputs(n, T, S, thetas, seed)
    bt = DiffBatch(y, h0, n_stein=S, checkpoint_every=C, noise={"z": z, "xi": xi}, gamma=gamma, paper_sign=sign,
                   drift_budget_gb=drift_budget_gb)
    return bt, (y, h0, z, xi)


def cfg_of(n, S, th, gamma=0.0, sign=0):
    mu, rho, sz, nu = th
    return config.plain_config(n=n, n_stein=S, mu=mu, rho=rho, sigma_z=sz, nu=nu, gamma=gamma, stein_sign=sign)


def oracle_pred_ll(cfg, y, h0, z, xi):
    o = oracle.OracleFilter(cfg, h0)
    out = o.run(y, {"z_pred": z, "xi": xi})
    return out["pred_loglik"], o


@pytest.mark.parametrize("n", [48, 300, 1024])
def test_forward_matches_oracle(torch_cuda, n):
    T, S = 12, 5
    bt, (y, h0, z, xi) = make_batch(n, T, S, THETAS, C=5)
    pl = bt.forward(torch.tensor(THETAS, device="cuda")).cpu().numpy()
    hT = bt.hT.cpu().numpy()
    for b, th in enumerate(THETAS):
        ref, o = oracle_pred_ll(cfg_of(n, S, th), y[b], h0[b], z[b], xi[b])
        assert np.max(np.abs(pl[b] - ref)) <= 1e-4 * max(1.0, np.abs(ref).max()), b
        # a tight cloud (sigma_z = 0.08) of few particles leaves RMSProp's v near 0, where the step
        # eps phi / sqrt(v + 1e-6) amplifies FP32 rounding of phi: the filter engine deviates by the
        # same 5e-4 there, so small n gets 1e-3
        assert np.max(np.abs(hT[b] - o.h)) <= (1e-4 if n >= 256 else 1e-3), b


def torch_grad(n, S, th, y, h0, z, xi, gamma=0.0, sign=0):
    cfg = cfg_of(n, S, th, gamma, sign)
    dt = torch.float64
    theta = [torch.tensor(v, dtype=dt, requires_grad=True) for v in th]
    pl, _, _ = torch_svpf.run(theta, torch.tensor(y, dtype=dt), torch.tensor(z, dtype=dt), torch.tensor(xi, dtype=dt),
                              torch.tensor(h0, dtype=dt), cfg)
    (-pl.sum()).backward()
    return np.array([t.grad.item() for t in theta])


def rel_err(g, ref):
    return np.max(np.abs(g - ref)) / max(1e-12, np.max(np.abs(ref)))


@pytest.mark.parametrize("gamma,sign", [(0.0, 0), (-0.3, 1)])
def test_adjoint_matches_torch_autograd(torch_cuda, gamma, sign):
    n, T, S, C = 40, 10, 5, 4  # three checkpoint segments, the last one partial
    bt, (y, h0, z, xi) = make_batch(n, T, S, GRAD_THETAS, C, gamma=gamma, sign=bool(sign), seed=3)
    theta = torch.tensor(GRAD_THETAS, device="cuda", requires_grad=True)
    loss = -predictive_loglik(theta, bt).sum()
    loss.backward()
    g = theta.grad.cpu().numpy()
    for b, th in enumerate(GRAD_THETAS):
        ref = torch_grad(n, S, th, y[b], h0[b], z[b], xi[b], gamma, sign)
        assert rel_err(g[b], ref) <= 1e-3, (b, g[b], ref)


def test_adjoint_at_the_cfg4_shape(torch_cuda):
    """One filter at N = 1024 over T = 96 ticks with checkpoints every 32 (three full segments):
    the HIP adjoint against float64 autograd of the dense restatement, run on the GPU (its
    autograd graph holds about 30 GB).

    The bandwidth is the median of N(N-1)/2 = 523,776 pairwise distances, and its gradient
    flows through the one selected pair.  Neighbouring order statistics lie about 1e-6 apart,
    so float32 and float64 arithmetic select different pairs of nearly equal distance, whose
    gradients differ: torch's own float32 autograd of this restatement is 6.5e-2 off its
    float64 gradient here, while the float64 gradient moves by 1e-8 when h0 moves by one
    float32 ulp.  The reference therefore takes the median pairs the float32 forward
    selected (DiffBatch(record_pairs=True)); everything else is its own float64 arithmetic."""
    n, T, S, C = 1024, 96, 5, 32
    th = GRAD_THETAS[0]
    y, h0, z, xi = inputs(n, T, S, [th], seed=21)
    bt = DiffBatch(y, h0, n_stein=S, checkpoint_every=C, noise={"z": z, "xi": xi}, record_pairs=True)
    theta = torch.tensor([th], device="cuda", requires_grad=True)
    pl32 = predictive_loglik(theta, bt)
    (-pl32.sum()).backward()
    g = theta.grad.cpu().numpy()[0]
    pairs = bt.mpair[0].cpu().numpy()
    assert (pairs[:, 0] >= 0).all() and (pairs[:, 0] < pairs[:, 1]).all()
    pl32 = pl32.detach().cpu().numpy()[0]
    del bt
    cfg = cfg_of(n, S, th)
    dt, dev = torch.float64, "cuda"
    tt = [torch.tensor(v, dtype=dt, device=dev, requires_grad=True) for v in th]
    pl, _, _ = torch_svpf.run(tt, torch.tensor(y[0], dtype=dt, device=dev), torch.tensor(z[0], dtype=dt, device=dev),
                              torch.tensor(xi[0], dtype=dt, device=dev), torch.tensor(h0[0], dtype=dt, device=dev),
                              cfg, pairs=[(int(a), int(b)) for a, b in pairs])
    (-pl.sum()).backward()
    ref = np.array([t.grad.item() for t in tt])
    assert np.max(np.abs(pl32 - pl.detach().cpu().numpy())) <= 1e-4 * max(1.0, float(np.abs(pl32).max()))
    del pl, tt
    torch.cuda.empty_cache()
    assert rel_err(g, ref) <= 1e-3, (g, ref)


def test_adjoint_matches_oracle_central_differences(torch_cuda):
    n, T, S = 64, 8, 5
    th = THETAS[0]
    bt, (y, h0, z, xi) = make_batch(n, T, S, [th], C=3, seed=5)
    theta = torch.tensor([th], device="cuda", requires_grad=True)
    (-predictive_loglik(theta, bt).sum()).backward()
    g = theta.grad.cpu().numpy()[0]
    fd = []
    for k, name in enumerate(("mu", "rho", "sigma_z", "nu")):
        # the loss is strongly curved (RMSProp's normalisation): a 1e-4 step is 5% off on rho
        # here, a 1e-6 step in FP64 converges to the derivative
        eps = 1e-6
        cp = cfg_of(n, S, th)
        cm = dict(cp)
        cp[name] += eps
        cm[name] -= eps
        lp = -oracle_pred_ll(cp, y[0], h0[0], z[0], xi[0])[0].sum()
        lm = -oracle_pred_ll(cm, y[0], h0[0], z[0], xi[0])[0].sum()
        fd.append((lp - lm) / (2 * eps))
    fd = np.array(fd)
    assert rel_err(g, fd) <= 1e-3, (g, fd)


@pytest.mark.parametrize("n,T,runs", [
    (128, 9, ((1, 32.0), (2, 32.0), (4, 32.0), (9, 32.0), (2, 0.0), (4, 0.0))),
    (300, 8, ((8, 32.0), (3, 32.0), (3, 0.0))),  # 512-thread workgroups
    (1000, 8, ((8, 32.0), (3, 32.0), (3, 0.0))),  # 1024-thread workgroups, one row per lane
    (2048, 12, ((12, 32.0), (4, 32.0), (4, 0.0))),  # two rows per lane
])
def test_gradient_independent_of_checkpoint_interval(torch_cuda, n, T, runs):
    """Bit-identical gradients and median pairs for every checkpoint interval, with the
    backward replaying positions from the stored Stein drift or (no drift buffer)
    recomputing the transport: the recompute reproduces the forward's positions bit for
    bit, so the pair the forward's bandwidth came from is found again at every tick."""
    S = 5
    grads, pairs = [], []
    for C, budget in runs:
        y, h0, z, xi = inputs(n, T, S, GRAD_THETAS, seed=7)
        bt = DiffBatch(y, h0, n_stein=S, checkpoint_every=C, noise={"z": z, "xi": xi},
                       drift_budget_gb=budget, record_pairs=True)
        assert (bt.drift is None) == (budget == 0.0)
        theta = torch.tensor(GRAD_THETAS, device="cuda", requires_grad=True)
        (-predictive_loglik(theta, bt).sum()).backward()
        grads.append(theta.grad.cpu().numpy())
        pairs.append(bt.mpair.cpu().numpy())
        bw = np.sqrt(bt.ell2.cpu().numpy())
        inside = (bw > 0.01) & (bw < 2.0)
        assert (pairs[-1][:, :, 0][inside] >= 0).all(), (C, budget)
        del bt
    for g, p in zip(grads[1:], pairs[1:]):
        assert np.array_equal(g, grads[0])
        assert np.array_equal(p, pairs[0])


def test_counter_generator_forward_and_backward(torch_cuda):
    n, T, S = 256, 6, 5
    y, h0, _, _ = inputs(n, T, S, THETAS, seed=9)
    bt = DiffBatch(y, h0, n_stein=S, checkpoint_every=2)
    theta = torch.tensor(THETAS, device="cuda", requires_grad=True)
    pl = predictive_loglik(theta, bt)
    (-pl.sum()).backward()
    assert torch.isfinite(pl).all() and torch.isfinite(theta.grad).all()
    # the same seeds give the same forward and gradient in a second workspace
    bt2 = DiffBatch(y, h0, n_stein=S, checkpoint_every=3)
    theta2 = torch.tensor(THETAS, device="cuda", requires_grad=True)
    pl2 = predictive_loglik(theta2, bt2)
    (-pl2.sum()).backward()
    assert torch.equal(pl, pl2) and torch.equal(theta.grad, theta2.grad)
