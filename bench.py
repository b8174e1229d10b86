"""Benchmark of the SVPF step on one MI355X.

Headline (value): filter-steps/s of the batched workload -- B=4096 independent
filters, per-filter N ~ U{128, 256, 512, 1024, 2048}, per-filter theta
(mu ~ U(-2,0), rho ~ U(.90,.995), sigma_z ~ U(.05,.4), nu ~ U(3,30)), each
filter's own synthetic series, 5 Stein iterations, FP32.  A step is one tick
of all B filters; K steps run in persistent tile-group launches with the
inputs resident in HBM.

Also reported: the same rate for the plain one-filter-per-launch path on the
same GPU (per-tick C-ABI step of single filters, measured on a stratified
subset and weighted to the workload's N mix) and the ratio; the single-filter
N=512 device-loop microseconds per step and the interactive per-tick latency;
the roofline of the dominant kernel (the N=2048 tile group) against the
FP32 vector peak; the differentiable filter (B=256, N=1024, T=500): forward
and forward + adjoint with checkpoints every 32 ticks and with the full tape;
the dense PyTorch restatement timed on the host cores as CPU baseline; the
full-feature single filter at N=4096 and N=16384 (regime-shift series, T=2000)
with the FP64 oracle as its CPU baseline.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "newton-stein-variational-particle-filter_amd")
for _p in (ROOT, PKG):
    if _p not in sys.path:
        sys.path.insert(0, _p)

import numpy as np  # noqa: E402

PEAK_FP32 = 157.3e12  # MI355X FP32 vector, FLOP/s
S_ITERS = 5
FLOP_PER_PAIR = 23 * S_ITERS + 9  # per filter-step: N^2 (23 S + 9)


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_workload(B, T, seed=0):
    from svpf_amd import config, data

    n, mu, rho, sz, nu = data.batched_thetas(B, seed=seed)
    cfgs, ys, h0s = [], np.zeros((B, T), np.float32), []
    rng = np.random.default_rng(seed + 1)
    for b in range(B):
        cfgs.append(config.plain_config(n=int(n[b]), n_stein=S_ITERS, mu=mu[b], rho=rho[b], sigma_z=sz[b], nu=nu[b]))
        # per-filter series: AR(1) log-vol with Student-t returns (vectorised over t)
        e = rng.standard_normal(T)
        h = np.empty(T)
        h[0] = mu[b] + sz[b] / np.sqrt(1 - rho[b] ** 2) * e[0]
        for t in range(1, T):
            h[t] = mu[b] + rho[b] * (h[t - 1] - mu[b]) + sz[b] * e[t]
        ys[b] = np.exp(h / 2) * rng.standard_t(nu[b], size=T)
        sd = np.sqrt(sz[b] ** 2 / (1 - rho[b] ** 2 + 1e-6))
        h0s.append(np.clip(mu[b] + sd * rng.standard_normal(int(n[b])), -15, 5).astype(np.float32))
    return cfgs, ys, h0s, n


def time_batched(torch, cfgs, ys, h0s, warmup, steps):
    from svpf_amd.batched import Batch

    bt = Batch(cfgs, ys, h0s)
    torch.cuda.synchronize()
    if warmup:
        bt.run(0, warmup)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bt.run(warmup, warmup + steps)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    st = bt.fstates()
    assert (st["status"] == 0).all(), "kernel rejected a filter"
    out = bt.outputs()[:, warmup:warmup + steps]
    assert np.isfinite(out).all()
    executed = out[:, :, 6].mean()
    return wall, bt, executed


def time_dominant_kernel(torch, cfgs, ys, h0s, n, warmup, steps):
    """The N=2048 tile group alone, timed with HIP events on its stream."""
    from svpf_amd.batched import Batch

    idx = [b for b in range(len(cfgs)) if n[b] == n.max()]
    bt = Batch([cfgs[b] for b in idx], ys[idx], [h0s[b] for b in idx])
    assert bt.n_groups == 1
    if warmup:
        bt.run(0, warmup)
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    bt.run(warmup, warmup + steps)
    e1.record(s)
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1)
    flops = float(len(idx)) * float(n.max()) ** 2 * FLOP_PER_PAIR * steps
    return ms, flops, len(idx), bt.groups


def time_plain(torch, cfgs, ys, h0s, n, ticks=8):
    """Plain path: one filter per launch per tick through svpf_step_graph."""
    from svpf_amd.filter import Filter

    per_bucket = {}
    for N in sorted(set(n.tolist())):
        b = int(np.nonzero(n == N)[0][0])
        f = Filter(cfgs[b])
        f.set_particles(h0s[b])
        f.step(float(ys[b, 0]), 0.0)  # capture the tick graph
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for t in range(1, ticks + 1):
            f.step(float(ys[b, t]), float(ys[b, t - 1]))
        per_bucket[N] = (time.perf_counter() - t0) / ticks
        f.close()
    # expected seconds per filter-step under the workload's N mix
    mix = np.mean([per_bucket[int(N)] for N in n])
    return 1.0 / mix, per_bucket


def time_single(torch, T=1000, warm=50):
    """Single filter N=512: device loop over T ticks, and the per-tick interactive latency."""
    from svpf_amd import config, data
    from svpf_amd.filter import Filter

    cfg = config.plain_config(n=512)
    y, _ = data.sv_series(T + warm, seed=0)
    h0 = data.initial_particles(512, -1.0, 0.97, 0.15, seed=2)
    f = Filter(cfg)
    f.set_particles(h0)
    dy = torch.tensor(y, device="cuda")
    f.run_device(dy[:warm])
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    f.run_device(dy[warm:])
    torch.cuda.synchronize()
    us_device = (time.perf_counter() - t0) / T * 1e6
    lat = []
    yp = float(y[warm - 1])
    for t in range(warm, warm + 200):
        t1 = time.perf_counter()
        f.step(float(y[t]), yp)
        lat.append(time.perf_counter() - t1)
        yp = float(y[t])
    f.close()
    lat = np.array(lat[10:]) * 1e6
    return us_device, float(np.percentile(lat, 50)), float(np.percentile(lat, 99))


JOINT_FLOP_PER_PAIR = 31  # 4 differences, 4 FMA distance, reciprocal, K^2, 8 FMA sums, mass


def time_joint(torch, T=1000, warm=50):
    """The joint state-parameter filter (include/svpf_joint.h) at the reference defaults
    (512 particles, 5 Stein iterations): svpf_joint_step with diagnostics per tick (the
    reference test's loop, p50 / p99), and the device loop over T ticks in one launch, at
    N = 512 and 2048; the FP64 oracle on one thread as CPU baseline."""
    import oracle
    from svpf_amd import joint

    y = np.random.default_rng(3).standard_normal(T + warm).astype(np.float32) * np.float32(np.exp(-1.75))
    out = {"workload": "one joint filter, reference defaults (5 Stein iterations), synthetic N(0, e^-3.5) series"}
    dy = torch.tensor(y, device="cuda")
    for n in (512, 2048):
        f = joint.JointFilter(seed=1, n_particles=n)
        f.run_device(dy[:warm])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        f.run_device(dy[warm:])
        torch.cuda.synchronize()
        us = (time.perf_counter() - t0) / T * 1e6
        flops = 5.0 * n * n * JOINT_FLOP_PER_PAIR
        out[f"n{n}_us_per_step_device_loop"] = us
        out[f"n{n}_pair_tflops"] = flops / (us * 1e-6) / 1e12
        if n == 512:
            lat = []
            for t in range(200):
                t1 = time.perf_counter()
                f.step(float(y[t]))
                lat.append(time.perf_counter() - t1)
            lat = np.array(lat[10:]) * 1e6
            out["n512_step_with_diag_us_p50"] = float(np.percentile(lat, 50))
            out["n512_step_with_diag_us_p99"] = float(np.percentile(lat, 99))
        f.close()
    c = joint.default_config()
    ocfg = {"n": 512, "n_stein_steps": 5}
    ocfg.update({k: c[k] for k in joint.CONFIG_FIELDS})
    o = oracle.JointOracle(ocfg, seed=1)
    t0 = time.perf_counter()
    for t in range(20):
        o.step(float(y[t]), 1, t)
    out["cpu_baseline"] = {"s_per_step_n512": (time.perf_counter() - t0) / 20, "cores": 1, "kind": "port",
                           "sample": "FP64 C oracle (svpfo_joint_step), N=512, 20 steps"}
    return out


def time_fused_v2(torch, n=512, T=1000, warm=100):
    """The single-workgroup fused step (svpf_fully_fused_step_v2) with the
    reference driver's arguments (10 Stein iterations over 5 stages, full
    Newton, guide, MIM, rejuvenation; counter generator): back-to-back calls
    (us/call from the wall clock around T calls), and the driver's loop, which
    reads h_mean / vol / ksd back after every call (p50 / p99)."""
    from svpf_amd.fused_v2 import FusedV2, v2_args

    rng = np.random.default_rng(0)
    y = (np.exp(-2.25) * rng.standard_normal(warm + max(T, 300))).astype(np.float32)
    f = FusedV2(np.full(n, -4.5, np.float32), mu_state=-4.5)
    a = v2_args(n=n)
    for t in range(warm):
        f.step(dict(a, timestep=t, y_t=float(y[t])), None)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for t in range(warm, warm + T):
        f.step(dict(a, timestep=t, y_t=float(y[t])), None)
    torch.cuda.synchronize()
    us_call = (time.perf_counter() - t0) / T * 1e6
    lat = []
    prev = dict(y_prev=0.0, h_mean_prev=-4.5, vol_prev=float(np.exp(-2.25)), ksd_prev=0.1)
    for t in range(warm, warm + 300):
        t1 = time.perf_counter()
        f.step(dict(a, timestep=t, y_t=float(y[t]), **prev), None)
        o = f.outputs()
        lat.append(time.perf_counter() - t1)
        prev = dict(y_prev=float(y[t]), h_mean_prev=o["h_mean"], vol_prev=o["vol"], ksd_prev=o["ksd"])
    lat = np.array(lat[20:]) * 1e6
    return {"n": n, "us_per_call_back_to_back": us_call, "driver_loop_latency_us_p50": float(np.percentile(lat, 50)),
            "driver_loop_latency_us_p99": float(np.percentile(lat, 99)),
            "arguments": "test/test_fully_fused_v2.cu:91-113 (10 Stein iterations, 5 stages, full Newton)"}


def time_extended(torch, B=512, n=512, warm=2, T=10):
    """The batched kernel's extended instantiation: B filters of n particles
    with the plain configuration, with the Heun step, with the Student-t state
    (nu_s = 5) and with both, each as one batched run (filter-steps/s); then one
    N=4096 filter on the grid path in the same four variants (device loop, us/step)."""
    from svpf_amd import config, data

    y = np.stack([data.sv_series(warm + T, seed=b)[0] for b in range(B)])
    h0s = [data.initial_particles(n, -1.0, 0.97, 0.15, seed=100 + b) for b in range(B)]
    out = {"workload": f"B={B} filters, N={n}, plain configuration (5 Stein iterations, local Newton, SVLD)"}
    for name, feats in (("plain", {}), ("heun", dict(use_heun=1)),
                        ("student_t_state", dict(use_student_t_state=1, nu_state=5.0)),
                        ("heun_student_t_state", dict(use_heun=1, use_student_t_state=1, nu_state=5.0))):
        cfgs = [config.plain_config(n=n, **feats) for _ in range(B)]
        wall, _, _ = time_batched(torch, cfgs, y, h0s, warm, T)
        out[name + "_filter_steps_per_s"] = B * T / wall
    # one filter on the grid path (device loop): plain, Heun, Student-t state, both
    from svpf_amd.filter import Filter

    N, TT = 4096, 200
    ys, _ = data.sv_series(TT + 20, seed=7)
    dy = torch.tensor(ys, device="cuda")
    for name, feats in (("plain", {}), ("heun", dict(use_heun=1)),
                        ("student_t_state", dict(use_student_t_state=1, nu_state=5.0)),
                        ("heun_student_t_state", dict(use_heun=1, use_student_t_state=1, nu_state=5.0))):
        f = Filter(config.plain_config(n=N, **feats))
        f.set_particles(data.initial_particles(N, -1.0, 0.97, 0.15, seed=8))
        assert f.path() == "grid"
        f.run_device(dy[:20])
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        f.run_device(dy[20:])
        torch.cuda.synchronize()
        out[f"grid_n{N}_{name}_us_per_step"] = (time.perf_counter() - t0) / TT * 1e6
        f.close()
    return out


def time_large(torch, n, T=2000):
    """One full-feature filter (MIM, asymmetric rho, EKF guide, SVLD, KSD-adaptive
    beta, early stop, at most 8 Stein iterations) on the regime-shift series:
    device loop over T ticks (timed), then the same series tick by tick for the
    executed Stein iterations."""
    from svpf_amd import config, data
    from svpf_amd.filter import Filter

    cfg = config.full_feature_config(n=n)
    y, _ = data.regime_shift_series(T=T, seed=0)
    h0 = data.initial_particles(n, -1.0, 0.98, 0.2, seed=1)
    f = Filter(cfg)
    path = f.path()
    dy = torch.tensor(y, device="cuda")
    ll = torch.zeros(T, device="cuda")
    f.set_particles(h0)
    f.run_device(dy[:20], ll[:20])  # warm-up
    f.set_particles(h0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    f.run_device(dy, ll)
    torch.cuda.synchronize()
    us = (time.perf_counter() - t0) / T * 1e6
    assert torch.isfinite(ll).all()
    f.set_particles(h0)
    steps = np.zeros(T)
    yp = 0.0
    for t in range(T):
        f.step(float(y[t]), yp)
        steps[t] = f.last_outputs()[6]
        yp = float(y[t])
    f.close()
    it = float(steps.mean())
    flops = float(n) ** 2 * (23 * it + 9)
    return {"us_per_step": us, "path": path, "executed_iters_mean": it, "tflops": flops / (us * 1e-6) / 1e12}


def cpu_large_baseline(n=4096, ticks=2):
    """The FP64 oracle (one thread) on the full-feature filter: seconds per tick."""
    import oracle
    from svpf_amd import config, data

    cfg = config.full_feature_config(n=n)
    y, _ = data.regime_shift_series(T=ticks, seed=0)
    h0 = data.initial_particles(n, -1.0, 0.98, 0.2, seed=1)
    f = oracle.OracleFilter(cfg, h0)
    t0 = time.perf_counter()
    yp = 0.0
    for t in range(ticks):
        f.tick(float(y[t]), yp, seed=42)
        yp = float(y[t])
    return (time.perf_counter() - t0) / ticks


def measured_traffic(ticks):
    """HBM bytes of one launch of the dominant kernel over `ticks` ticks, from the
    committed PMC summary (FETCH_SIZE and WRITE_SIZE runs of `bench.py --only
    dominant`, each scaled by what the counter reads for a 256 MiB copy).  A work
    item is one tick of one filter, so the filters' state is loaded and stored
    once per tick and the traffic grows with the tick count."""
    path = os.path.join(ROOT, "profiles", "svpf_batched_kernel_n2048_group_trace.md")
    try:
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic_bytes_per_tick"):
                    return float(line.split("=")[1]) * ticks
    except (OSError, ValueError):
        pass
    return None


def cpu_threads():
    """Host threads for the CPU baselines: the box's CPU share (at most 16)."""
    return max(1, min(16, os.cpu_count() or 1))


def cpu_info():
    """The host CPU the baselines ran on: model name, logical CPUs, CPUs this process may use."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        avail = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        avail = None
    return {"cpu_model": model, "os_cpu_count": os.cpu_count(), "cpus_available": avail}


def cpu_baseline(cfgs, ys, h0s, n, ticks=3):
    """The dense PyTorch restatement of the filter tick (float32, all host
    threads): one filter per N bucket, one warm-up tick then `ticks` timed
    ticks, weighted to the workload's N mix."""
    import torch

    from oracle import torch_svpf

    torch.set_num_threads(cpu_threads())
    ticks = max(1, min(ticks, ys.shape[1] - 1))
    per_bucket = {}
    rng = np.random.default_rng(7)
    for N in sorted(set(n.tolist())):
        b = int(np.nonzero(n == N)[0][0])
        c = cfgs[b]
        theta = [torch.tensor(c[k], dtype=torch.float32) for k in ("mu", "rho", "sigma_z", "nu")]
        x = torch.tensor(h0s[b], dtype=torch.float32)
        v = torch.zeros_like(x)
        z = torch.tensor(rng.standard_normal((ticks + 1, N)), dtype=torch.float32)
        xi = torch.tensor(rng.standard_normal((ticks + 1, S_ITERS, N)), dtype=torch.float32)
        y = torch.tensor(ys[b], dtype=torch.float32)
        with torch.no_grad():
            x, v, _ = torch_svpf.tick(x, v, y[0], torch.tensor(0.0), z[0], xi[0], theta, c)
            t0 = time.perf_counter()
            for t in range(1, ticks + 1):
                x, v, _ = torch_svpf.tick(x, v, y[t], y[t - 1], z[t], xi[t], theta, c)
            per_bucket[N] = (time.perf_counter() - t0) / ticks
    mix = np.mean([per_bucket[int(N)] for N in n])
    return 1.0 / mix, per_bucket


def oracle_fp64_rate(cfgs, ys, h0s, n, ticks=2):
    """Secondary reference: the scalar FP64 C oracle on one thread."""
    import oracle

    per_bucket = {}
    for N in sorted(set(n.tolist())):
        b = int(np.nonzero(n == N)[0][0])
        f = oracle.OracleFilter(cfgs[b], h0s[b])
        t0 = time.perf_counter()
        yp = 0.0
        for t in range(ticks):
            f.tick(float(ys[b, t]), yp, seed=b, filter_id=b)
            yp = float(ys[b, t])
        per_bucket[N] = (time.perf_counter() - t0) / ticks
    mix = np.mean([per_bucket[int(N)] for N in n])
    return 1.0 / mix


def diff_workload(B=256, N=1024, T=500):
    """Differentiable workload: B filters of N particles over T ticks with per-filter theta and series."""
    from svpf_amd import data

    _, mu, rho, sz, nu = data.batched_thetas(B, seed=3)
    theta = np.stack([mu, rho, sz, nu], 1).astype(np.float32)
    ys = np.zeros((B, T), np.float32)
    h0 = np.zeros((B, N), np.float32)
    for b in range(B):
        ys[b], _ = data.sv_series(T, mu=mu[b], rho=rho[b], sigma_z=sz[b], nu=nu[b], seed=1000 + b)
        h0[b] = data.initial_particles(N, mu[b], rho[b], sz[b], seed=2000 + b)
    return theta, ys, h0


def time_diff(torch, theta, ys, h0, checkpoint_every):
    """Forward, and forward + adjoint, of the differentiable filter (HIP events on the launch stream)."""
    from svpf_amd.diff import DiffBatch

    bt = DiffBatch(ys, h0, n_stein=S_ITERS, checkpoint_every=checkpoint_every)
    th = torch.tensor(theta, device="cuda")
    g = -torch.ones((bt.B, bt.T), device="cuda")
    bt.forward(th)
    bt.backward(g)
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
    ev[0].record(s)
    bt.forward(th)
    ev[1].record(s)
    bt.backward(g)
    ev[2].record(s)
    torch.cuda.synchronize()
    fwd = ev[0].elapsed_time(ev[1])
    bwd = ev[1].elapsed_time(ev[2])
    assert torch.isfinite(bt.gtheta).all()
    tape_gb = bt.tape.numel() * 4 / 1e9
    del bt
    torch.cuda.empty_cache()
    return fwd, bwd, tape_gb


def cpu_diff_baseline(theta, ys, h0, ticks=2):
    """Dense PyTorch restatement with autograd (float32, host threads): one filter, forward and forward+backward."""
    import torch

    from oracle import torch_svpf
    from svpf_amd import config

    torch.set_num_threads(cpu_threads())
    N = h0.shape[1]
    cfg = config.plain_config(n=N, n_stein=S_ITERS)
    rng = np.random.default_rng(11)
    z = torch.tensor(rng.standard_normal((ticks, N)), dtype=torch.float32)
    xi = torch.tensor(rng.standard_normal((ticks, S_ITERS, N)), dtype=torch.float32)
    y = torch.tensor(ys[0, :ticks], dtype=torch.float32)
    h = torch.tensor(h0[0], dtype=torch.float32)
    th = [torch.tensor(float(v), requires_grad=True) for v in theta[0]]
    # one untimed tick first: the host thread pool and allocator warm up on it
    pw, _, _ = torch_svpf.run(th, y[:1], z[:1], xi[:1], h, cfg)
    (-pw.sum()).backward()
    for t in th:
        t.grad = None
    t0 = time.perf_counter()
    with torch.no_grad():
        torch_svpf.run(th, y, z, xi, h, cfg)
    fwd = (time.perf_counter() - t0) / ticks
    t0 = time.perf_counter()
    pl, _, _ = torch_svpf.run(th, y, z, xi, h, cfg)
    (-pl.sum()).backward()
    fb = (time.perf_counter() - t0) / ticks
    return fwd, fb


TWO_FACTOR_THETA = (-1.0, 0.80, 0.30, 0.995, 0.08, 5.0)  # mu, rho_f, sigma_f, rho_s, sigma_s, nu


def twofactor_workload(B, N, T, seed=5):
    """B two-factor filters of N particles; each filter's own series from the generating theta."""
    from svpf_amd import data

    mu, rf, sf, rs, ss, nu = TWO_FACTOR_THETA
    ys = np.zeros((B, T), np.float32)
    x0 = np.zeros((B, 2, N), np.float32)
    rng = np.random.default_rng(seed)
    for b in range(B):
        ys[b] = data.two_factor_series(T, mu=mu, rho_f=rf, sigma_f=sf, rho_s=rs, sigma_s=ss, nu=nu, seed=seed + b)[0]
        x0[b, 0] = rng.standard_normal(N) * sf / np.sqrt(1 - rf * rf)
        x0[b, 1] = rng.standard_normal(N) * ss / np.sqrt(1 - rs * rs)
    return ys, x0


def time_twofactor(torch, B=1024, N=2048, T=32, C=32, drift_budget_gb=150.0):
    """Two-factor filter at the cfg shape: forward and forward + adjoint with checkpoints every C
    ticks (32 for this workload; at T <= C the forward writes the whole tape), the Stein drift
    and the prior's shift of every iteration kept for the backward when they fit drift_budget_gb
    (126 GB at T=1000: MI355X holds it).
    Also the plain forward (checkpoints only, no drift, no tape) that the ratio is set against."""
    from svpf_amd.twofactor import TwoFactorBatch

    ys, x0 = twofactor_workload(B, N, T)
    th = torch.tensor([TWO_FACTOR_THETA] * B, device="cuda")
    g = -torch.ones((B, T), device="cuda")
    s = torch.cuda.current_stream()

    def timed(bt, backward):
        if T <= 64:  # warm-up pass for the short horizon
            bt.forward(th)
            if backward:
                bt.backward(g)
            torch.cuda.synchronize()
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(3)]
        ev[0].record(s)
        bt.forward(th)
        ev[1].record(s)
        if backward:
            bt.backward(g)
        ev[2].record(s)
        torch.cuda.synchronize()
        return ev[0].elapsed_time(ev[1]), (ev[1].elapsed_time(ev[2]) if backward else 0.0)

    bt = TwoFactorBatch(ys, x0, n_stein=S_ITERS, checkpoint_every=C, drift_budget_gb=drift_budget_gb)
    drift = bt.drift is not None
    fwd, bwd = timed(bt, True)
    assert torch.isfinite(bt.gtheta).all() and torch.isfinite(bt.out).all()
    del bt
    torch.cuda.empty_cache()
    # the plain forward: checkpoints only (an interval below T, so no tape), no drift buffer
    bt = TwoFactorBatch(ys, x0, n_stein=S_ITERS, checkpoint_every=C if C < T else max(1, T // 4),
                        drift_budget_gb=0.0)
    plain, _ = timed(bt, False)
    del bt
    torch.cuda.empty_cache()
    flops = float(N) ** 2 * 39 * S_ITERS * B * T  # 2-D pair work: 21 flops transport + 18 prior per pair-iteration
    return {"workload": f"two-factor B={B} N={N} T={T} S={S_ITERS}, checkpoints every {C} ticks"
                        + (" (one segment: the forward writes the whole tape)" if C >= T else "")
                        + (", Stein drift kept for the backward" if drift else ""),
            "filter_steps_per_s": B * T / (plain * 1e-3), "plain_fwd_ms": plain, "fwd_ms": fwd,
            "fwd_bwd_ms": fwd + bwd, "fwd_bwd_over_fwd": (fwd + bwd) / fwd,
            "fwd_bwd_over_plain_fwd": (fwd + bwd) / plain, "fwd_tflops": flops / (plain * 1e-3) / 1e12}


def calibration_demo(torch, B=128, N=256, T=600, epochs=200, lr=1e-2):
    """Adam on one theta shared by B two-factor filters (each filter runs its own series
    drawn from the generating theta) from a perturbed start; reports the recovered theta and
    the summed loss at the generating, starting and fitted theta: a fitted loss below the
    generating theta's says the objective's minimum is elsewhere (the objective is at fault),
    one above it says the optimiser stopped short."""
    from svpf_amd.twofactor import THETA_NAMES, TwoFactorBatch, calibrate

    ys, x0 = twofactor_workload(B, N, T, seed=77)
    bt = TwoFactorBatch(ys, x0, n_stein=S_ITERS, checkpoint_every=32)
    true = np.array(TWO_FACTOR_THETA, np.float32)
    start = true.copy()
    start[0] += 0.5
    start[1] = 0.6
    start[2] *= 1.5
    start[3] = 0.98
    start[4] *= 1.5
    start[5] = 10.0
    t0 = time.perf_counter()
    th, hist = calibrate(bt, torch.tensor(start, device="cuda"), epochs=epochs, lr=lr, shared=True,
                         log=lambda ep, l: log(f"calibration epoch {ep}: mean loss {l:.3f}"))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    est = th.cpu().numpy()
    # the summed loss at the generating, starting and fitted theta (same series and noise)
    loss_at = {}
    for name, v in (("true", true), ("start", start), ("estimate", est)):
        pl = bt.forward(torch.tensor(np.tile(v, (B, 1)), device="cuda"))[:, :, 0]
        loss_at[name] = float(-pl.sum())
    verdict = ("objective: the fitted theta scores better than the generating one"
               if loss_at["estimate"] < loss_at["true"] else
               "optimiser: the generating theta scores better than the fitted one")
    return {"workload": f"one theta shared by B={B} filters (own series each), N={N}, T={T}, S={S_ITERS}, "
                        f"Adam lr {lr}, {epochs} epochs",
            "true": dict(zip(THETA_NAMES, true.tolist())), "start": dict(zip(THETA_NAMES, start.tolist())),
            "estimate": dict(zip(THETA_NAMES, est.tolist())),
            "loss_first": float(hist[0].sum()), "loss_last": float(hist[-1].sum()), "loss_at": loss_at,
            "recovery_limited_by": verdict, "wall_s": wall, "s_per_epoch": wall / epochs}


def profile_only(torch, what, B, K, W, n_large=16384):
    """One workload for a profiler run; prints what ran."""
    if what == "dominant":
        src = torch.ones(64 * 1024 * 1024, device="cuda")  # 256 MiB
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        cfgs, ys, h0s, n = build_workload(B, W + K + 1)
        ms, flops, nf, groups = time_dominant_kernel(torch, cfgs, ys, h0s, n, W, K)
        log(f"copy 256 MiB; dominant group: {nf} filters, {K} ticks, {ms:.2f} ms, {flops / ms / 1e9:.2f} TFLOP/s")
    elif what == "large":
        log(json.dumps(time_large(torch, n_large, T=K)))
    elif what == "calib":
        print(json.dumps(calibration_demo(torch)), flush=True)
    elif what == "calib5":
        print(json.dumps(calibration_demo(torch, B=1024, N=2048, T=1000, epochs=K)), flush=True)
    elif what == "twofactor":
        log(json.dumps(time_twofactor(torch)))
    elif what == "extended":
        log(json.dumps(time_extended(torch, T=K)))
    elif what == "v2":
        log(json.dumps(time_fused_v2(torch, T=K, warm=W)))
    elif what == "joint":
        log(json.dumps(time_joint(torch, T=K, warm=W)))
    elif what == "diff":
        theta, dys, dh0 = diff_workload()
        fwd, bwd, _ = time_diff(torch, theta, dys, dh0, 32)
        log(f"diff: fwd {fwd:.1f} ms, bwd {bwd:.1f} ms")
    else:
        us, p50, p99 = time_single(torch, T=200, warm=20)
        log(f"single N=512: {us:.1f} us/step device loop, tick p50 {p50:.1f} us")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--B", type=int, default=4096)
    ap.add_argument("--no-extras", action="store_true", help="headline only")
    ap.add_argument("--n", type=int, default=16384, help="particles of --only large (default 16384)")
    ap.add_argument("--only", choices=["dominant", "diff", "single", "twofactor", "calib", "calib5", "large", "v2",
                                       "extended", "joint"],
                    default=None,
                    help="run one workload once and exit (for profiler runs): the N=2048 tile group of the "
                         "batched workload preceded by a 256 MiB device copy (traffic calibration), the "
                         "differentiable filter, the single N=512 filter, the two-factor filter, the "
                         "two-factor calibration (prints its JSON; calib5: at B=1024, N=2048, T=1000 for "
                         "--steps epochs), the full-feature filter (N from --n, default 16384) "
                         "over --steps ticks, the fused single-workgroup step over --steps calls, or the "
                         "Heun / Student-t batched runs over --steps ticks")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: filters are independent and this build runs on one MI355X; "
              "--gpus above 1 is not supported.", file=sys.stderr)
        sys.exit(2)

    import torch

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    from svpf_amd import _lib
    from svpf_amd.batched import TILES

    _lib.load()
    B, K, W = args.B, args.steps, args.warmup
    if args.only:
        profile_only(torch, args.only, B, K, W, n_large=args.n)
        return
    log(f"building workload B={B} T={W + K}")
    cfgs, ys, h0s, n = build_workload(B, W + K + 1)
    log("batched run")
    wall, bt, executed = time_batched(torch, cfgs, ys, h0s, W, K)
    value = B * K / wall
    ms_per_step = wall / K * 1e3
    log(f"batched: {value:.1f} filter-steps/s ({ms_per_step:.3f} ms/step)")

    log("dominant kernel (N=2048 group)")
    ms_dom, flops_dom, nf_dom, groups = time_dominant_kernel(torch, cfgs, ys, h0s, n, W, K)
    achieved = flops_dom / (ms_dom * 1e-3)
    total_flops = float((n.astype(np.float64) ** 2).sum()) * FLOP_PER_PAIR * K

    extras = {}
    if not args.no_extras:
        log("plain one-filter-per-launch path")
        plain_rate, plain_bucket = time_plain(torch, cfgs, ys, h0s, n)
        log("single filter N=512")
        us_dev, p50, p99 = time_single(torch)
        extras = {
            "plain_filter_steps_per_s": plain_rate,
            "batched_vs_plain": value / plain_rate,
            "plain_s_per_step_by_n": {str(k): v for k, v in plain_bucket.items()},
            "single_n512_us_per_step_device_loop": us_dev,
            "single_n512_tick_latency_us_p50": p50,
            "single_n512_tick_latency_us_p99": p99,
        }
        log("extended features (Heun step, Student-t state), B=512 N=512")
        extras["extended_features"] = time_extended(torch)
        log("fused single-workgroup step (svpf_fully_fused_step_v2), N=512")
        extras["fused_v2"] = time_fused_v2(torch)
        log("joint state-parameter filter, N=512 and 2048")
        extras["joint"] = time_joint(torch)
        log("differentiable filter B=256 N=1024 T=500")
        theta, dys, dh0 = diff_workload()
        T_d = dys.shape[1]
        fwd32, bwd32, tape32 = time_diff(torch, theta, dys, dh0, 32)
        fwdA, bwdA, tapeA = time_diff(torch, theta, dys, dh0, T_d)
        log(f"diff: fwd {fwd32:.1f} ms, fwd+bwd {fwd32 + bwd32:.1f} ms (ckpt 32); "
            f"fwd {fwdA:.1f} ms, fwd+bwd {fwdA + bwdA:.1f} ms (full tape)")
        log("cpu baseline of the differentiable filter (dense torch + autograd)")
        cfwd, cfb = cpu_diff_baseline(theta, dys, dh0)
        diff_flops = float(dh0.shape[1]) ** 2 * 23 * S_ITERS * dh0.shape[0] * T_d
        extras["diff"] = {
            "workload": "B=256 filters, N=1024, T=500, S=5, per-filter theta and series, loss -sum_t pred_ll",
            "fwd_ms": fwd32, "fwd_bwd_ms": fwd32 + bwd32, "fwd_us_per_step": fwd32 * 1e3 / T_d,
            "fwd_bwd_us_per_step": (fwd32 + bwd32) * 1e3 / T_d, "fwd_bwd_over_fwd": (fwd32 + bwd32) / fwd32,
            "checkpoint_every": 32, "tape_gb": tape32,
            "full_tape": {"fwd_ms": fwdA, "fwd_bwd_ms": fwdA + bwdA, "fwd_bwd_over_fwd": (fwdA + bwdA) / fwdA,
                          "tape_gb": tapeA},
            "fwd_tflops": diff_flops / (fwd32 * 1e-3) / 1e12,
            "cpu_baseline": {"fwd_s_per_filter_step": cfwd, "fwd_bwd_s_per_filter_step": cfb,
                             "cores": cpu_threads(), "kind": "port",
                             "sample": "dense torch float32 restatement, 1 filter N=1024, 2 ticks, autograd"},
        }
        extras["fwd_bwd_us_per_step"] = (fwd32 + bwd32) * 1e3 / T_d
        log("two-factor filter B=1024 N=2048")
        extras["twofactor"] = time_twofactor(torch)
        log("two-factor filter B=1024 N=2048 over T=1000 (forward + adjoint, checkpoints every 32)")
        extras["twofactor_t1000"] = time_twofactor(torch, T=1000, C=32)
        log("full-feature single filters N=4096 and N=16384, T=2000")
        large = {str(N): time_large(torch, N) for N in (4096, 16384)}
        log("cpu baseline of the full-feature filter (FP64 oracle, N=4096, one thread)")
        large["cpu_baseline"] = {"s_per_step_n4096": cpu_large_baseline(), "cores": 1, "kind": "port",
                                 "sample": "FP64 C oracle, full-feature filter N=4096, 2 ticks"}
        large["workload"] = ("one filter, regime-shift series T=2000 (mu -1 -> 1.5 at t=1000, jumps p=0.02 of +2, "
                             "rho up 0.90 / down 0.98, sigma_z 0.2, nu 4), MIM, EKF guide, SVLD T=1, "
                             "KSD-adaptive beta, early stop, at most 8 Stein iterations")
        extras["large_filter"] = large
    log("cpu baseline (dense torch restatement)")
    cpu_rate, cpu_bucket = cpu_baseline(cfgs, ys, h0s, n, ticks=8)
    log("secondary reference (FP64 oracle, one thread)")
    oracle_rate = oracle_fp64_rate(cfgs, ys, h0s, n)

    rec = {
        "metric": "filter-steps/s at B=4096 batched vs plain; us/step at N=512; fwd+bwd us/step",
        "value": value,
        "unit": "filter-steps/s",
        "n_gpus": 1,
        "steps": K,
        "warmup": W,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"batched SVPF, B={B} filters, N~U{{128..2048}}, per-filter theta and series, "
                        f"S={S_ITERS} Stein iterations, median IMQ bandwidth, local Newton, SVLD",
            "B": B, "stein_iters": S_ITERS, "executed_iters_mean": float(executed),
            "particles_total": int(n.sum()), "tile_groups": int(bt.n_groups),
        },
        "achieved_tflops_whole_batch": total_flops / wall / 1e12,
        "roofline": {
            # the FP32 vector peak: the pair loops run on the VALU (no MFMA: a 1-D pairwise
            # interaction is not a dense contraction); 157.3 TFLOP/s is also the FP32 matrix peak
            "bound": "valu",
            "achieved": achieved / 1e12,
            "peak": PEAK_FP32 / 1e12,
            "unit": "TFLOP/s",
            "frac": achieved / PEAK_FP32,
            "traffic": measured_traffic(K),
            "traffic_source": "rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE of the 5-tick launch, per tick x the "
                              "launch's ticks, profiles/svpf_batched_kernel_n2048_group_trace.md",
            "kernel": f"svpf_batched_kernel<{int(groups['bs'][0])},{TILES[int(groups['tile'][0])][0]},"
                      f"{TILES[int(groups['tile'][0])][1]},false> (N=2048 group, {nf_dom} filters, grid {int(groups['grid'][0])})",
            "kernel_ms": ms_dom,
            "flops_per_launch": flops_dom,
        },
        "cpu_baseline": {
            "value": cpu_rate,
            "unit": "filter-steps/s",
            "cores": cpu_threads(),
            "kind": "port",
            "sample": "dense PyTorch float32 restatement of the tick (oracle/torch_svpf.py), one filter per N "
                      "bucket {128..2048} x 8 ticks after 1 warm-up, weighted to the workload N mix",
            "s_per_step_by_n": {str(k): v for k, v in cpu_bucket.items()},
            "oracle_fp64_1thread_filter_steps_per_s": oracle_rate,
            **cpu_info(),
        },
    }
    rec.update(extras)
    print(json.dumps(rec), flush=True)


if __name__ == "__main__":
    main()
