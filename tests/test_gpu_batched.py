"""Batched form: ragged filters in persistent tile-group launches, each filter
checked against the FP64 oracle on identical inputs."""
import numpy as np
import pytest

import oracle
from svpf_amd import config, data
from svpf_amd.batched import O_HMEAN, O_LOGLIK, O_STEPS, O_VOL, Batch

pytestmark = pytest.mark.gpu

SIZES = [1, 64, 100, 128, 256, 300, 512, 1000, 2048, 130, 257, 40]


def make_batch_inputs(B, T, sizes, seed=0, rows=5):
    n, mu, rho, sz, nu = data.batched_thetas(B, seed=seed, sizes=sizes)
    n = np.array([sizes[b % len(sizes)] for b in range(B)], dtype=np.int32)
    cfgs, ys, h0s = [], [], []
    for b in range(B):
        cfgs.append(config.plain_config(n=int(n[b]), mu=mu[b], rho=rho[b], sigma_z=sz[b], nu=nu[b]))
        y, _ = data.sv_series(T, mu=mu[b], rho=rho[b], sigma_z=sz[b], nu=nu[b], seed=1000 + b)
        ys.append(y)
        h0s.append(data.initial_particles(int(n[b]), mu[b], rho[b], sz[b], seed=2000 + b))
    return cfgs, np.stack(ys), h0s


def test_batched_predrawn_matches_oracle(torch_cuda):
    B, T = len(SIZES), 6
    cfgs, y, h0s = make_batch_inputs(B, T, SIZES)
    noise = {"z_pred": [], "xi": []}
    for b in range(B):
        nz = data.predraw_noise(T, cfgs[b]["n"], 5, seed=3000 + b)
        noise["z_pred"].append(nz["z_pred"])
        noise["xi"].append(nz["xi"])
    bt = Batch(cfgs, y, h0s, noise=noise)
    bt.run(0, T)
    torch_cuda.cuda.synchronize()
    out = bt.outputs()
    assert (bt.fstates()["status"] == 0).all()
    assert (bt.fstates()["timestep"] == T).all()
    for b in range(B):
        o = oracle.OracleFilter(cfgs[b], h0s[b])
        ref = o.run(y[b], {"z_pred": noise["z_pred"][b], "xi": noise["xi"][b]})
        assert np.sqrt(np.mean((out[b, :, O_HMEAN] - ref["h_mean"]) ** 2)) <= 1e-4, b
        assert np.max(np.abs(out[b, :, O_VOL] - ref["vol"])) <= 1e-4, b
        assert np.max(np.abs(out[b, :, O_LOGLIK] - ref["loglik"])) <= 1e-3, b
        assert np.array_equal(out[b, :, O_STEPS].astype(int), ref["steps"])
        assert np.max(np.abs(bt.particles(b) - o.h)) <= 2e-4, b


def test_batched_counter_generator_and_split_ticks(torch_cuda):
    """Counter generator keyed (seed=b, filter=b); two launches [0,3) + [3,6)
    equal one launch [0,6) bit for bit, and both match the oracle."""
    B, T = 8, 6
    sizes = [64, 128, 256, 512, 96, 200, 700, 33]
    cfgs, y, h0s = make_batch_inputs(B, T, sizes, seed=5)
    a = Batch(cfgs, y, h0s)
    a.run(0, 3)
    a.run(3, 6)
    b_ = Batch(cfgs, y, h0s)
    b_.run(0, 6)
    torch_cuda.cuda.synchronize()
    oa, ob = a.outputs(), b_.outputs()
    assert np.array_equal(oa, ob)
    for b in range(B):
        o = oracle.OracleFilter(cfgs[b], h0s[b])
        hm = []
        yp = 0.0
        for t in range(T):
            hm.append(o.tick(float(y[b, t]), yp, seed=b, filter_id=b)["h_mean"])
            yp = float(y[b, t])
        assert np.sqrt(np.mean((oa[b, :, O_HMEAN] - np.array(hm)) ** 2)) <= 1e-4, b


def every_tile():
    from svpf_amd import batched

    return [(bs, t) for bs in batched.BLOCK_SIZES for t in batched.valid_tiles(bs)]


@pytest.mark.parametrize("bs,tile", every_tile())
def test_every_tile_matches_oracle(torch_cuda, bs, tile):
    """Each instantiated tile shape, at full and ragged occupancy, against the oracle."""
    from svpf_amd import batched

    cap = batched.tile_capacity(tile, bs)
    full = min(cap, 2048)
    sizes = sorted({full, max(2, cap // 2 + 3), max(2, cap // 4 + 1)})
    sizes = [s for s in sizes if cap // 4 < s <= cap] or [full]
    T = 3
    cfgs, y, h0s = make_batch_inputs(len(sizes), T, sizes, seed=40 + tile)
    noise = {"z_pred": [], "xi": []}
    for b in range(len(sizes)):
        nz = data.predraw_noise(T, cfgs[b]["n"], 5, seed=4000 + b)
        noise["z_pred"].append(nz["z_pred"])
        noise["xi"].append(nz["xi"])
    bt = Batch(cfgs, y, h0s, noise=noise, tile_table={s: (bs, tile) for s in sizes})
    assert set(zip(bt.groups["bs"].tolist(), bt.groups["tile"].tolist())) == {(bs, tile)}
    bt.run(0, T)
    torch_cuda.cuda.synchronize()
    out = bt.outputs()
    assert (bt.fstates()["status"] == 0).all()
    for b in range(len(sizes)):
        o = oracle.OracleFilter(cfgs[b], h0s[b])
        ref = o.run(y[b], {"z_pred": noise["z_pred"][b], "xi": noise["xi"][b]})
        assert np.sqrt(np.mean((out[b, :, O_HMEAN] - ref["h_mean"]) ** 2)) <= 1e-4, (b, sizes[b])
        assert np.max(np.abs(bt.particles(b) - o.h)) <= 2e-4, (b, sizes[b])
