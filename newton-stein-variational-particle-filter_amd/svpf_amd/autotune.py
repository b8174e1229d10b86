"""Tile autotuner of the batched kernel.

For every particle-count bucket, every instantiated (block size, tile) whose
capacity holds the bucket without leaving more than 3/4 of its rows idle is
timed on a batch of filters of that size (HIP events around a persistent
launch, after a warm-up tick).  The fastest per bucket is written to
``tuned_tiles.json`` next to this module, which ``Batch`` reads when it
assigns tiles (a filter of n particles takes the entry of the smallest bucket
>= n).

    python -m svpf_amd.autotune --sizes 128 256 512 1024 2048 4096 --B 820 --T 3
"""
from __future__ import annotations

import argparse
import json
import os

import numpy as np

from . import batched, config, data

DEFAULT_SIZES = (128, 256, 512, 1024, 2048, 4096)


def candidates(n):
    out = []
    for bs in batched.BLOCK_SIZES:
        for t in batched.valid_tiles(bs):
            cap = batched.tile_capacity(t, bs)
            if n <= cap <= max(4 * n, 64):
                out.append((bs, t))
    return out


def _inputs(n, B, T, seed):
    _, mu, rho, sz, nu = data.batched_thetas(B, seed=seed)
    cfgs = [config.plain_config(n=n, mu=mu[b], rho=rho[b], sigma_z=sz[b], nu=nu[b]) for b in range(B)]
    y = np.stack([data.sv_series(T, mu=mu[b], rho=rho[b], sigma_z=sz[b], nu=nu[b], seed=500 + b)[0]
                  for b in range(B)])
    h0 = [data.initial_particles(n, mu[b], rho[b], sz[b], seed=900 + b) for b in range(B)]
    return cfgs, y, h0


def time_tile(torch, n, bs, tile, cfgs, y, h0, T):
    bt = batched.Batch(cfgs, y, h0, tile_table={n: (bs, tile)})
    assert bt.n_groups == 1 and bt.groups["tile"][0] == tile and bt.groups["bs"][0] == bs
    bt.run(0, 1)
    torch.cuda.synchronize()
    s = torch.cuda.current_stream()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(s)
    bt.run(1, 1 + T)
    e1.record(s)
    torch.cuda.synchronize()
    st = bt.fstates()
    if (st["status"] != 0).any():
        raise RuntimeError(f"tile ({bs}, {tile}) rejected a filter at n={n}")
    ms = e0.elapsed_time(e1)
    return len(cfgs) * T / (ms * 1e-3), ms


def tune(sizes=DEFAULT_SIZES, B=820, T=3, out=None, log=print):
    import torch

    table, report = {}, {}
    for n in sizes:
        cfgs, y, h0 = _inputs(n, B, T + 1, seed=n)
        res = []
        for bs, tile in candidates(n):
            rate, ms = time_tile(torch, n, bs, tile, cfgs, y, h0, T)
            res.append((rate, bs, tile, ms))
            log(f"n={n:5d} bs={bs:4d} tile={tile} ({batched.TILES[tile][0]} rows/lane, {batched.TILES[tile][1]} "
                f"lanes/row): {rate:10.1f} filter-steps/s ({ms:.2f} ms for {T} ticks of {B} filters)")
        rate, bs, tile, _ = max(res)
        table[n] = (bs, tile)
        report[n] = [dict(bs=b, tile=t, rate=r, ms=m) for r, b, t, m in sorted(res, reverse=True)]
    if out is not None:
        with open(out, "w") as f:
            json.dump({str(k): list(v) for k, v in table.items()}, f, indent=1)
    return table, report


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--sizes", type=int, nargs="+", default=list(DEFAULT_SIZES))
    ap.add_argument("--B", type=int, default=820)
    ap.add_argument("--T", type=int, default=3)
    ap.add_argument("--out", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_tiles.json"))
    ap.add_argument("--report", default=None, help="also write every candidate's rate as JSON here")
    args = ap.parse_args()
    table, report = tune(args.sizes, args.B, args.T, args.out)
    print(json.dumps({str(k): list(v) for k, v in table.items()}))
    if args.report:
        with open(args.report, "w") as f:
            json.dump(report, f, indent=1)


if __name__ == "__main__":
    main()
