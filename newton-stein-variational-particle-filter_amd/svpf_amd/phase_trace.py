"""Phase split of the single-filter cooperative tick (latency analysis).

Runs a device loop of one filter on libsvpf_amd_trace.so (``make -C
newton-stein-variational-particle-filter_amd/csrc trace``), whose cooperative
kernel stamps the 100 MHz wall clock at its phase boundaries (workgroup 0's
view), and prints the median time of each phase per tick:

    python -m svpf_amd.phase_trace [--n 512] [--ticks 200]

Stamps of tick t (slot t mod 256 of svpf_trace_buf, 64 per tick): 0 start,
1 predict done, 2 predictive log-likelihood, 50 particles sorted, 3 bandwidth
(median selection) done, 4 mixture means, then per Stein iteration k
8 + 3k gradient done, 9 + 3k records gathered, 10 + 3k transport done, then
40 outputs reduced, 41 final positions and partials gathered, 42 the tick
finished; inside the median selection 52 + r at the end of round r and 62
when the rounds are done.  With the median workgroup (the default for the
median bandwidth) the sort and the selection run there: it stamps 44 when its
predict is done and 45 when the median is published, and workgroup 0's wait for
it falls into the first iteration's exchange.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os

import numpy as np

from . import _lib

TRACE_LIB = os.path.join(os.path.dirname(_lib.LIB_PATH), "libsvpf_amd_trace.so")
CLOCK_HZ = 100e6


def phase_split(buf: np.ndarray, ticks: range, iters: int, median_workgroup: bool = False) -> dict:
    """Median microseconds per phase over the given ticks (and the mean and 90th percentile of
    the tick); buf [256, 64] uint64 stamps."""
    us = lambda a, b: (buf[ticks.start:ticks.stop, b].astype(np.int64) - buf[ticks.start:ticks.stop, a].astype(np.int64)) / CLOCK_HZ * 1e6
    if median_workgroup:
        out = {
            "predict": us(0, 1),
            "predictive_loglik": us(1, 2),
            "mixture_means": us(3, 4),
            "median_workgroup_predict_done_after_tick_start": us(0, 44),
            "median_workgroup_sort_and_selection": us(44, 45),
            "median_published_after_tick_start": us(0, 45),
        }
    else:
        out = {
            "predict": us(0, 1),
            "predictive_loglik": us(1, 2),
            "sort": us(2, 50),
            "median_selection": us(50, 3),
            "mixture_means": us(3, 4),
        }
    prev = 4
    grad, exch, tran = [], [], []
    for k in range(iters):
        grad.append(us(prev, 8 + 3 * k))
        exch.append(us(8 + 3 * k, 9 + 3 * k))
        tran.append(us(9 + 3 * k, 10 + 3 * k))
        prev = 10 + 3 * k
    out["gradient_per_iteration"] = np.mean(grad, axis=0)
    out["exchange_per_iteration"] = np.mean(exch, axis=0)
    out["transport_per_iteration"] = np.mean(tran, axis=0)
    out["stein_loop_total"] = np.sum(grad, axis=0) + np.sum(exch, axis=0) + np.sum(tran, axis=0)
    out["outputs_reduction"] = us(prev, 40)
    out["final_exchange"] = us(40, 41)  # final positions and partials, one gather
    out["finish"] = us(41, 42)
    out["tick"] = us(0, 42)
    res = {k: float(np.median(v)) for k, v in out.items()}
    res["tick_mean"] = float(np.mean(out["tick"]))
    res["tick_p90"] = float(np.quantile(out["tick"], 0.9))
    if median_workgroup:
        res["median_published_mean"] = float(np.mean(out["median_published_after_tick_start"]))
        return res
    # median selection: rounds (slots 52..61 stamped at the end of each) and the candidate phase
    sel = buf[ticks.start:ticks.stop]
    rounds = []
    for row in sel:
        k = 0
        while k < 10 and row[52 + k] > row[50] and (k == 0 or row[52 + k] > row[51 + k]):
            k += 1
        rounds.append(k)
    res["median_rounds"] = float(np.median(rounds))
    res["median_first_round"] = float(np.median(us(50, 52)))
    res["median_candidates_and_rank"] = float(np.median(us(62, 3)))
    return res


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--n", type=int, default=512)
    ap.add_argument("--ticks", type=int, default=200)
    ap.add_argument("--median-workgroup", type=int, default=1, choices=(0, 1),
                    help="1: the median in a workgroup of its own (the default); 0: in every row workgroup")
    args = ap.parse_args(argv)
    if not os.path.exists(TRACE_LIB):
        raise SystemExit(f"{TRACE_LIB} is not built: make -C newton-stein-variational-particle-filter_amd/csrc trace")
    _lib.LIB_PATH = TRACE_LIB
    L = _lib.load()
    import torch

    from . import config, data
    from .filter import Filter

    L.svpf_trace_dump.restype = ctypes.c_int
    L.svpf_trace_dump.argtypes = [ctypes.c_void_p]
    T = min(args.ticks, 256)
    cfg = config.plain_config(n=args.n)
    y, _ = data.sv_series(T, seed=0)
    f = Filter(cfg, median_workgroup=bool(args.median_workgroup))
    f.set_particles(data.initial_particles(args.n, -1.0, 0.97, 0.15, seed=2))
    path = f.path()
    if path != "coop":
        raise SystemExit(f"n={args.n} runs on the {path} path; the phase stamps are the cooperative kernel's")
    f.run_device(torch.tensor(y, device="cuda"))
    torch.cuda.synchronize()
    buf = np.zeros((256, 64), dtype=np.uint64)
    rc = L.svpf_trace_dump(buf.ctypes.data)
    if rc:
        raise SystemExit(f"svpf_trace_dump failed ({rc})")
    iters = int(cfg.get("stein_max_steps", 5))
    # skip the first ticks (the Stein budget's warm-up)
    split = phase_split(buf, range(10, T - 1), iters, median_workgroup=bool(args.median_workgroup))
    f.close()
    print(json.dumps({"n": args.n, "ticks": T, "stein_iterations": iters, "median_workgroup": args.median_workgroup,
                      "median_us": split}, indent=1))


if __name__ == "__main__":
    main()
