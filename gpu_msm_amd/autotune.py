"""Autotuner: measured sweeps of (window c, bucket split seg) per
(curve, log2 N, batch size) on homogeneous batches, the best-plan table, and
the planner MLP trained on the sweeps.

    python -m gpu_msm_amd.autotune [--out gpu_msm_amd/tuning/sweep.json] [--quick] [--train-only]

Each cell times `batch` MSMs of size N in one pipeline run (gm_msm_batch,
host window combine) for every candidate plan; the per-MSM time is the run
time divided by the batch size (minimum over repetitions).
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import planner as pl
from . import workloads as wl

CAP_POINTS = 1 << 22


def candidates(curve: int, log2n: int):
    """Plans timed per cell: widths in the band [h - 3, h + 2] around the
    heuristic's h whose top window is not sparse, times segments 4..32, plus
    the heuristic's own plan, (0, 0): the library's choice, recorded with
    seg 0."""
    import gpu_msm_amd as gm

    h, _, _ = gm.plan_query(curve, 1 << log2n, 5, 299)
    cs = [c for c in range(h - 3, h + 3) if pl.C_MIN <= c <= pl.C_MAX and pl.window_ok(c)]
    out = [(0, 0)]
    for c in cs:
        for s in pl.SEGS:
            if s <= (1 << (c - 1)):
                out.append((c, s))
    return out, h


def time_cell(gm, bases, scalars, pool_n: int, curve: int, log2n: int, batch: int, plan, reps: int, rng):
    n = 1 << log2n
    specs = []
    for k in range(batch):
        specs.append((n, int(rng.integers(0, pool_n - n + 1)), k * n, plan[0], plan[1]))
    gm.msm_batch(bases, scalars, specs, scalar_bits=299)  # warm-up
    best = 1e30
    for _ in range(reps):
        t0 = time.perf_counter()
        gm.msm_batch(bases, scalars, specs, scalar_bits=299)
        best = min(best, time.perf_counter() - t0)
    return best * 1e3 / batch


def sweep(curves=(1, 2), log2ns=range(8, 21), batches=(1, 8, 64, 256), reps: int = 3, log=print,
          on_cell=None):
    """Timings of every candidate plan per cell; `on_cell(records)` is called
    after each cell (the caller can checkpoint partial sweeps)."""
    import gpu_msm_amd as gm

    rng = np.random.default_rng(11)
    pool_n = 1 << max(log2ns)
    pool_k = wl.uniform_k(rng, pool_n)
    sc = gm.Scalars.upload(wl.uniform_scalars(rng, CAP_POINTS))
    records = []
    for curve in curves:
        bases = gm.Bases.from_k(curve, pool_k)
        for lg in log2ns:
            for b in batches:
                if (b << lg) > CAP_POINTS:
                    continue
                cands, h = candidates(curve, lg)
                for plan in cands:
                    ms = time_cell(gm, bases, sc, pool_n, curve, lg, b, plan, reps, rng)
                    rec = {"curve": curve, "log2n": lg, "batch": b, "c": plan[0] or h, "seg": plan[1],
                           "ms_per_msm": round(ms, 5), "heuristic_c": h}
                    if plan == (0, 0):
                        rec["heuristic"] = True  # the library's own plan (seg chosen per batch)
                    records.append(rec)
                cell = [r for r in records if r["curve"] == curve and r["log2n"] == lg and r["batch"] == b
                        and not r.get("heuristic")]
                best = min(cell, key=lambda r: r["ms_per_msm"])
                log(f"autotune: curve {curve} N=2^{lg} batch {b}: best c={best['c']} seg={best['seg']} "
                    f"{best['ms_per_msm']:.4f} ms/MSM over {len(cell)} plans")
                if on_cell:
                    on_cell(records)
        bases.free()
    sc.free()
    return records


def best_table(records):
    cells = {}
    for r in records:
        if not r.get("heuristic"):
            cells.setdefault((r["curve"], r["log2n"], r["batch"]), []).append(r)
    best = []
    for (cv, lg, b), rs in sorted(cells.items()):
        r = min(rs, key=lambda q: q["ms_per_msm"])
        best.append({"curve": cv, "log2n": lg, "batch": b, "c": r["c"], "seg": r["seg"], "ms_per_msm": r["ms_per_msm"]})
    return best


def training_set(records, tol: float = 0.02):
    """One sample per plan within `tol` of its cell's best time, so flat
    optima give averaged (not arbitrary) targets. Widths with a sparse top
    window are left out: the planner never picks them."""
    cells = {}
    for r in records:
        if pl.window_ok(r["c"]) and not r.get("heuristic"):
            cells.setdefault((r["curve"], r["log2n"], r["batch"]), []).append(r)
    xs, ts = [], []
    for (cv, lg, b), rs in sorted(cells.items()):
        t0 = min(q["ms_per_msm"] for q in rs)
        for q in rs:
            if q["ms_per_msm"] <= t0 * (1 + tol):
                xs.append(pl.features(lg, b, cv - 1))
                ts.append(pl.targets(q["c"], q["seg"], q["heuristic_c"]))
    return np.array(xs, dtype=np.float32), np.array(ts, dtype=np.float32)


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=os.path.join(pl.TUNING_DIR, "sweep.json"))
    ap.add_argument("--mlp-out", default=os.path.join(pl.TUNING_DIR, "mlp.json"))
    ap.add_argument("--quick", action="store_true")
    ap.add_argument("--epochs", type=int, default=4000)
    ap.add_argument("--train-only", action="store_true", help="retrain the MLP on the existing sweep file")
    args = ap.parse_args()
    log = lambda s: print(s, file=sys.stderr, flush=True)  # noqa: E731
    os.makedirs(os.path.dirname(os.path.abspath(args.out)), exist_ok=True)

    def save(recs):
        with open(args.out, "w") as f:
            json.dump({"records": recs, "best": best_table(recs)}, f, indent=0)

    if args.train_only:
        with open(args.out) as f:
            recs = json.load(f)["records"]
    elif args.quick:
        recs = sweep(curves=(1,), log2ns=(8, 12, 16), batches=(1, 8), reps=1, log=log, on_cell=save)
    else:
        recs = sweep(log=log, on_cell=save)
    table = best_table(recs)
    if not args.train_only:
        save(recs)
    x, t = training_set(recs)
    params, curve = pl.mlp_train(x, t, epochs=args.epochs)
    with open(args.mlp_out, "w") as f:
        json.dump({"params": [float(v) for v in params], "loss_first": float(curve[0]), "loss_last": float(curve[-1]),
                   "samples": int(len(x)), "epochs": args.epochs}, f)
    log(f"autotune: {len(recs)} timings, {len(table)} cells, mlp loss {curve[0]:.4f} -> {curve[-1]:.5f}")
    print(json.dumps({"cells": len(table), "timings": len(recs), "mlp_loss_first": float(curve[0]),
                      "mlp_loss_last": float(curve[-1])}))


if __name__ == "__main__":
    main()
