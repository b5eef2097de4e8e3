"""What the trained planner MLP gains over the analytic heuristic, read off
the committed autotuner sweep (gpu_msm_amd/tuning/sweep.json: measured
per-MSM times of every candidate (c, seg) per (curve, log2 N, batch) cell on
an MI355X) for the committed network (gpu_msm_amd/tuning/mlp.json).

For each cell the heuristic's time is the sweep's own measurement of the
library's plan (window 0, seg 0: the record marked "heuristic"); the MLP's
plan is the network's output decoded exactly as MlpPlanner does, looked up
among the cell's measured candidates. The network sees the batch size; the
heuristic does not, which is where it gains (batched small MSMs, whose best
width grows with the batch). tests/test_gpu_planner_gain.py measures the same
comparison live. CPU only: the network's forward pass here is the float64
oracle."""
import json
import math
import os

import numpy as np
import pytest

import gpu_msm_amd as gm
from gpu_msm_amd import planner as pl
from oracle import mlp_oracle as mo

SWEEP = os.path.join(pl.TUNING_DIR, "sweep.json")
MLP = os.path.join(pl.TUNING_DIR, "mlp.json")


def cell_gains():
    with open(SWEEP) as f:
        recs = json.load(f)["records"]
    with open(MLP) as f:
        params = np.array(json.load(f)["params"], dtype=np.float64)
    cells, heur = {}, {}
    for r in recs:
        key = (r["curve"], r["log2n"], r["batch"])
        if r.get("heuristic"):
            heur[key] = r["ms_per_msm"]
        else:
            cells.setdefault(key, {})[(r["c"], r["seg"])] = r["ms_per_msm"]
    mlp = pl.MlpPlanner(params.astype(np.float32))
    out = []
    for (cv, lg, b), times in sorted(cells.items()):
        if (cv, lg, b) not in heur:
            continue
        y, _ = mo.forward(params, pl.features(np.array([lg]), np.array([b]), np.array([cv - 1])).astype(np.float64))
        w, s = mlp.plans_from_outputs(cv, np.array([1 << lg]), y.astype(np.float32))
        mplan = (int(w[0]), int(s[0]))
        if mplan in times:
            out.append(((cv, lg, b), mplan, heur[(cv, lg, b)], times[mplan], min(times.values())))
    return out, len(cells)


def test_sweep_covers_the_plans():
    rows, ncells = cell_gains()
    assert ncells >= 80
    assert len(rows) >= 0.9 * ncells  # the heuristic and the MLP's plan were measured


def test_mlp_gains_over_heuristic_on_measured_sweep():
    rows, _ = cell_gains()
    gains = [th / tm for _, _, th, tm, _ in rows]
    geo = math.exp(sum(math.log(g) for g in gains) / len(gains))
    batched = [th / tm for (cell, _, th, tm, _) in rows if cell[2] >= 8 and cell[1] <= 12]
    geo_b = math.exp(sum(math.log(g) for g in batched) / len(batched))
    worst = min(gains)
    for cell, mp, th, tm, best in rows:
        print(f"curve {cell[0]} N=2^{cell[1]} batch {cell[2]}: heuristic {th:.4f} ms, mlp {mp} {tm:.4f} ms, "
              f"best {best:.4f} ms, gain {th / tm:.3f}")
    print(f"geometric-mean gain {geo:.3f} over {len(rows)} cells; {geo_b:.3f} on batched N <= 2^12; worst {worst:.3f}")
    assert geo > 1.02
    assert geo_b > 1.04
    assert worst > 0.8
