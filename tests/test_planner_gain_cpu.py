"""What the planner MLP gains over the analytic heuristic on cells it was
NOT trained on, read off the committed autotuner sweep
(gpu_msm_amd/tuning/sweep.json: measured per-MSM times of every candidate
(c, seg) per (curve, log2 N, batch) cell on an MI355X).

Five-fold cross-validation over the sweep's cells: the network is trained
(the oracle's Adam, the same full-batch recipe the library's GPU trainer
runs) on the plans within 2% of the best of four fifths of the cells, and
each held-out cell's MLP plan is looked up among that cell's measured
candidates and compared with the heuristic's measured plan (window 0, seg 0:
the library's own choice, recorded by the sweep). The network predicts a
correction to the heuristic's width (planner.targets); the same test with
absolute widths as targets gains nothing held out (geometric mean 1.000).
The committed network (gpu_msm_amd/tuning/mlp.json, trained on every cell)
is checked for its decoding and in-sample fit. CPU only."""
import json
import math
import os

import numpy as np

from gpu_msm_amd import autotune as at
from gpu_msm_amd import planner as pl
from oracle import mlp_oracle as mo

SWEEP = os.path.join(pl.TUNING_DIR, "sweep.json")
MLP = os.path.join(pl.TUNING_DIR, "mlp.json")


def load():
    with open(SWEEP) as f:
        recs = json.load(f)["records"]
    cells, heur = {}, {}
    for r in recs:
        key = (r["curve"], r["log2n"], r["batch"])
        if r.get("heuristic"):
            heur[key] = r["ms_per_msm"]
        else:
            cells.setdefault(key, {})[(r["c"], r["seg"])] = r["ms_per_msm"]
    return recs, cells, heur


def mlp_gains(params, keys, cells, heur):
    """(cell, heuristic ms / MLP-plan ms) for every key whose MLP plan was
    measured."""
    mlp = pl.MlpPlanner(np.asarray(params, dtype=np.float32))
    out = []
    for cv, lg, b in keys:
        y, _ = mo.forward(np.asarray(params), pl.features(np.array([lg]), np.array([b]),
                                                          np.array([cv - 1])).astype(np.float64))
        w, s = mlp.plans_from_outputs(cv, np.array([1 << lg]), y.astype(np.float32))
        t = cells[(cv, lg, b)].get((int(w[0]), int(s[0])))
        if t is not None:
            out.append(((cv, lg, b), heur[(cv, lg, b)] / t))
    return out


def geo(xs):
    return math.exp(sum(math.log(x) for x in xs) / len(xs))


def test_sweep_covers_the_plans():
    _, cells, heur = load()
    assert len(cells) >= 80 and len([k for k in cells if k in heur]) >= 80


def test_mlp_gains_over_heuristic_on_held_out_cells():
    recs, cells, heur = load()
    keys = sorted(k for k in cells if k in heur)
    order = np.random.default_rng(0).permutation(len(keys))
    rows = []
    for fold in range(5):
        held = {keys[i] for i in order[fold::5]}
        x, t = at.training_set([r for r in recs if (r["curve"], r["log2n"], r["batch"]) not in held])
        params, _ = mo.adam_train(mo.init(0), x, t, 4000, 3e-3)
        rows += mlp_gains(params, sorted(held), cells, heur)
    gains = [g for _, g in rows]
    batched = [g for (cell, g) in rows if cell[2] >= 8 and cell[1] <= 12]
    for cell, g in rows:
        print(f"held out: curve {cell[0]} N=2^{cell[1]} batch {cell[2]}: gain {g:.3f}")
    print(f"held-out geometric-mean gain {geo(gains):.3f} over {len(gains)} cells; {geo(batched):.3f} on "
          f"batched N <= 2^12; worst {min(gains):.3f}")
    assert len(rows) >= 75
    assert geo(gains) > 1.01
    assert geo(batched) > 1.03


def test_committed_network_decodes_and_fits_the_sweep():
    _, cells, heur = load()
    with open(MLP) as f:
        params = np.array(json.load(f)["params"], dtype=np.float64)
    keys = sorted(k for k in cells if k in heur)
    rows = mlp_gains(params, keys, cells, heur)
    assert len(rows) >= 0.9 * len(keys)  # its plans are measured candidates
    print(f"in-sample geometric-mean gain {geo([g for _, g in rows]):.3f} over {len(rows)} cells")
    assert geo([g for _, g in rows]) > 1.0
