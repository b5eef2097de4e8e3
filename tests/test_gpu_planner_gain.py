"""What the trained planner MLP gains over the analytic heuristic, measured
live: homogeneous batches of MSMs of one size run once with the heuristic's
plan (window 0, seg 0: the library's own choice) and once with the plan the
committed network (gpu_msm_amd/tuning/mlp.json) predicts through the HIP
forward kernel. Both runs must give bit-identical results (a plan only splits
the work); the per-MSM time of each is the best of REPS runs.

The cells are the sweep's batched small-MSM cells, where the heuristic (which
does not see the batch size) leaves the most on the table, plus single MSMs
of medium size as a control. The test prints each cell and the geometric-mean
gain (1.02-1.03 on an MI355X when the network was committed) and asserts only
that the network is not grossly slower: single cells swing by tens of percent
between runs. tests/test_planner_gain_cpu.py asserts the gain itself over
every cell of the committed sweep."""
import math
import time

import numpy as np
import pytest

import gpu_msm_amd as gm
from gpu_msm_amd import planner as pl
from gpu_msm_amd import workloads as wl

pytestmark = pytest.mark.gpu

POOL = 1 << 18
REPS = 9
CELLS = [  # (curve, log2 N, batch)
    (1, 8, 8), (1, 9, 8), (1, 10, 8), (1, 11, 64), (1, 12, 8), (1, 12, 64), (1, 16, 1), (1, 15, 8),
    (2, 8, 64), (2, 8, 256), (2, 9, 1), (2, 9, 256), (2, 10, 64), (2, 13, 8), (2, 14, 1),
]


def _time_pair(bases, scalars, a, b):
    """Best-of-REPS per-MSM times of two spec lists, runs interleaved so that
    clock or load drift hits both alike; returns (ta, tb, result_a, result_b)."""
    ra = gm.msm_batch(bases, scalars, a, scalar_bits=299)
    rb = gm.msm_batch(bases, scalars, b, scalar_bits=299)
    ta = tb = 1e30
    for _ in range(REPS):
        t0 = time.perf_counter()
        ra = gm.msm_batch(bases, scalars, a, scalar_bits=299)
        ta = min(ta, time.perf_counter() - t0)
        t0 = time.perf_counter()
        rb = gm.msm_batch(bases, scalars, b, scalar_bits=299)
        tb = min(tb, time.perf_counter() - t0)
    return ta * 1e3 / len(a), tb * 1e3 / len(b), ra, rb


def test_mlp_plans_beat_heuristic_live():
    mlp = pl.MlpPlanner.load(f"{pl.TUNING_DIR}/mlp.json")
    rng = np.random.default_rng(77)
    pool_k = wl.uniform_k(rng, POOL)
    sc = gm.Scalars.upload(wl.uniform_scalars(rng, POOL))
    gains, lines = [], []
    for curve in (1, 2):
        bases = gm.Bases.from_k(curve, pool_k)
        for cv, lg, b in CELLS:
            if cv != curve:
                continue
            n = 1 << lg
            assert b * n <= POOL
            offs = [(int(rng.integers(0, POOL - n + 1)), k * n) for k in range(b)]
            w, s = mlp.plan(curve, np.full(b, n), b)
            h = [(n, bo, so) for bo, so in offs]
            m = [(n, bo, so, int(w[k]), int(s[k])) for k, (bo, so) in enumerate(offs)]
            print(f"cell curve {cv} N=2^{lg} batch {b}", flush=True)
            th, tm, rh, rm = _time_pair(bases, sc, h, m)
            assert rh.tobytes() == rm.tobytes(), (cv, lg, b)
            hc = gm.plan_query(curve, n, 5, 299)[0]
            gains.append(th / tm)
            lines.append(f"curve {cv} N=2^{lg} batch {b}: heuristic c={hc} {th:.4f} ms/MSM, "
                         f"mlp (c={w[0]}, seg={s[0]}) {tm:.4f} ms/MSM, gain {th / tm:.3f}")
        bases.free()
    sc.free()
    geo = math.exp(sum(math.log(g) for g in gains) / len(gains))
    print("\n".join(lines))
    print(f"geometric-mean gain of the MLP over the heuristic: {geo:.3f} over {len(gains)} cells")
    assert mlp.fallbacks == 0
    # single cells swing with timing noise; the committed sweep pins the gain
    # itself, this guards against a network that has gone wrong
    assert geo > 0.9
