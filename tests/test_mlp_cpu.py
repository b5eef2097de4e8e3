"""Pins the planner-MLP oracle (oracle/mlp_oracle.py) by central finite
differences, and the planner's feature/decode helpers. CPU only."""
import numpy as np

import gpu_msm_amd as gm
from gpu_msm_amd import planner as pl
from oracle import mlp_oracle as mo


def test_param_count_matches_library_layout():
    assert mo.NP == 1250


def test_oracle_gradient_matches_finite_differences():
    rng = np.random.default_rng(1)
    p = mo.init(3).astype(np.float64) + rng.normal(0, 0.05, mo.NP)
    x = rng.normal(0, 1, (17, 3))
    t = rng.normal(0, 1, (17, 2))
    _, g = mo.loss_grad(p, x, t)
    eps = 1e-6
    idx = rng.choice(mo.NP, 80, replace=False)
    for i in idx:
        pp, pm = p.copy(), p.copy()
        pp[i] += eps
        pm[i] -= eps
        num = (mo.loss_grad(pp, x, t)[0] - mo.loss_grad(pm, x, t)[0]) / (2 * eps)
        assert abs(num - g[i]) <= 1e-6 + 1e-5 * abs(num), (i, num, g[i])


def test_oracle_adam_reduces_loss():
    rng = np.random.default_rng(2)
    x = rng.normal(0, 1, (40, 3))
    t = np.stack([np.sin(x[:, 0]), x[:, 1] * x[:, 2]], axis=1)
    _, curve = mo.adam_train(mo.init(0), x, t, 300, 3e-3)
    assert curve[-1] < 0.3 * curve[0]


def test_planner_features_and_decode_roundtrip():
    x = pl.features(20, 1, 0)
    assert np.allclose(x, [1.4, -1.0, 0.0])
    y = pl.targets([16, 9, 4], [8, 4, 32], [15, 10, 4])
    dec = pl.decode(y)
    assert dec == [(1.0, 8), (-1.0, 4), (0.0, 32)]  # width corrections to the heuristic's
    assert pl.decode(np.array([[np.nan, 0.0]], dtype=np.float32)) == [None]
    assert pl.decode(np.array([[11.0, 0.0]], dtype=np.float32)) == [None]  # correction of 22: out of range
    mlp = pl.MlpPlanner(np.zeros(1250, dtype=np.float32))
    h = gm.plan_query(1, 1 << 12, 5, 299)[0]
    w, s = mlp.plans_from_outputs(1, np.array([1 << 12]), pl.targets([h + 2], [8], [h]))
    assert (w[0] == h + 2 or not pl.window_ok(h + 2)) and s[0] == 8


def test_table_planner_nearest_cell():
    tp = pl.TablePlanner([{"curve": 1, "log2n": 10, "batch": 1, "c": 9, "seg": 4},
                          {"curve": 1, "log2n": 16, "batch": 64, "c": 14, "seg": 8},
                          {"curve": 2, "log2n": 10, "batch": 1, "c": 8, "seg": 4}])
    w, s = tp.plan(1, [1 << 10, 1 << 15, 1 << 17], 64)
    assert list(w) == [9, 14, 14] and list(s) == [4, 8, 8]
