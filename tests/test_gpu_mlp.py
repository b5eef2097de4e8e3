"""HIP planner-MLP kernels (forward, backward, Adam training in a hipGraph)
against the numpy oracle. Tolerances are FP32: the GPU accumulates in float,
the oracle in float64."""
import ctypes

import numpy as np
import pytest

import gpu_msm_amd as gm
from gpu_msm_amd import planner as pl
from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu
RTOL, ATOL = 2e-4, 2e-6


def _data(seed, n):
    rng = np.random.default_rng(seed)
    x = rng.normal(0, 1, (n, 3)).astype(np.float32)
    t = np.stack([np.tanh(x[:, 0] - x[:, 2]), 0.5 * x[:, 1]], axis=1).astype(np.float32)
    return x, t


def test_forward_matches_oracle():
    p = mo.init(5)
    x, _ = _data(6, 100)
    y = np.zeros((100, 2), dtype=np.float32)
    assert gm.lib().gm_mlp_forward(p.ctypes.data, x.ctypes.data, 100, y.ctypes.data, 0) == 0
    want, _ = mo.forward(p, x.astype(np.float64))
    np.testing.assert_allclose(y, want, rtol=RTOL, atol=ATOL)


@pytest.mark.parametrize("n", [1, 37, 257])
def test_gradient_matches_oracle(n):
    p = mo.init(7) + np.random.default_rng(8).normal(0, 0.02, mo.NP).astype(np.float32)
    x, t = _data(9, n)
    g = np.zeros(mo.NP, dtype=np.float32)
    loss = np.zeros(1, dtype=np.float32)
    assert gm.lib().gm_mlp_grad(p.ctypes.data, x.ctypes.data, t.ctypes.data, n, g.ctypes.data, loss.ctypes.data, 0) == 0
    wl, wg = mo.loss_grad(p, x.astype(np.float64), t.astype(np.float64))
    np.testing.assert_allclose(loss[0], wl, rtol=RTOL)
    np.testing.assert_allclose(g, wg, rtol=1e-3, atol=1e-6)


def test_training_follows_oracle_adam():
    x, t = _data(10, 64)
    p0 = mo.init(11)
    p, curve = pl.mlp_train(x, t, epochs=200, lr=3e-3, init=p0)
    wp, wcurve = mo.adam_train(p0, x.astype(np.float64), t.astype(np.float64), 200, 3e-3)
    np.testing.assert_allclose(curve[:20], wcurve[:20], rtol=1e-3)
    assert curve[-1] < 0.5 * curve[0]
    np.testing.assert_allclose(p, wp, rtol=5e-2, atol=5e-3)
