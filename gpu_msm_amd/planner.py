"""MSM plan selection: window width c and bucket split (buckets per
bucket-reduction lane, `seg`) for each MSM of a batch.

Three planners share one interface, ``plan(curve, sizes, batch) -> (windows,
segs)``:

- ``HeuristicPlanner``: the library's analytic cost model (window 0, seg 0).
- ``TablePlanner``: the autotuner's measured optimum per (curve, log2 N,
  batch size), nearest neighbour in (log2 N, log2 batch).
- ``MlpPlanner``: an FP32 MLP (inputs log2 N, batch size, curve; two hidden
  layers of 32) trained on the autotuner's sweeps by the library's own HIP
  kernels. It predicts a correction to the heuristic's width and the
  bucket split; the width h + correction is rounded to the nearest valid c
  inside the sweeps' candidate band (heuristic c - 3 .. + 2, where the
  network was trained) and off widths with a sparse top window; a non-finite
  prediction falls back to the heuristic.

Every planner only changes how the work is split, never the result: any
valid (c, seg) gives the same bit-exact MSM.
"""
from __future__ import annotations

import json
import math
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TUNING_DIR = os.path.join(HERE, "tuning")

C_MIN, C_MAX = 3, 20
SEGS = (4, 8, 16, 32)


def features(log2n, batch, curve) -> np.ndarray:
    """MLP inputs: (log2 N - 13)/5, (log2 batch - 3)/3, curve (0 G1, 1 G2)."""
    log2n = np.asarray(log2n, dtype=np.float64)
    lb = np.log2(np.maximum(np.asarray(batch, dtype=np.float64), 1.0))
    cv = np.asarray(curve, dtype=np.float64)
    log2n, lb, cv = np.broadcast_arrays(log2n, lb, cv)
    return np.stack([(log2n - 13.0) / 5.0, (lb - 3.0) / 3.0, cv], axis=-1).astype(np.float32)


def targets(c, seg, h) -> np.ndarray:
    """MLP outputs for a plan: the width as a correction to the heuristic's
    width h, (c - h)/2, and (log2 seg - 3)/1.5. The network learns where the
    analytic model is off (notably with the batch size, which it does not
    see), not the width itself: trained on all but a fifth of the sweep's
    cells, absolute targets gain nothing on the held-out fifth, corrections do
    (tests/test_planner_gain_cpu.py)."""
    c = np.asarray(c, dtype=np.float64)
    h = np.asarray(h, dtype=np.float64)
    ls = np.log2(np.asarray(seg, dtype=np.float64))
    return np.stack([(c - h) / 2.0, (ls - 3.0) / 1.5], axis=-1).astype(np.float32)


def decode(y: np.ndarray):
    """MLP outputs -> (width correction dc, seg), or None if unusable."""
    dc = y[..., 0] * 2.0
    ls = y[..., 1] * 1.5 + 3.0
    out = []
    for di, li in zip(np.atleast_1d(dc), np.atleast_1d(ls)):
        if not (math.isfinite(di) and math.isfinite(li)) or abs(di) > C_MAX:
            out.append(None)
            continue
        seg = 2 ** int(round(min(max(li, 2.0), 5.0)))
        out.append((float(di), seg))
    return out


def window_ok(c: int, bits: int = 299) -> bool:
    """Same rule as the library's window_ok: the top window keeps at least
    4/5 of its c bits (a sparse top window sends a whole window to a few
    buckets and serialises the bucket phase)."""
    nw = (bits + c) // c
    return 5 * (bits + 1 - (nw - 1) * c) >= 4 * c


def snap_window(c: float, h: int, bits: int = 299) -> int:
    """Predicted width -> the nearest width in the sweeps' candidate band
    [h - 3, h + 2] around the heuristic h whose top window is not sparse."""
    lo, hi = max(C_MIN, h - 3), min(C_MAX, h + 2)
    ct = min(max(float(c), lo), hi)
    best = h
    for k in range(lo, hi + 1):
        if window_ok(k, bits) and abs(k - ct) < abs(best - ct):
            best = k
    return best


class HeuristicPlanner:
    name = "heuristic"

    def plan(self, curve: int, sizes, batch: int):
        n = len(sizes)
        return np.zeros(n, dtype=np.int32), np.zeros(n, dtype=np.int32)


class TablePlanner:
    name = "autotuned"

    def __init__(self, table: list[dict]):
        # entries: {"curve": 1|2, "log2n": int, "batch": int, "c": int, "seg": int}
        self.entries = table

    @classmethod
    def load(cls, path: str) -> "TablePlanner":
        with open(path) as f:
            return cls(json.load(f)["best"])

    def lookup(self, curve: int, log2n: float, batch: int):
        best, bd = None, 1e30
        lb = math.log2(max(batch, 1))
        for e in self.entries:
            if e["curve"] != curve:
                continue
            d = (e["log2n"] - log2n) ** 2 + 0.25 * (math.log2(e["batch"]) - lb) ** 2
            if d < bd:
                bd, best = d, e
        return (best["c"], best["seg"]) if best else (0, 0)

    def plan(self, curve: int, sizes, batch: int):
        w = np.zeros(len(sizes), dtype=np.int32)
        s = np.zeros(len(sizes), dtype=np.int32)
        for k, n in enumerate(sizes):
            w[k], s[k] = self.lookup(curve, math.log2(max(int(n), 1)), batch)
        return w, s


class MlpPlanner:
    name = "mlp"

    def __init__(self, params: np.ndarray, gpu_index: int = 0):
        self.params = np.ascontiguousarray(params, dtype=np.float32)
        self.gpu = gpu_index
        self.fallbacks = 0

    @classmethod
    def load(cls, path: str) -> "MlpPlanner":
        with open(path) as f:
            return cls(np.array(json.load(f)["params"], dtype=np.float32))

    def predict(self, x: np.ndarray) -> np.ndarray:
        """Forward pass on the GPU (the library's k_mlp_forward)."""
        from . import _check, _ptr, lib

        x = np.ascontiguousarray(x, dtype=np.float32)
        y = np.zeros((len(x), 2), dtype=np.float32)
        _check(lib().gm_mlp_forward(_ptr(self.params), _ptr(x), len(x), _ptr(y), self.gpu), "mlp forward")
        return y

    def plan(self, curve: int, sizes, batch: int):
        sizes = np.asarray(sizes)
        x = features(np.log2(np.maximum(sizes, 1)), np.full(len(sizes), batch), np.full(len(sizes), curve - 1))
        return self.plans_from_outputs(curve, sizes, self.predict(x))

    def plans_from_outputs(self, curve: int, sizes, y: np.ndarray):
        """Network outputs (one row per MSM) -> (windows, segs); a row that
        decodes to nothing usable leaves that MSM to the heuristic (0, 0)."""
        import gpu_msm_amd as gm

        dec = decode(y)
        w = np.zeros(len(sizes), dtype=np.int32)
        s = np.zeros(len(sizes), dtype=np.int32)
        heur = {}
        for k, d in enumerate(dec):
            if d is None:
                self.fallbacks += 1  # heuristic for this MSM
                continue
            n = int(sizes[k])
            if n not in heur:
                heur[n] = gm.plan_query(curve, max(n, 1), 5, 299)[0]
            c = snap_window(heur[n] + d[0], heur[n])
            w[k], s[k] = c, min(d[1], 1 << (c - 1))
        return w, s


def mlp_train(x: np.ndarray, t: np.ndarray, epochs: int = 3000, lr: float = 3e-3, seed: int = 0,
              gpu_index: int = 0, init: np.ndarray | None = None):
    """Trains the planner MLP on the GPU (gm_mlp_train); returns (params, loss curve)."""
    from . import _check, _ptr, lib

    if init is None:
        rng = np.random.default_rng(seed)
        parts = []
        for shape in ((32, 3), (32,), (32, 32), (32,), (2, 32), (2,)):
            if len(shape) == 2:
                parts.append(rng.normal(0, np.sqrt(2.0 / shape[1]), size=shape).ravel())
            else:
                parts.append(np.zeros(shape))
        init = np.concatenate(parts)
    p = np.ascontiguousarray(init, dtype=np.float32).copy()
    x = np.ascontiguousarray(x, dtype=np.float32)
    t = np.ascontiguousarray(t, dtype=np.float32)
    curve = np.zeros(epochs, dtype=np.float32)
    _check(lib().gm_mlp_train(_ptr(p), _ptr(x), _ptr(t), len(x), epochs, lr, _ptr(curve), gpu_index), "mlp train")
    return p, curve
