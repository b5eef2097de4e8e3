"""Seeded synthetic MSM workloads (shared by bench.py, the autotuner, the
server and the tests). Inputs only; no reference data.

- ``uniform_scalars``: uniform in [0, r) by rejection (5 x 64-bit draws, top
  limb masked to 43 bits).
- ``batch_workload``: the mixed-size batch — MSM sizes log-uniform in
  [2^lo, 2^hi], bases as slices of one resident pool at seeded offsets, one
  fresh scalar per point.
"""
from __future__ import annotations

import numpy as np

R_LIMBS = [0x0428001400040001, 0x7BB9B0E8D8CA3461, 0xD04C98CCC4C050BC, 0x7995B34995830FA4, 0x00000511B70539F2]


def uniform_scalars(rng: np.random.Generator, n: int) -> np.ndarray:
    out = np.zeros((0, 5), dtype=np.uint64)
    r = np.array(R_LIMBS, dtype=np.uint64)
    while len(out) < n:
        m = int((n - len(out)) * 1.7) + 64
        v = np.frombuffer(rng.bytes(40 * m), dtype=np.uint64).reshape(m, 5).copy()
        v[:, 4] &= np.uint64((1 << 43) - 1)
        lt = np.zeros(m, dtype=bool)
        eq = np.ones(m, dtype=bool)
        for j in range(4, -1, -1):
            lt |= eq & (v[:, j] < r[j])
            eq &= v[:, j] == r[j]
        out = np.concatenate([out, v[lt]])
    return np.ascontiguousarray(out[:n])


def uniform_k(rng: np.random.Generator, n: int) -> np.ndarray:
    return np.frombuffer(rng.bytes(8 * n), dtype=np.uint64).copy()


def log_uniform_sizes(rng: np.random.Generator, count: int, lo: int, hi: int) -> np.ndarray:
    return np.floor(2.0 ** rng.uniform(lo, hi, size=count)).astype(np.int64)


def batch_workload(seed: int = 4, count: int = 256, lo: int = 8, hi: int = 18, pool_log2: int = 18):
    """Returns (pool_k, sizes, base_offsets, scalar_offsets, scalars)."""
    rng = np.random.default_rng(seed)
    pool_n = 1 << pool_log2
    pool_k = uniform_k(rng, pool_n)
    sizes = log_uniform_sizes(rng, count, lo, hi)
    sizes = np.minimum(sizes, pool_n)
    base_off = np.array([int(rng.integers(0, pool_n - n + 1)) for n in sizes], dtype=np.int64)
    scal_off = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    scalars = uniform_scalars(rng, int(sizes.sum()))
    return pool_k, sizes, base_off, scal_off, scalars


def specs_from(sizes, base_off, scal_off, windows=None, segs=None):
    """(n, base_offset, scalar_offset, window, seg) tuples."""
    out = []
    for k in range(len(sizes)):
        w = 0 if windows is None else int(windows[k])
        s = 0 if segs is None else int(segs[k])
        out.append((int(sizes[k]), int(base_off[k]), int(scal_off[k]), w, s))
    return out
