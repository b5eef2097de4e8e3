"""Seeded synthetic MSM workloads (shared by bench.py, the autotuner, the
server and the tests). Inputs only; no reference data.

- ``uniform_scalars``: uniform in [0, r) by rejection (5 x 64-bit draws, top
  limb masked to 43 bits).
- ``batch_workload``: the mixed-size batch — MSM sizes log-uniform in
  [2^lo, 2^hi], bases as slices of one resident pool at seeded offsets, one
  fresh scalar per point.
- ``closed_form_scalars``: for bases P_i = k_i G, the exponent
  t = sum_i s_i k_i mod r of each MSM, so that its result is t G.
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


R_ORDER = sum(v << (64 * j) for j, v in enumerate(R_LIMBS))


def closed_form_scalars(k: np.ndarray, s: np.ndarray, sizes) -> list:
    """t_j = sum_i s_i k_i mod r for consecutive MSMs: MSM j owns the next
    sizes[j] rows of k (u64) and s (rows of u64 limbs). Exact in numpy: every
    64 x 64-bit limb product is formed from 32-bit halves, each partial
    product split into 32-bit halves whose column sums stay below 2^64 for
    fewer than 2^32 rows."""
    k = np.ascontiguousarray(k, dtype=np.uint64)
    s = np.ascontiguousarray(s, dtype=np.uint64).reshape(len(k), -1)
    sizes = np.asarray(sizes, dtype=np.int64)
    assert sizes.sum() == len(k)
    starts = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    nz = sizes > 0
    m32, sh32 = np.uint64(0xFFFFFFFF), np.uint64(32)
    kl, kh = k & m32, k >> sh32
    acc = [0] * len(sizes)

    def seg_sum(v: np.ndarray) -> np.ndarray:
        out = np.zeros(len(sizes), dtype=np.uint64)
        if len(v):
            out[nz] = np.add.reduceat(v, starts[nz])
        return out

    for j in range(s.shape[1]):
        sl, sh = s[:, j] & m32, s[:, j] >> sh32
        for a, b, w in ((kl, sl, 0), (kl, sh, 32), (kh, sl, 32), (kh, sh, 64)):
            p = a * b
            lo, hi = seg_sum(p & m32), seg_sum(p >> sh32)
            for q in range(len(sizes)):
                acc[q] += (int(lo[q]) + (int(hi[q]) << 32)) << (64 * j + w)
    return [v % R_ORDER for v in acc]
