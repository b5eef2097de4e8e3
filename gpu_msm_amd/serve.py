"""Served mixed MSM stream (continuous batching) — load generator and report.

Poisson arrivals at a fixed rate; each request is G1 with probability 0.8,
else G2; N log-uniform in [2^lo, 2^hi]; bases are slices of resident pools
(one per curve); scalars uniform in [0, r) taken from a resident host pool.

    python -m gpu_msm_amd.serve [--rate 200] [--requests 400] [--lo 10] [--hi 20]

Reports MSMs/s, server-side p50/p99 latency (submit -> result), client-side
p50/p99 from the scheduled arrival time, batching statistics, and checks a
sample of results against their closed form computed on the host and against
the same MSM run alone.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import threading
import time

import numpy as np

import gpu_msm_amd as gm
from . import workloads as wl


def make_stream(seed: int, count: int, lo: int, hi: int, pool_n: int, g2_frac: float = 0.2):
    rng = np.random.default_rng(seed)
    curves = np.where(rng.random(count) < g2_frac, gm.CURVE_G2, gm.CURVE_G1)
    sizes = np.minimum(wl.log_uniform_sizes(rng, count, lo, hi), pool_n)
    boff = np.array([int(rng.integers(0, pool_n - n + 1)) for n in sizes])
    return rng, curves, sizes, boff


class Stream:
    def __init__(self, log2_pool: int = 20, scalar_pool: int = 1 << 21, seed: int = 21):
        rng = np.random.default_rng(seed)
        self.pool_n = 1 << log2_pool
        k = wl.uniform_k(rng, self.pool_n)
        self.k = k  # pool bases are k_i G: results have the closed form (sum s_i k_i) G
        self.g1 = gm.Bases.from_k(gm.CURVE_G1, k)
        self.g2 = gm.Bases.from_k(gm.CURVE_G2, k)
        self.scalars = wl.uniform_scalars(rng, scalar_pool)

    def pool(self, curve):
        return self.g1 if curve == gm.CURVE_G1 else self.g2

    def scalar_slice(self, rng, n):
        off = int(rng.integers(0, len(self.scalars) - n + 1)) if n < len(self.scalars) else 0
        reps = -(-n // len(self.scalars))
        if reps > 1:
            return np.ascontiguousarray(np.tile(self.scalars, (reps, 1))[:n])
        return self.scalars[off:off + n]


def run(stream: Stream, rate: float, count: int, lo: int, hi: int, seed: int = 5, slots: int = 3,
        max_batch: int = 64, budget: int = 48 << 30, planner: int = gm.PLANNER_HEURISTIC, mlp_params=None,
        check: int = 16, log=None, results_out: list | None = None, small_max_n: int = 0) -> dict:
    """Serves `count` requests; with results_out, appends (curve, n,
    base_offset, scalars, rc, affine record) per request for the caller to
    check."""
    rng, curves, sizes, boff = make_stream(seed, count, lo, hi, stream.pool_n)
    arr = np.cumsum(rng.exponential(1.0 / rate, size=count))  # scheduled arrivals (s)
    scal = [stream.scalar_slice(rng, int(n)) for n in sizes]
    srv = gm.Server(stream.g1, stream.g2, slots=slots, max_batch=max_batch, hbm_budget_bytes=budget,
                    planner=planner, mlp_params=mlp_params, small_max_n=small_max_n)
    server_lat = np.zeros(count)
    tickets = [None] * count
    ready = [threading.Event() for _ in range(count)]
    results = [None] * count
    client_lat = np.zeros(count)
    t0 = time.perf_counter()

    def submitter():
        for i in range(count):
            d = t0 + arr[i] - time.perf_counter()
            if d > 0:
                time.sleep(d)
            tickets[i] = srv.submit(int(curves[i]), int(boff[i]), scal[i], scalar_bits=299, borrow=True)
            ready[i].set()

    def waiter():
        for i in range(count):
            ready[i].wait()
            rc, res, lat = srv.wait(tickets[i], int(curves[i]))
            client_lat[i] = (time.perf_counter() - (t0 + arr[i])) * 1e3
            server_lat[i] = lat
            results[i] = (rc, res)

    ts, tw = threading.Thread(target=submitter), threading.Thread(target=waiter)
    ts.start()
    tw.start()
    ts.join()
    tw.join()
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    ok = sum(1 for rc, _ in results if rc == 0)
    if results_out is not None:
        for i in range(count):
            results_out.append((int(curves[i]), int(sizes[i]), int(boff[i]), scal[i], results[i][0], results[i][1]))
    # sampled checks: the closed form (sum_i s_i k_i mod r) G computed on the
    # host, and the same MSM run alone on the GPU
    idx = np.linspace(0, count - 1, min(check, count)).astype(int)
    same = closed = True
    for i in idx:
        n, b, cv = int(sizes[i]), int(boff[i]), int(curves[i])
        t = wl.closed_form_scalars(stream.k[b:b + n], scal[i], [n])[0]
        closed &= results[i][0] == 0 and results[i][1].tobytes() == gm.generator_multiple(cv, t).tobytes()
        sc = gm.Scalars.upload(scal[i])
        alone = gm.msm_resident(stream.pool(cv), sc, n=n, base_offset=b, scalar_bits=299)
        sc.free()
        same &= results[i][0] == 0 and alone.tobytes() == results[i][1].tobytes()
    # latency per curve and size class (the server's small class: N <= 2^16)
    small_n = (1 << 16) if small_max_n == 0 else small_max_n
    by_class = {}
    for cv, cname in ((gm.CURVE_G1, "g1"), (gm.CURVE_G2, "g2")):
        for small, sname in ((True, f"n_le_{small_n}"), (False, f"n_gt_{small_n}")):
            sel = (curves == cv) & ((sizes <= small_n) == small)
            if sel.any():
                v = server_lat[sel]
                by_class[f"{cname}_{sname}"] = {"requests": int(sel.sum()),
                                                "p50_ms": round(float(np.percentile(v, 50)), 3),
                                                "p99_ms": round(float(np.percentile(v, 99)), 3)}
    return {
        "requests": count, "rate_per_s": rate, "completed_ok": ok, "wall_s": round(wall, 3),
        "latency_by_class": by_class,
        "msms_per_s": round(ok / wall, 1),
        "p50_ms": round(st["p50_ms"], 3), "p99_ms": round(st["p99_ms"], 3), "max_ms": round(st["max_ms"], 3),
        "client_p50_ms": round(float(np.percentile(client_lat, 50)), 3),
        "client_p99_ms": round(float(np.percentile(client_lat, 99)), 3),
        "batches": st["batches"], "mean_batch": round(st["mean_batch"], 2),
        "peak_concurrent_batches": st["peak_concurrent_batches"],
        "peak_inflight_GB": round(st["peak_inflight_bytes"] / 1e9, 3),
        "g2_fraction": round(float(np.mean(curves == gm.CURVE_G2)), 3),
        "mean_n": int(np.mean(sizes)),
        "sampled_closed_form": bool(closed), "sampled_same_alone": bool(same), "sampled": int(len(idx)),
    }


def plain(stream: Stream, count: int, lo: int, hi: int, seed: int = 5) -> dict:
    """The same request list run one MSM at a time (no batching), back to back."""
    rng, curves, sizes, boff = make_stream(seed, count, lo, hi, stream.pool_n)
    rng.exponential(1.0, size=count)
    scal = [stream.scalar_slice(rng, int(n)) for n in sizes]
    lat = []
    t0 = time.perf_counter()
    for i in range(count):
        t1 = time.perf_counter()
        sc = gm.Scalars.upload(scal[i])
        gm.msm_resident(stream.pool(int(curves[i])), sc, n=int(sizes[i]), base_offset=int(boff[i]), scalar_bits=299)
        sc.free()
        lat.append((time.perf_counter() - t1) * 1e3)
    wall = time.perf_counter() - t0
    return {"requests": count, "msms_per_s": round(count / wall, 1), "wall_s": round(wall, 3),
            "service_p50_ms": round(float(np.percentile(lat, 50)), 3),
            "service_p99_ms": round(float(np.percentile(lat, 99)), 3)}


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--rate", type=float, default=200.0)
    ap.add_argument("--requests", type=int, default=400)
    ap.add_argument("--lo", type=int, default=10)
    ap.add_argument("--hi", type=int, default=20)
    ap.add_argument("--slots", type=int, default=3)
    ap.add_argument("--max-batch", type=int, default=64)
    ap.add_argument("--plain", action="store_true")
    args = ap.parse_args()
    st = Stream()
    out = {"served": run(st, args.rate, args.requests, args.lo, args.hi, slots=args.slots,
                         max_batch=args.max_batch)}
    if args.plain:
        out["plain"] = plain(st, args.requests, args.lo, args.hi)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
