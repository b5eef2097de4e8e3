"""The two multi-MSM workloads at their stated sizes, every result checked
against its closed form computed on the host by the oracle.

- Batched: 256 G1 MSMs, N log-uniform in [2^8, 2^18], bases slices of one
  resident pool of 2^18 points k_i G, scalars uniform in [0, r), run as one
  hipGraph-captured pipeline (replayed), under the analytic heuristic's plan,
  the autotuned table's and the planner MLP's.
- Served: a Poisson stream of 48 requests, 80% G1 / 20% G2, N log-uniform in
  [2^10, 2^20], over resident pools of 2^20 bases per curve, through the
  continuous-batching server.

For bases P_i = k_i G the MSM equals t G with t = sum_i s_i k_i mod r; t
comes from exact integer sums (workloads.closed_form_scalars, itself checked
against Python integers in test_workloads_cpu.py) and t G from the oracle's
scalar multiple, so no GPU result is compared with another GPU result."""
import os

import numpy as np
import pytest

import gpu_msm_amd as gm
import hipmem as hm
from gpu_msm_amd import planner as pl
from gpu_msm_amd import serve
from gpu_msm_amd import workloads as wl
from oracle import oracle as o

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def gpu():
    hm.require_gpu()


def closed_forms(curve: str, ts) -> list:
    g = o.g1_generator() if curve == "g1" else o.g2_generator()
    return [o.scalar_mul(curve, g, t).tobytes() for t in ts]


def test_batched_256_mixed_msms_vs_host_closed_form():
    pool_k, sizes, boff, soff, sc = wl.batch_workload()
    assert len(sizes) == 256 and sizes.min() >= 1 << 8 and sizes.max() <= 1 << 18
    pool = gm.Bases.from_k(gm.CURVE_G1, pool_k)
    s = gm.Scalars.upload(sc)
    idx = np.concatenate([np.arange(b, b + n) for n, b in zip(sizes, boff)])
    want = closed_forms("g1", wl.closed_form_scalars(pool_k[idx], sc, sizes))
    planners = [pl.HeuristicPlanner()]
    for name, cls in (("sweep.json", pl.TablePlanner), ("mlp.json", pl.MlpPlanner)):
        path = os.path.join(pl.TUNING_DIR, name)
        if os.path.exists(path):
            planners.append(cls.load(path))
    for p in planners:
        w, sg = p.plan(gm.CURVE_G1, sizes, len(sizes))
        print(f"batch of 256 under the {p.name} plan (hipGraph)", flush=True)
        bt = gm.Batch(pool, s, wl.specs_from(sizes, boff, soff, w, sg), scalar_bits=299, use_graph=True)
        for replay in range(2):
            got = bt.run()
            bad = [j for j in range(len(sizes)) if got[j:j + 1].tobytes() != want[j]]
            assert not bad, (p.name, replay, bad[:8])
        bt.free()
    pool.free()
    s.free()


def test_served_mixed_stream_vs_host_closed_form():
    st = serve.Stream()
    assert st.pool_n == 1 << 20
    out = []
    print("served stream: 48 requests, 80% G1 / 20% G2, N in [2^10, 2^20]", flush=True)
    rep = serve.run(st, 400.0, 48, 10, 20, seed=17, check=4, results_out=out)
    assert rep["completed_ok"] == 48 and rep["sampled_closed_form"] and rep["sampled_same_alone"]
    assert len(out) == 48 and max(r[1] for r in out) > 1 << 17
    assert {r[0] for r in out} == {gm.CURVE_G1, gm.CURVE_G2}
    for curve, n, b, scal, rc, rec in out:
        assert rc == 0
        t = wl.closed_form_scalars(st.k[b:b + n], scal, [n])[0]
        name = "g1" if curve == gm.CURVE_G1 else "g2"
        assert rec.tobytes() == closed_forms(name, [t])[0], (name, n, b)
    st.g1.free()
    st.g2.free()
