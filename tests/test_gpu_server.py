"""The continuous-batching request server: every result equals the same MSM
run alone (and the oracle for small ones), HBM admission, overlap of
batches, clean failure of an injected device error."""
import threading

import numpy as np
import pytest

import gpu_msm_amd as gm
from oracle import oracle as o

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def pools():
    pool_n = 1 << 14
    ks = o.gen_u64(300, pool_n)
    return ks, gm.Bases.from_k(gm.CURVE_G1, ks), gm.Bases.from_k(gm.CURVE_G2, ks)


def _requests(seed, count, pool_n, lo=4, hi=12):
    rng = np.random.default_rng(seed)
    reqs = []
    for _ in range(count):
        curve = gm.CURVE_G2 if rng.random() < 0.2 else gm.CURVE_G1
        n = int(2 ** rng.uniform(lo, hi))
        bo = int(rng.integers(0, pool_n - n + 1))
        sc = o.gen_scalars_modr(int(rng.integers(1, 1 << 30)), n)
        reqs.append((curve, bo, sc))
    return reqs


def test_mixed_stream_results_equal_alone(pools):
    ks, g1, g2 = pools
    reqs = _requests(1, 60, len(ks))
    srv = gm.Server(g1, g2, slots=3, max_batch=16)
    tickets = [srv.submit(c, bo, sc, scalar_bits=299) for c, bo, sc in reqs]
    outs = [srv.wait(t, c) for t, (c, _, _) in zip(tickets, reqs)]
    st = srv.stats()
    srv.close()
    assert st["completed"] == 60 and st["failed"] == 0
    assert st["batches"] < 60  # requests were batched
    for (c, bo, sc), (rc, res, lat) in zip(reqs, outs):
        assert rc == 0 and lat > 0
        pool = g1 if c == gm.CURVE_G1 else g2
        s = gm.Scalars.upload(sc)
        alone = gm.msm_resident(pool, s, n=len(sc), base_offset=bo, scalar_bits=299)
        assert alone.tobytes() == res.tobytes()
        if len(sc) <= 64:
            curve = "g1" if c == gm.CURVE_G1 else "g2"
            want = o.msm(curve, o.bases_from_k(curve, ks[bo:bo + len(sc)]), sc)
            assert want.tobytes() == res.tobytes()


def test_overlapping_batches_and_concurrent_submitters(pools):
    ks, g1, g2 = pools
    reqs = _requests(2, 96, len(ks), lo=8, hi=13)
    srv = gm.Server(g1, g2, slots=4, max_batch=8)
    tickets = [None] * len(reqs)

    def submit(part):
        for i in part:
            c, bo, sc = reqs[i]
            tickets[i] = srv.submit(c, bo, sc, scalar_bits=299, borrow=True)

    th = [threading.Thread(target=submit, args=(range(k, len(reqs), 3),)) for k in range(3)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    outs = [srv.wait(tickets[i], reqs[i][0]) for i in range(len(reqs))]
    st = srv.stats()
    srv.close()
    assert all(rc == 0 for rc, _, _ in outs)
    assert st["peak_concurrent_batches"] >= 2
    # the same request alone through a fresh server gives the same point
    srv1 = gm.Server(g1, g2, slots=1, max_batch=1)
    for i in (0, 17, 95):
        c, bo, sc = reqs[i]
        rc, res, _ = srv1.wait(srv1.submit(c, bo, sc, scalar_bits=299), c)
        assert rc == 0 and res.tobytes() == outs[i][1].tobytes()
    srv1.close()


def test_hbm_admission_budget(pools):
    ks, g1, g2 = pools
    reqs = _requests(3, 24, len(ks), lo=10, hi=14)
    budget = 96 << 20
    srv = gm.Server(g1, g2, slots=4, max_batch=64, hbm_budget_bytes=budget)
    tickets = [srv.submit(c, bo, sc, scalar_bits=299) for c, bo, sc in reqs]
    outs = [srv.wait(t, c) for t, (c, _, _) in zip(tickets, reqs)]
    st = srv.stats()
    srv.close()
    assert all(rc == 0 for rc, _, _ in outs)
    _, _, ws = gm.plan_query(gm.CURVE_G2, 1 << 14, 5, 299)
    # in-flight footprint stays within the budget unless one request alone exceeds it
    assert st["peak_inflight_bytes"] <= max(budget, ws + (1 << 14) * 40 + (1 << 20))


def test_injected_device_failure_fails_cleanly(pools):
    ks, g1, g2 = pools
    reqs = _requests(4, 10, len(ks), lo=4, hi=8)
    srv = gm.Server(g1, g2, slots=1, max_batch=64)
    srv.inject_failures(1)
    tickets = [srv.submit(c, bo, sc, scalar_bits=299) for c, bo, sc in reqs]
    outs = [srv.wait(t, c) for t, (c, _, _) in zip(tickets, reqs)]
    # later requests on the same server succeed
    c, bo, sc = reqs[0]
    rc, res, _ = srv.wait(srv.submit(c, bo, sc, scalar_bits=299), c)
    st = srv.stats()
    srv.close()
    assert any(r == -4 for r, _, _ in outs)
    assert all(r in (0, -4) for r, _, _ in outs)
    assert rc == 0 and st["failed"] >= 1


def test_scalar_bound_error_and_bad_arguments(pools):
    ks, g1, g2 = pools
    srv = gm.Server(g1, g2, slots=2, max_batch=8)
    big = np.tile(o.int_to_limbs(1 << 305, 5), (3, 1))
    t = srv.submit(gm.CURVE_G1, 0, big, scalar_bits=299)
    rc, _, _ = srv.wait(t, gm.CURVE_G1)
    assert rc == -5
    with pytest.raises(gm.MsmError):
        srv.submit(gm.CURVE_G1, len(ks) - 2, big, scalar_bits=299)  # past the end of the pool
    srv.close()


def test_small_requests_pass_large_ones(pools):
    """Size classes: with small_max_n = 256 and two slots, large requests
    (G2, 2^14 points, one per batch) hold at most one slot at a time, so a
    small request submitted behind four of them takes the other slot at once
    and completes before the last large one, whose batch waits for the three
    before it; every result equals the same MSM alone."""
    ks, g1, g2 = pools
    n_big = 1 << 14
    big = [o.gen_scalars_modr(500 + k, n_big) for k in range(4)]
    small = o.gen_scalars_modr(510, 64)
    srv = gm.Server(g1, g2, slots=2, max_batch=1, small_max_n=256)
    tb = [srv.submit(gm.CURVE_G2, 0, sc, scalar_bits=299) for sc in big]
    ts = srv.submit(gm.CURVE_G1, 100, small, scalar_bits=299)
    outs_b = [srv.wait(t, gm.CURVE_G2) for t in tb]
    rc_s, res_s, lat_s = srv.wait(ts, gm.CURVE_G1)
    st = srv.stats()
    srv.close()
    assert rc_s == 0 and all(rc == 0 for rc, _, _ in outs_b)
    print(f"small request {lat_s:.2f} ms; large requests {[round(l, 2) for _, _, l in outs_b]} ms", flush=True)
    assert lat_s < outs_b[-1][2]
    assert st["peak_concurrent_batches"] <= 2
    s = gm.Scalars.upload(small)
    assert gm.msm_resident(g1, s, n=64, base_offset=100, scalar_bits=299).tobytes() == res_s.tobytes()
    want = o.msm("g1", o.bases_from_k("g1", ks[100:164]), small)
    assert want.tobytes() == res_s.tobytes()
    s2 = gm.Scalars.upload(big[3])
    assert gm.msm_resident(g2, s2, n=n_big, base_offset=0, scalar_bits=299).tobytes() == outs_b[3][1].tobytes()
