"""Writes tests/golden/golden.json from the independent big-integer model.

Nothing here uses the oracle or the product library: every expected value is
computed with Python integers (tests/pymodel.py) from the mathematics, or is a
known answer from the reference's own tests (reference tests/test_fp.cu,
tests/test_fp2.cu, tests/test_msm.cu). Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

import pymodel as m  # noqa: E402

P = m.P
R = m.R_ORDER


def hx(v: int) -> str:
    return hex(v)


def pt1(p):
    return None if p is None else [hx(p[0]), hx(p[1])]


def pt2(p):
    return None if p is None else [[hx(p[0][0]), hx(p[0][1])], [hx(p[1][0]), hx(p[1][1])]]


# Workload definitions shared with the tests (the same seeds are used by the
# GPU tests and the bench; the generator is mt19937_64).
def workload(curve: str, log2n: int):
    """Seeds of the seeded MSM workloads: bases k_i*G (k_i one 64-bit draw),
    scalars uniform in [0, r)."""
    base = 1000 if curve == "g1" else 3000
    return {"seed_k": base + log2n, "seed_s": base + 1000 + log2n}


def g2_zero_positions(n: int, seed: int):
    g = m.MT64(seed)
    z = []
    seen = set()
    while len(z) < n // 100:
        v = g.next() % n
        if v not in seen:
            seen.add(v)
            z.append(v)
    return z


def closed_form(curve: str, ks, ss):
    t = sum(k * s for k, s in zip(ks, ss)) % R
    if curve == "g1":
        return t, m.ec_mul(m.FP_OPS, m.G1_STD, t)
    return t, m.ec_mul(m.FP2_OPS, m.G2_STD, t)


def main() -> None:
    out = {"p": hx(P), "r": hx(R)}

    # ---- Fp random vectors (Montgomery-form operands, canonical outputs)
    g = m.MT64(42)
    fpv = []
    for _ in range(32):
        a = sum(g.next() << (64 * k) for k in range(7)) % P
        b = sum(g.next() << (64 * k) for k in range(7)) % P
        fpv.append({
            "a": hx(a), "b": hx(b),
            "add": hx((a + b) % P), "sub": hx((a - b) % P), "neg": hx(-a % P),
            "mont_mul": hx(m.mont_mul(a, b)), "to_mont": hx(m.to_mont(a)),
            "from_mont": hx(m.from_mont(a)),
            "mont_inv": hx(pow(a * m.RMONT_INV, -1, P) * m.RMONT % P if a else 0),
        })
    # edge operands
    for a, b in [(0, 0), (1, P - 1), (P - 1, P - 1), (P - 2, 2), (1 << 445, (1 << 445) + 12345)]:
        fpv.append({
            "a": hx(a), "b": hx(b),
            "add": hx((a + b) % P), "sub": hx((a - b) % P), "neg": hx(-a % P),
            "mont_mul": hx(m.mont_mul(a, b)), "to_mont": hx(m.to_mont(a)),
            "from_mont": hx(m.from_mont(a)),
            "mont_inv": hx(pow(a * m.RMONT_INV, -1, P) * m.RMONT % P if a else 0),
        })
    out["fp_vectors"] = fpv

    # ---- known answers of reference tests/test_fp.cu (values in standard form)
    out["fp_known"] = {
        "one_plus_one": hx(2),                      # test_fp.cu:262-282
        "five_times_three": hx(15),                 # test_fp.cu:306-326, 375-408
        "pm1_plus_1": hx(0),                        # test_fp.cu:942-966
        "fifty_minus_hundred": hx(P - 50),          # test_fp.cu:999-1022
        "pm1_squared": hx(1),                       # test_fp.cu:1062-1085
        "R_mod_p": hx(m.RMONT % P),
    }

    # ---- Fp2 random vectors (Montgomery-form components)
    g = m.MT64(43)
    f2v = []
    for _ in range(16):
        a = tuple(sum(g.next() << (64 * k) for k in range(7)) % P for _ in range(2))
        b = tuple(sum(g.next() << (64 * k) for k in range(7)) % P for _ in range(2))
        prod = m.f2_mul(a, b)  # (a*b) on Montgomery reps = R^2 * real product
        f2v.append({
            "a": [hx(a[0]), hx(a[1])], "b": [hx(b[0]), hx(b[1])],
            "add": [hx(v) for v in m.f2_add(a, b)], "sub": [hx(v) for v in m.f2_sub(a, b)],
            "mont_mul": [hx(v * m.RMONT_INV % P) for v in prod],
        })
    out["fp2_vectors"] = f2v
    out["fp2_known"] = {"one_plus_i_squared": [hx(0), hx(2)], "i_squared": [hx(P - 1), hx(0)]}

    # ---- MSM closed forms from reference tests/test_msm.cu:42-48,132-592:
    # bases all G, scalars 1..N, expected (N(N+1)/2) * G
    tri = {}
    for n in (10, 100):
        tri[f"g1_{n}"] = pt1(m.ec_mul(m.FP_OPS, m.G1_STD, n * (n + 1) // 2))
    tri["g2_10"] = pt2(m.ec_mul(m.FP2_OPS, m.G2_STD, 55))
    out["triangular"] = tri

    # ---- C1: G1, N = 1024, host path, 8 zero and 8 one scalars
    n = 1024
    w = workload("g1", 10)
    ks = m.gen_u64(w["seed_k"], n)
    ss = m.gen_scalars_modr(w["seed_s"], n)
    zero_idx = [128 * j + 3 for j in range(8)]
    one_idx = [128 * j + 67 for j in range(8)]
    for i in zero_idx:
        ss[i] = 0
    for i in one_idx:
        ss[i] = 1
    t, res = closed_form("g1", ks, ss)
    out["c1_g1_1024"] = {**w, "n": n, "zero_idx": zero_idx, "one_idx": one_idx,
                         "k0": hx(ks[0]), "s0": hx(ss[0]), "s1": hx(ss[1]),
                         "base0": pt1(m.ec_mul(m.FP_OPS, m.G1_STD, ks[0])),
                         "t": hx(t), "result": pt1(res)}

    # ---- C2 / C3 closed forms (seeded full-size workloads)
    sizes = {"g1": [12, 14, 16, 18, 20], "g2": [10, 12, 14, 16, 20]}
    if os.environ.get("GOLDEN_QUICK"):
        sizes = {"g1": [12], "g2": [10]}
    for curve in ("g1", "g2"):
        for lg in sizes[curve]:
            n = 1 << lg
            w = workload(curve, lg)
            ks = m.gen_u64(w["seed_k"], n)
            ss = m.gen_scalars_modr(w["seed_s"], n)
            rec = {**w, "n": n}
            if curve == "g2":
                zs = g2_zero_positions(n, 5000 + lg)
                for i in zs:
                    ss[i] = 0
                rec["zero_seed"] = 5000 + lg
                rec["n_zero"] = len(zs)
            t, res = closed_form(curve, ks, ss)
            rec["t"] = hx(t)
            rec["result"] = pt1(res) if curve == "g1" else pt2(res)
            rec["s_last"] = hx(ss[-1])
            out[f"{curve}_2_{lg}"] = rec
            print(curve, lg, "done", flush=True)

    path = os.path.join(HERE, "golden.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)


if __name__ == "__main__":
    main()
