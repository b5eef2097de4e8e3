"""GPU parity of the point wire formats and validation (include/gpu_msm_ext.h,
"point formats"; the caller's side of the MSM, reference
rust-api/src/conversions.rs) against the big-integer model in
tests/pymodel.py.

The byte layout follows ark-serialize's CanonicalSerialize for
short-Weierstrass points as written in pymodel.encode_point; tfhe-zk-pok and
arkworks are not available offline, so agreement with their own encoder is
unpinned here. Everything else is pinned by the mathematics: encodings are
compared byte for byte with the model, decoding inverts encoding, decoded
coordinates satisfy the curve equation, flags select the documented root, and
invalid inputs get the documented status."""
import numpy as np
import pytest

import gpu_msm_amd as gm
import pymodel as m
from oracle import oracle as o

pytestmark = pytest.mark.gpu

CURVE = {"g1": gm.CURVE_G1, "g2": gm.CURVE_G2}


def std_point(curve, rec):
    return m.g1_from_abi(rec) if curve == "g1" else m.g2_from_abi(rec)


def sample_points(curve, n, seed):
    pts = o.bases_from_k(curve, o.gen_u64(seed, n))
    pts[3]["inf"] = 1
    pts[3]["x"] = 0
    pts[3]["y"] = 0
    return pts


def fields(r):
    return r["x"].tobytes() + r["y"].tobytes() + r["inf"].tobytes()


def rec_from_std(curve, pt):
    dt = gm.G1_DTYPE if curve == "g1" else gm.G2_DTYPE
    r = np.zeros(1, dtype=dt)
    if pt is None:
        r["inf"] = 1
        return r
    if curve == "g1":
        r["x"][0] = o.int_to_limbs(m.to_mont(pt[0]))
        r["y"][0] = o.int_to_limbs(m.to_mont(pt[1]))
    else:
        r["x"][0] = np.concatenate([o.int_to_limbs(m.to_mont(v)) for v in pt[0]])
        r["y"][0] = np.concatenate([o.int_to_limbs(m.to_mont(v)) for v in pt[1]])
    return r


@pytest.mark.parametrize("curve", ["g1", "g2"])
@pytest.mark.parametrize("compressed", [True, False])
def test_serialize_matches_model_and_round_trips(curve, compressed):
    pts = sample_points(curve, 300, 4100 + (curve == "g2"))
    enc = gm.serialize_points(CURVE[curve], pts, compressed)
    for i in range(len(pts)):
        want = m.encode_point(curve, std_point(curve, pts[i]), compressed)
        assert enc[i].tobytes() == want, i
    back, st = gm.deserialize_points(CURVE[curve], enc, compressed)
    assert not st.any(), np.nonzero(st)
    for i in range(len(pts)):
        assert fields(back[i:i + 1]) == fields(pts[i:i + 1]), i


@pytest.mark.parametrize("curve", ["g1", "g2"])
def test_compressed_flag_selects_the_documented_root(curve):
    """Both roots of one x: flag 0x80 gives the larger of {y, -y}, no flag the
    smaller; each satisfies y^2 = x^3 + b."""
    p = sample_points(curve, 8, 4200)[5:6]
    x, y = std_point(curve, p[0])
    ops = m.FP_OPS if curve == "g1" else m.FP2_OPS
    ny = ops[4](y)
    enc = m.encode_point(curve, (x, y), True)
    raw = bytearray(enc)
    for flag, want_larger in ((0x00, False), (0x80, True)):
        raw[-1] = (raw[-1] & 0x3F) | flag
        got, st = gm.deserialize_points(CURVE[curve], np.frombuffer(bytes(raw), np.uint8), True)
        assert st[0] == 0
        gx, gy = std_point(curve, got[0])
        assert gx == x and gy in (y, ny)
        larger = m.larger_g1(gy) if curve == "g1" else m.larger_g2(gy)
        assert larger == want_larger
        assert (m.on_curve_g1 if curve == "g1" else m.on_curve_g2)((gx, gy))


def test_invalid_encodings_g1():
    """Non-canonical x, both flag bits, a non-square x^3 + 1, an off-curve
    uncompressed point, and the 2-torsion point (-1, 0): on the curve, not in
    the prime-order subgroup."""
    good = m.encode_point("g1", std_point("g1", sample_points("g1", 4, 4300)[1]), True)
    cases = []
    cases.append((m.fp_bytes(m.P), True, 1))                   # x = p
    both = bytearray(good)
    both[-1] |= 0xC0
    cases.append((bytes(both), True, 1))                        # both flags
    x = 2
    while m.is_square_fp(x ** 3 + 1):
        x += 1
    cases.append((m.fp_bytes(x), True, 2))                      # no square root
    pt = std_point("g1", sample_points("g1", 4, 4301)[2])
    off = m.encode_point("g1", (pt[0], (pt[1] + 1) % m.P), False)
    cases.append((off, False, 2))                               # off the curve
    tor = m.encode_point("g1", (m.P - 1, 0), False)
    cases.append((tor, False, 3))                               # 2-torsion
    for data, comp, want in cases:
        got, st = gm.deserialize_points(gm.CURVE_G1, np.frombuffer(data, np.uint8), comp)
        assert st[0] == want, (want, st[0])
        assert got["inf"][0] == 1
    # without the checks, the off-curve and torsion encodings decode as given
    for data, want_pt in ((off, (pt[0], (pt[1] + 1) % m.P)), (tor, (m.P - 1, 0))):
        got, st = gm.deserialize_points(gm.CURVE_G1, np.frombuffer(data, np.uint8), False, checks=0)
        assert st[0] == 0 and std_point("g1", got[0]) == want_pt


def test_invalid_encodings_g2():
    """G2: a non-square x^3 + (1 + i), and a point of E'(Fp2) outside the
    order-r subgroup (a random curve point: the cofactor is large)."""
    x = (3, 5)
    while m.is_square_fp2(m.f2_add(m.f2_mul(m.f2_mul(x, x), x), (1, 1))):
        x = (x[0] + 1, x[1])
    got, st = gm.deserialize_points(gm.CURVE_G2, np.frombuffer(m.fp_bytes(x[0]) + m.fp_bytes(x[1]), np.uint8), True)
    assert st[0] == 2
    x = (7, 11)
    while not m.is_square_fp2(m.f2_add(m.f2_mul(m.f2_mul(x, x), x), (1, 1))):
        x = (x[0] + 1, x[1])
    enc = np.frombuffer(m.fp_bytes(x[0]) + m.fp_bytes(x[1]), np.uint8)
    got, st = gm.deserialize_points(gm.CURVE_G2, enc, True, checks=0)
    assert st[0] == 0
    q = std_point("g2", got[0])
    assert q[0] == x and m.on_curve_g2(q)
    got, st = gm.deserialize_points(gm.CURVE_G2, enc, True)
    assert st[0] == 3


@pytest.mark.parametrize("curve", ["g1", "g2"])
def test_check_points(curve):
    """gm_points_check: subgroup points pass, a perturbed y fails the curve
    equation (b = 1 / 1 + i, not the reference's constants 4 / 4i)."""
    pts = sample_points(curve, 64, 4400)
    st = gm.check_points(CURVE[curve], pts)
    assert not st.any()
    bad = pts.copy()
    bad[7]["y"][0] ^= np.uint64(1 << 40)
    st = gm.check_points(CURVE[curve], bad)
    assert st[7] == 2 and np.count_nonzero(st) == 1
    if curve == "g1":
        tor = rec_from_std("g1", (m.P - 1, 0))
        assert gm.check_points(gm.CURVE_G1, tor)[0] == 3
        assert gm.check_points(gm.CURVE_G1, tor, checks=gm.CHECK_ON_CURVE)[0] == 0


@pytest.mark.parametrize("curve", ["g1", "g2"])
def test_canonical_limbs_round_trip(curve):
    """Montgomery records <-> standard-form records (from_sign_and_limbs /
    into_bigint limbs) against the oracle's conversion."""
    pts = sample_points(curve, 200, 4500)
    std = gm.to_canonical(CURVE[curve], pts)
    for i in range(len(pts)):
        if pts[i]["inf"]:
            assert std[i]["inf"] == 1 and not std[i]["x"].any() and not std[i]["y"].any()
        else:
            assert fields(std[i:i + 1]) == fields(o.from_mont(curve, pts[i:i + 1]))
    back = gm.from_canonical(CURVE[curve], std)
    assert fields(back) == fields(pts)
