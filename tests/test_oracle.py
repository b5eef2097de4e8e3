"""Pins the CPU oracle (oracle/) against golden vectors, the reference tests'
known answers and closed forms. CPU only."""
import numpy as np
import pytest

import pymodel as m
from oracle import oracle as o

P = m.P


def L(v):
    return o.int_to_limbs(v)


def I(a):
    return o.limbs_to_int(a)


@pytest.fixture(scope="module", autouse=True)
def _built():
    o.build()


def test_constants():
    assert o.P == m.P and o.R_ORDER == m.R_ORDER
    r = np.zeros(5, dtype=np.uint64)
    o.lib().or_scalar_order(o.ptr(r))
    assert I(r) == m.R_ORDER
    one = np.zeros(7, dtype=np.uint64)
    o.lib().or_fp_one_mont(o.ptr(one))
    assert I(one) == (1 << 448) % P


def test_fp_golden_vectors(golden):
    for v in golden["fp_vectors"]:
        a, b = L(int(v["a"], 16)), L(int(v["b"], 16))
        assert I(o.fp_binop("or_fp_add", a, b)) == int(v["add"], 16)
        assert I(o.fp_binop("or_fp_sub", a, b)) == int(v["sub"], 16)
        assert I(o.fp_binop("or_fp_mont_mul", a, b)) == int(v["mont_mul"], 16)
        assert I(o.fp_unop("or_fp_neg", a)) == int(v["neg"], 16)
        assert I(o.fp_unop("or_fp_to_mont", a)) == int(v["to_mont"], 16)
        assert I(o.fp_unop("or_fp_from_mont", a)) == int(v["from_mont"], 16)
        assert I(o.fp_unop("or_fp_mont_inv", a)) == int(v["mont_inv"], 16)


def _std_op(name, *vals):
    """Apply a Montgomery-domain oracle op to standard-form ints, return standard form."""
    ms = [o.fp_unop("or_fp_to_mont", L(v)) for v in vals]
    r = o.fp_binop(name, *ms) if len(ms) == 2 else o.fp_unop(name, *ms)
    return I(o.fp_unop("or_fp_from_mont", r))


def test_fp_reference_known_answers(golden):
    # reference tests/test_fp.cu:262-282, 306-408, 942-966, 999-1022, 1062-1085
    k = golden["fp_known"]
    assert I(o.fp_binop("or_fp_add", L(1), L(1))) == int(k["one_plus_one"], 16)
    assert _std_op("or_fp_mont_mul", 5, 3) == int(k["five_times_three"], 16)
    assert I(o.fp_binop("or_fp_add", L(P - 1), L(1))) == int(k["pm1_plus_1"], 16)
    r = o.fp_binop("or_fp_sub", L(50), L(100))
    assert I(r) == int(k["fifty_minus_hundred"], 16)
    # explicit limbs of p - 50 from reference tests/test_fp.cu:1007-1008
    assert list(r) == [0x311c0026aab0aa79, 0x56ee4528c573b5cc, 0x824e6dc3e23acdee, 0xf75a64bbac71602,
                       0x95a4b78a02fe32, 0x200fc34965aad640, 0x3cdee0fb28c5e535]
    assert _std_op("or_fp_mont_mul", P - 1, P - 1) == int(k["pm1_squared"], 16)


def test_fp_random_against_bigint():
    rng = np.random.default_rng(7)
    for _ in range(200):
        a = int.from_bytes(rng.bytes(56), "little") % P
        b = int.from_bytes(rng.bytes(56), "little") % P
        assert I(o.fp_binop("or_fp_mont_mul", L(a), L(b))) == m.mont_mul(a, b)
        assert I(o.fp_binop("or_fp_sub", L(a), L(b))) == (a - b) % P
        assert I(o.fp_binop("or_fp_add", L(a), L(b))) == (a + b) % P
    # Montgomery round trip and inverse on a few values
    for v in (1, 2, P - 1, 12345678901234567890):
        mv = o.fp_unop("or_fp_to_mont", L(v))
        assert I(o.fp_unop("or_fp_from_mont", mv)) == v
        iv = o.fp_unop("or_fp_mont_inv", mv)
        assert I(o.fp_binop("or_fp_mont_mul", mv, iv)) == (1 << 448) % P


def test_fp2_golden_and_known(golden):
    def f2(pair):
        return np.concatenate([L(int(pair[0], 16)), L(int(pair[1], 16))])

    for v in golden["fp2_vectors"]:
        a, b = f2(v["a"]), f2(v["b"])
        for name, key in (("or_fp2_add", "add"), ("or_fp2_sub", "sub"), ("or_fp2_mont_mul", "mont_mul")):
            r = np.zeros(14, dtype=np.uint64)
            getattr(o.lib(), name)(o.ptr(r), o.ptr(a), o.ptr(b))
            assert [I(r[:7]), I(r[7:])] == [int(x, 16) for x in v[key]]
    # (1+i)^2 = 2i, i^2 = -1 (reference tests/test_fp2.cu:196-244), in Montgomery form
    one = o.fp_unop("or_fp_to_mont", L(1))
    zero = L(0)
    opi = np.concatenate([one, one])
    i_ = np.concatenate([zero, one])
    r = np.zeros(14, dtype=np.uint64)
    o.lib().or_fp2_mont_mul(o.ptr(r), o.ptr(opi), o.ptr(opi))
    assert [m.from_mont(I(r[:7])), m.from_mont(I(r[7:]))] == [0, 2]
    o.lib().or_fp2_mont_mul(o.ptr(r), o.ptr(i_), o.ptr(i_))
    assert [m.from_mont(I(r[:7])), m.from_mont(I(r[7:]))] == [P - 1, 0]
    # inverse
    a = np.concatenate([o.fp_unop("or_fp_to_mont", L(3)), o.fp_unop("or_fp_to_mont", L(7))])
    ia = np.zeros(14, dtype=np.uint64)
    o.lib().or_fp2_mont_inv(o.ptr(ia), o.ptr(a))
    want = m.f2_inv((3, 7))
    assert (m.from_mont(I(ia[:7])), m.from_mont(I(ia[7:]))) == want


def test_generators_on_curve_and_order():
    g1 = o.g1_generator()
    g2 = o.g2_generator()
    assert o.lib().or_g1_is_on_curve(o.ptr(g1))
    assert o.lib().or_g2_is_on_curve(o.ptr(g2))
    assert m.g1_from_abi(g1[0]) == m.G1_STD
    assert m.g2_from_abi(g2[0]) == m.G2_STD
    assert o.scalar_mul("g1", g1, m.R_ORDER)[0]["inf"] == 1
    assert o.scalar_mul("g2", g2, m.R_ORDER)[0]["inf"] == 1
    # (r-1)*G = -G
    ng = o.scalar_mul("g1", g1, m.R_ORDER - 1)
    assert m.g1_from_abi(ng[0]) == m.ec_neg(m.FP_OPS, m.G1_STD)


def _g1_rec(x, y):
    g = np.zeros(1, dtype=o.G1_DT)
    g[0]["x"] = o.fp_unop("or_fp_to_mont", L(x % P))
    g[0]["y"] = o.fp_unop("or_fp_to_mont", L(y % P))
    return g


def _g2_rec(x, y):
    g = np.zeros(1, dtype=o.G2_DT)
    g[0]["x"] = np.concatenate([o.fp_unop("or_fp_to_mont", L(v % P)) for v in x])
    g[0]["y"] = np.concatenate([o.fp_unop("or_fp_to_mont", L(v % P)) for v in y])
    return g


def test_is_on_curve_rejects_points_off_the_curve():
    """The on-curve tests are not constant: points off y^2 = x^3 + 1 (G1) and
    y^2 = x^3 + (1 + i) (G2) are rejected, among them a point of the
    reference's y^2 = x^3 + 4, and a point of the real curve other than the
    generator is accepted."""
    lib = o.lib()
    assert lib.or_g1_is_on_curve(o.ptr(_g1_rec(0, 1)))  # 1 = 0 + 1
    assert lib.or_g1_is_on_curve(o.ptr(_g1_rec(0, P - 1)))
    assert not lib.or_g1_is_on_curve(o.ptr(_g1_rec(0, 2)))  # on y^2 = x^3 + 4 only
    gx, gy = m.G1_STD
    assert not lib.or_g1_is_on_curve(o.ptr(_g1_rec(gx, gy + 1)))
    assert not lib.or_g1_is_on_curve(o.ptr(_g1_rec(gx + 1, gy)))
    (x0, x1), (y0, y1) = m.G2_STD
    assert lib.or_g2_is_on_curve(o.ptr(_g2_rec((x0, x1), (P - y0, P - y1))))  # -G2
    assert not lib.or_g2_is_on_curve(o.ptr(_g2_rec((x0, x1), (y0 + 1, y1))))
    assert not lib.or_g2_is_on_curve(o.ptr(_g2_rec((x0, x1), (y0, y1 + 1))))
    assert not lib.or_g2_is_on_curve(o.ptr(_g2_rec((x1, x0), (y0, y1))))
    # infinity records are on the curve whatever their coordinates
    inf = _g1_rec(0, 2)
    inf[0]["inf"] = 1
    assert lib.or_g1_is_on_curve(o.ptr(inf))


def test_scalar_mul_against_bigint():
    g1, g2 = o.g1_generator(), o.g2_generator()
    for k in (1, 2, 3, 0xFFFFFFFFFFFFFFFF, m.R_ORDER // 3 + 17):
        assert m.g1_from_abi(o.scalar_mul("g1", g1, k)[0]) == m.ec_mul(m.FP_OPS, m.G1_STD, k)
    for k in (1, 5, 0x123456789ABCDEF):
        assert m.g2_from_abi(o.scalar_mul("g2", g2, k)[0]) == m.ec_mul(m.FP2_OPS, m.G2_STD, k)


def test_projective_add_double_special_cases():
    g = o.g1_generator()
    gp = np.zeros(1, dtype=o.G1P_DT)
    o.lib().or_g1_affine_to_proj(o.ptr(gp), o.ptr(g))
    # P + P through add takes the doubling branch
    s = np.zeros(1, dtype=o.G1P_DT)
    d = np.zeros(1, dtype=o.G1P_DT)
    o.lib().or_g1_proj_add(o.ptr(s), o.ptr(gp), o.ptr(gp))
    o.lib().or_g1_proj_double(o.ptr(d), o.ptr(gp))
    assert m.g1_from_abi(o.to_affine("g1", s)[0]) == m.ec_mul(m.FP_OPS, m.G1_STD, 2)
    assert m.g1_from_abi(o.to_affine("g1", d)[0]) == m.ec_mul(m.FP_OPS, m.G1_STD, 2)
    # P + (-P) = O (Z = 0)
    ng = g.copy()
    ng[0]["y"] = o.fp_unop("or_fp_neg", g[0]["y"].copy())
    ngp = np.zeros(1, dtype=o.G1P_DT)
    o.lib().or_g1_affine_to_proj(o.ptr(ngp), o.ptr(ng))
    o.lib().or_g1_proj_add(o.ptr(s), o.ptr(gp), o.ptr(ngp))
    assert I(s[0]["Z"]) == 0
    # infinity conventions: affine infinity -> (0,0,0)
    inf = np.zeros(1, dtype=o.G1_DT)
    inf[0]["inf"] = 1
    o.lib().or_g1_affine_to_proj(o.ptr(s), o.ptr(inf))
    assert I(s[0]["Z"]) == 0 and I(s[0]["X"]) == 0


def test_mt19937_64_known_answer():
    # C++ [rand.predef]: the 10000th output of a default-seeded mt19937_64.
    v = o.gen_u64(5489, 10000)
    assert int(v[-1]) == 9981545732273789042
    assert [int(x) for x in o.gen_u64(1234, 50)] == m.gen_u64(1234, 50)
    s = o.gen_scalars_modr(99, 20)
    assert [o.limbs_to_int(r) for r in s] == m.gen_scalars_modr(99, 20)


def test_msm_triangular_closed_form(golden):
    # reference tests/test_msm.cu:132-248, 487-592: bases G, scalars 1..N
    for curve, n in (("g1", 10), ("g1", 100), ("g2", 10)):
        g = o.g1_generator() if curve == "g1" else o.g2_generator()
        bases = np.repeat(g, n)
        scalars = np.arange(1, n + 1, dtype=np.uint64).reshape(-1, 1)
        res = o.msm(curve, bases, scalars, threads=2)
        want = golden["triangular"][f"{curve}_{n}"]
        got = m.g1_from_abi(res[0]) if curve == "g1" else m.g2_from_abi(res[0])
        if curve == "g1":
            assert [hex(got[0]), hex(got[1])] == want
        else:
            assert [[hex(got[0][0]), hex(got[0][1])], [hex(got[1][0]), hex(got[1][1])]] == want


def test_msm_c1_closed_form(golden):
    """C1: G1, N=1024, bases k_i*G, scalars in [0, r) with 8 zeros and 8 ones."""
    c = golden["c1_g1_1024"]
    n = c["n"]
    ks = o.gen_u64(c["seed_k"], n)
    ss = o.gen_scalars_modr(c["seed_s"], n)
    ss[c["zero_idx"]] = 0
    for i in c["one_idx"]:
        ss[i] = [1, 0, 0, 0, 0]
    assert hex(int(ks[0])) == c["k0"] and hex(o.limbs_to_int(ss[0])) == c["s0"]
    bases = o.bases_from_k("g1", ks)
    assert [hex(v) for v in m.g1_from_abi(bases[0])] == c["base0"]
    res = o.msm("g1", bases, ss, threads=8)
    got = m.g1_from_abi(res[0])
    assert [hex(got[0]), hex(got[1])] == c["result"]
    # a different window width gives the same affine point
    res2 = o.msm("g1", bases, ss, c=5, threads=8)
    assert bytes(res2.tobytes()) == bytes(res.tobytes())


def test_msm_small_bruteforce_with_infinity():
    rng = np.random.default_rng(3)
    n = 24
    ks = rng.integers(1, 2**63, size=n, dtype=np.uint64)
    bases = o.bases_from_k("g2", ks)
    bases[5]["inf"] = 1
    bases[5]["x"] = 0
    bases[5]["y"] = 0
    ss = [int(rng.integers(0, 2**62)) * (1 << 200) + int(rng.integers(0, 2**62)) for _ in range(n)]
    ss[3] = 0
    ss[7] = 1
    sc = np.array([o.int_to_limbs(s, 5) for s in ss], dtype=np.uint64)
    res = o.msm("g2", bases, sc, threads=2)
    acc = None
    for i in range(n):
        if i == 5:
            continue
        acc = m.ec_add(m.FP2_OPS, acc, m.ec_mul(m.FP2_OPS, m.G2_STD, int(ks[i]) * ss[i]))
    assert m.g2_from_abi(res[0]) == acc
