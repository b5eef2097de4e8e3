"""GPU parity of the reference's C++ device-pointer surface: include/curve.h
(point operations, Montgomery batch conversion, device-pointer MSM), fp.h /
fp2.h (stream element operations) and fp_kernels.h / fp2_kernels.h (arrays and
host-pointer operations), called by mangled name through ctypes, plus the
caller harnesses (the Rust layer's C ABI flow, a C++ device-API caller, and
the reference's own tests/test_msm.cu flow). Device buffers and the stream
come from the HIP runtime directly (tests/hipmem.py).

Expected values come from the big-integer model (tests/pymodel.py: textbook
affine group law, Python modular arithmetic) and the oracle's MSM; every
comparison is bit-exact on the ABI records."""
import ctypes
import os
import subprocess
import sys

import numpy as np
import pytest

import gpu_msm_amd as gm
import hipmem as hm
import pymodel as m
from cxx_surface import I32, P, U32, U64, Surface
from oracle import oracle as o

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "harness"))
import build as harness  # noqa: E402

pytestmark = pytest.mark.gpu

MASK64 = (1 << 64) - 1

@pytest.fixture(scope="module")
def S():
    hm.require_gpu()
    return Surface()


stream = hm.stream
to_dev = hm.to_dev
to_host = hm.to_host
empty_dev = hm.empty_dev


# ------------------------------------------------------------ ABI records

def limbs(v: int, n: int = 7) -> np.ndarray:
    return o.int_to_limbs(v, n)


def rec_g1(pt, mont: bool = True) -> np.ndarray:
    r = np.zeros(1, dtype=o.G1_DT)
    if pt is None:
        r["inf"] = 1
        return r
    f = m.to_mont if mont else (lambda v: v)
    r["x"][0] = limbs(f(pt[0]))
    r["y"][0] = limbs(f(pt[1]))
    return r


def rec_g2(pt, mont: bool = True) -> np.ndarray:
    r = np.zeros(1, dtype=o.G2_DT)
    if pt is None:
        r["inf"] = 1
        return r
    f = m.to_mont if mont else (lambda v: v)
    r["x"][0] = np.concatenate([limbs(f(pt[0][0])), limbs(f(pt[0][1]))])
    r["y"][0] = np.concatenate([limbs(f(pt[1][0])), limbs(f(pt[1][1]))])
    return r


CURVES = {
    "g1": dict(ops=m.FP_OPS, gen=m.G1_STD, rec=rec_g1, dt=o.G1_DT, T="G1Point"),
    "g2": dict(ops=m.FP2_OPS, gen=m.G2_STD, rec=rec_g2, dt=o.G2_DT, T="G2Point"),
}


def fields(r: np.ndarray) -> bytes:
    """The bytes a caller reads: x, y and the infinity flag (padding excluded)."""
    return r["x"].tobytes() + r["y"].tobytes() + r["inf"].tobytes()


def point_call(S, curve, name, out_rec, *args, extra=()):
    T = CURVES[curve]["T"]
    f = S.fn(f"{name}<{T}>", [P, U32, P] + [P] * len(args) + [a[0] for a in extra])
    d_out = to_dev(out_rec)
    devs = [to_dev(a) for a in args]
    f(stream(), 0, d_out.data_ptr(), *[d.data_ptr() for d in devs], *[a[1] for a in extra])
    return to_host(d_out, CURVES[curve]["dt"])


# --------------------------------------------------------- point operations

@pytest.mark.parametrize("curve", ["g1", "g2"])
def test_point_add_double_neg_vs_group_law(S, curve):
    """reference src/curve.cu:72-227 (affine add / double / negate) through
    point_add / point_double / point_neg (curve.cu:2386-2429)."""
    c = CURVES[curve]
    ops, rec = c["ops"], c["rec"]
    p1 = m.ec_mul(ops, c["gen"], 0x1234567890ABCDEF)
    p2 = m.ec_mul(ops, c["gen"], 0xFEDCBA0987654321)
    neg1 = m.ec_neg(ops, p1)
    garbage = np.zeros(1, dtype=c["dt"])
    for name in ("point_add", "point_add_async"):
        cases = [(p1, p2), (p1, p1), (p1, neg1), (None, p2), (p1, None), (None, None)]
        for a, b in cases:
            got = point_call(S, curve, name, garbage, rec(a), rec(b))
            assert fields(got) == fields(rec(m.ec_add(ops, a, b))), (name, a is None, b is None)
    for name in ("point_double", "point_double_async"):
        for a in (p1, p2, None):
            got = point_call(S, curve, name, garbage, rec(a))
            assert fields(got) == fields(rec(m.ec_add(ops, a, a)))
    for name in ("point_neg", "point_neg_async"):
        for a in (p1, None):
            got = point_call(S, curve, name, garbage, rec(a))
            assert fields(got) == fields(rec(m.ec_neg(ops, a)))


def test_point_double_two_torsion_gives_infinity(S):
    """y = 0 doubles to O (reference src/curve.cu:90-93): (-1, 0) is on
    y^2 = x^3 + 1 and has order 2."""
    t = (m.P - 1, 0)
    assert m.on_curve_g1(t)
    got = point_call(S, "g1", "point_double", np.zeros(1, dtype=o.G1_DT), rec_g1(t))
    assert fields(got) == fields(rec_g1(None))


def test_point_add_infinity_copies_operand_as_stored(S):
    """O + P returns P's record unchanged and O + O keeps the second operand's
    coordinates with the flag set (reference src/curve.cu:178-189)."""
    q = np.zeros(1, dtype=o.G1_DT)
    q["x"][0] = limbs(5)
    q["y"][0] = limbs(7)
    inf = rec_g1(None)
    got = point_call(S, "g1", "point_add", np.zeros(1, dtype=o.G1_DT), inf, q)
    assert fields(got) == fields(q)
    q_inf = q.copy()
    q_inf["inf"] = 1
    got = point_call(S, "g1", "point_add", np.zeros(1, dtype=o.G1_DT), inf, q_inf)
    assert fields(got) == fields(q_inf)


@pytest.mark.parametrize("curve", ["g1", "g2"])
def test_point_scalar_mul_vs_group_law(S, curve):
    """reference src/curve.cu:231-268 through point_scalar_mul_u64 and the
    multi-limb point_scalar_mul (curve.cu:2476-2503); k = r gives O."""
    c = CURVES[curve]
    ops, rec = c["ops"], c["rec"]
    p = m.ec_mul(ops, c["gen"], 987654321)
    out = np.zeros(1, dtype=c["dt"])
    for k in (0, 1, 2, 3, 0xDEADBEEFCAFEBABE, MASK64):
        for name in ("point_scalar_mul_u64", "point_scalar_mul_u64_async"):
            got = point_call(S, curve, name, out, rec(p), extra=[(U64, k)])
            assert fields(got) == fields(rec(m.ec_mul(ops, p, k))), (name, k)
    got = point_call(S, curve, "point_scalar_mul_u64", out, rec(None), extra=[(U64, 5)])
    assert fields(got) == fields(rec(None))
    T = c["T"]
    for k, nl in ((m.R_ORDER, 5), (m.R_ORDER + 1, 5), (0x1234 << 200 | 0x77, 5), (123, 8)):
        ks = to_dev(limbs(k, nl))
        for name in ("point_scalar_mul", "point_scalar_mul_async"):
            f = S.fn(f"{name}<{T}>", [P, U32, P, P, P, I32])
            d_out = to_dev(out)
            d_p = to_dev(rec(p))
            f(stream(), 0, d_out.data_ptr(), d_p.data_ptr(), ks.data_ptr(), nl)
            got = to_host(d_out, c["dt"])
            assert fields(got) == fields(rec(m.ec_mul(ops, p, k % (1 << (64 * nl))))), (name, hex(k))


@pytest.mark.parametrize("curve", ["g1", "g2"])
def test_point_montgomery_conversions(S, curve):
    """point_to/from_montgomery (reference src/curve.cu:2305-2331, 2446-2474),
    point_at_infinity, and the in-place batch conversion, which leaves records
    at infinity untouched (curve.cu:2333-2352)."""
    c = CURVES[curve]
    ops, rec = c["ops"], c["rec"]
    p = m.ec_mul(ops, c["gen"], 31337)
    out = np.zeros(1, dtype=c["dt"])
    for name in ("point_to_montgomery", "point_to_montgomery_async"):
        got = point_call(S, curve, name, out, rec(p, mont=False))
        assert fields(got) == fields(rec(p))
        got = point_call(S, curve, name, out, rec(None))
        assert fields(got) == fields(rec(None))
    for name in ("point_from_montgomery", "point_from_montgomery_async"):
        got = point_call(S, curve, name, out, rec(p))
        assert fields(got) == fields(rec(p, mont=False))
    for name in ("point_at_infinity", "point_at_infinity_async"):
        got = point_call(S, curve, name, rec(p))
        assert fields(got) == fields(rec(None))
    pts = [m.ec_mul(ops, c["gen"], 1000 + 7 * i) for i in range(300)]
    batch = np.concatenate([rec(q, mont=False) for q in pts])
    for i in (0, 17, 299):
        batch[i]["inf"] = 1
        batch[i]["x"] = limbs(11, len(batch[i]["x"]))
    before = batch.copy()
    d = to_dev(batch)
    for name in ("point_to_montgomery_batch", "point_to_montgomery_batch_async"):
        if name.endswith("async"):
            d = to_dev(batch)
        f = S.fn(f"{name}<{c['T']}>", [P, U32, P, I32])
        f(stream(), 0, d.data_ptr(), len(pts))
        got = to_host(d, c["dt"])
        for i in range(len(pts)):
            want = before[i:i + 1] if before[i]["inf"] else rec(pts[i])
            assert fields(got[i:i + 1]) == fields(want), i


# -------------------------------------------------------- device-pointer MSM

@pytest.mark.parametrize("curve,n", [("g1", 3000), ("g2", 500)])
def test_point_msm_u64_device_api_vs_oracle(S, curve, n):
    """point_msm_u64_{g1,g2} and their _async forms (reference
    src/curve.cu:2538-2758) on device buffers, caller scratch sized by the
    query and NULL scratch, against the oracle's MSM."""
    ks = o.gen_u64(900 + n, n)
    bases = o.bases_from_k(curve, ks)
    bases[5]["inf"] = 1
    bases[5]["x"] = 0
    bases[5]["y"] = 0
    sc = o.gen_u64(901 + n, n)
    sc[9] = 0
    want = o.msm(curve, bases, sc.reshape(-1, 1))
    q = S.fn(f"point_msm_u64_scratch_bytes_{curve}", [I32], ctypes.c_size_t)
    nbytes = q(n)
    assert nbytes > 0
    d_b, d_s = to_dev(bases), to_dev(sc)
    scratch = empty_dev(nbytes)
    proj_dt = o.G1P_DT if curve == "g1" else o.G2P_DT
    for name in (f"point_msm_u64_{curve}", f"point_msm_u64_async_{curve}"):
        f = S.fn(name, [P, U32, P, P, P, P, I32])
        for use_scratch in (True, False):
            d_out = empty_dev(proj_dt.itemsize)
            f(stream(), 0, d_out.data_ptr(), d_b.data_ptr(), d_s.data_ptr(),
              scratch.data_ptr() if use_scratch else None, n)
            got = o.to_affine(curve, to_host(d_out, proj_dt))
            assert fields(got) == fields(want), (name, use_scratch)
        d_out = to_dev(np.full(proj_dt.itemsize, 0xAB, dtype=np.uint8))
        f(stream(), 0, d_out.data_ptr(), d_b.data_ptr(), d_s.data_ptr(), None, 0)
        assert not to_host(d_out, np.uint8).any(), "n = 0 must give the all-zero (Z = 0) point"


@pytest.mark.parametrize("curve", ["g1", "g2"])
def test_point_msm_u64_async_back_to_back(S, curve):
    """Six point_msm_u64_async calls enqueued on one stream with no host wait
    between them (one shared scratch buffer, then NULL scratch), each into its
    own result record: the stream-ordered host finishes (pinned slots taken and
    released in stream order) must give every MSM its own oracle result."""
    sizes = [700, 1, 64, 513, 2, 1000]
    ks = o.gen_u64(930, max(sizes))
    bases = o.bases_from_k(curve, ks)
    sc = o.gen_u64(931, max(sizes))
    q = S.fn(f"point_msm_u64_scratch_bytes_{curve}", [I32], ctypes.c_size_t)
    scratch = empty_dev(q(max(sizes)))
    d_b, d_s = to_dev(bases), to_dev(sc)
    proj_dt = o.G1P_DT if curve == "g1" else o.G2P_DT
    f = S.fn(f"point_msm_u64_async_{curve}", [P, U32, P, P, P, P, I32])
    for use_scratch in (True, False):
        d_out = empty_dev(len(sizes) * proj_dt.itemsize)
        for k, n in enumerate(sizes):
            f(stream(), 0, d_out.data_ptr() + k * proj_dt.itemsize, d_b.data_ptr(), d_s.data_ptr(),
              scratch.data_ptr() if use_scratch else None, n)
        got = to_host(d_out, proj_dt)
        for k, n in enumerate(sizes):
            want = o.msm(curve, bases[:n], sc[:n].reshape(-1, 1))
            assert fields(o.to_affine(curve, got[k:k + 1])) == fields(want), (use_scratch, k, n)


@pytest.mark.parametrize("curve,n", [("g1", 2000), ("g2", 300)])
def test_point_msm_multi_limb_device_api_vs_oracle(S, curve, n):
    """point_msm<PointType> / point_msm_async (reference
    src/curve.cu:2763-2826): 5-limb scalars below r, affine result."""
    ks = o.gen_u64(1900 + n, n)
    bases = o.bases_from_k(curve, ks)
    sc = o.gen_scalars_modr(1901 + n, n)
    want = o.msm(curve, bases, sc)
    c = CURVES[curve]
    q = S.fn(f"point_msm_scratch_bytes<{c['T']}>", [I32, I32], ctypes.c_size_t)
    scratch = empty_dev(q(n, 5))
    d_b, d_s = to_dev(bases), to_dev(sc)
    for name in ("point_msm", "point_msm_async"):
        f = S.fn(f"{name}<{c['T']}>", [P, U32, P, P, P, I32, P, I32])
        for use_scratch in (True, False):
            d_out = empty_dev(c["dt"].itemsize)
            f(stream(), 0, d_out.data_ptr(), d_b.data_ptr(), d_s.data_ptr(), 5,
              scratch.data_ptr() if use_scratch else None, n)
            got = to_host(d_out, c["dt"])
            assert fields(got) == fields(want), (name, use_scratch)
        d_out = empty_dev(c["dt"].itemsize)
        f(stream(), 0, d_out.data_ptr(), d_b.data_ptr(), d_s.data_ptr(), 5, None, 0)
        assert fields(to_host(d_out, c["dt"])) == fields(c["rec"](None))


# ------------------------------------------------------------ field elements

def rand_fp(rng, n):
    return [int.from_bytes(rng.bytes(56), "little") % m.P for _ in range(n)]


def fp_arr(vals) -> np.ndarray:
    return np.stack([limbs(v) for v in vals]).astype(np.uint64)


def fp2_arr(vals) -> np.ndarray:
    return np.stack([np.concatenate([limbs(a), limbs(b)]) for a, b in vals]).astype(np.uint64)


def ints(arr: np.ndarray):
    return [o.limbs_to_int(r) for r in arr.reshape(-1, 7)]


def f2_ints(arr: np.ndarray):
    v = ints(arr)
    return list(zip(v[0::2], v[1::2]))


def test_fp_stream_ops_vs_bigint(S):
    """fp.h stream API (reference src/fp_kernels.cu:753-906): add, sub, mul
    and div in normal form, mont_mul, neg, inv, copy, zero, one, Montgomery
    conversions; inputs include 0, 1 and p - 1."""
    rng = np.random.default_rng(5)
    vals = rand_fp(rng, 12) + [0, 1, m.P - 1]
    pairs = list(zip(vals, reversed(vals)))
    rinv = m.RMONT_INV
    bin_ops = {
        "fp_add": lambda a, b: (a + b) % m.P,
        "fp_sub": lambda a, b: (a - b) % m.P,
        "fp_mul": lambda a, b: a * b % m.P,
        "fp_div": lambda a, b: a * pow(b, -1, m.P) % m.P if b else 0,
        "fp_mont_mul": lambda a, b: a * b * rinv % m.P,
    }
    un_ops = {
        "fp_neg": lambda a: -a % m.P,
        "fp_inv": lambda a: pow(a, -1, m.P) if a else 0,
        "fp_copy": lambda a: a,
        "fp_to_montgomery": lambda a: a * m.RMONT % m.P,
        "fp_from_montgomery": lambda a: a * rinv % m.P,
    }
    out = empty_dev(56)
    for base, fn in bin_ops.items():
        for name in (base, base + "_async"):
            f = S.fn(name, [P, U32, P, P, P])
            for a, b in pairs:
                da, db = to_dev(limbs(a)), to_dev(limbs(b))
                f(stream(), 0, out.data_ptr(), da.data_ptr(), db.data_ptr())
                assert ints(to_host(out, np.uint64))[0] == fn(a, b), (name, a, b)
    for base, fn in un_ops.items():
        for name in (base, base + "_async"):
            f = S.fn(name, [P, U32, P, P])
            for a in vals:
                da = to_dev(limbs(a))
                f(stream(), 0, out.data_ptr(), da.data_ptr())
                assert ints(to_host(out, np.uint64))[0] == fn(a), (name, a)
    for name, want in (("fp_zero", 0), ("fp_one", 1), ("fp_zero_async", 0), ("fp_one_async", 1)):
        f = S.fn(name, [P, U32, P])
        f(stream(), 0, out.data_ptr())
        assert ints(to_host(out, np.uint64))[0] == want


def test_fp_host_pointer_ops_vs_bigint(S):
    """fp_kernels.h host-pointer operations (reference src/fp_kernels.cu:
    297-751): the predicates, cmov, sqrt (a^((p+1)/4)), Euler's criterion and
    pow, plus the arithmetic ones."""
    rng = np.random.default_rng(6)
    vals = rand_fp(rng, 10) + [0, 1, m.P - 1, 4]
    h = lambda v: limbs(v)  # noqa: E731
    ptr = lambda a: a.ctypes.data  # noqa: E731
    s = stream()
    out = np.zeros(7, dtype=np.uint64)
    f_cmp = S.fn("fp_cmp_gpu", [P, U32, P, P], I32)
    f_zero = S.fn("fp_is_zero_gpu", [P, U32, P], ctypes.c_bool)
    f_one = S.fn("fp_is_one_gpu", [P, U32, P], ctypes.c_bool)
    f_sqrt = S.fn("fp_sqrt_gpu", [P, U32, P, P], ctypes.c_bool)
    f_qr = S.fn("fp_is_quadratic_residue_gpu", [P, U32, P], ctypes.c_bool)
    f_pow = S.fn("fp_pow_u64_gpu", [P, U32, P, P, U64])
    f_cmov = S.fn("fp_cmov_gpu", [P, U32, P, P, U64])
    f_mul = S.fn("fp_mul_gpu", [P, U32, P, P, P])
    f_inv = S.fn("fp_inv_gpu", [P, U32, P, P])
    euler = (m.P - 1) // 2
    for a in vals:
        b = vals[(vals.index(a) + 3) % len(vals)]
        ha, hb = h(a), h(b)
        assert f_cmp(s, 0, ptr(ha), ptr(hb)) == (a > b) - (a < b)
        assert f_zero(s, 0, ptr(ha)) == (a == 0)
        assert f_one(s, 0, ptr(ha)) == (a == 1)
        qr = a == 0 or pow(a, euler, m.P) == 1
        assert f_qr(s, 0, ptr(ha)) == qr
        ok = f_sqrt(s, 0, ptr(out), ptr(ha))
        assert ok == qr
        assert o.limbs_to_int(out) == (pow(a, (m.P + 1) // 4, m.P) if qr else 0)
        if qr:
            assert o.limbs_to_int(out) ** 2 % m.P == a
        for e in (0, 1, 65537, MASK64):
            f_pow(s, 0, ptr(out), ptr(ha), e)
            assert o.limbs_to_int(out) == pow(a, e, m.P)
        f_mul(s, 0, ptr(out), ptr(ha), ptr(hb))
        assert o.limbs_to_int(out) == a * b % m.P
        f_inv(s, 0, ptr(out), ptr(ha))
        assert o.limbs_to_int(out) == (pow(a, -1, m.P) if a else 0)
        for cond in (0, 1, 2, 3):
            dst = h(b).copy()
            f_cmov(s, 0, ptr(dst), ptr(ha), cond)
            assert o.limbs_to_int(dst) == (a if cond & 1 else b)
    assert sum(1 for a in vals if pow(a, euler, m.P) == m.P - 1) >= 2, "both residue classes exercised"


def test_fp_array_kernels_vs_bigint(S):
    """fp_add/mul_array_{host,device} (reference src/fp_kernels.cu:151-295)
    over 5000 elements."""
    rng = np.random.default_rng(7)
    n = 5000
    a, b = rand_fp(rng, n), rand_fp(rng, n)
    ha, hb = fp_arr(a), fp_arr(b)
    s = stream()
    for name, fn in (("fp_add_array", lambda x, y: (x + y) % m.P), ("fp_mul_array", lambda x, y: x * y % m.P)):
        c = np.zeros_like(ha)
        S.fn(name + "_host", [P, U32, P, P, P, I32])(s, 0, c.ctypes.data, ha.ctypes.data, hb.ctypes.data, n)
        want = [fn(x, y) for x, y in zip(a, b)]
        assert ints(c) == want, name
        da, db, dc = to_dev(ha), to_dev(hb), empty_dev(ha.nbytes)
        S.fn(name + "_device", [P, U32, P, P, P, I32])(s, 0, dc.data_ptr(), da.data_ptr(), db.data_ptr(), n)
        assert ints(to_host(dc, np.uint64)) == want, name


def test_fp2_ops_vs_bigint(S):
    """fp2.h stream API and fp2_kernels.h (reference src/fp2_kernels.cu,
    formulas src/fp2.cu): Karatsuba mul and square in normal form, the norm
    inverse, division, conjugation (= Frobenius), multiplication by i, the
    predicates and cmov; i^2 = -1 and (1 + i)^2 = 2i as known answers."""
    rng = np.random.default_rng(8)
    P_ = m.P
    vals = list(zip(rand_fp(rng, 6), rand_fp(rng, 6))) + [(0, 0), (1, 0), (0, 1), (1, 1), (P_ - 1, 0)]
    mul = lambda x, y: ((x[0] * y[0] - x[1] * y[1]) % P_, (x[0] * y[1] + x[1] * y[0]) % P_)  # noqa: E731

    def inv(x):
        if x == (0, 0):
            return (0, 0)
        nrm = pow((x[0] * x[0] + x[1] * x[1]) % P_, -1, P_)
        return (x[0] * nrm % P_, -x[1] * nrm % P_)

    bin_ops = {
        "fp2_add": lambda x, y: ((x[0] + y[0]) % P_, (x[1] + y[1]) % P_),
        "fp2_sub": lambda x, y: ((x[0] - y[0]) % P_, (x[1] - y[1]) % P_),
        "fp2_mul": mul,
        "fp2_div": lambda x, y: mul(x, inv(y)),
    }
    un_ops = {
        "fp2_square": lambda x: mul(x, x),
        "fp2_neg": lambda x: (-x[0] % P_, -x[1] % P_),
        "fp2_conjugate": lambda x: (x[0], -x[1] % P_),
        "fp2_inv": inv,
        "fp2_copy": lambda x: x,
    }
    assert mul((0, 1), (0, 1)) == (P_ - 1, 0) and mul((1, 1), (1, 1)) == (0, 2)
    s = stream()
    out = empty_dev(112)
    for base, fn in bin_ops.items():
        f_dev = S.fn(base, [P, U32, P, P, P])
        f_gpu = S.fn(base + "_gpu", [P, P, P])
        for x, y in zip(vals, reversed(vals)):
            da, db = to_dev(fp2_arr([x])), to_dev(fp2_arr([y]))
            f_dev(s, 0, out.data_ptr(), da.data_ptr(), db.data_ptr())
            assert f2_ints(to_host(out, np.uint64))[0] == fn(x, y), (base, x, y)
            ho = np.zeros(14, dtype=np.uint64)
            hx, hy = fp2_arr([x]), fp2_arr([y])  # kept alive across the call
            f_gpu(ho.ctypes.data, hx.ctypes.data, hy.ctypes.data)
            assert f2_ints(ho)[0] == fn(x, y), (base + "_gpu", x, y)
    gpu_only = {"fp2_mul_by_i": lambda x: (-x[1] % P_, x[0]), "fp2_frobenius": lambda x: (x[0], -x[1] % P_)}
    for base, fn in list(un_ops.items()) + list(gpu_only.items()):
        f_gpu = S.fn(base + "_gpu", [P, P])
        f_dev = None if base in gpu_only else S.fn(base + "_async", [P, U32, P, P])
        for x in vals:
            ho = np.zeros(14, dtype=np.uint64)
            hx = fp2_arr([x])
            f_gpu(ho.ctypes.data, hx.ctypes.data)
            assert f2_ints(ho)[0] == fn(x), (base, x)
            if f_dev:
                da = to_dev(fp2_arr([x]))
                f_dev(s, 0, out.data_ptr(), da.data_ptr())
                assert f2_ints(to_host(out, np.uint64))[0] == fn(x), (base + "_async", x)
    f_cmp = S.fn("fp2_cmp_gpu", [P, P], I32)
    f_zero = S.fn("fp2_is_zero_gpu", [P], ctypes.c_bool)
    f_one = S.fn("fp2_is_one_gpu", [P], ctypes.c_bool)
    f_cmov = S.fn("fp2_cmov_gpu", [P, P, U64])
    for x in vals:
        y = vals[(vals.index(x) + 2) % len(vals)]
        hx, hy = fp2_arr([x]), fp2_arr([y])
        assert f_cmp(hx.ctypes.data, hy.ctypes.data) == ((x > y) - (x < y))
        assert f_zero(hx.ctypes.data) == (x == (0, 0))
        assert f_one(hx.ctypes.data) == (x == (1, 0))
        for cond in (0, 1):
            dst = hy.copy()
            f_cmov(dst.ctypes.data, hx.ctypes.data, cond)
            assert f2_ints(dst)[0] == (x if cond else y)
    n = 2000
    xs = list(zip(rand_fp(rng, n), rand_fp(rng, n)))
    ys = list(zip(rand_fp(rng, n), rand_fp(rng, n)))
    hx, hy = fp2_arr(xs), fp2_arr(ys)
    for name, fn in (("fp2_add_array_host", bin_ops["fp2_add"]), ("fp2_mul_array_host", mul)):
        c = np.zeros_like(hx)
        S.fn(name, [P, U32, P, P, P, I32])(s, 0, c.ctypes.data, hx.ctypes.data, hy.ctypes.data, n)
        assert f2_ints(c) == [fn(x, y) for x, y in zip(xs, ys)], name


# ----------------------------------------------------------------- harnesses

def _write(tmp_path, name, arr):
    p = tmp_path / name
    np.ascontiguousarray(arr).tofile(p)
    return str(p)


@pytest.mark.parametrize("curve,n", [("g1", 1024), ("g2", 256), ("g1", 0)])
def test_rust_flow_harness_vs_oracle(tmp_path, curve, n):
    """The reference Rust layer's sequence over the C ABI (types.rs:345-376 /
    492-532, then to_affine and from_montgomery), as a C program."""
    exe = harness.build("rust_flow")
    ks = o.gen_u64(3100 + n, max(n, 1))[:n]
    bases = o.bases_from_k(curve, ks)
    sc = o.gen_u64(3101 + n, max(n, 1))[:n]
    given = np.stack([o.from_mont(curve, bases[i:i + 1])[0] for i in range(n)]) if curve == "g1" and n else bases
    out = tmp_path / "out.bin"
    print(f"rust_flow {curve} n={n}", flush=True)
    r = subprocess.run([exe, curve, str(n), _write(tmp_path, "p.bin", given), _write(tmp_path, "s.bin", sc),
                        str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    got = np.fromfile(out, dtype=CURVES[curve]["dt"])
    if n == 0:
        want = CURVES[curve]["rec"](None)
    else:
        want = o.from_mont(curve, o.msm(curve, bases, sc.reshape(-1, 1)))
    assert fields(got) == fields(want)


@pytest.mark.parametrize("curve,n", [("g1", 4096), ("g2", 512)])
def test_device_api_harness_vs_oracle(tmp_path, curve, n):
    """A C++ caller of include/curve.h: batch Montgomery conversion, then
    point_msm_u64_* sync / async with sized and NULL scratch (the program
    checks the four agree), normalised with the C ABI helper."""
    exe = harness.build("device_api")
    ks = o.gen_u64(3200 + n, n)
    bases = o.bases_from_k(curve, ks)
    sc = o.gen_u64(3201 + n, n)
    std = np.stack([o.from_mont(curve, bases[i:i + 1])[0] for i in range(n)])
    out = tmp_path / "out.bin"
    print(f"device_api {curve} n={n}", flush=True)
    r = subprocess.run([exe, curve, str(n), _write(tmp_path, "p.bin", std), _write(tmp_path, "s.bin", sc),
                        str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "agree=1" in r.stdout
    got = np.fromfile(out, dtype=CURVES[curve]["dt"])
    want = o.msm(curve, bases, sc.reshape(-1, 1))
    assert fields(got) == fields(want)


def test_msm_flow_harness_reference_tests(tmp_path):
    """The reference's tests/test_msm.cu flow call for call (host
    g1_generator / g2_generator, fp_to_montgomery in place, scratch sized as
    (blocks + 1) * MSM_BUCKET_COUNT points, point_msm_u64_*, the expected
    N(N+1)/2 * G by point_scalar_mul_u64, host projective_to_affine, device
    point_from_montgomery, fp_cmp): N = 10 (G1, G2), 100 (G1), then 1000 and
    5000 on both curves, whose reference-sized scratch is far too small for
    this pipeline; guard buffers around it must stay untouched."""
    exe = harness.build("msm_flow")
    print("msm_flow default cases", flush=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ALL PASS" in r.stdout and r.stdout.count("PASS (") == 7, r.stdout


@pytest.mark.parametrize("curve,n", [("g1", 3000), ("g2", 700)])
def test_point_msm_u64_reference_sized_scratch_vs_oracle(S, curve, n):
    """A d_scratch sized the reference's way ((blocks + 1) * 16 projective
    points, reference tests/test_msm.cu:146-155) is shorter than this
    pipeline's workspace: the MSM must not write past it (it takes
    stream-ordered memory of its own) and still equal the oracle. A buffer
    allocated right after the scratch must come back unchanged."""
    ks = o.gen_u64(3300 + n, n)
    bases = o.bases_from_k(curve, ks)
    sc = o.gen_u64(3301 + n, n)
    want = o.msm(curve, bases, sc.reshape(-1, 1))
    proj_dt = o.G1P_DT if curve == "g1" else o.G2P_DT
    ref_bytes = ((n + 255) // 256 + 1) * 16 * proj_dt.itemsize
    need = S.fn(f"point_msm_u64_scratch_bytes_{curve}", [I32], ctypes.c_size_t)(n)
    assert ref_bytes < need
    d_b, d_s = to_dev(bases), to_dev(sc)
    scratch = empty_dev(ref_bytes, 0x5A)
    guard = empty_dev(1 << 20, 0x5A)
    for name in (f"point_msm_u64_{curve}", f"point_msm_u64_async_{curve}"):
        f = S.fn(name, [P, U32, P, P, P, P, I32])
        d_out = empty_dev(proj_dt.itemsize)
        print(f"{name} n={n} reference-sized scratch {ref_bytes} B", flush=True)
        f(stream(), 0, d_out.data_ptr(), d_b.data_ptr(), d_s.data_ptr(), scratch.data_ptr(), n)
        got = o.to_affine(curve, to_host(d_out, proj_dt))
        assert fields(got) == fields(want), name
    assert (to_host(guard, np.uint8) == 0x5A).all()
