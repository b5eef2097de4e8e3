"""The lazily reduced field products of the bucket kernels at the edges of
their input ranges (gm_field_ops_check, include/gpu_msm_ext.h), against big
integers: every result must be the exact Montgomery product (value mod p) and
lie in [0, 2p) with normalised 28-bit limbs, the invariant the next product
relies on.

These products reduce lazily: a sum of limb products (signed where an operand
is a raw difference) is reduced once, then brought to [0, 2p) by as few
corrections as its value range allows. G1 (one Fp element per lane): a
product of values below 2p needs none, a raw product or difference of
products one conditional addition, the fused X3 one pass. G2 (one Fp2 element
per lane pair, each lane reducing its own component): a plain product takes
-p R into part 1's sum so that one conditional addition suffices; raw
products, differences of products and the fused X3 take one pass each. Those
passes subtract k 2p with the quotient k read off the top limb, which can tie
with a multiple of 2p and then takes a correcting branch. Whether the ranges
hold is decided by operands near their bounds (components near 2p, raw
differences near +-2p), which random field elements almost never are; the
operands here are drawn mostly from such edge values, after the worst-case
sums of each product, and for the fused X3 also values exactly at, just below
and just above multiples of 2p. Both forms run: the accumulation's
single-chain column sums (chain = 1) and the fix-up and tail kernels'
(chain = 0)."""
import numpy as np
import pytest

import gpu_msm_amd as gm
import hipmem as hm
import pymodel as m

pytestmark = pytest.mark.gpu

NL = 16
M28 = (1 << 28) - 1
P = m.P
RINV = pow(1 << 448, -1, P)
T = 2 * P - 1
EDGE = [0, 1, 2, P - 2, P - 1, P, P + 1, 2 * P - 2, T, (1 << 420) - 1, (1 << 446) - 1 - (1 << 420),
        (1 << 448) % P]
N = 3072
OPS = {0: "mul", 1: "mul_raw", 2: "mul_sub", 3: "mul_sub_raw", 4: "sqr_raw", 5: "sqr_sub2"}
# which of a, b, c, d are raw differences for each op
RAW = {0: (0, 0, 0, 0), 1: (1, 1, 0, 0), 2: (0, 0, 0, 0), 3: (1, 1, 0, 0), 4: (1, 0, 0, 0), 5: (1, 0, 0, 0)}
COMPS = {gm.CURVE_G1: 1, gm.CURVE_G2: 2}


def limbs(v: int) -> list[int]:
    return [(v >> (28 * i)) & M28 for i in range(NL)]


def raw_limbs(v: int) -> list[int]:
    """Limbs of a raw value v in (-2p, 2p) as x - y with x = max(v, 0)."""
    x, y = (v, 0) if v >= 0 else (0, -v)
    return [(a - b) & 0xFFFFFFFF for a, b in zip(limbs(x), limbs(y))]


def value(ls) -> int:
    return sum(int(x) << (28 * i) for i, x in enumerate(ls))


def pick(rng) -> int:
    """A value in [0, 2p): an edge value most of the time."""
    if rng.random() < 0.75:
        return EDGE[int(rng.integers(len(EDGE)))]
    return int.from_bytes(rng.bytes(57), "little") % (2 * P)


def operand(rng, raw: bool):
    """(value, 16 uint32 limbs) of one component: in [0, 2p), or a raw
    difference x - y of two such values taken limb by limb."""
    x = pick(rng)
    if not raw:
        return x, limbs(x)
    y = pick(rng)
    return x - y, [(a - b) & 0xFFFFFFFF for a, b in zip(limbs(x), limbs(y))]


def fmul(curve: int, a, b):
    if curve == gm.CURVE_G1:
        return (a[0] * b[0],)
    return (a[0] * b[0] - a[1] * b[1], a[0] * b[1] + a[1] * b[0])


def expected(curve: int, op: int, a, b, c, d):
    if op in (0, 1):
        s = fmul(curve, a, b)
    elif op in (2, 3):
        s = tuple(u - v for u, v in zip(fmul(curve, a, b), fmul(curve, c, d)))
    else:
        s = fmul(curve, a, a)
    e = tuple(v * RINV % P for v in s)
    if op == 5:
        e = tuple((e[k] - b[k] - 2 * c[k]) % P for k in range(len(e)))
    return e


def extremes(curve: int, op: int):
    """Deterministic worst cases of each product's sums: (a, b, c, d) values,
    raw ones of magnitude up to 2p - 1 where the op takes raw operands."""
    t = T
    if curve == gm.CURVE_G1:
        z = (0,)
        if op in (0, 2):  # every operand in [0, 2p)
            return [((t,), (t,), (t,), (t,)), ((0,), (t,), (t,), (t,)), ((t,), (t,), z, z)]
        if op in (1, 3):  # a, b raw; c, d in [0, 2p)
            return [((t,), (-t,), (t,), (t,)), ((-t,), (-t,), z, z), ((-t,), (t,), (0,), (t,))]
        base = [((t,), (t,), (t,), z), ((-t,), z, z, z), ((0,), (t,), (t,), z), ((t,), z, z, z)]
        if op == 5:  # a = 0: 6p - b - 2c exactly at, just below and just above the multiples of 2p
            base += [(z, (b,), (c,), z) for b, c in ((2, t), (3, t), (1, t), (0, P), (1, P), (0, 0), (1, 0))]
        return base
    z = (0, 0)
    if op in (0, 2):
        return [((t, t), (t, t), (t, t), (t, t)), ((0, t), (0, t), (t, 0), (t, 0)),
                ((t, 0), (t, 0), (0, t), (0, t)), ((t, t), (t, t), z, z)]
    if op in (1, 3):
        return [((t, -t), (t, -t), (t, t), (t, t)), ((t, -t), (t, t), (0, t), (0, t)),
                ((-t, -t), (-t, -t), (t, 0), (t, 0)), ((-t, t), (t, -t), z, z)]
    base = [((t, -t), (t, t), (t, t), z), ((-t, -t), z, z, z), ((t, t), (t, t), (t, t), z),
            ((-t, t), (t, 0), (0, t), z)]
    if op == 5:  # a = 0: 8p - b - 2c exactly at, just below and just above the multiples of 2p
        base += [(z, (2, 3), (t, t), z), (z, (1, 0), (t, 0), z), (z, (0, 1), (P, P), z),
                 (z, (0, 0), (0, 0), z), (z, (1, 2), (0, t), z)]
    return base


def build_case(curve: int, op: int, seed: int):
    rng = np.random.default_rng(seed)
    k_n = COMPS[curve]
    arrs = np.zeros((4, N, k_n, NL), dtype=np.uint32)
    ext = extremes(curve, op)
    vals = []
    for g in range(N):
        elem = []
        for j in range(4):
            comp = []
            for k in range(k_n):
                if g < len(ext):
                    v = ext[g][j][k]
                    ls = raw_limbs(v) if RAW[op][j] else limbs(v)
                else:
                    v, ls = operand(rng, bool(RAW[op][j]))
                arrs[j, g, k] = ls
                comp.append(v)
            elem.append(tuple(comp))
        vals.append(elem)
    return arrs, vals


@pytest.mark.parametrize("chain", [1, 0])
@pytest.mark.parametrize("op", sorted(OPS))
@pytest.mark.parametrize("curve", [gm.CURVE_G1, gm.CURVE_G2])
def test_field_products_at_range_edges(curve, op, chain):
    hm.require_gpu()
    arrs, vals = build_case(curve, op, 100 + 10 * curve + op)
    dev = [hm.to_dev(arrs[j]) for j in range(4)]
    out = hm.empty_dev(arrs[0].nbytes)
    rc = gm.lib().gm_field_ops_check(curve, chain, op, *[b.data_ptr() for b in dev], out.data_ptr(), N, hm.stream())
    assert rc == 0, f"gm_field_ops_check rc {rc}"
    k_n = COMPS[curve]
    res = hm.to_host(out, np.uint32).reshape(N, k_n, NL)
    bad = []
    for g in range(N):
        want = expected(curve, op, *vals[g])
        for k in range(k_n):
            ls = res[g, k]
            v = value(ls)
            if int(ls.max()) > M28 or v >= 2 * P or v % P != want[k]:
                bad.append((g, k, hex(v)))
    assert not bad, f"G{curve} {OPS[op]} chain={chain}: {len(bad)} wrong of {k_n * N}, first {bad[:3]}"


def test_field_check_rejects_bad_arguments():
    hm.require_gpu()
    fn = gm.lib().gm_field_ops_check
    buf = hm.empty_dev(2 * NL * 8)
    ptrs = [buf.data_ptr()] * 5
    assert fn(gm.CURVE_G2, 0, 6, *ptrs, 1, hm.stream()) != 0  # no op 6
    assert fn(3, 0, 0, *ptrs, 1, hm.stream()) != 0  # no curve 3
    assert fn(gm.CURVE_G1, 0, 0, *ptrs, 0, hm.stream()) == 0  # nothing to do
