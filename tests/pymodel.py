"""Independent big-integer model of BLS12-446 used to pin the oracle.

Written from the mathematics, not from the reference code: Python ints mod p,
Fp2 = Fp[i]/(i^2 + 1), the textbook affine short-Weierstrass group law
(a = 0), and the BLS12 parameterisation p = (x-1)^2 (x^4 - x^2 + 1)/3 + x,
r = x^4 - x^2 + 1 for x = -0x6008204000000020001.
"""
from __future__ import annotations

X_SEED = -0x6008204000000020001
P = (X_SEED - 1) ** 2 * (X_SEED**4 - X_SEED**2 + 1) // 3 + X_SEED
R_ORDER = X_SEED**4 - X_SEED**2 + 1
RMONT = 1 << 448
RMONT_INV = pow(RMONT, -1, P)


def to_mont(v: int) -> int:
    return v * RMONT % P


def from_mont(v: int) -> int:
    return v * RMONT_INV % P


def mont_mul(a: int, b: int) -> int:
    return a * b * RMONT_INV % P


# ---------------------------------------------------------------- Fp2 (ints pairs)

def f2_add(a, b):
    return ((a[0] + b[0]) % P, (a[1] + b[1]) % P)


def f2_sub(a, b):
    return ((a[0] - b[0]) % P, (a[1] - b[1]) % P)


def f2_mul(a, b):
    return ((a[0] * b[0] - a[1] * b[1]) % P, (a[0] * b[1] + a[1] * b[0]) % P)


def f2_inv(a):
    n = pow(a[0] * a[0] + a[1] * a[1], -1, P)
    return (a[0] * n % P, -a[1] * n % P)


def f2_neg(a):
    return (-a[0] % P, -a[1] % P)


# ------------------------------------------------------------ affine group law
# A point is None (infinity) or (x, y) with coordinates in the field given by
# the ops tuple (add, sub, mul, inv, neg, zero, small).

FP_OPS = (
    lambda a, b: (a + b) % P,
    lambda a, b: (a - b) % P,
    lambda a, b: a * b % P,
    lambda a: pow(a, -1, P),
    lambda a: -a % P,
    0,
    lambda k: k % P,
)
FP2_OPS = (f2_add, f2_sub, f2_mul, f2_inv, f2_neg, (0, 0), lambda k: (k % P, 0))


def ec_add(ops, p1, p2):
    add, sub, mul, inv, neg, zero, small = ops
    if p1 is None:
        return p2
    if p2 is None:
        return p1
    x1, y1 = p1
    x2, y2 = p2
    if x1 == x2:
        if add(y1, y2) == zero:
            return None
        lam = mul(mul(small(3), mul(x1, x1)), inv(mul(small(2), y1)))
    else:
        lam = mul(sub(y2, y1), inv(sub(x2, x1)))
    x3 = sub(sub(mul(lam, lam), x1), x2)
    y3 = sub(mul(lam, sub(x1, x3)), y1)
    return (x3, y3)


def ec_neg(ops, p):
    return None if p is None else (p[0], ops[4](p[1]))


def ec_mul(ops, p, k: int):
    if k < 0:
        return ec_mul(ops, ec_neg(ops, p), -k)
    acc = None
    while k:
        if k & 1:
            acc = ec_add(ops, acc, p)
        p = ec_add(ops, p, p)
        k >>= 1
    return acc


def on_curve_g1(p) -> bool:
    if p is None:
        return True
    x, y = p
    return (y * y - x * x * x - 1) % P == 0


def on_curve_g2(p) -> bool:
    if p is None:
        return True
    x, y = p
    lhs = f2_mul(y, y)
    rhs = f2_add(f2_mul(f2_mul(x, x), x), (1, 1))
    return lhs == rhs


# Generators in standard form (the constants of reference src/curve.cu:786-814,
# as integers).
def _limbs(*ls):
    return sum(v << (64 * k) for k, v in enumerate(ls))


G1_STD = (
    _limbs(0x3BF9166C8236F4FA, 0x8BC02B7CBE6A9E8D, 0x11C1E56B3E4BC80B, 0x6B20D782901A6F62,
           0x2CE8C34265BF3841, 0x11B73D3D76AE9851, 0x326ED6BD777FC6A3),
    _limbs(0xFE6F792612016B30, 0x22DB0CE6034A9DB9, 0xB9093F32002756DA, 0x39D7F424B6660204,
           0xF843C947AA57F571, 0xD6D62D244E413636, 0x1A7CAF4A4D3887A6),
)
G2_STD = (
    (
        _limbs(0x0E529EE4DCE9991D, 0xD6EBAF149094F1CC, 0x043C6BF16312D638, 0x062B61439640E885,
               0x18DAD8ED784DD225, 0xA57C0038441F7D15, 0x21F8D4A76F74541A),
        _limbs(0xCAF5185423A7D23A, 0x7CEF6ACB145B6413, 0x2879DD439B019B8B, 0x71449CDECA4F0007,
               0xDEBAF4A2C5534527, 0xA1B4E791D1B86560, 0x1E0F563C601BB8DC),
    ),
    (
        _limbs(0x274315837455B919, 0x82039E4221FF3507, 0x00346CEBAD16A036, 0x0177BFD6654E681E,
               0xDDFF621B5DB3F897, 0x0CC61570301497A7, 0x115EA2305A78F646),
        _limbs(0x392236E9CF2976C2, 0xD8AB17C84B9F03CD, 0x8A8E6755F9D82FD1, 0x7532834528CD5A64,
               0x0B0BCC3FB6F2161C, 0x76A2FFCB7D47679D, 0x25ED2192B203C1FE),
    ),
)


# -------------------------------------------------- conversion from ABI arrays

def limbs_to_int(a) -> int:
    return sum(int(x) << (64 * k) for k, x in enumerate(list(a)))


def g1_from_abi(rec, mont: bool = True):
    """(x, y) standard-form ints, or None, from one G1Point record."""
    if rec["inf"]:
        return None
    x, y = limbs_to_int(rec["x"]), limbs_to_int(rec["y"])
    return (from_mont(x), from_mont(y)) if mont else (x, y)


def g2_from_abi(rec, mont: bool = True):
    if rec["inf"]:
        return None
    xs = rec["x"]
    ys = rec["y"]
    x = (limbs_to_int(xs[:7]), limbs_to_int(xs[7:]))
    y = (limbs_to_int(ys[:7]), limbs_to_int(ys[7:]))
    if mont:
        x = (from_mont(x[0]), from_mont(x[1]))
        y = (from_mont(y[0]), from_mont(y[1]))
    return (x, y)


# ------------------------------------------------------------- mt19937_64 RNG

class MT64:
    """Pure-Python mt19937_64 (Matsumoto & Nishimura), the std::mt19937_64 engine."""

    def __init__(self, seed: int):
        M = (1 << 64) - 1
        self.mt = [0] * 312
        self.mt[0] = seed & M
        for i in range(1, 312):
            prev = self.mt[i - 1]
            self.mt[i] = (6364136223846793005 * (prev ^ (prev >> 62)) + i) & M
        self.idx = 312

    def next(self) -> int:
        M = (1 << 64) - 1
        if self.idx >= 312:
            mt = self.mt
            for i in range(312):
                x = (mt[i] & 0xFFFFFFFF80000000) | (mt[(i + 1) % 312] & 0x7FFFFFFF)
                xa = x >> 1
                if x & 1:
                    xa ^= 0xB5026F5AA96619E9
                mt[i] = mt[(i + 156) % 312] ^ xa
            self.idx = 0
        y = self.mt[self.idx]
        self.idx += 1
        y ^= (y >> 29) & 0x5555555555555555
        y ^= (y << 17) & 0x71D67FFFEDA60000 & M
        y ^= (y << 37) & 0xFFF7EEE000000000 & M
        y ^= y >> 43
        return y & M


def gen_u64(seed: int, n: int):
    g = MT64(seed)
    return [g.next() for _ in range(n)]


def gen_scalars_modr(seed: int, n: int):
    """Uniform in [0, r): 5 draws, top limb masked to 43 bits, reject >= r."""
    g = MT64(seed)
    out = []
    while len(out) < n:
        v = [g.next() for _ in range(5)]
        v[4] &= (1 << 43) - 1
        s = sum(x << (64 * k) for k, x in enumerate(v))
        if s < R_ORDER:
            out.append(s)
    return out


# ------------------------------------------------ point encoding (wire format)
# The layout of ark-serialize's CanonicalSerialize for short-Weierstrass
# affine points (the format tfhe-zk-pok uses; reference
# rust-api/src/conversions.rs converts to those points): canonical integers,
# 56 little-endian bytes per Fp (446 bits leave the top two bits of the last
# byte free), Fp2 as c0 then c1; compressed = x, uncompressed = x then y; flag
# 0x40 infinity, 0x80 "y is the larger of {y, -y}" in the encoding's last byte.
# Written from that description, independently of the device code.

HALF = (P - 1) // 2


def fp_bytes(v: int) -> bytes:
    return int(v).to_bytes(56, "little")


def larger_g1(y: int) -> bool:
    return y > HALF


def larger_g2(y) -> bool:
    return y[1] > HALF if y[1] else y[0] > HALF


def encode_point(curve: str, pt, compressed: bool) -> bytes:
    """pt: standard-form coordinates (ints for G1, pairs for G2) or None."""
    esz = 56 if curve == "g1" else 112
    size = esz if compressed else 2 * esz
    if pt is None:
        b = bytearray(size)
        b[-1] |= 0x40
        return bytes(b)
    x, y = pt
    enc = (lambda v: fp_bytes(v)) if curve == "g1" else (lambda v: fp_bytes(v[0]) + fp_bytes(v[1]))
    b = bytearray(enc(x) if compressed else enc(x) + enc(y))
    if (larger_g1(y) if curve == "g1" else larger_g2(y)):
        b[-1] |= 0x80
    return bytes(b)


def is_square_fp(a: int) -> bool:
    return a % P == 0 or pow(a, (P - 1) // 2, P) == 1


def is_square_fp2(a) -> bool:
    """a is a square in Fp2 = Fp[i]/(i^2 + 1) iff its norm a0^2 + a1^2 is a
    square in Fp."""
    return is_square_fp((a[0] * a[0] + a[1] * a[1]) % P)
