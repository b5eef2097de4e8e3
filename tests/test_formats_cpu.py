"""The wire-format model (tests/pymodel.py, point encoding) without a GPU:
the flag bits never collide with value bits, every non-zero coordinate has
exactly one "larger" member in {y, -y}, the curve generators encode to their
coordinates, and the square-root criteria used to build invalid test inputs
agree with brute force on a small field."""
import random

import pymodel as m


def test_flag_bits_are_free_in_canonical_values():
    assert (m.P - 1).bit_length() == 446
    assert m.fp_bytes(m.P - 1)[-1] & 0xC0 == 0


def test_larger_is_one_of_each_pair():
    rng = random.Random(7)
    for _ in range(200):
        y = rng.randrange(1, m.P)
        assert m.larger_g1(y) != m.larger_g1(-y % m.P)
        y2 = (rng.randrange(m.P), rng.randrange(1, m.P) if rng.random() < 0.8 else 0)
        if y2 == (0, 0):
            continue
        ny2 = (-y2[0] % m.P, -y2[1] % m.P)
        assert m.larger_g2(y2) != m.larger_g2(ny2)


def test_generator_encodings():
    g1 = m.G1_STD
    enc = m.encode_point("g1", g1, True)
    assert len(enc) == 56 and int.from_bytes(enc, "little") & ((1 << 446) - 1) == g1[0]
    assert (enc[-1] & 0x80 != 0) == (g1[1] > (m.P - 1) // 2)
    enc = m.encode_point("g2", m.G2_STD, False)
    assert len(enc) == 224
    assert int.from_bytes(enc[:56], "little") == m.G2_STD[0][0]
    assert int.from_bytes(enc[168:], "little") & ((1 << 446) - 1) == m.G2_STD[1][1]
    assert m.encode_point("g1", None, False) == bytes(111) + b"\x40"


def test_square_criteria_match_brute_force_small_field():
    """Euler's criterion and the Fp2 norm criterion, checked on p = 43
    (3 mod 4, like the curve's p) by enumerating all squares."""
    p = 43
    squares = {(a * a) % p for a in range(p)}
    for a in range(p):
        assert (a == 0 or pow(a, (p - 1) // 2, p) == 1) == (a in squares)
    sq2 = set()
    for a0 in range(p):
        for a1 in range(p):
            sq2.add(((a0 * a0 - a1 * a1) % p, (2 * a0 * a1) % p))
    for a0 in range(p):
        for a1 in range(p):
            n = (a0 * a0 + a1 * a1) % p
            crit = n == 0 or pow(n, (p - 1) // 2, p) == 1
            assert crit == ((a0, a1) in sq2)
