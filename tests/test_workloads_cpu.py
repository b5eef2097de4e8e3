"""The closed-form checker's pieces on the CPU: the exact exponent sums of
workloads.closed_form_scalars against Python integers (random rows, empty
and one-row MSMs, and 2^20 rows of maximal values, whose sum has the closed
form n (2^64 - 1)(r - 1)), and the library's host generator multiple
(gm_generator_multiple) against the oracle's scalar multiple and r G = O."""
import numpy as np

import gpu_msm_amd as gm
from gpu_msm_amd import workloads as wl
from oracle import oracle as o

R = wl.R_ORDER


def test_r_limbs_are_the_group_order():
    assert R == gm.R_ORDER
    x = -0x6008204000000020001
    assert R == x**4 - x**2 + 1


def test_closed_form_scalars_vs_python_integers():
    rng = np.random.default_rng(3)
    k = wl.uniform_k(rng, 900)
    s = wl.uniform_scalars(rng, 900)
    sizes = [0, 1, 300, 0, 599]
    got = wl.closed_form_scalars(k, s, sizes)
    at, want = 0, []
    for n in sizes:
        acc = sum(int(k[i]) * sum(int(w) << (64 * j) for j, w in enumerate(s[i])) for i in range(at, at + n))
        want.append(acc % R)
        at += n
    assert got == want


def test_closed_form_scalars_maximal_values_full_size():
    n = 1 << 20
    k = np.full(n, np.uint64(0xFFFFFFFFFFFFFFFF))
    rm1 = R - 1
    s = np.tile(np.array([(rm1 >> (64 * j)) & 0xFFFFFFFFFFFFFFFF for j in range(5)], dtype=np.uint64), (n, 1))
    assert wl.closed_form_scalars(k, s, [n]) == [n * 0xFFFFFFFFFFFFFFFF * rm1 % R]
    assert wl.closed_form_scalars(k, s, [n - 5, 5]) == [(n - 5) * 0xFFFFFFFFFFFFFFFF * rm1 % R,
                                                        5 * 0xFFFFFFFFFFFFFFFF * rm1 % R]


def test_generator_multiple_vs_oracle():
    for curve, name in ((gm.CURVE_G1, "g1"), (gm.CURVE_G2, "g2")):
        g = o.g1_generator() if name == "g1" else o.g2_generator()
        for k in [1, 2, 3, 12345, R - 1, (1 << 298) + 77, 0x1234567890ABCDEF << 200]:
            assert gm.generator_multiple(curve, k).tobytes() == o.scalar_mul(name, g, k).tobytes(), (name, k)
        inf = gm.generator_multiple(curve, R)  # r G = O: (0, 0, infinity)
        assert inf["inf"][0] == 1 and not inf["x"].any() and not inf["y"].any()
        assert gm.generator_multiple(curve, 0).tobytes() == inf.tobytes()
