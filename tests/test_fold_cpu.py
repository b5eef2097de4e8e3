"""The one-pass fold of the bucket kernels' lazily reduced field products
(fold_est, gpu_msm_amd/csrc/field.hpp), built for the host: for every range
it serves ((-2p, 4p), (-4p, 4p), (0, 8p), (0, 12p)) the quotient floor(v / 2p)
is read off the top limb against the multiples of 2p, and a tie in the top
limb takes a correcting conditional subtraction. Values exactly at, next to
and half a top-limb unit around every multiple of 2p in range (negative ones
included), plus random ones, must come out as v - floor(v / 2p) 2p, in
[0, 2p) with normalised limbs, against Python integers."""
import os
import random
import shutil
import subprocess
import sys

import pytest

import pymodel as m

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "harness"))
import build as harness  # noqa: E402

M28 = (1 << 28) - 1
RANGES = [(-1, 1), (-2, 1), (0, 3), (0, 5)]


def encode(v: int):
    top = v >> 420  # floor, also for negative v
    low = v - (top << 420)
    return [(low >> (28 * i)) & M28 for i in range(15)], top


def cases():
    P2 = 2 * m.P
    out = []
    for kmin, kmax in RANGES:
        lo, hi = kmin * P2, (kmax + 1) * P2
        vals = set()
        for j in range(kmin, kmax + 2):
            for d in (-2, -1, 0, 1, 2, -(1 << 419), 1 << 419):
                if lo <= j * P2 + d < hi:
                    vals.add(j * P2 + d)
        rnd = random.Random(1000 * kmin + kmax)
        vals.update(rnd.randrange(lo, hi) for _ in range(200))
        out += [(kmin, kmax, v) for v in sorted(vals)]
    return out


@pytest.mark.skipif(shutil.which(os.path.join(harness.ROCM, "bin", "hipcc")) is None, reason="hipcc not found")
def test_fold_est_at_and_around_every_multiple_of_2p():
    exe = harness.build("fold_host")
    cs = cases()
    text = "".join(f"{a} {b} " + " ".join(map(str, encode(v)[0])) + f" {encode(v)[1]}\n" for a, b, v in cs)
    res = subprocess.run([exe], input=text, capture_output=True, text=True, check=True).stdout.split("\n")
    bad = []
    for (a, b, v), line in zip(cs, res):
        ls = [int(x) for x in line.split()]
        r = sum(x << (28 * i) for i, x in enumerate(ls))
        if len(ls) != 16 or max(ls) > M28 or r != v % (2 * m.P):
            bad.append((a, b, v))
    assert len(res) >= len(cs) and not bad, f"{len(bad)} of {len(cs)} wrong, first {bad[:2]}"
