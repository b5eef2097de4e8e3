"""gm_point_msm_noalloc (the MSM entry point that allocates nothing) against
the entry point it varies, point_msm_async<PointType> (its own stream-ordered
scratch), through the C ABI: two calls in a row on one stream, the second
smaller than the first (the fixed workspace keeps the first call's contents),
bit-exact affine records, plus n = 0 (O) and the size limit (-6, nothing
launched)."""
import ctypes

import numpy as np
import pytest

import gpu_msm_amd as gm
import hipmem as hm
from cxx_surface import I32, P, U32, Surface
from oracle import oracle as o

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def S():
    hm.require_gpu()
    return Surface()


def noalloc():
    f = gm.lib().gm_point_msm_noalloc
    f.argtypes = [P, U32, I32, P, P, P, I32, I32]
    f.restype = I32
    return f


@pytest.mark.parametrize("curve,sizes", [("g1", [1 << 15, 1000]), ("g2", [1 << 14, 300])])
def test_noalloc_matches_point_msm_async(S, curve, sizes):
    cv = gm.CURVE_G1 if curve == "g1" else gm.CURVE_G2
    T = "G1Point" if curve == "g1" else "G2Point"
    dt = o.G1_DT if curve == "g1" else o.G2_DT
    n0 = max(sizes)
    ks = o.gen_u64(4100 + n0, n0)
    b = gm.Bases.from_k(cv, ks)
    bases = b.download(0, n0)
    b.free()
    sc = o.gen_scalars_modr(4101 + n0, n0)
    d_b, d_s = hm.to_dev(bases), hm.to_dev(sc)
    ref = S.fn(f"point_msm_async<{T}>", [P, U32, P, P, P, I32, P, I32])
    f = noalloc()
    s = hm.stream()
    outs = []
    for n in sizes:  # two calls in a row, the second smaller
        print(f"{curve} noalloc n={n}", flush=True)
        d_out = hm.empty_dev(dt.itemsize, 0xAB)
        assert f(s, 0, cv, d_out.data_ptr(), d_b.data_ptr(), d_s.data_ptr(), 5, n) == 0
        outs.append(d_out)
    got = [hm.to_host(d, dt) for d in outs]
    for n, g in zip(sizes, got):
        d_ref = hm.empty_dev(dt.itemsize)
        ref(s, 0, d_ref.data_ptr(), d_b.data_ptr(), d_s.data_ptr(), 5, None, n)
        want = hm.to_host(d_ref, dt)
        assert g.tobytes() == want.tobytes(), (curve, n)
    small = min(sizes)
    assert got[1].tobytes() == o.msm(curve, bases[:small], sc[:small]).tobytes()
    d_out = hm.empty_dev(dt.itemsize, 0xAB)
    assert f(s, 0, cv, d_out.data_ptr(), d_b.data_ptr(), d_s.data_ptr(), 5, 0) == 0
    inf = hm.to_host(d_out, dt)
    assert inf["inf"][0] == 1 and not inf["x"].any() and not inf["y"].any()


def test_noalloc_rejects_sizes_beyond_its_workspace():
    hm.require_gpu()
    n = 1 << 17  # G1 with 5-limb scalars: a 177 MB workspace against the fixed 144 MiB
    d_b = hm.empty_dev(n * o.G1_DT.itemsize)
    d_s = hm.empty_dev(n * 40)
    d_out = hm.empty_dev(o.G1_DT.itemsize, 0xAB)
    f = noalloc()
    assert f(hm.stream(), 0, gm.CURVE_G1, d_out.data_ptr(), d_b.data_ptr(), d_s.data_ptr(), 5, n) == -6
    assert f(hm.stream(), 0, 3, d_out.data_ptr(), d_b.data_ptr(), d_s.data_ptr(), 5, 16) == -1
    assert f(hm.stream(), 0, gm.CURVE_G1, d_out.data_ptr(), d_b.data_ptr(), d_s.data_ptr(), 9, 16) == -1
    assert (hm.to_host(d_out, np.uint8) == 0xAB).all()  # nothing was written
