"""C ABI checks that need no GPU: the library loads, exports every declared
symbol, and its host helper functions (reference rust-api/src/c_wrapper.cu:
12-59, 261-288) agree with the oracle."""
import ctypes
import os
import re

import numpy as np
import pytest

import gpu_msm_amd as gm
import pymodel as m
from oracle import oracle as o

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for h in ("rust_wrapper.h", "gpu_msm_ext.h"):
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for mm in re.finditer(r"^[A-Za-z_][\w\s\*]*?\b(\w+)\s*\(", txt, flags=re.M):
            if mm.group(1) not in ("if", "defined"):
                names.add(mm.group(1))
    return names


@pytest.fixture(scope="module")
def L():
    if not os.path.exists(gm.LIB_PATH):
        pytest.skip("library not built (python -m gpu_msm_amd.build)")
    return gm.lib()


def test_exports_every_declared_symbol(L):
    names = declared_symbols()
    assert len(names) >= 28
    for n in names:
        assert hasattr(L, n), n
    assert names <= set(gm.EXPORTS)


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(gm, "_lib", None)
    monkeypatch.setattr(gm, "LIB_PATH", "/nonexistent/libgpu_msm_amd.so")
    with pytest.raises(gm.MsmError):
        gm.lib()


def test_montgomery_helpers_match_oracle(L):
    # fp_to_montgomery_wrapper is one host Montgomery product (v * R^2 R^-1):
    # random values and the edges 0, 1, p - 1, p - 2 and values with every
    # 64-bit word at its maximum below p pin the host product against Python
    rng = np.random.default_rng(11)
    vals = [0, 1, 2, m.P - 1, m.P - 2, (1 << 445), m.P - (1 << 64), (1 << 64) - 1]
    vals += [int.from_bytes(rng.bytes(56), "little") % m.P for _ in range(2000)]
    for v in vals:
        a = o.int_to_limbs(v)
        out = np.zeros(7, dtype=np.uint64)
        L.fp_to_montgomery_wrapper(out.ctypes.data, a.ctypes.data)
        assert o.limbs_to_int(out) == m.to_mont(v), v
    g = o.g1_generator()  # Montgomery
    r = np.zeros(1, dtype=gm.G1_DTYPE)
    L.g1_from_montgomery_wrapper(r.ctypes.data, g.ctypes.data)
    assert (o.limbs_to_int(r[0]["x"]), o.limbs_to_int(r[0]["y"])) == m.G1_STD
    g2 = o.g2_generator()
    r2 = np.zeros(1, dtype=gm.G2_DTYPE)
    L.g2_from_montgomery_wrapper(r2.ctypes.data, g2.ctypes.data)
    assert m.g2_from_abi(r2[0], mont=False) == m.G2_STD
    inf = np.zeros(1, dtype=gm.G1_DTYPE)
    inf["inf"] = 1
    inf["x"][0][0] = 5
    L.g1_from_montgomery_wrapper(r.ctypes.data, inf.ctypes.data)
    assert r[0]["inf"] == 1 and o.limbs_to_int(r[0]["x"]) == 0


def test_projective_affine_helpers(L):
    g = o.g1_generator()
    # a projective representative with Z != 1: the oracle's 5*G in projective form
    gp = np.zeros(1, dtype=o.G1P_DT)
    o.lib().or_g1_affine_to_proj(o.ptr(gp), o.ptr(g))
    acc = gp.copy()
    for _ in range(4):
        o.lib().or_g1_proj_add(o.ptr(acc), o.ptr(acc), o.ptr(gp))
    aff = np.zeros(1, dtype=gm.G1_DTYPE)
    L.projective_to_affine_g1_wrapper(aff.ctypes.data, acc.ctypes.data)
    assert m.g1_from_abi(aff[0]) == m.ec_mul(m.FP_OPS, m.G1_STD, 5)
    back = np.zeros(1, dtype=gm.G1P_DTYPE)
    L.affine_to_projective_g1_wrapper(back.ctypes.data, aff.ctypes.data)
    assert o.limbs_to_int(back[0]["Z"]) == (1 << 448) % m.P
    # infinity conventions
    z = np.zeros(1, dtype=gm.G1P_DTYPE)
    L.projective_to_affine_g1_wrapper(aff.ctypes.data, z.ctypes.data)
    assert aff[0]["inf"] == 1
    inf = np.zeros(1, dtype=gm.G1_DTYPE)
    L.g1_point_at_infinity_wrapper(inf.ctypes.data)
    assert bool(L.g1_is_infinity_wrapper(inf.ctypes.data))
    L.affine_to_projective_g1_wrapper(back.ctypes.data, inf.ctypes.data)
    assert not back.view(np.uint64).any()
    # G2
    g2 = o.g2_generator()
    g2p = np.zeros(1, dtype=o.G2P_DT)
    o.lib().or_g2_affine_to_proj(o.ptr(g2p), o.ptr(g2))
    o.lib().or_g2_proj_double(o.ptr(g2p), o.ptr(g2p))
    a2 = np.zeros(1, dtype=gm.G2_DTYPE)
    L.projective_to_affine_g2_wrapper(a2.ctypes.data, g2p.ctypes.data)
    assert m.g2_from_abi(a2[0]) == m.ec_mul(m.FP2_OPS, m.G2_STD, 2)


def test_msm_wrapper_argument_errors_without_gpu(L):
    # n <= 0 is rejected before any device work (reference c_wrapper.cu:71-73)
    out = np.zeros(1, dtype=gm.G1P_DTYPE)
    pts = np.zeros(1, dtype=gm.G1_DTYPE)
    sc = np.zeros(1, dtype=np.uint64)
    assert L.g1_msm_wrapper(out.ctypes.data, pts.ctypes.data, sc.ctypes.data, 0, 0) == -1
    assert L.g2_msm_wrapper(out.ctypes.data, pts.ctypes.data, sc.ctypes.data, -5, 0) == -1
    # Python mirror: empty input gives infinity without a call
    assert gm.G1Projective.msm([], []).rec.view(np.uint64).sum() == 0
    with pytest.raises(gm.MsmError):
        gm.G1Projective.msm(np.zeros(2, dtype=gm.G1_DTYPE), np.zeros(3, dtype=np.uint64))


def test_plan_query(L):
    c, nw, ws = gm.plan_query(gm.CURVE_G1, 1 << 20, 5, 299)
    assert 8 <= c <= 20 and nw * c >= 300 and ws > 0
    c2, nw2, _ = gm.plan_query(gm.CURVE_G1, 1 << 20, 1, 64, window=4)
    assert (c2, nw2) == (4, 17)
