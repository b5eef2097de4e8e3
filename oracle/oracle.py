"""ctypes loader for the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py use this as the checker. The product package
``gpu_msm_amd`` never imports it.

Arrays use the reference ABI layouts (reference include/curve.h:13-43):
G1Point 120 B, G2Point 232 B, G1ProjectivePoint 168 B, G2ProjectivePoint 336 B.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

G1_DT = np.dtype([("x", "<u8", 7), ("y", "<u8", 7), ("inf", "u1"), ("pad", "u1", 7)])
G2_DT = np.dtype([("x", "<u8", 14), ("y", "<u8", 14), ("inf", "u1"), ("pad", "u1", 7)])
G1P_DT = np.dtype([("X", "<u8", 7), ("Y", "<u8", 7), ("Z", "<u8", 7)])
G2P_DT = np.dtype([("X", "<u8", 14), ("Y", "<u8", 14), ("Z", "<u8", 14)])
assert G1_DT.itemsize == 120 and G2_DT.itemsize == 232
assert G1P_DT.itemsize == 168 and G2P_DT.itemsize == 336

P = 0x3CDEE0FB28C5E535200FC34965AAD6400095A4B78A02FE320F75A64BBAC71602824E6DC3E23ACDEE56EE4528C573B5CC311C0026AAB0AAAB
R_ORDER = 0x511B70539F27995B34995830FA4D04C98CCC4C050BC7BB9B0E8D8CA34610428001400040001


def build() -> str:
    """Compile liboracle.so with gcc (oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        vp, i, u64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint64
        sig = {
            "or_fp_cmp": ([vp, vp], i),
            "or_fp_add": ([vp, vp, vp], None),
            "or_fp_sub": ([vp, vp, vp], None),
            "or_fp_neg": ([vp, vp], None),
            "or_fp_mont_mul": ([vp, vp, vp], None),
            "or_fp_to_mont": ([vp, vp], None),
            "or_fp_from_mont": ([vp, vp], None),
            "or_fp_mont_inv": ([vp, vp], None),
            "or_fp_one_mont": ([vp], None),
            "or_fp2_add": ([vp, vp, vp], None),
            "or_fp2_sub": ([vp, vp, vp], None),
            "or_fp2_mont_mul": ([vp, vp, vp], None),
            "or_fp2_mont_inv": ([vp, vp], None),
            "or_g1_generator": ([vp], None),
            "or_g2_generator": ([vp], None),
            "or_g1_to_mont": ([vp, vp], None),
            "or_g2_to_mont": ([vp, vp], None),
            "or_g1_from_mont": ([vp, vp], None),
            "or_g2_from_mont": ([vp, vp], None),
            "or_g1_affine_to_proj": ([vp, vp], None),
            "or_g2_affine_to_proj": ([vp, vp], None),
            "or_g1_proj_to_affine": ([vp, vp], None),
            "or_g2_proj_to_affine": ([vp, vp], None),
            "or_g1_proj_add": ([vp, vp, vp], None),
            "or_g2_proj_add": ([vp, vp, vp], None),
            "or_g1_proj_double": ([vp, vp], None),
            "or_g2_proj_double": ([vp, vp], None),
            "or_g1_is_on_curve": ([vp], ctypes.c_bool),
            "or_g2_is_on_curve": ([vp], ctypes.c_bool),
            "or_g1_scalar_mul": ([vp, vp, vp, i], None),
            "or_g2_scalar_mul": ([vp, vp, vp, i], None),
            "or_g1_bases_from_k": ([vp, vp, i, i], None),
            "or_g2_bases_from_k": ([vp, vp, i, i], None),
            "or_g1_msm": ([vp, vp, vp, i, i, i, i], None),
            "or_g2_msm": ([vp, vp, vp, i, i, i, i], None),
            "or_gen_u64": ([u64, i, vp], None),
            "or_gen_scalars_modr": ([u64, i, vp], None),
            "or_scalar_order": ([vp], None),
        }
        for name, (args, res) in sig.items():
            fn = getattr(_lib, name)
            fn.argtypes = args
            fn.restype = res
    return _lib


def ptr(a: np.ndarray) -> int:
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data


# ---------------------------------------------------------------- conversions

def int_to_limbs(v: int, n: int = 7) -> np.ndarray:
    return np.array([(v >> (64 * k)) & 0xFFFFFFFFFFFFFFFF for k in range(n)], dtype=np.uint64)


def limbs_to_int(a) -> int:
    return sum(int(x) << (64 * k) for k, x in enumerate(np.asarray(a, dtype=np.uint64).ravel()))


def fp(v: int) -> np.ndarray:
    return int_to_limbs(v % P)


# ------------------------------------------------------------------ wrappers

def fp_binop(name: str, a: np.ndarray, b: np.ndarray) -> np.ndarray:
    out = np.zeros(a.shape[-1], dtype=np.uint64)
    getattr(lib(), name)(ptr(out), ptr(np.ascontiguousarray(a)), ptr(np.ascontiguousarray(b)))
    return out


def fp_unop(name: str, a: np.ndarray) -> np.ndarray:
    out = np.zeros(a.shape[-1], dtype=np.uint64)
    getattr(lib(), name)(ptr(out), ptr(np.ascontiguousarray(a)))
    return out


def g1_generator(mont: bool = True) -> np.ndarray:
    g = np.zeros(1, dtype=G1_DT)
    lib().or_g1_generator(ptr(g))
    if mont:
        lib().or_g1_to_mont(ptr(g), ptr(g))
    return g


def g2_generator(mont: bool = True) -> np.ndarray:
    g = np.zeros(1, dtype=G2_DT)
    lib().or_g2_generator(ptr(g))
    if mont:
        lib().or_g2_to_mont(ptr(g), ptr(g))
    return g


def gen_u64(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, dtype=np.uint64)
    lib().or_gen_u64(seed, n, ptr(out))
    return out


def gen_scalars_modr(seed: int, n: int) -> np.ndarray:
    out = np.zeros((n, 5), dtype=np.uint64)
    lib().or_gen_scalars_modr(seed, n, ptr(out))
    return out


def bases_from_k(curve: str, k: np.ndarray, threads: int = 8) -> np.ndarray:
    k = np.ascontiguousarray(k, dtype=np.uint64)
    dt = G1_DT if curve == "g1" else G2_DT
    out = np.zeros(len(k), dtype=dt)
    getattr(lib(), f"or_{curve}_bases_from_k")(ptr(out), ptr(k), len(k), threads)
    return out


def msm(curve: str, bases: np.ndarray, scalars: np.ndarray, c: int = 0, threads: int = 8) -> np.ndarray:
    """Host Pippenger MSM; scalars shaped (n, limbs). Returns one affine point (Montgomery)."""
    scalars = np.ascontiguousarray(scalars, dtype=np.uint64)
    if scalars.ndim == 1:
        scalars = scalars.reshape(-1, 1)
    n, limbs = scalars.shape
    dt = G1_DT if curve == "g1" else G2_DT
    out = np.zeros(1, dtype=dt)
    getattr(lib(), f"or_{curve}_msm")(ptr(out), ptr(np.ascontiguousarray(bases)), ptr(scalars), limbs, n, c, threads)
    return out


def scalar_mul(curve: str, base: np.ndarray, k: int) -> np.ndarray:
    limbs = max(1, (k.bit_length() + 63) // 64)
    kk = int_to_limbs(k, limbs)
    dt = G1_DT if curve == "g1" else G2_DT
    out = np.zeros(1, dtype=dt)
    getattr(lib(), f"or_{curve}_scalar_mul")(ptr(out), ptr(np.ascontiguousarray(base)), ptr(kk), limbs)
    return out


def to_affine(curve: str, proj: np.ndarray) -> np.ndarray:
    dt = G1_DT if curve == "g1" else G2_DT
    out = np.zeros(1, dtype=dt)
    getattr(lib(), f"or_{curve}_proj_to_affine")(ptr(out), ptr(np.ascontiguousarray(proj)))
    return out


def from_mont(curve: str, pt: np.ndarray) -> np.ndarray:
    out = np.zeros_like(pt)
    getattr(lib(), f"or_{curve}_from_mont")(ptr(out), ptr(np.ascontiguousarray(pt)))
    return out
