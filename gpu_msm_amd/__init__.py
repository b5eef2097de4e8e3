"""gpu_msm_amd — BLS12-446 G1/G2 multi-scalar multiplication on MI355X (gfx950).

Python host mirror of the reference's Rust-facing API (reference
rust-api/src/types.rs: G1Affine, G1Projective::msm, G2Affine,
G2Projective::msm) over the library's C ABI (include/rust_wrapper.h,
include/gpu_msm_ext.h), loaded with ctypes from the in-tree
``libgpu_msm_amd.so``. There is no CPU fallback: if the HIP library is not
built, every call raises.

Arrays of points use the ABI record layouts as numpy structured dtypes
(``G1_DTYPE`` 120 B, ``G2_DTYPE`` 232 B, ``G1P_DTYPE`` 168 B, ``G2P_DTYPE`` 336 B).
"""
from __future__ import annotations

import ctypes
import os
from typing import Sequence

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgpu_msm_amd.so")

G1_DTYPE = np.dtype([("x", "<u8", 7), ("y", "<u8", 7), ("inf", "u1"), ("pad", "u1", 7)])
G2_DTYPE = np.dtype([("x", "<u8", 14), ("y", "<u8", 14), ("inf", "u1"), ("pad", "u1", 7)])
G1P_DTYPE = np.dtype([("X", "<u8", 7), ("Y", "<u8", 7), ("Z", "<u8", 7)])
G2P_DTYPE = np.dtype([("X", "<u8", 14), ("Y", "<u8", 14), ("Z", "<u8", 14)])

CURVE_G1 = 1
CURVE_G2 = 2
POINTS_MONTGOMERY = 0x1
SCALARS_LT_R = 0x2
BATCH_GRAPH = 0x1
BATCH_DEVICE_COMBINE = 0x2
R_ORDER = 0x511B70539F27995B34995830FA4D04C98CCC4C050BC7BB9B0E8D8CA34610428001400040001
R_BITS = 299

ERRORS = {
    -1: "invalid size or argument",
    -2: "device or stream setup failed",
    -3: "device allocation failed",
    -4: "HIP error during conversion or MSM",
    -5: "scalar exceeds the declared bit bound",
}

# Exported C symbols (include/rust_wrapper.h + include/gpu_msm_ext.h).
EXPORTS = [
    "affine_to_projective_g1_wrapper", "affine_to_projective_g2_wrapper",
    "projective_to_affine_g1_wrapper", "projective_to_affine_g2_wrapper",
    "g1_point_at_infinity_wrapper", "g2_point_at_infinity_wrapper",
    "g1_projective_point_at_infinity_wrapper", "g2_projective_point_at_infinity_wrapper",
    "g1_is_infinity_wrapper", "g2_is_infinity_wrapper",
    "g1_msm_wrapper", "g2_msm_wrapper",
    "g1_from_montgomery_wrapper", "g2_from_montgomery_wrapper", "fp_to_montgomery_wrapper",
    "init_device_generators", "g1_msm_multi_wrapper", "g2_msm_multi_wrapper", "gm_plan_query",
    "gm_bases_upload", "gm_bases_from_k", "gm_bases_download", "gm_bases_count", "gm_bases_free",
    "gm_scalars_upload", "gm_scalars_free", "gm_msm_resident", "gm_msm_batch",
    "gm_batch_create", "gm_batch_launch", "gm_batch_wait", "gm_batch_run", "gm_batch_count",
    "gm_batch_workspace_bytes", "gm_batch_free", "gm_mlp_forward", "gm_mlp_grad", "gm_mlp_train",
    "gm_server_create", "gm_server_submit", "gm_server_wait", "gm_server_get_stats", "gm_server_inject_failures",
    "gm_server_destroy",
    "gm_set_profiling", "gm_profile_last", "gm_last_msm_info",
    "gm_point_encoded_size", "gm_points_serialize", "gm_points_deserialize", "gm_points_check",
    "gm_points_to_canonical", "gm_points_from_canonical", "gm_generator_multiple",
    "gm_point_msm_noalloc", "gm_field_ops_check",
]


class MsmError(RuntimeError):
    pass


_lib = None


def lib() -> ctypes.CDLL:
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MsmError(f"HIP library not built: {LIB_PATH} (run python -m gpu_msm_amd.build)")
    L = ctypes.CDLL(LIB_PATH)
    vp, i, u32, sz = ctypes.c_void_p, ctypes.c_int, ctypes.c_uint32, ctypes.c_size_t
    sig = {
        "affine_to_projective_g1_wrapper": ([vp, vp], None),
        "affine_to_projective_g2_wrapper": ([vp, vp], None),
        "projective_to_affine_g1_wrapper": ([vp, vp], None),
        "projective_to_affine_g2_wrapper": ([vp, vp], None),
        "g1_point_at_infinity_wrapper": ([vp], None),
        "g2_point_at_infinity_wrapper": ([vp], None),
        "g1_projective_point_at_infinity_wrapper": ([vp], None),
        "g2_projective_point_at_infinity_wrapper": ([vp], None),
        "g1_is_infinity_wrapper": ([vp], ctypes.c_bool),
        "g2_is_infinity_wrapper": ([vp], ctypes.c_bool),
        "g1_msm_wrapper": ([vp, vp, vp, i, u32], i),
        "g2_msm_wrapper": ([vp, vp, vp, i, u32], i),
        "g1_from_montgomery_wrapper": ([vp, vp], None),
        "g2_from_montgomery_wrapper": ([vp, vp], None),
        "fp_to_montgomery_wrapper": ([vp, vp], None),
        "init_device_generators": ([vp, u32], None),
        "g1_msm_multi_wrapper": ([vp, vp, vp, i, i, u32, u32], i),
        "g2_msm_multi_wrapper": ([vp, vp, vp, i, i, u32, u32], i),
        "gm_plan_query": ([i, i, i, i, i, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(sz)], i),
        "gm_bases_upload": ([ctypes.POINTER(vp), i, vp, i, i, u32], i),
        "gm_bases_from_k": ([ctypes.POINTER(vp), i, vp, i, u32], i),
        "gm_bases_download": ([vp, vp, i, i], i),
        "gm_bases_count": ([vp], i),
        "gm_bases_free": ([vp], None),
        "gm_scalars_upload": ([ctypes.POINTER(vp), vp, i, i, u32], i),
        "gm_scalars_free": ([vp], None),
        "gm_msm_resident": ([vp, vp, i, vp, i, i, i, i, i], i),
        "gm_msm_batch": ([vp, vp, vp, vp, i, i, i], i),
        "gm_batch_create": ([ctypes.POINTER(vp), vp, vp, vp, i, i, i], i),
        "gm_batch_launch": ([vp], i),
        "gm_batch_wait": ([vp, vp], i),
        "gm_batch_run": ([vp, vp], i),
        "gm_batch_count": ([vp], i),
        "gm_batch_workspace_bytes": ([vp], sz),
        "gm_batch_free": ([vp], None),
        "gm_mlp_forward": ([vp, vp, i, vp, u32], i),
        "gm_mlp_grad": ([vp, vp, vp, i, vp, vp, u32], i),
        "gm_mlp_train": ([vp, vp, vp, i, i, ctypes.c_float, vp, u32], i),
        "gm_server_create": ([ctypes.POINTER(vp), vp, vp, vp], i),
        "gm_server_submit": ([vp, i, i, i, vp, i, i, u32, ctypes.POINTER(ctypes.c_uint64)], i),
        "gm_server_wait": ([vp, ctypes.c_uint64, vp, ctypes.POINTER(ctypes.c_double), i], i),
        "gm_server_get_stats": ([vp, vp], i),
        "gm_server_inject_failures": ([vp, i], i),
        "gm_server_destroy": ([vp], None),
        "gm_set_profiling": ([u32, i], i),
        "gm_profile_last": ([u32, ctypes.POINTER(ctypes.c_float), i], i),
        "gm_last_msm_info": ([u32, ctypes.POINTER(i), ctypes.POINTER(i), ctypes.POINTER(ctypes.c_longlong)], i),
        "gm_point_encoded_size": ([i, i], i),
        "gm_points_serialize": ([i, vp, vp, i, i, u32], i),
        "gm_points_deserialize": ([i, vp, vp, i, i, i, vp, u32], i),
        "gm_points_check": ([i, vp, i, i, vp, u32], i),
        "gm_points_to_canonical": ([i, vp, vp, i, u32], i),
        "gm_points_from_canonical": ([i, vp, vp, i, u32], i),
        "gm_generator_multiple": ([i, vp, i, vp], i),
        "gm_field_ops_check": ([i, i, i, vp, vp, vp, vp, vp, i, vp], i),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def _ptr(a: np.ndarray) -> int:
    if not a.flags["C_CONTIGUOUS"]:
        raise ValueError("array must be C-contiguous")
    return a.ctypes.data


def _check(rc: int, what: str) -> None:
    if rc != 0:
        raise MsmError(f"{what} failed with error code {rc}: {ERRORS.get(rc, 'unknown')}")


def int_to_limbs(v: int, n: int = 7) -> np.ndarray:
    return np.array([(v >> (64 * k)) & 0xFFFFFFFFFFFFFFFF for k in range(n)], dtype=np.uint64)


def limbs_to_int(a) -> int:
    return sum(int(x) << (64 * k) for k, x in enumerate(np.asarray(a, dtype=np.uint64).ravel()))


def scalars_to_limbs(scalars: Sequence[int], limbs: int = 5) -> np.ndarray:
    """Python ints -> (n, limbs) little-endian u64 array."""
    out = np.zeros((len(scalars), limbs), dtype=np.uint64)
    for i, s in enumerate(scalars):
        out[i] = int_to_limbs(int(s), limbs)
    return out


# ------------------------------------------------------------ typed points

class _Affine:
    DTYPE: np.dtype
    CURVE: int

    def __init__(self, rec: np.ndarray):
        self.rec = np.ascontiguousarray(rec, dtype=self.DTYPE).reshape(1)

    @classmethod
    def infinity(cls):
        r = np.zeros(1, dtype=cls.DTYPE)
        r["inf"] = 1
        return cls(r)

    def is_infinity(self) -> bool:
        fn = "g1_is_infinity_wrapper" if self.CURVE == CURVE_G1 else "g2_is_infinity_wrapper"
        return bool(getattr(lib(), fn)(_ptr(self.rec)))

    def __eq__(self, other) -> bool:
        return isinstance(other, type(self)) and self.rec.tobytes() == other.rec.tobytes()

    def __repr__(self) -> str:
        return f"{type(self).__name__}({'inf' if self.rec['inf'][0] else hex(limbs_to_int(self.rec['x'][0]))})"


class _Projective:
    DTYPE: np.dtype
    AFF: type
    CURVE: int

    def __init__(self, rec: np.ndarray):
        self.rec = np.ascontiguousarray(rec, dtype=self.DTYPE).reshape(1)

    @classmethod
    def infinity(cls):
        return cls(np.zeros(1, dtype=cls.DTYPE))

    def to_affine(self):
        """reference types.rs:331-341 -> projective_to_affine_*_wrapper."""
        out = np.zeros(1, dtype=self.AFF.DTYPE)
        fn = "projective_to_affine_g1_wrapper" if self.CURVE == CURVE_G1 else "projective_to_affine_g2_wrapper"
        getattr(lib(), fn)(_ptr(out), _ptr(self.rec))
        return self.AFF(out)

    @classmethod
    def msm(cls, points, scalars, gpu_index: int = 0):
        """sum_i scalars[i] * points[i] with one u64 scalar per point
        (reference types.rs:345-376 / 492-532): G1 points in standard form,
        G2 points in Montgomery form; empty input gives infinity."""
        pts = _points_array(points, cls.AFF.DTYPE)
        sc = np.ascontiguousarray(np.asarray(scalars, dtype=np.uint64))
        if len(pts) != len(sc):
            raise MsmError(f"Points and scalars must have the same length: {len(pts)} != {len(sc)}")
        if len(pts) == 0:
            return cls.infinity()
        out = np.zeros(1, dtype=cls.DTYPE)
        fn = "g1_msm_wrapper" if cls.CURVE == CURVE_G1 else "g2_msm_wrapper"
        _check(getattr(lib(), fn)(_ptr(out), _ptr(pts), _ptr(sc), len(pts), gpu_index), "MSM computation")
        return cls(out)

    @classmethod
    def msm_multi(cls, points, scalars: np.ndarray, gpu_index: int = 0, montgomery: bool = False,
                  scalars_lt_r: bool = False):
        """Multi-limb scalars shaped (n, limbs) (reference point_msm<T>,
        src/curve.cu:2763-2826) with an explicit input form."""
        pts = _points_array(points, cls.AFF.DTYPE)
        sc = np.ascontiguousarray(np.asarray(scalars, dtype=np.uint64))
        if sc.ndim == 1:
            sc = sc.reshape(-1, 1)
        if len(pts) != len(sc):
            raise MsmError(f"Points and scalars must have the same length: {len(pts)} != {len(sc)}")
        if len(pts) == 0:
            return cls.infinity()
        flags = (POINTS_MONTGOMERY if montgomery else 0) | (SCALARS_LT_R if scalars_lt_r else 0)
        out = np.zeros(1, dtype=cls.DTYPE)
        fn = "g1_msm_multi_wrapper" if cls.CURVE == CURVE_G1 else "g2_msm_multi_wrapper"
        _check(getattr(lib(), fn)(_ptr(out), _ptr(pts), _ptr(sc), sc.shape[1], len(pts), gpu_index, flags),
               "MSM computation")
        return cls(out)


def _points_array(points, dtype) -> np.ndarray:
    if isinstance(points, np.ndarray):
        return np.ascontiguousarray(points, dtype=dtype)
    return np.ascontiguousarray(np.concatenate([p.rec for p in points]) if len(points) else np.zeros(0, dtype))


class G1Affine(_Affine):
    DTYPE = G1_DTYPE
    CURVE = CURVE_G1


class G2Affine(_Affine):
    DTYPE = G2_DTYPE
    CURVE = CURVE_G2


class G1Projective(_Projective):
    DTYPE = G1P_DTYPE
    AFF = G1Affine
    CURVE = CURVE_G1


class G2Projective(_Projective):
    DTYPE = G2P_DTYPE
    AFF = G2Affine
    CURVE = CURVE_G2


def from_montgomery(curve: int, pts: np.ndarray) -> np.ndarray:
    """reference c_wrapper.cu:261-283, elementwise over an array of records."""
    out = np.zeros_like(pts)
    fn = "g1_from_montgomery_wrapper" if curve == CURVE_G1 else "g2_from_montgomery_wrapper"
    f = getattr(lib(), fn)
    size = pts.dtype.itemsize
    for k in range(len(pts)):
        f(out.ctypes.data + k * size, pts.ctypes.data + k * size)
    return out


# ---------------------------------------------------------- point formats
# Batch conversions on the GPU (include/gpu_msm_ext.h, "point formats"): the
# arkworks CanonicalSerialize layout of tfhe-zk-pok points (reference
# rust-api/src/conversions.rs), validation, and canonical limbs.

CHECK_ON_CURVE = 0x1
CHECK_SUBGROUP = 0x2
STATUS = {0: "ok", 1: "invalid encoding", 2: "not on the curve", 3: "not in the prime-order subgroup"}


def _curve_dtype(curve: int):
    return G1_DTYPE if curve == CURVE_G1 else G2_DTYPE


def encoded_size(curve: int, compressed: bool = True) -> int:
    return lib().gm_point_encoded_size(curve, int(compressed))


def serialize_points(curve: int, points: np.ndarray, compressed: bool = True, gpu_index: int = 0) -> np.ndarray:
    """ABI affine records (Montgomery) -> (n, size) uint8 encodings."""
    pts = np.ascontiguousarray(points, dtype=_curve_dtype(curve))
    size = encoded_size(curve, compressed)
    out = np.zeros((len(pts), size), dtype=np.uint8)
    rc = lib().gm_points_serialize(curve, _ptr(out), _ptr(pts), len(pts), int(compressed), gpu_index)
    _check(rc, "point serialization")
    return out


def deserialize_points(curve: int, data: np.ndarray, compressed: bool = True,
                       checks: int = CHECK_ON_CURVE | CHECK_SUBGROUP, gpu_index: int = 0):
    """(n, size) uint8 encodings -> (ABI records in Montgomery form, int32
    status per point). Invalid points come back as the point at infinity."""
    size = encoded_size(curve, compressed)
    raw = np.ascontiguousarray(data, dtype=np.uint8).reshape(-1, size)
    out = np.zeros(len(raw), dtype=_curve_dtype(curve))
    status = np.zeros(len(raw), dtype=np.int32)
    rc = lib().gm_points_deserialize(curve, _ptr(out), _ptr(raw), len(raw), int(compressed), checks, _ptr(status),
                                     gpu_index)
    if rc < 0:
        _check(rc, "point deserialization")
    return out, status


def check_points(curve: int, points: np.ndarray, checks: int = CHECK_ON_CURVE | CHECK_SUBGROUP,
                 gpu_index: int = 0) -> np.ndarray:
    pts = np.ascontiguousarray(points, dtype=_curve_dtype(curve))
    status = np.zeros(len(pts), dtype=np.int32)
    rc = lib().gm_points_check(curve, _ptr(pts), len(pts), checks, _ptr(status), gpu_index)
    if rc < 0:
        _check(rc, "point check")
    return status


def to_canonical(curve: int, points: np.ndarray, gpu_index: int = 0) -> np.ndarray:
    pts = np.ascontiguousarray(points, dtype=_curve_dtype(curve))
    out = np.zeros_like(pts)
    _check(lib().gm_points_to_canonical(curve, _ptr(out), _ptr(pts), len(pts), gpu_index), "to canonical")
    return out


def from_canonical(curve: int, points: np.ndarray, gpu_index: int = 0) -> np.ndarray:
    pts = np.ascontiguousarray(points, dtype=_curve_dtype(curve))
    out = np.zeros_like(pts)
    _check(lib().gm_points_from_canonical(curve, _ptr(out), _ptr(pts), len(pts), gpu_index), "from canonical")
    return out


# ---------------------------------------------------------- resident data

class Bases:
    """Affine bases resident in HBM in the engine's format."""

    def __init__(self, handle: int, curve: int, n: int):
        self.h, self.curve, self.n = handle, curve, n

    @classmethod
    def upload(cls, curve: int, points: np.ndarray, montgomery: bool = True, gpu_index: int = 0) -> "Bases":
        pts = np.ascontiguousarray(points, dtype=G1_DTYPE if curve == CURVE_G1 else G2_DTYPE)
        h = ctypes.c_void_p()
        _check(lib().gm_bases_upload(ctypes.byref(h), curve, _ptr(pts), len(pts), int(montgomery), gpu_index),
               "bases upload")
        return cls(h.value, curve, len(pts))

    @classmethod
    def from_k(cls, curve: int, k: np.ndarray, gpu_index: int = 0) -> "Bases":
        """P_i = k_i * G computed on the device."""
        k = np.ascontiguousarray(k, dtype=np.uint64)
        h = ctypes.c_void_p()
        _check(lib().gm_bases_from_k(ctypes.byref(h), curve, _ptr(k), len(k), gpu_index), "bases generation")
        return cls(h.value, curve, len(k))

    def download(self, offset: int = 0, n: int | None = None) -> np.ndarray:
        n = self.n - offset if n is None else n
        out = np.zeros(n, dtype=G1_DTYPE if self.curve == CURVE_G1 else G2_DTYPE)
        _check(lib().gm_bases_download(self.h, _ptr(out), offset, n), "bases download")
        return out

    def free(self) -> None:
        if self.h:
            lib().gm_bases_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Scalars:
    """Scalars resident in HBM, (n, limbs) little-endian u64."""

    def __init__(self, handle: int, n: int, limbs: int):
        self.h, self.n, self.limbs = handle, n, limbs

    @classmethod
    def upload(cls, scalars: np.ndarray, gpu_index: int = 0) -> "Scalars":
        sc = np.ascontiguousarray(scalars, dtype=np.uint64)
        if sc.ndim == 1:
            sc = sc.reshape(-1, 1)
        h = ctypes.c_void_p()
        _check(lib().gm_scalars_upload(ctypes.byref(h), _ptr(sc), sc.shape[0], sc.shape[1], gpu_index),
               "scalars upload")
        return cls(h.value, sc.shape[0], sc.shape[1])

    def free(self) -> None:
        if self.h:
            lib().gm_scalars_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def msm_resident(bases: Bases, scalars: Scalars, n: int | None = None, base_offset: int = 0, scalar_offset: int = 0,
                 scalar_bits: int = 0, window: int = 0, device_combine: bool = False) -> np.ndarray:
    """MSM over resident data; returns one affine ABI record (Montgomery form)."""
    n = min(bases.n - base_offset, scalars.n - scalar_offset) if n is None else n
    out = np.zeros(1, dtype=G1_DTYPE if bases.curve == CURVE_G1 else G2_DTYPE)
    _check(lib().gm_msm_resident(_ptr(out), bases.h, base_offset, scalars.h, scalar_offset, n, scalar_bits, window,
                                 int(device_combine)), "resident MSM")
    return out


SPEC_DTYPE = np.dtype([("n", "<i4"), ("base_offset", "<i4"), ("scalar_offset", "<i4"), ("window", "<i4"),
                       ("seg", "<i4")])


def msm_batch(bases: Bases, scalars: Scalars, specs, scalar_bits: int = 0, device_combine: bool = False) -> np.ndarray:
    """Many MSMs in one pipeline run. specs: array of SPEC_DTYPE or sequence of
    (n, base_offset, scalar_offset[, window[, seg]]). Returns affine ABI records."""
    sp = _spec_array(specs)
    out = np.zeros(len(sp), dtype=G1_DTYPE if bases.curve == CURVE_G1 else G2_DTYPE)
    _check(lib().gm_msm_batch(_ptr(out), bases.h, scalars.h, _ptr(sp), len(sp), scalar_bits, int(device_combine)),
           "batched MSM")
    return out


def _spec_array(specs) -> np.ndarray:
    if isinstance(specs, np.ndarray) and specs.dtype == SPEC_DTYPE:
        return np.ascontiguousarray(specs)
    sp = np.zeros(len(specs), dtype=SPEC_DTYPE)
    for k, s in enumerate(specs):
        sp[k] = tuple(s) + (0,) * (5 - len(s))
    return sp


class Batch:
    """A fixed batch of MSMs over resident data, optionally captured in a
    hipGraph once and replayed on every run."""

    def __init__(self, bases: Bases, scalars: Scalars, specs, scalar_bits: int = 0, use_graph: bool = True,
                 device_combine: bool = False):
        self.specs = _spec_array(specs)
        self.bases, self.scalars = bases, scalars  # keep the handles alive
        self.curve = bases.curve
        h = ctypes.c_void_p()
        flags = (BATCH_GRAPH if use_graph else 0) | (BATCH_DEVICE_COMBINE if device_combine else 0)
        _check(lib().gm_batch_create(ctypes.byref(h), bases.h, scalars.h, _ptr(self.specs), len(self.specs),
                                     scalar_bits, flags), "batch create")
        self.h = h.value
        self.out = np.zeros(len(self.specs), dtype=G1_DTYPE if self.curve == CURVE_G1 else G2_DTYPE)

    def run(self) -> np.ndarray:
        _check(lib().gm_batch_run(self.h, _ptr(self.out)), "batch run")
        return self.out.copy()

    def launch(self) -> None:
        _check(lib().gm_batch_launch(self.h), "batch launch")

    def wait(self) -> np.ndarray:
        _check(lib().gm_batch_wait(self.h, _ptr(self.out)), "batch wait")
        return self.out.copy()

    @property
    def workspace_bytes(self) -> int:
        return int(lib().gm_batch_workspace_bytes(self.h))

    def free(self) -> None:
        if getattr(self, "h", None):
            lib().gm_batch_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def generator_multiple(curve: int, k: int) -> np.ndarray:
    """k * G for the curve's generator, computed on the host (one affine ABI
    record, Montgomery form). With bases P_i = k_i G an MSM equals
    (sum_i s_i k_i mod r) G, so this is an independent host check of a GPU
    result (workloads.closed_form_scalar gives the exponent)."""
    kk = int_to_limbs(k % R_ORDER, 5)
    out = np.zeros(1, dtype=_curve_dtype(curve))
    _check(lib().gm_generator_multiple(curve, _ptr(kk), 5, _ptr(out)), "gm_generator_multiple")
    return out


def plan_query(curve: int, n: int, limbs: int, bits: int = 0, window: int = 0):
    c, nw, ws = ctypes.c_int(), ctypes.c_int(), ctypes.c_size_t()
    _check(lib().gm_plan_query(curve, n, limbs, bits, window, ctypes.byref(c), ctypes.byref(nw), ctypes.byref(ws)),
           "plan query")
    return c.value, nw.value, ws.value


def set_profiling(on: bool, gpu_index: int = 0) -> None:
    _check(lib().gm_set_profiling(gpu_index, int(on)), "set profiling")


STAGES = ["hist", "scan", "scatter", "accumulate", "fixup", "bucket_reduce", "window_sum"]


def profile_last(gpu_index: int = 0) -> dict:
    buf = (ctypes.c_float * len(STAGES))()
    _check(lib().gm_profile_last(gpu_index, buf, len(STAGES)), "profile read")
    return dict(zip(STAGES, list(buf)))


def last_msm_info(gpu_index: int = 0) -> dict:
    c, nw, e = ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong()
    _check(lib().gm_last_msm_info(gpu_index, ctypes.byref(c), ctypes.byref(nw), ctypes.byref(e)), "msm info")
    return {"window": c.value, "windows": nw.value, "entries": e.value}


# ------------------------------------------------------------------ server

class _ServerConfig(ctypes.Structure):
    _fields_ = [("gpu_index", ctypes.c_uint32), ("slots", ctypes.c_int), ("max_batch", ctypes.c_int),
                ("hbm_budget_bytes", ctypes.c_size_t), ("planner", ctypes.c_int),
                ("mlp_params", ctypes.c_void_p), ("small_max_n", ctypes.c_int)]


class _ServerStats(ctypes.Structure):
    _fields_ = [("submitted", ctypes.c_uint64), ("completed", ctypes.c_uint64), ("failed", ctypes.c_uint64),
                ("batches", ctypes.c_uint64), ("mean_batch", ctypes.c_double), ("p50_ms", ctypes.c_double),
                ("p99_ms", ctypes.c_double), ("max_ms", ctypes.c_double), ("peak_inflight_bytes", ctypes.c_size_t),
                ("peak_concurrent_batches", ctypes.c_int)]


PLANNER_HEURISTIC = 0
PLANNER_MLP = 1
SUBMIT_BORROW = 0x1


class Server:
    """Continuous-batching MSM request loop over resident G1/G2 base pools
    (gm_server_* in the library)."""

    def __init__(self, g1_pool: Bases | None, g2_pool: Bases | None, slots: int = 3, max_batch: int = 64,
                 hbm_budget_bytes: int = 32 << 30, planner: int = PLANNER_HEURISTIC, mlp_params=None,
                 gpu_index: int = 0, small_max_n: int = 0):
        """small_max_n: requests of at most this many points form their own
        batches, go first and keep one slot to themselves (0: 2^16, < 0: no
        small class)."""
        self.pools = (g1_pool, g2_pool)
        self._mlp = None if mlp_params is None else np.ascontiguousarray(mlp_params, dtype=np.float32)
        cfg = _ServerConfig(gpu_index, slots, max_batch, hbm_budget_bytes, planner,
                            None if self._mlp is None else self._mlp.ctypes.data, small_max_n)
        h = ctypes.c_void_p()
        _check(lib().gm_server_create(ctypes.byref(h), ctypes.byref(cfg), g1_pool.h if g1_pool else None,
                                      g2_pool.h if g2_pool else None), "server create")
        self.h = h.value
        self._borrowed = {}

    def submit(self, curve: int, base_offset: int, scalars: np.ndarray, scalar_bits: int = 0,
               borrow: bool = False) -> int:
        sc = np.ascontiguousarray(scalars, dtype=np.uint64)
        if sc.ndim == 1:
            sc = sc.reshape(-1, 1)
        t = ctypes.c_uint64()
        _check(lib().gm_server_submit(self.h, curve, base_offset, sc.shape[0], _ptr(sc), sc.shape[1], scalar_bits,
                                      SUBMIT_BORROW if borrow else 0, ctypes.byref(t)), "submit")
        if borrow:
            self._borrowed[t.value] = sc  # keep alive until waited
        return t.value

    def wait(self, ticket: int, curve: int, timeout_ms: int = -1):
        """Returns (rc, affine record or None, latency ms)."""
        out = np.zeros(1, dtype=G1_DTYPE if curve == CURVE_G1 else G2_DTYPE)
        lat = ctypes.c_double()
        rc = lib().gm_server_wait(self.h, ticket, _ptr(out), ctypes.byref(lat), timeout_ms)
        if rc != -6:
            self._borrowed.pop(ticket, None)
        return rc, (out if rc == 0 else None), lat.value

    def stats(self) -> dict:
        st = _ServerStats()
        _check(lib().gm_server_get_stats(self.h, ctypes.byref(st)), "server stats")
        return {k: getattr(st, k) for k, _ in _ServerStats._fields_}

    def inject_failures(self, batches: int) -> None:
        _check(lib().gm_server_inject_failures(self.h, batches), "inject failures")

    def close(self) -> None:
        if getattr(self, "h", None):
            lib().gm_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
