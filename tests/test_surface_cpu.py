"""The C++ surface and the caller harnesses without a GPU: every function the
C++ headers declare is exported under the C++ name a caller links to, both
harnesses compile and link (rust_flow as strict C11 against rust_wrapper.h,
device_api as C++17 against curve.h), argument checks abort with the
reference's PANIC behaviour before any device work, and the Rust-layer flow
handles empty input the way types.rs does."""
import os
import subprocess
import sys

import numpy as np
import pytest

import gpu_msm_amd as gm
from cxx_surface import Surface, declared_functions
from oracle import oracle as o

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), "harness"))
import build as harness  # noqa: E402


@pytest.fixture(scope="module")
def S():
    if not os.path.exists(gm.LIB_PATH):
        pytest.skip("library not built (python -m gpu_msm_amd.build)")
    return Surface()


def test_every_declared_cxx_function_is_exported(S):
    """Every declaration of the C++ headers has its own exported symbol: an
    overloaded name needs as many symbols as it has declarations."""
    from collections import Counter

    names = Counter(declared_functions())
    assert sum(names.values()) >= 190
    missing = {n: (k, len(S.matches(n))) for n, k in names.items() if len(S.matches(n)) < k}
    assert not missing, missing


def test_reference_host_surface_names_are_declared():
    """The host-callable functions of the reference's include/fp.h:34-148,
    include/fp2.h:20-89 and include/curve.h:50-134, name for name."""
    want = """fp_cmp fp_is_zero fp_is_one fp_zero fp_one fp_one_montgomery fp_copy fp_add_raw fp_sub_raw fp_add
    fp_sub fp_mul_raw fp_reduce fp_mont_reduce fp_mont_mul fp_to_montgomery fp_from_montgomery
    fp_to_montgomery_batch fp_from_montgomery_batch fp_mul fp_neg fp_inv fp_mont_inv fp_div fp_pow fp_pow_u64
    fp_sqrt fp_is_quadratic_residue fp_cmov fp_modulus
    fp2_cmp fp2_is_zero fp2_is_one fp2_zero fp2_one fp2_copy fp2_add fp2_sub fp2_mul fp2_mont_mul fp2_square
    fp2_neg fp2_conjugate fp2_inv fp2_mont_inv fp2_div fp2_cmov fp2_frobenius fp2_mul_by_i
    fp2_mul_by_non_residue
    curve_b_g1 curve_b_g2 is_on_curve_g1 is_on_curve_g2 g1_point_at_infinity g2_point_at_infinity
    g1_is_infinity g2_is_infinity g1_generator g2_generator g1_projective_point_at_infinity
    g2_projective_point_at_infinity affine_to_projective projective_to_affine_g1 projective_to_affine_g2
    projective_point_add projective_point_double""".split()
    declared = set(declared_functions())
    assert not [n for n in want if n not in declared]
    for t in ("point_add", "point_double", "point_neg", "point_scalar_mul"):
        assert f"{t}<G1Point>" in declared and f"{t}<G2Point>" in declared


def test_templates_are_the_point_type_instantiations(S):
    """C++ callers link to point_msm<G1Point> etc. (reference
    src/curve.cu:2833-2875 instantiates both PointTypes)."""
    for t in ("G1Point", "G2Point"):
        sym = S.mangled(f"point_msm<{t}>")
        assert sym.startswith("_Z9point_msmI") and f"I7{t}E" in sym


@pytest.mark.parametrize("name", ["rust_flow", "device_api", "msm_flow"])
def test_harness_compiles_and_links(S, name):
    exe = harness.build(name)
    assert os.access(exe, os.X_OK)


PANIC_SNIPPET = r"""
import sys
sys.path.insert(0, {tests!r})
from cxx_surface import Surface, P, U32, I32
s = Surface()
{call}
print("returned")
"""


@pytest.mark.parametrize("call,message", [
    ("s.fn('point_add_async<G1Point>', [P, U32, P, P, P])(None, 0, None, None, None)",
     "point_add_async: null pointer argument"),
    ("s.fn('point_msm_u64_g1', [P, U32, P, P, P, P, I32])(None, 0, 1, 1, 1, None, -1)",
     "point_msm_u64_g1: invalid size n"),
    ("s.fn('fp_add_array_device', [P, U32, P, P, P, I32])(None, 0, 8, 8, 8, -3)",
     "fp_add_array_device: invalid size n"),
    ("s.fn('point_to_montgomery_batch<G2Point>', [P, U32, P, I32])(None, 0, None, 5)",
     "point_to_montgomery_batch: null pointer argument"),
])
def test_argument_checks_abort_like_reference_panic(S, call, message):
    """reference include/device.h:25-39: PANIC_IF_FALSE prints and abort()s.
    The checks run before any HIP call, so no device is needed."""
    tests = os.path.dirname(os.path.abspath(__file__))
    code = PANIC_SNIPPET.format(tests=tests, call=call)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == -6, (r.returncode, r.stdout, r.stderr)
    assert message in r.stderr
    assert "returned" not in r.stdout


def test_rust_flow_empty_input_is_infinity(S, tmp_path):
    """types.rs:351-353: an empty MSM is the point at infinity, with no FFI
    MSM call; the harness then normalises through the C ABI helpers."""
    exe = harness.build("rust_flow")
    for curve, dt in (("g1", o.G1_DT), ("g2", o.G2_DT)):
        pts, sc, out = tmp_path / "p.bin", tmp_path / "s.bin", tmp_path / "o.bin"
        pts.write_bytes(b"")
        sc.write_bytes(b"")
        r = subprocess.run([exe, curve, "0", str(pts), str(sc), str(out)], capture_output=True, text=True,
                           timeout=60)
        assert r.returncode == 0, r.stderr
        got = np.fromfile(out, dtype=dt)
        assert got["inf"][0] == 1 and not got["x"].any() and not got["y"].any()
