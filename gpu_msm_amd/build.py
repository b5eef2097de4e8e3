"""Builds gpu_msm_amd/libgpu_msm_amd.so for gfx950 with hipcc (in-tree).

Usage: python -m gpu_msm_amd.build   (or gpu_msm_amd.build.build())
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgpu_msm_amd.so")
SOURCES = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("GPU_MSM_ARCH", "gfx950")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _deps(path: str, seen: set | None = None) -> set:
    """Quoted #include dependencies of a source, recursively."""
    import re

    seen = set() if seen is None else seen
    for m in re.finditer(r'^\s*#include\s+"([^"]+)"', open(path).read(), flags=re.M):
        dep = os.path.normpath(os.path.join(os.path.dirname(path), m.group(1)))
        if dep not in seen and os.path.exists(dep):
            seen.add(dep)
            _deps(dep, seen)
    return seen


def _inputs_mtime(src_path: str) -> float:
    return max(os.path.getmtime(p) for p in {src_path, *_deps(src_path)})


def _compile(src: str, incremental: bool = False) -> str:
    obj = os.path.join(BUILD, src.replace(".hip", ".o"))
    if incremental and os.path.exists(obj):
        if os.path.getmtime(obj) > _inputs_mtime(os.path.join(CSRC, src)):
            return obj
    cmd = [HIPCC, *FLAGS, "-I", os.path.join(os.path.dirname(HERE), "include"), "-c",
           os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed on {src}")
    return obj


def build(verbose: bool = True, incremental: bool = False) -> str:
    """Compile every HIP source for gfx950 and link the shared library.
    incremental=True skips sources whose object is newer than all inputs."""
    os.makedirs(BUILD, exist_ok=True)
    workers = min(len(SOURCES), max(1, os.cpu_count() or 1))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        objs = list(ex.map(lambda s: _compile(s, incremental), SOURCES))
    tmp = LIB + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("link failed")
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


if __name__ == "__main__":
    build(incremental="--incremental" in sys.argv)
