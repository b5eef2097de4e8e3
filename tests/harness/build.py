"""Compiles the caller harnesses against the built library:
rust_flow (C11 with gcc, on include/rust_wrapper.h), device_api and msm_flow
(C++17 with g++ on include/curve.h and the HIP runtime API headers); and
fold_host, a host build of the field code's one-pass fold (hipcc, no library). Binaries go to
tests/harness/_build/ and find libgpu_msm_amd.so through an $ORIGIN rpath, so
the tree can move.

    python tests/harness/build.py
"""
import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
OUT = os.path.join(HERE, "_build")
INC = os.path.join(ROOT, "include")
LIBDIR = os.path.join(ROOT, "gpu_msm_amd")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

LINK = ["-L", LIBDIR, "-lgpu_msm_amd", "-Wl,-rpath,$ORIGIN/../../../gpu_msm_amd",
        f"-Wl,-rpath-link,{ROCM}/lib"]
COMMANDS = {
    "rust_flow": ["gcc", "-std=c11", "-O1", "-Wall", "-Wextra", "-Werror", "-I", INC,
                  os.path.join(HERE, "rust_flow.c"), "-o", os.path.join(OUT, "rust_flow"), *LINK],
    "device_api": ["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-D__HIP_PLATFORM_AMD__",
                   "-I", INC, "-I", f"{ROCM}/include", os.path.join(HERE, "device_api.cpp"), "-o",
                   os.path.join(OUT, "device_api"), *LINK, "-L", f"{ROCM}/lib", "-lamdhip64"],
    "msm_flow": ["g++", "-std=c++17", "-O1", "-Wall", "-Wextra", "-Werror", "-D__HIP_PLATFORM_AMD__",
                 "-I", INC, "-I", f"{ROCM}/include", os.path.join(HERE, "msm_flow.cpp"), "-o",
                 os.path.join(OUT, "msm_flow"), *LINK, "-L", f"{ROCM}/lib", "-lamdhip64"],
}
COMMANDS["fold_host"] = [os.path.join(ROCM, "bin", "hipcc"), "-std=c++17", "-O1", "--offload-arch=gfx950",
                         "-I", os.path.join(ROOT, "gpu_msm_amd", "csrc"), "-I", INC,
                         os.path.join(HERE, "fold_host.cpp"), "-o", os.path.join(OUT, "fold_host")]
SOURCES = {"rust_flow": "rust_flow.c", "device_api": "device_api.cpp", "msm_flow": "msm_flow.cpp",
           "fold_host": "fold_host.cpp"}


def build(name: str) -> str:
    """Path of the harness binary, rebuilt when missing or older than its
    source, its headers or the library."""
    exe = os.path.join(OUT, name)
    deps = [os.path.join(HERE, SOURCES[name]), os.path.join(LIBDIR, "libgpu_msm_amd.so")]
    deps += [os.path.join(INC, h) for h in os.listdir(INC)]
    if name == "fold_host":  # the field code itself, not the library
        deps = [os.path.join(HERE, SOURCES[name]), os.path.join(ROOT, "gpu_msm_amd", "csrc", "field.hpp")]
    if os.path.exists(exe) and os.path.getmtime(exe) >= max(os.path.getmtime(d) for d in deps):
        return exe
    os.makedirs(OUT, exist_ok=True)
    r = subprocess.run(COMMANDS[name], capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"building {name} failed:\n{r.stderr}")
    return exe


if __name__ == "__main__":
    for n in COMMANDS:
        print(build(n))
