"""Device buffers and a stream for the GPU tests, straight from the HIP
runtime (libamdhip64 through ctypes): no framework in between, so a GPU test
that cannot reach the runtime or a device fails instead of skipping."""
from __future__ import annotations

import ctypes
import os

import numpy as np

_H2D, _D2H = 1, 2
_rt = None


def runtime() -> ctypes.CDLL:
    global _rt
    if _rt is None:
        rocm = os.environ.get("ROCM_PATH", "/opt/rocm")
        for name in ("libamdhip64.so", os.path.join(rocm, "lib", "libamdhip64.so")):
            try:
                _rt = ctypes.CDLL(name)
                break
            except OSError:
                continue
        if _rt is None:
            raise RuntimeError("libamdhip64.so not found: the HIP runtime is required for GPU tests")
        vp, sz = ctypes.c_void_p, ctypes.c_size_t
        for fn, args in (("hipMalloc", [ctypes.POINTER(vp), sz]), ("hipFree", [vp]),
                         ("hipMemcpy", [vp, vp, sz, ctypes.c_int]), ("hipMemset", [vp, ctypes.c_int, sz]),
                         ("hipStreamCreate", [ctypes.POINTER(vp)]), ("hipStreamSynchronize", [vp]),
                         ("hipDeviceSynchronize", []), ("hipGetDeviceCount", [ctypes.POINTER(ctypes.c_int)]),
                         ("hipSetDevice", [ctypes.c_int])):
            f = getattr(_rt, fn)
            f.argtypes = args
            f.restype = ctypes.c_int
    return _rt


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


def device_count() -> int:
    n = ctypes.c_int(0)
    check(runtime().hipGetDeviceCount(ctypes.byref(n)), "hipGetDeviceCount")
    return n.value


def require_gpu() -> None:
    """Raises (the test fails) unless a HIP device is visible."""
    if device_count() < 1:
        raise RuntimeError("no HIP device visible")
    check(runtime().hipSetDevice(0), "hipSetDevice")


class DevBuf:
    """A hipMalloc'd buffer, freed when the object goes."""

    def __init__(self, nbytes: int):
        self.nbytes = max(int(nbytes), 8)
        p = ctypes.c_void_p()
        check(runtime().hipMalloc(ctypes.byref(p), self.nbytes), f"hipMalloc({self.nbytes})")
        self.ptr = p.value

    def data_ptr(self) -> int:
        return self.ptr

    def __del__(self):
        if getattr(self, "ptr", None) and _rt is not None:
            _rt.hipFree(self.ptr)
            self.ptr = None


_stream = None


def stream() -> int:
    """One stream shared by the tests of a process."""
    global _stream
    if _stream is None:
        s = ctypes.c_void_p()
        check(runtime().hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        _stream = s.value
    return _stream


def to_dev(arr: np.ndarray) -> DevBuf:
    a = np.ascontiguousarray(arr)
    b = DevBuf(a.nbytes)
    if a.nbytes:
        check(runtime().hipMemcpy(b.ptr, a.ctypes.data, a.nbytes, _H2D), "hipMemcpy H2D")
    return b


def empty_dev(nbytes: int, fill: int = 0) -> DevBuf:
    b = DevBuf(nbytes)
    check(runtime().hipMemset(b.ptr, fill, b.nbytes), "hipMemset")
    return b


def to_host(b: DevBuf, dtype, nbytes: int | None = None) -> np.ndarray:
    """Waits for the shared stream and the device, then copies back."""
    check(runtime().hipStreamSynchronize(stream()), "hipStreamSynchronize")
    check(runtime().hipDeviceSynchronize(), "hipDeviceSynchronize")
    n = b.nbytes if nbytes is None else nbytes
    out = np.zeros(n, dtype=np.uint8)
    check(runtime().hipMemcpy(out.ctypes.data, b.ptr, n, _D2H), "hipMemcpy D2H")
    return out.view(dtype)
