"""ctypes access to the library's C++ surface (include/curve.h, fp.h, fp2.h,
fp_kernels.h, fp2_kernels.h) by mangled symbol name.

The names come from the built library's dynamic symbol table (nm, mangled and
demangled side by side), so a test calls exactly the symbol a C++ caller
compiled against those headers links to."""
import ctypes
import os
import re
import subprocess

import gpu_msm_amd as gm

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADERS = ("curve.h", "fp.h", "fp2.h", "fp_kernels.h", "fp2_kernels.h")

P = ctypes.c_void_p
U32 = ctypes.c_uint32
U64 = ctypes.c_uint64
I32 = ctypes.c_int


def symbol_table(path: str) -> dict:
    """demangled signature -> mangled symbol, for the library's defined
    dynamic symbols."""
    raw = subprocess.run(["nm", "-D", "--defined-only", path], check=True, capture_output=True,
                         text=True).stdout.splitlines()
    dem = subprocess.run(["nm", "-DC", "--defined-only", path], check=True, capture_output=True,
                         text=True).stdout.splitlines()
    assert len(raw) == len(dem)
    out = {}
    for a, b in zip(raw, dem):
        pa, pb = a.split(None, 2), b.split(None, 2)
        if len(pa) == 3 and len(pb) == 3:
            out[pb[2]] = pa[2]
    return out


def declared_functions() -> list:
    """Function names (with the PointType instantiations for templates) that
    the C++ headers declare."""
    names = []
    for h in HEADERS:
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        txt = re.sub(r"//[^\n]*", "", txt)
        for mm in re.finditer(r"(template\s*<typename PointType>\s*)?(?:void|bool|int|size_t)\s+(\w+)\s*\(", txt):
            if mm.group(1):
                names += [f"{mm.group(2)}<G1Point>", f"{mm.group(2)}<G2Point>"]
            else:
                names.append(mm.group(2))
    return names


class Surface:
    def __init__(self):
        self.lib = gm.lib()
        self.table = symbol_table(gm.LIB_PATH)
        self._cache = {}

    def mangled(self, name: str) -> str:
        pat = re.compile(r"(^|\s)" + re.escape(name) + r"\(")
        hits = [m for d, m in self.table.items() if pat.search(d)]
        if len(hits) != 1:
            raise KeyError(f"{name}: {len(hits)} matching symbols")
        return hits[0]

    def fn(self, name: str, argtypes, restype=None):
        key = (name, tuple(argtypes), restype)
        if key not in self._cache:
            f = getattr(self.lib, self.mangled(name))
            f.argtypes = list(argtypes)
            f.restype = restype
            self._cache[key] = f
        return self._cache[key]
