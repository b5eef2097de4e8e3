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
    the C++ headers declare, once per declaration: an overloaded name (the
    host and the stream form of fp_add, the G1 and G2 forms of
    projective_point_add) appears once per overload."""
    names = []
    ret = r"(?:const\s+\w+\s*&|void|bool|int|size_t|uint64_t)"
    for h in HEADERS:
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        txt = re.sub(r"//[^\n]*", "", txt)
        txt = re.sub(r'extern\s+"C"[^;]*;', "", txt)  # C symbols (rust_wrapper.h)
        for mm in re.finditer(r"(template\s*<typename PointType>\s*)?" + ret + r"\s+(\w+)\s*\(", txt):
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

    def matches(self, name: str) -> list:
        """(demangled, mangled) of every exported overload of `name`."""
        pat = re.compile(r"(^|\s)" + re.escape(name) + r"\(")
        return [(d, m) for d, m in self.table.items() if pat.search(d)]

    def mangled(self, name: str, contains: str | None = None) -> str:
        """The one symbol of `name` whose demangled signature holds
        `contains`; without it, an overloaded name resolves to its stream
        form (the overload taking ihipStream_t*)."""
        hits = self.matches(name)
        if contains is not None:
            hits = [h for h in hits if contains in h[0]]
        elif len(hits) > 1:
            hits = [h for h in hits if "ihipStream_t" in h[0]]
        if len(hits) != 1:
            raise KeyError(f"{name}: {len(hits)} matching symbols")
        return hits[0][1]

    def fn(self, name: str, argtypes, restype=None, contains: str | None = None):
        key = (name, tuple(argtypes), restype, contains)
        if key not in self._cache:
            f = getattr(self.lib, self.mangled(name, contains))
            f.argtypes = list(argtypes)
            f.restype = restype
            self._cache[key] = f
        return self._cache[key]
