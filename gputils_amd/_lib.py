"""ctypes binding of libgputils_hip.so.

The argument and return types of every entry point are read from
include/gputils_hip.h, so the binding always matches the C ABI.  Loading
fails loudly (ImportError) when the shared library has not been built:
there is no fallback implementation.
"""
from __future__ import annotations

import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "gputils_hip.h")
LIB_PATH = os.path.join(ROOT, "gputils_amd", "lib", "libgputils_hip.so")

_CTYPES = {
    "int64_t": ctypes.c_int64,
    "int": ctypes.c_int,
    "double": ctypes.c_double,
    "float": ctypes.c_float,
    "size_t": ctypes.c_size_t,
    "unsigned char": ctypes.c_ubyte,
    "gputils_gels_plan_t": ctypes.c_void_p,
    "gputils_graph_t": ctypes.c_void_p,
    "void": None,
}

_DECL = re.compile(r"^\s*(?P<ret>(?:const\s+)?[\w ]+?\s*\*?)\s*(?P<name>gputils_\w+)\s*\((?P<args>[^)]*)\)\s*;",
                   re.MULTILINE | re.DOTALL)


def _ctype(decl: str):
    d = " ".join(decl.replace("const", " ").split())
    if "*" in d:
        base = d.replace("*", "").strip()
        if base == "char":
            return ctypes.c_char_p
        return ctypes.c_void_p
    # strip the parameter name
    parts = d.split()
    typ = " ".join(parts[:-1]) if len(parts) > 1 and parts[-1] not in ("int", "char") else d
    if typ not in _CTYPES:
        raise ValueError(f"unmapped C type in {decl!r}")
    return _CTYPES[typ]


def parse_header(path: str = HEADER) -> dict:
    """{name: (restype, [argtypes])} for every gputils_* declaration."""
    text = open(path).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.DOTALL)
    out = {}
    for m in _DECL.finditer(text):
        ret = m.group("ret").strip()
        args = [a.strip() for a in m.group("args").split(",") if a.strip()]
        if args == ["void"]:
            args = []
        restype = ctypes.c_char_p if "char" in ret and "*" in ret else _ctype(ret + " r")
        out[m.group("name")] = (restype, [_ctype(a) for a in args])
    return out


_lib = None


def lib():
    """Load libgputils_hip.so and declare every entry point of the header."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m gputils_amd.build` "
                              "(there is no fallback implementation)")
        L = ctypes.CDLL(LIB_PATH)
        for name, (restype, argtypes) in parse_header().items():
            fn = getattr(L, name)
            fn.restype = restype
            fn.argtypes = argtypes
        _lib = L
    return _lib


class GputilsError(RuntimeError):
    pass


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().gputils_last_error().decode(errors="replace")
        raise GputilsError(f"{what or 'gputils call'} failed with status {rc}: {msg}")
