"""Loader of the in-tree gfx950 library ``mitrace_amd/lib/libmitrace_gpu.so``.

There is no fallback: if the library is missing the import of the bindings fails loudly.
"""
import ctypes as C
import os

from . import abi

LIB_PATH = os.environ.get("MITRACE_LIB_PATH") or os.path.join(
    os.path.dirname(os.path.abspath(__file__)), "lib", "libmitrace_gpu.so")

_lib = None


class MiTraceError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"mitrace_gpu error {code}: {msg}")
        self.code = code


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `python -m mitrace_amd.build` "
                          "(there is no CPU fallback)")
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in abi.SIGNATURES.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    if L.mt_abi_version() != 3:
        raise ImportError("libmitrace_gpu ABI version mismatch")
    _lib = L
    return L


def check(rc):
    if rc != 0:
        msg = load().mt_last_error()
        raise MiTraceError(rc, msg.decode(errors="replace") if msg else "")
    return rc
