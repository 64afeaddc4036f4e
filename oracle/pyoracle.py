"""ctypes binding of the CPU oracle (test infrastructure only).

Only tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg import
this module.  It loads oracle/_build/liboracle.so (bounds-checked) or
liboracle_fast.so (the timed CPU baseline) and wraps it in the same
`Maestro` driver the HIP library uses, so one call sequence runs on both.
"""
import ctypes as C
import os

from maestroex_amd import _abi
from maestroex_amd.driver import Maestro

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBS = {}


def lib_path(fast=False):
    return os.path.join(_HERE, "_build", "liboracle_fast.so" if fast else "liboracle.so")


def load(fast=False):
    key = bool(fast)
    if key not in _LIBS:
        path = lib_path(fast)
        if not os.path.exists(path):
            raise FileNotFoundError(f"oracle library not built: {path} (run make -C oracle)")
        lib = C.CDLL(path)
        fns = _abi.bind(lib, "mxo_")
        lib.mxo_omp_threads.argtypes = [C.c_int]
        lib.mxo_omp_threads.restype = C.c_int
        fns["omp_threads"] = lib.mxo_omp_threads
        _LIBS[key] = (lib, fns)
    return _LIBS[key]


def copy_params(p):
    q = _abi.Params()
    C.memmove(C.byref(q), C.byref(p), C.sizeof(_abi.Params))
    return q


def oracle_maestro(params, levels, fast=False):
    lib, fns = load(fast)
    return Maestro(fns, copy_params(params), levels, name="oracle", keepalive=lib)


def set_threads(n, fast=True):
    lib, fns = load(fast)
    return fns["omp_threads"](n)
