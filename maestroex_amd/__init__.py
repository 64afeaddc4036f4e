"""MI355X-native advection hot path of MAESTROeX.

The compute lives in libmaestroex_amd.so (hand-written HIP kernels for
gfx950 behind the C-ABI of include/mx_adv.h).  This package loads it, mirrors
the reference's Maestro member functions in Python (`Maestro`), and builds
synthetic workloads.  There is no CPU fallback: `load()` raises if the
library or a GPU is missing.
"""
import ctypes as C
import os

from . import _abi
from .driver import Maestro
from .fabs import Backend, BaseState, FaceMultiFab, Geometry, MultiFab, make_params, split_domain

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libmaestroex_amd.so")
_LIB = None


def _torch_first():
    """Initialise torch's HIP runtime before this library's.  torch ships
    its own HIP runtime; when libmaestroex_amd.so had initialised the device
    first, torch in the same process reported no GPU.  Python callers that
    pair the library with torch tensors therefore start torch first."""
    try:
        import torch
    except ImportError:
        return
    torch.cuda.is_available()


def load():
    """Load libmaestroex_amd.so and bind the C-ABI; raises if it is missing."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is not built; run `python -m maestroex_amd.build` "
                "(the advection path has no CPU fallback)")
        _torch_first()
        lib = C.CDLL(LIB_PATH)
        fns = _abi.bind(lib, "mx_adv_", extra=_abi.LIB_ONLY)
        _LIB = (lib, fns)
    return _LIB


def exported_symbols():
    """Names the C-ABI header declares (used by the export test)."""
    return ["mx_adv_" + n for n in list(_abi.SIGNATURES) + list(_abi.LIB_ONLY)]


def gpu_maestro(params, levels, stream=None):
    """A `Maestro` bound to the HIP library on the current device."""
    lib, fns = load()
    q = _abi.Params()
    C.memmove(C.byref(q), C.byref(params), C.sizeof(_abi.Params))
    m = Maestro(fns, q, levels, name="mx_adv", keepalive=lib)
    if stream is not None:
        rc = fns["set_stream"](m.ctx, C.c_void_p(stream))
        if rc != _abi.MX_OK:
            raise _abi.MxError(rc, "set_stream")
    return m


def device_backend(device="cuda:0"):
    import torch
    if not torch.cuda.is_available():
        raise RuntimeError("no GPU visible: the advection path runs only on the GPU")
    return Backend("torch", device)


__all__ = [
    "Backend", "BaseState", "FaceMultiFab", "Geometry", "Maestro", "MultiFab", "make_params",
    "split_domain", "load", "gpu_maestro", "device_backend", "exported_symbols", "LIB_PATH",
]
