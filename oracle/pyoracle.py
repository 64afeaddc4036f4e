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
        fns = _abi.bind(lib, "mxo_", names=[n for n in _abi.SIGNATURES if n not in _abi.NO_ORACLE])
        lv = lib.mxo_advance_pass_levels
        lv.argtypes = _abi.LIB_ONLY["advance_pass_levels"]
        lv.restype = C.c_int
        fns["advance_pass_levels"] = lv
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


def _dptr(a):
    import numpy as np
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(C.c_double))


def step_op_count(s, u, dt, dx, rel_eps):
    """FP64 operations per cell of the differentiable step for one component
    s (N, N, N), counted by the oracle's op-counting instantiation: a dict of
    the phases (ppm, predictor, corner, edge, flux_update) and the cost of one
    van Leer slope and one PPM type 1 face value."""
    import numpy as np
    lib, _ = load(False)
    s = np.ascontiguousarray(s, dtype=np.float64)
    u = [np.ascontiguousarray(x, dtype=np.float64) for x in u]
    out = np.zeros(7)
    f = lib.mxo_step_op_count
    f.restype = C.c_int
    rc = f(s.shape[0], C.c_double(dt), (C.c_double * 3)(*dx), C.c_double(rel_eps), _dptr(s),
           *[_dptr(x) for x in u], _dptr(out))
    if rc:
        raise _abi.MxError(rc, "oracle.step_op_count")
    keys = ["ppm", "predictor", "corner", "edge", "flux_update", "slope", "face"]
    return dict(zip(keys, out.tolist()))


def scal_step_bc(s, u, dt, dx, rel_eps, periodic, ds=None, du=None, ppm_type=1, fast=False,
                 conservative=False):
    """The differentiable scalar step (oracle_step.cpp) on the cells of a
    domain whose direction d is periodic (periodic[d]) or outflow (foextrap),
    ppm_type 1 or 2, MakeEdgeScal in the advective or (conservative) the
    conservative form.
    s: (ncomp, nz, ny, nx) indexed [c, k, j, i]; u: three face arrays, u[d]
    one point longer along d when d is outflow (faces 0..n[d]).  Returns
    snew, or (snew, dsnew) when the direction ds (and optionally du) is
    given."""
    import numpy as np
    lib, _ = load(fast)
    s = np.ascontiguousarray(s, dtype=np.float64)
    ncomp = s.shape[0]
    n = [s.shape[3], s.shape[2], s.shape[1]]
    per = [1 if p else 0 for p in periodic]
    u = [np.ascontiguousarray(x, dtype=np.float64) for x in u]
    for d in range(3):
        want = [n[2], n[1], n[0]]
        if not per[d]:
            want[2 - d] += 1
        if list(u[d].shape) != want:
            raise ValueError(f"u[{d}] has shape {u[d].shape}, expected {tuple(want)}")
    null = C.POINTER(C.c_double)()
    snew = np.empty_like(s)
    dsnew = None
    dsp = null
    dup = [null, null, null]
    keep = []
    if ds is not None:
        ds = np.ascontiguousarray(ds, dtype=np.float64)
        keep.append(ds)
        dsp = _dptr(ds)
        dsnew = np.empty_like(s)
        for d in range(3):
            if du is not None and du[d] is not None:
                a = np.ascontiguousarray(du[d], dtype=np.float64)
                keep.append(a)
                dup[d] = _dptr(a)
    f = lib.mxo_scal_step_bc
    f.restype = C.c_int
    rc = f((C.c_int * 3)(*n), (C.c_int * 3)(*per), ncomp, ppm_type, C.c_double(dt),
           (C.c_double * 3)(*dx), C.c_double(rel_eps), _dptr(s), *[_dptr(x) for x in u],
           _dptr(snew), dsp, *dup, _dptr(dsnew) if dsnew is not None else null,
           1 if conservative else 0)
    if rc:
        raise _abi.MxError(rc, "oracle.scal_step_bc")
    return snew if ds is None else (snew, dsnew)


def scal_step(s, u, dt, dx, rel_eps, ds=None, du=None, ppm_type=1, fast=False):
    """The differentiable scalar step on a periodic N^3 box (oracle_step.cpp),
    ppm_type 1 or 2.
    s: (ncomp, N, N, N) indexed [c, k, j, i]; u: three (N, N, N) face arrays
    (face index = the cell above it; face N is face 0).  Returns snew, or
    (snew, dsnew) when the direction ds (and optionally du) is given."""
    import numpy as np
    lib, _ = load(fast)
    s = np.ascontiguousarray(s, dtype=np.float64)
    u = [np.ascontiguousarray(x, dtype=np.float64) for x in u]
    ncomp, n = s.shape[0], s.shape[1]
    dxa = (C.c_double * 3)(*dx)
    snew = np.empty_like(s)
    if ds is None:
        f = lib.mxo_scal_step
        f.restype = C.c_int
        rc = f(n, ncomp, ppm_type, C.c_double(dt), dxa, C.c_double(rel_eps), _dptr(s),
               *[_dptr(x) for x in u], _dptr(snew))
        if rc:
            raise _abi.MxError(rc, "oracle.scal_step")
        return snew
    ds = np.ascontiguousarray(ds, dtype=np.float64)
    dun = [None if (du is None or du[d] is None) else np.ascontiguousarray(du[d], dtype=np.float64)
           for d in range(3)]
    dsnew = np.empty_like(s)
    f = lib.mxo_scal_step_jvp
    f.restype = C.c_int
    null = C.POINTER(C.c_double)()
    rc = f(n, ncomp, ppm_type, C.c_double(dt), dxa, C.c_double(rel_eps), _dptr(s),
           *[_dptr(x) for x in u], _dptr(ds), *[(_dptr(x) if x is not None else null) for x in dun],
           _dptr(snew), _dptr(dsnew))
    if rc:
        raise _abi.MxError(rc, "oracle.scal_step_jvp")
    return snew, dsnew
