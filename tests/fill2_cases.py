"""Two-level cases for FillPatch above level 0 (shared by the GPU tests):
level 0 of 16^3 in two boxes, level 1 (ratio 2) with boxes side by side, on
the low z side and in the high x / low y / high z corner, periodic or with
every non-periodic boundary rule; valid cells random, ghost cells a
sentinel so that every cell the fill leaves is compared too."""
import numpy as np

from maestroex_amd import _abi
from maestroex_amd.fabs import Geometry, MultiFab, make_params

N0 = 16
B0 = [((0, 0, 0), (7, 15, 15)), ((8, 0, 0), (15, 15, 15))]
B1 = [((8, 8, 8), (15, 23, 15)), ((16, 8, 8), (23, 15, 15)), ((4, 24, 0), (11, 31, 9)),
      ((24, 0, 20), (31, 7, 31))]
SENTINEL = 1.0e300

P = _abi
CASES = {
    "periodic": dict(periodic=True, bcs=None, vel=0),
    # per component (lo, hi) BCRecs: every rule, one-sided slopes (hoextrap)
    "walls": dict(periodic=False, vel=0, bcs=[
        [(P.BC_FOEXTRAP, P.BC_HOEXTRAP, P.BC_REFLECT_EVEN),
         (P.BC_HOEXTRAP, P.BC_FOEXTRAP, P.BC_HOEXTRAP)],
        [(P.BC_REFLECT_ODD, P.BC_REFLECT_EVEN, P.BC_FOEXTRAP),
         (P.BC_REFLECT_EVEN, P.BC_REFLECT_ODD, P.BC_HOEXTRAP)],
        [(P.BC_HOEXTRAP, P.BC_HOEXTRAP, P.BC_HOEXTRAP),
         (P.BC_HOEXTRAP, P.BC_HOEXTRAP, P.BC_HOEXTRAP)]]),
    # velocity: ext_dir sides are zeroed, one-sided slopes next to them
    "velocity_ext_dir": dict(periodic=False, vel=1, bcs=[
        [(P.BC_EXT_DIR, P.BC_FOEXTRAP, P.BC_EXT_DIR), (P.BC_EXT_DIR, P.BC_HOEXTRAP, P.BC_FOEXTRAP)],
        [(P.BC_FOEXTRAP, P.BC_EXT_DIR, P.BC_REFLECT_ODD), (P.BC_HOEXTRAP, P.BC_EXT_DIR, P.BC_EXT_DIR)],
        [(P.BC_REFLECT_EVEN, P.BC_HOEXTRAP, P.BC_EXT_DIR), (P.BC_EXT_DIR, P.BC_FOEXTRAP, P.BC_EXT_DIR)]]),
}


def params(periodic):
    bc = _abi.PHYS_INTERIOR if periodic else _abi.PHYS_OUTFLOW
    per = [1, 1, 1] if periodic else [0, 0, 0]
    g0 = Geometry(3, [N0] * 3, [1.0 / N0] * 3, per, [bc] * 3, [bc] * 3)
    g1 = Geometry(3, [2 * N0] * 3, [0.5 / N0] * 3, per, [bc] * 3, [bc] * 3)
    return make_params(g0, nspec=1, extra_levels=[g1])


def bcs_for(case, ncomp):
    if case["bcs"] is None:
        return [[(P.BC_INT_DIR,) * 3, (P.BC_INT_DIR,) * 3]] * ncomp
    return [case["bcs"][n % len(case["bcs"])] for n in range(ncomp)]


def make(ncomp, ng, seed, backend):
    """(crse, fine) numpy-backed MultiFabs: random valid cells, sentinel
    ghosts; `backend` is unused (the callers copy to the device)."""
    from maestroex_amd.fabs import Backend
    NP = Backend("numpy")
    rng = np.random.default_rng(seed)
    crse = MultiFab(B0, ncomp, 1, 3, NP, fill=SENTINEL)
    fine = MultiFab(B1, ncomp, ng, 3, NP, fill=SENTINEL)
    for mf in (crse, fine):
        for b in range(len(mf.arrays)):
            reg = (slice(None),) + mf.valid_region(b)
            mf.arrays[b][reg] = rng.normal(size=mf.arrays[b][reg].shape)
    return crse, fine
