"""Level-0 FillPatch on the GPU against the oracle, bit for bit: random
valid data and sentinel ghosts, 2-D and 3-D, periodic and bounded
directions, every BCType the fill applies (foextrap, hoextrap, reflect_even,
reflect_odd, ext_dir for velocity and Pi), uneven boxes, two bounded
directions meeting at corners, a copy from another MultiFab with component
offsets, and the error returns."""
import numpy as np
import pytest

import pyoracle
from fill_cases import (EXT_DIR, FOEXTRAP, HOEXTRAP, INT_DIR, REFLECT_EVEN, REFLECT_ODD,
                        layout, new_mf, random_field)
from maestroex_amd import _abi
from maestroex_amd.fabs import Backend

pytestmark = pytest.mark.gpu
NP = Backend("numpy")

CASES = {
    # name: dim, n, max_grid, periodic, ng, ncomp, variable_type, BCRec per component
    "3d_outflow_z": (3, 16, 8, (1, 1, 0), 4, 3, 0,
                     [((INT_DIR, INT_DIR, FOEXTRAP), (INT_DIR, INT_DIR, FOEXTRAP))] * 3),
    "3d_hoextrap_z_uneven": (3, 12, (5, 7, 4), (1, 1, 0), 3, 2, 0,
                             [((INT_DIR, INT_DIR, HOEXTRAP), (INT_DIR, INT_DIR, HOEXTRAP))] * 2),
    "3d_walls_vel": (3, 12, 6, (0, 1, 0), 3, 3, 1,
                     [((REFLECT_ODD, INT_DIR, EXT_DIR), (REFLECT_ODD, INT_DIR, FOEXTRAP)),
                      ((REFLECT_EVEN, INT_DIR, EXT_DIR), (REFLECT_EVEN, INT_DIR, FOEXTRAP)),
                      ((REFLECT_EVEN, INT_DIR, EXT_DIR), (REFLECT_EVEN, INT_DIR, HOEXTRAP))]),
    "2d_corners": (2, 16, 8, (0, 0), 4, 4, 0,
                   [((FOEXTRAP, HOEXTRAP, 0), (HOEXTRAP, FOEXTRAP, 0)),
                    ((REFLECT_EVEN, REFLECT_ODD, 0), (REFLECT_ODD, REFLECT_EVEN, 0)),
                    ((HOEXTRAP, REFLECT_EVEN, 0), (FOEXTRAP, REFLECT_ODD, 0)),
                    ((INT_DIR, FOEXTRAP, 0), (FOEXTRAP, INT_DIR, 0))]),
    "2d_periodic_x": (2, 16, (8, 4), (1, 0), 3, 2, 1,
                      [((INT_DIR, EXT_DIR, 0), (INT_DIR, REFLECT_ODD, 0))] * 2),
}


def run(case, seed, src_offsets=None):
    import maestroex_amd as mx
    dim, n, mg, per, ng, nc, vt, rec = CASES[case]
    params, boxes = layout(dim, n=n, max_grid=mg, periodic=per)
    outs = []
    for be, mk in ((NP, pyoracle.oracle_maestro), (mx.device_backend(), mx.gpu_maestro)):
        m = mk(params, [boxes])
        if src_offsets is None:
            mf = new_mf(boxes, nc, ng, dim, be, random_field(seed))
            m.FillPatch(0, mf, mf, 0, 0, nc, 0, rec, vt)
        else:
            scomp, dcomp = src_offsets
            src = new_mf(boxes, nc + scomp, 0, dim, be, random_field(seed))
            mf = new_mf(boxes, nc + dcomp, ng, dim, be, random_field(seed + 1))
            m.FillPatch(0, mf, src, scomp, dcomp, nc, 0, rec, vt)
        m.sync()
        outs.append(mf.to_numpy())
    return outs


@pytest.mark.parametrize("case", sorted(CASES))
def test_fill_patch_parity(case):
    print(f"[parity] fill_patch {case}", flush=True)
    ref, got = run(case, 3)
    for b, (x, y) in enumerate(zip(ref, got)):
        assert np.array_equal(x, y), f"{case}: FAB {b}"


def test_fill_patch_from_other_multifab():
    print("[parity] fill_patch src != mf", flush=True)
    ref, got = run("3d_outflow_z", 4, src_offsets=(1, 2))
    for b, (x, y) in enumerate(zip(ref, got)):
        assert np.array_equal(x, y), f"FAB {b}"


def test_fill_patch_errors():
    import maestroex_amd as mx
    print("[errors] fill_patch", flush=True)
    params, boxes = layout(3, n=8, max_grid=4, periodic=(1, 1, 0))
    m = mx.gpu_maestro(params, [boxes])
    mf = new_mf(boxes, params.nscal, 2, 3, mx.device_backend(), random_field(1))
    before = mf.to_numpy()
    rec = [((INT_DIR, INT_DIR, EXT_DIR), (INT_DIR, INT_DIR, FOEXTRAP))]
    with pytest.raises(_abi.MxError) as e:  # a scalar other than Pi with ext_dir
        m.FillPatch(0, mf, mf, 0, 0, 1, 0, rec, 0)
    assert e.value.code == _abi.MX_ERR_BC
    with pytest.raises(_abi.MxError) as e:  # an unknown BCType
        m.FillPatch(0, mf, mf, 0, 0, 1, 0, [((0, 0, 9), (0, 0, 2))], 0)
    assert e.value.code == _abi.MX_ERR_BC
    with pytest.raises(_abi.MxError) as e:  # one MultiFab, overlapping component ranges
        m.FillPatch(0, mf, mf, 0, 1, 2, 0, rec * 2, 1)
    assert e.value.code == _abi.MX_ERR_ARG
    m.sync()
    for x, y in zip(before, mf.to_numpy()):
        assert np.array_equal(x, y)


def test_fill_patch_plan_reuse():
    """Later calls on one context reuse the level's copy plan with other
    MultiFabs: each call still matches the oracle"""
    import maestroex_amd as mx
    print("[parity] fill_patch plan reuse", flush=True)
    dim, n, mg, per, ng, nc, vt, rec = CASES["3d_walls_vel"]
    params, boxes = layout(dim, n=n, max_grid=mg, periodic=per)
    mo = pyoracle.oracle_maestro(params, [boxes])
    mg_ = mx.gpu_maestro(params, [boxes])
    for seed in (5, 6, 7):
        outs = []
        for be, m in ((NP, mo), (mx.device_backend(), mg_)):
            mf = new_mf(boxes, nc, ng, dim, be, random_field(seed))
            m.FillPatch(0, mf, mf, 0, 0, nc, 0, rec, vt)
            m.sync()
            outs.append(mf.to_numpy())
        for b, (x, y) in enumerate(zip(*outs)):
            assert np.array_equal(x, y), f"seed {seed}: FAB {b}"
