"""GPU parity of MakeEdgeState1d: the HIP kernel against the CPU oracle,
bit for bit, on random base states with w0 of both signs (and at and below
rel_eps), planar one- and two-level, spherical, every ppm_type and slope
order.  Entries a chunk does not cover must stay untouched (NaN)."""
import numpy as np
import pytest

import pyoracle
from edge1d_cases import dummy_boxes, params, random_fields, run, two_level_geom
from maestroex_amd import _abi

pytestmark = pytest.mark.gpu

CASES = [(ppm, so) for ppm, so in [(0, 0), (0, 2), (0, 4), (1, 4), (2, 4)]]


def both(ppm, sph, so, rg, s, f, w0):
    import maestroex_amd as mx
    p = params(ppm, sph, so)
    out = []
    for kind in ("cpu", "gpu"):
        if kind == "cpu":
            m = pyoracle.oracle_maestro(p, [dummy_boxes()])
            bk = pyoracle_backend()
        else:
            m = mx.gpu_maestro(p, [dummy_boxes()])
            bk = mx.device_backend()
        out.append(run(m, rg, s, f, w0, bk))
    return out


def pyoracle_backend():
    from maestroex_amd.fabs import Backend
    return Backend("numpy")


def assert_same(a, b, what):
    same = (a == b) | (np.isnan(a) & np.isnan(b))
    assert same.all(), f"{what}: {int((~same).sum())} of {a.size} values differ"


@pytest.mark.parametrize("ppm,so", CASES)
@pytest.mark.parametrize("sph", [0, 1])
@pytest.mark.parametrize("nr", [16, 1536])
def test_edge_state_1d_single_level(ppm, so, sph, nr):
    print(f"[parity] edge1d ppm={ppm} slope_order={so} spherical={sph} nr={nr}", flush=True)
    rg = _abi.radial_geom(nr, 1.0 / nr)
    s, f, w0 = random_fields(1, nr, 7 + nr)
    cpu, gpu = both(ppm, sph, so, rg, s, f, w0)
    assert_same(cpu, gpu, "sedge")


@pytest.mark.parametrize("ppm,so", CASES)
def test_edge_state_1d_two_levels(ppm, so):
    print(f"[parity] edge1d two levels ppm={ppm} slope_order={so}", flush=True)
    rg = two_level_geom()
    s, f, w0 = random_fields(2, 64, 21)
    cpu, gpu = both(ppm, 0, so, rg, s, f, w0)
    assert_same(cpu, gpu, "sedge")
