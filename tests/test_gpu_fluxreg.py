"""GPU parity of the flux register (reflux_type 2): CrseInit, FineAdd and
Reflux on random two-level data (periodic and outflow domains, fine boxes
side by side and on the domain boundary), and the hierarchy pass with
reflux_type 2 (its fluxes and every level's registers), all bit-exact
against the oracle."""
import numpy as np
import pytest

import pyoracle
from fluxreg_cases import B0, B1, DT, N0, NCOMP, make_case, params, to_device
from harness import assert_bitexact, compare_face, compare_mf
from maestroex_amd.fabs import Backend
from maestroex_amd.workloads import Hierarchy

pytestmark = pytest.mark.gpu

AC = (1.0 / N0) ** 2
AF = (0.5 / N0) ** 2


def _ops(m, f0, f1, s0):
    for d in range(3):
        m.flux_reg_crse_init(1, f0[d], d, 0, 0, NCOMP, -1.0 * DT * AC)
    m.flux_reg_crse_init(1, f0[0], 0, 1, 3, 2, 0.5)  # other components, another mult
    for d in range(3):
        m.flux_reg_fine_add(1, f1[d], d, 0, 0, NCOMP, 1.0 * DT * AF)
    m.flux_reg_fine_add(1, f1[2], 2, 0, 4, 1, 0.25)
    m.flux_reg_reflux(1, s0, 1.0, 0, 0, NCOMP)
    m.flux_reg_reflux(1, s0, 0.5, 3, 1, 2)


@pytest.mark.parametrize("periodic", [True, False])
def test_flux_register_parity(periodic):
    import maestroex_amd as mx
    print(f"[parity] flux register periodic={periodic}", flush=True)
    dev = mx.device_backend()
    p = params(periodic=periodic)
    f0, f1, s0, _ = make_case(seed=11, dyadic_values=False)
    g0, g1, gs0 = to_device(f0, dev), to_device(f1, dev), to_device(s0, dev)
    mo = pyoracle.oracle_maestro(p, [B0, B1])
    mg = mx.gpu_maestro(p, [B0, B1])
    _ops(mo, f0, f1, s0)
    _ops(mg, g0, g1, gs0)
    mg.sync()
    ro, rg = mo.flux_reg_read(1), mg.flux_reg_read(1)
    assert ro.shape == rg.shape
    nd = int(np.sum(ro.view(np.int64) != rg.view(np.int64)))
    assert nd == 0, f"{nd} register values differ"
    assert_bitexact(compare_mf(s0, gs0, region="all"), "refluxed state")
    mo.flux_reg_setval(1, 2.5)
    mg.flux_reg_setval(1, 2.5)
    assert np.all(mg.flux_reg_read(1) == 2.5) and np.all(mo.flux_reg_read(1) == 2.5)


def test_pass_reflux_type_2_parity():
    import maestroex_amd as mx
    print("[parity] hierarchy pass reflux_type 2", flush=True)
    kw = dict(ppm_type=1, counts=(3, 3), n0=16, level0_edges=(8, 8), edges=((8, 12), (4, 8)),
              reflux_type=2)
    hc = Hierarchy(backend=Backend("numpy"), **kw)
    mo = pyoracle.oracle_maestro(hc.params, hc.boxes)
    mo.advance_pass_levels([hc.pass_fields(lev) for lev in range(hc.nlev)])
    hg = Hierarchy(backend=mx.device_backend(), **kw)
    mg = mx.gpu_maestro(hg.params, hg.boxes)
    mg.advance_pass_levels([hg.pass_fields(lev) for lev in range(hg.nlev)])
    mg.sync()
    comps = [0, 1] + list(range(4, hc.nscal))
    for lev in range(hc.nlev):
        fc, fg = hc.pass_fields(lev), hg.pass_fields(lev)
        assert_bitexact(compare_face(fc["sflux"], fg["sflux"], comps=comps), f"lev {lev} sflux")
        assert_bitexact(compare_face(fc["umac"], fg["umac"], region="all"), f"lev {lev} umac")
    for lev in range(1, hc.nlev):
        ro, rg = mo.flux_reg_read(lev), mg.flux_reg_read(lev)
        nd = int(np.sum(ro.view(np.int64) != rg.view(np.int64)))
        assert nd == 0, f"lev {lev}: {nd} register values differ"
