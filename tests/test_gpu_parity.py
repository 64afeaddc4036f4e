"""GPU parity: the HIP library against the CPU oracle on identical inputs.

The bar is bit-exactness (both builds keep IEEE semantics without FMA
contraction), which is stricter than the 1e-12 relative max-norm the path
is specified to; every test reports the differing-value count and fails on
any difference.  A limiter-branch disagreement between the two would show as
a differing value, so a zero count also means zero branch disagreements.
"""
import numpy as np
import pytest

from harness import (assert_bitexact, compare_face, compare_mf, gpu_for, oracle_for, pair)
from maestroex_amd import _abi
from maestroex_amd.workloads import uneven_boxes

pytestmark = pytest.mark.gpu

P = _abi
CASES = [
    ("3d_ppm1", dict(dim=3, n=16, max_grid=8, ppm_type=1)),
    ("3d_ppm2", dict(dim=3, n=16, max_grid=8, ppm_type=2)),
    ("3d_ppm0", dict(dim=3, n=16, max_grid=8, ppm_type=0)),
    ("2d_ppm1", dict(dim=2, n=32, max_grid=16, ppm_type=1)),
    ("2d_ppm2", dict(dim=2, n=32, max_grid=16, ppm_type=2)),
    ("2d_ppm0", dict(dim=2, n=32, max_grid=16, ppm_type=0)),
    ("3d_periodic_onebox", dict(dim=3, n=16, max_grid=16, ppm_type=1, all_periodic=True)),
    ("3d_uneven_boxes", dict(dim=3, n=24, max_grid=(16, 8, 12), ppm_type=2)),
    ("3d_trace_forces", dict(dim=3, n=16, max_grid=8, ppm_type=1, ppm_trace_forces=1)),
    ("3d_walls_ppm1", dict(dim=3, n=16, max_grid=8, ppm_type=1,
                           phys_bc={0: (P.PHYS_INFLOW, P.PHYS_OUTFLOW),
                                    1: (P.PHYS_SLIPWALL, P.PHYS_NOSLIPWALL),
                                    2: (P.PHYS_SYMMETRY, P.PHYS_OUTFLOW)})),
    ("3d_walls_ppm2", dict(dim=3, n=16, max_grid=8, ppm_type=2,
                           phys_bc={0: (P.PHYS_SLIPWALL, P.PHYS_INFLOW),
                                    1: (P.PHYS_NOSLIPWALL, P.PHYS_SYMMETRY),
                                    2: (P.PHYS_OUTFLOW, P.PHYS_SLIPWALL)})),
    ("3d_walls_ppm0", dict(dim=3, n=16, max_grid=8, ppm_type=0,
                           phys_bc={0: (P.PHYS_INFLOW, P.PHYS_SLIPWALL),
                                    1: (P.PHYS_SYMMETRY, P.PHYS_NOSLIPWALL),
                                    2: (P.PHYS_OUTFLOW, P.PHYS_INFLOW)})),
    ("2d_walls_ppm1", dict(dim=2, n=32, max_grid=16, ppm_type=1,
                           phys_bc={0: (P.PHYS_OUTFLOW, P.PHYS_INFLOW),
                                    1: (P.PHYS_SLIPWALL, P.PHYS_NOSLIPWALL)})),
    ("2d_walls_ppm2", dict(dim=2, n=32, max_grid=16, ppm_type=2,
                           phys_bc={0: (P.PHYS_INFLOW, P.PHYS_SYMMETRY),
                                    1: (P.PHYS_NOSLIPWALL, P.PHYS_OUTFLOW)})),
    ("3d_spherical_ppm1", dict(dim=3, n=16, max_grid=8, ppm_type=1, spherical=True)),
    ("3d_spherical_ppm2_uneven", dict(dim=3, n=24, max_grid=(16, 8, 12), ppm_type=2,
                                      spherical=True)),
    ("3d_spherical_ppm2_random_boxes", dict(dim=3, n=32, ppm_type=2, spherical=True,
                                            boxes=uneven_boxes(32, 3, 8, 14, seed=4))),
    ("3d_spherical_exact", dict(dim=3, n=16, max_grid=8, ppm_type=2, spherical=True,
                                use_exact_base_state=1)),
]


def _run_pass(m, w):
    m.advance_pass(0, w.pass_fields())
    m.sync()


@pytest.mark.parametrize("name,cfg", CASES, ids=[c[0] for c in CASES])
def test_pass_parity(name, cfg):
    print(f"[parity] pass {name} {cfg}", flush=True)
    wc, wg = pair(**cfg)
    _run_pass(oracle_for(wc), wc)
    _run_pass(gpu_for(wg), wg)
    assert_bitexact(compare_face(wc.utrans, wg.utrans, region="all"), f"{name} utrans")
    assert_bitexact(compare_face(wc.umac, wg.umac, region="all"), f"{name} umac")
    comps = [0, 1] + list(range(4, wc.nscal))
    assert_bitexact(compare_face(wc.sedge, wg.sedge, comps=comps), f"{name} sedge")
    assert_bitexact(compare_face(wc.sflux, wg.sflux, comps=comps), f"{name} sflux")
    assert_bitexact(compare_mf(wc.etarhoflux, wg.etarhoflux, region="all"), f"{name} etarhoflux")


KERNEL_CHOICE_CASES = [
    ("3d_ppm1", dict(dim=3, n=16, max_grid=8, ppm_type=1)),
    ("3d_uneven_ppm2", dict(dim=3, n=24, max_grid=(16, 8, 12), ppm_type=2)),
    ("3d_walls_ppm2", CASES[10][1]),
    ("3d_walls_ppm1_wide", dict(dim=3, n=40, max_grid=(40, 20, 40), ppm_type=1,
                                phys_bc={0: (P.PHYS_INFLOW, P.PHYS_OUTFLOW),
                                         1: (P.PHYS_SLIPWALL, P.PHYS_NOSLIPWALL),
                                         2: (P.PHYS_SYMMETRY, P.PHYS_OUTFLOW)})),
]


@pytest.mark.parametrize("fused,tz", [(0, 32), (1, 1), (1, 5), (1, 32)])
@pytest.mark.parametrize("name,cfg", KERNEL_CHOICE_CASES, ids=[c[0] for c in KERNEL_CHOICE_CASES])
def test_pass_kernel_choices(name, cfg, fused, tz):
    """The three-kernel and the fused z-marching MakeEdgeScal, the latter
    with z-chunks of 1, 5 and 32 planes (chunk edges inside and at the box
    faces, tiles narrower than a box and wider than one), all bit-exact."""
    print(f"[parity] kernel choice {name} fused={fused} tz={tz}", flush=True)
    wc, wg = pair(**cfg)
    _run_pass(oracle_for(wc), wc)
    m = gpu_for(wg)
    m.set_tuning("edge_scal_fused", fused)
    m.set_tuning("edge_scal_tz", tz)
    _run_pass(m, wg)
    comps = [0, 1] + list(range(4, wc.nscal))
    assert_bitexact(compare_face(wc.sedge, wg.sedge, comps=comps), f"{name} sedge")
    assert_bitexact(compare_face(wc.sflux, wg.sflux, comps=comps), f"{name} sflux")


@pytest.mark.parametrize("epilogue", [0, 1])
@pytest.mark.parametrize("ept,spt", [(0, 1), (1, 1), (2, 1), (3, 2), (4, 3), (5, 1), (6, 2),
                                     (1, 3)])
def test_pass_enthalpy_pred_types(ept, spt, epilogue):
    """The whole pass for every enthalpy_pred_type (with species_pred_type
    1-3): MakeRhoHFlux's planar branches that read only the RhoH edge state
    (ept 0, 1, 3) run as the fused edge kernel's epilogue by default, the
    others as the flux kernel; every output bit-exact, walls included."""
    cfg = dict(CASES[9][1], enthalpy_pred_type=ept, species_pred_type=spt)
    print(f"[parity] pass ept={ept} spt={spt} epilogue={epilogue}", flush=True)
    wc, wg = pair(**cfg)
    _run_pass(oracle_for(wc), wc)
    m = gpu_for(wg)
    m.set_tuning("rhoh_flux_epilogue", epilogue)
    _run_pass(m, wg)
    comps = [0, 1] + list(range(4, wc.nscal))
    assert_bitexact(compare_face(wc.sedge, wg.sedge, comps=comps), f"ept {ept} sedge")
    assert_bitexact(compare_face(wc.sflux, wg.sflux, comps=comps), f"ept {ept} sflux")
    assert_bitexact(compare_mf(wc.etarhoflux, wg.etarhoflux, region="all"), f"ept {ept} eta")


SPH_EPT_CASES = [CASES[14], CASES[16], CASES[17]]


@pytest.mark.parametrize("ept", [0, 1, 2, 3, 5])
@pytest.mark.parametrize("name,cfg", SPH_EPT_CASES, ids=[c[0] for c in SPH_EPT_CASES])
def test_pass_enthalpy_pred_types_spherical(name, cfg, ept):
    """Spherical passes through MakeRhoHFlux's spherical branches with the
    face base states (umac (rho0_edge h0_edge + sedge), umac sedge, umac
    rho_edge h_edge, umac (rho'_edge + rho0_edge)(h_edge + h0_edge)) and,
    with the exact base state, its branches that leave the x-faces as they
    are: every output bit-exact, uneven boxes included."""
    cfg = dict(cfg, enthalpy_pred_type=ept)
    print(f"[parity] spherical pass {name} ept={ept}", flush=True)
    wc, wg = pair(**cfg)
    _run_pass(oracle_for(wc), wc)
    _run_pass(gpu_for(wg), wg)
    comps = [0, 1] + list(range(4, wc.nscal))
    assert_bitexact(compare_face(wc.sedge, wg.sedge, comps=comps), f"{name} ept {ept} sedge")
    assert_bitexact(compare_face(wc.sflux, wg.sflux, comps=comps), f"{name} ept {ept} sflux")
    assert_bitexact(compare_mf(wc.etarhoflux, wg.etarhoflux, region="all"), f"{name} eta")


XCD_CASES = [CASES[0], CASES[3], CASES[7], CASES[10], CASES[16], CASES[17]]


@pytest.mark.parametrize("name,cfg", XCD_CASES, ids=[c[0] for c in XCD_CASES])
def test_pass_xcd_remap(name, cfg):
    """The XCD-aware block order of the staged kernels (VelPred, MakeUtrans
    from its traced states, the fluxes) computes every point once: the whole
    pass bit-exact, box grids that are not multiples of 8 blocks included."""
    print(f"[parity] xcd remap {name}", flush=True)
    wc, wg = pair(**cfg)
    _run_pass(oracle_for(wc), wc)
    m = gpu_for(wg)
    m.set_tuning("xcd_remap", 1)
    _run_pass(m, wg)
    assert_bitexact(compare_face(wc.utrans, wg.utrans, region="all"), f"{name} utrans")
    assert_bitexact(compare_face(wc.umac, wg.umac, region="all"), f"{name} umac")
    comps = [0, 1] + list(range(4, wc.nscal))
    assert_bitexact(compare_face(wc.sedge, wg.sedge, comps=comps), f"{name} sedge")
    assert_bitexact(compare_face(wc.sflux, wg.sflux, comps=comps), f"{name} sflux")
    assert_bitexact(compare_mf(wc.etarhoflux, wg.etarhoflux, region="all"), f"{name} etarhoflux")


@pytest.mark.parametrize("es_mode,vp_mode", [(0, 0), (2, 2), (1, 1)])
@pytest.mark.parametrize("name,cfg", XCD_CASES, ids=[c[0] for c in XCD_CASES])
def test_pass_fused_xcd_order(name, cfg, es_mode, vp_mode):
    """The block orders of the fused MakeEdgeScal and VelPred set
    explicitly (the other pass tests run the defaults: 1 for the edge
    kernel; for VelPred 2 on one level of equal boxes, else 1):
    0 the plain order (tile blockIdx.x, component blockIdx.z), 1 tiles round
    robin over the XCDs and 2 contiguous tile ranges per XCD, each with the
    component blocks of a tile back to back (tile counts that are not
    multiples of 8 included): the whole pass bit-exact."""
    print(f"[parity] fused block orders {name} edge {es_mode} velpred {vp_mode}", flush=True)
    wc, wg = pair(**cfg)
    _run_pass(oracle_for(wc), wc)
    m = gpu_for(wg)
    m.set_tuning("edge_scal_xcd", es_mode)
    m.set_tuning("velpred_fused_xcd", vp_mode)
    _run_pass(m, wg)
    assert_bitexact(compare_face(wc.utrans, wg.utrans, region="all"), f"{name} utrans")
    assert_bitexact(compare_face(wc.umac, wg.umac, region="all"), f"{name} umac")
    comps = [0, 1] + list(range(4, wc.nscal))
    assert_bitexact(compare_face(wc.sedge, wg.sedge, comps=comps), f"{name} sedge")
    assert_bitexact(compare_face(wc.sflux, wg.sflux, comps=comps), f"{name} sflux")
    assert_bitexact(compare_mf(wc.etarhoflux, wg.etarhoflux, region="all"), f"{name} etarhoflux")


VP_ALL_CASES = [CASES[0], CASES[1], CASES[7], CASES[9], CASES[10], CASES[16], CASES[8]]

def test_set_tuning_rejects_bad_values():
    """A tuning value outside a key's range, and an unknown key, are
    MX_ERR_ARG and change nothing; the range ends are accepted."""
    from maestroex_amd import _abi
    print("[parity] set_tuning ranges", flush=True)
    _, wg = pair(**CASES[0][1])
    m = gpu_for(wg)
    bad = [("utrans_line", 2), ("utrans_line", -1), ("utrans_chunks", 0), ("utrans_chunks", 65),
           ("velpred_fused_tile", -1), ("velpred_fused_tile", 6), ("velpred_fused_xcd", 3), ("velpred_fused_xcd", -2),
           ("edge_scal_xcd", -1), ("edge_scal_fused", 2), ("no_such_key", 1)]
    for key, val in bad:
        with pytest.raises(_abi.MxError) as e:
            m.set_tuning(key, val)
        assert e.value.code == _abi.MX_ERR_ARG, key
    for key, val in [("utrans_chunks", 1), ("utrans_chunks", 64), ("velpred_fused_tile", 0),
                     ("velpred_fused_tile", 5), ("utrans_line", 0), ("velpred_fused_xcd", 2),
                     ("velpred_fused_xcd", -1)]:
        m.set_tuning(key, val)


UT_CASES = [CASES[0], CASES[1], CASES[6], CASES[7], CASES[9], CASES[10], CASES[17],
            KERNEL_CHOICE_CASES[3]]


@pytest.mark.parametrize("line,chunks", [(0, 4), (1, 1), (1, 3), (1, 4), (1, 7)])
@pytest.mark.parametrize("name,cfg", UT_CASES, ids=[c[0] for c in UT_CASES])
def test_pass_utrans_forms(name, cfg, line, chunks):
    """MakeUtrans one face per thread (each cell traced for both of its
    faces) and in the line form (x-face rows through LDS, y/z columns
    marched in 1, 3, 4 and 7 chunks, so chunks of one face, chunk ends
    inside the box and empty chunks of small boxes all occur): utrans and
    the whole pass bit-exact, walls, inflow, spherical w0mac and uneven boxes
    included."""
    print(f"[parity] utrans line={line} chunks={chunks} {name}", flush=True)
    wc, wg = pair(**cfg)
    _run_pass(oracle_for(wc), wc)
    m = gpu_for(wg)
    m.set_tuning("utrans_line", line)
    m.set_tuning("utrans_chunks", chunks)
    _run_pass(m, wg)
    assert_bitexact(compare_face(wc.utrans, wg.utrans, region="all"), f"{name} utrans")
    assert_bitexact(compare_face(wc.umac, wg.umac, region="all"), f"{name} umac")
    comps = [0, 1] + list(range(4, wc.nscal))
    assert_bitexact(compare_face(wc.sedge, wg.sedge, comps=comps), f"{name} sedge")


@pytest.mark.parametrize("name,cfg", VP_ALL_CASES, ids=[c[0] for c in VP_ALL_CASES])
def test_pass_velpred_ppm_per_component(name, cfg):
    """VelPred's PPM with one velocity component per thread (the default
    computes all three in one thread; traced forces, ppm_type 0 and 2-D
    always take this kernel): the whole pass bit-exact."""
    print(f"[parity] velpred ppm per component {name}", flush=True)
    wc, wg = pair(**cfg)
    _run_pass(oracle_for(wc), wc)
    m = gpu_for(wg)
    m.set_tuning("velpred_fused", 0)
    m.set_tuning("velpred_ppm_all", 0)
    _run_pass(m, wg)
    assert_bitexact(compare_face(wc.utrans, wg.utrans, region="all"), f"{name} utrans")
    assert_bitexact(compare_face(wc.umac, wg.umac, region="all"), f"{name} umac")
    comps = [0, 1] + list(range(4, wc.nscal))
    assert_bitexact(compare_face(wc.sedge, wg.sedge, comps=comps), f"{name} sedge")
    assert_bitexact(compare_face(wc.sflux, wg.sflux, comps=comps), f"{name} sflux")


@pytest.mark.parametrize("name,cfg", VP_ALL_CASES, ids=[c[0] for c in VP_ALL_CASES])
def test_pass_velpred_trans_per_pair(name, cfg):
    """VelPred's transverse states one (direction, correction) pair per
    thread (the default computes all six of a cell in one thread): the
    whole pass bit-exact."""
    print(f"[parity] velpred trans per pair {name}", flush=True)
    wc, wg = pair(**cfg)
    _run_pass(oracle_for(wc), wc)
    m = gpu_for(wg)
    m.set_tuning("velpred_fused", 0)
    m.set_tuning("velpred_trans_all", 0)
    _run_pass(m, wg)
    assert_bitexact(compare_face(wc.utrans, wg.utrans, region="all"), f"{name} utrans")
    assert_bitexact(compare_face(wc.umac, wg.umac, region="all"), f"{name} umac")
    comps = [0, 1] + list(range(4, wc.nscal))
    assert_bitexact(compare_face(wc.sedge, wg.sedge, comps=comps), f"{name} sedge")
    assert_bitexact(compare_face(wc.sflux, wg.sflux, comps=comps), f"{name} sflux")


VP_FUSED_CASES = VP_ALL_CASES + [CASES[11], CASES[17], KERNEL_CHOICE_CASES[3]]


@pytest.mark.parametrize("fused,tz,comp,tile", [(0, 32, 1, 0), (1, 1, 1, 0), (1, 5, 1, 0),
                                                (1, 32, 1, 0), (1, 5, 0, 0), (1, 5, 1, 1),
                                                (1, 5, 1, 2), (1, 32, 1, 2), (1, 5, 1, 3),
                                                (1, 5, 1, 4), (1, 5, 1, 5), (1, 32, 0, 0)])
@pytest.mark.parametrize("name,cfg", VP_FUSED_CASES, ids=[c[0] for c in VP_FUSED_CASES])
def test_pass_velpred_fused(name, cfg, fused, tz, comp, tile):
    """VelPred as the fused z-marching kernel (the default for 3-D
    ppm_type 1/2 without traced forces) -- one velocity component per block
    (every tile shape and register budget) or all three -- with z-chunks of
    1, 5 and 32 planes, and as the staged kernels: the whole pass bit-exact,
    walls, inflow, spherical w0mac and partial tiles included."""
    print(f"[parity] velpred fused={fused} tz={tz} comp={comp} tile={tile} {name}", flush=True)
    wc, wg = pair(**cfg)
    _run_pass(oracle_for(wc), wc)
    m = gpu_for(wg)
    m.set_tuning("velpred_fused", fused)
    m.set_tuning("velpred_fused_tz", tz)
    m.set_tuning("velpred_fused_comp", comp)
    m.set_tuning("velpred_fused_tile", tile)
    _run_pass(m, wg)
    assert_bitexact(compare_face(wc.utrans, wg.utrans, region="all"), f"{name} utrans")
    assert_bitexact(compare_face(wc.umac, wg.umac, region="all"), f"{name} umac")
    comps = [0, 1] + list(range(4, wc.nscal))
    assert_bitexact(compare_face(wc.sedge, wg.sedge, comps=comps), f"{name} sedge")
    assert_bitexact(compare_face(wc.sflux, wg.sflux, comps=comps), f"{name} sflux")
    assert_bitexact(compare_mf(wc.etarhoflux, wg.etarhoflux, region="all"), f"{name} etarhoflux")


@pytest.mark.parametrize("tile", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("name,cfg", KERNEL_CHOICE_CASES, ids=[c[0] for c in KERNEL_CHOICE_CASES])
def test_pass_fused_tile_shapes(name, cfg, tile):
    """The fused kernel's tile shapes and block sizes (32 x 8 and 16 x 16 on
    384 threads, 16 x 12 and 16 x 11 on 256, 16 x 8 on 192), with 5-plane
    z-chunks: bit-exact, partial tiles included."""
    print(f"[parity] fused tile {tile} {name}", flush=True)
    wc, wg = pair(**cfg)
    _run_pass(oracle_for(wc), wc)
    m = gpu_for(wg)
    m.set_tuning("edge_scal_tile", tile)
    m.set_tuning("edge_scal_tz", 5)
    _run_pass(m, wg)
    comps = [0, 1] + list(range(4, wc.nscal))
    assert_bitexact(compare_face(wc.sedge, wg.sedge, comps=comps), f"{name} sedge")
    assert_bitexact(compare_face(wc.sflux, wg.sflux, comps=comps), f"{name} sflux")


@pytest.mark.parametrize("is_vel,is_cons,ppm", [(1, 0, 1), (1, 0, 2), (0, 1, 1), (0, 1, 2), (1, 0, 0),
                                                (0, 1, 0)])
@pytest.mark.parametrize("dim,fused,tile", [(2, 0, 0), (3, 0, 0), (3, 1, 0), (3, 1, 1), (3, 1, 2),
                                            (3, 1, 3)])
def test_edge_scal_variants(dim, fused, tile, is_vel, is_cons, ppm):
    """MakeEdgeScal on the velocity (VelocityAdvance: is_vel, bcs_u) and in
    conservative form, with walls on every side; in 3-D with the three-kernel
    and with the fused kernel (ppm_type 0 always takes the three kernels)."""
    n = 32 if dim == 2 else 16
    bc = {0: (P.PHYS_SLIPWALL, P.PHYS_INFLOW), 1: (P.PHYS_OUTFLOW, P.PHYS_NOSLIPWALL)}
    if dim == 3:
        bc[2] = (P.PHYS_SYMMETRY, P.PHYS_OUTFLOW)
    cfg = dict(dim=dim, n=n, max_grid=n // 2, ppm_type=ppm, phys_bc=bc)
    print(f"[parity] edge_scal dim={dim} fused={fused} tile={tile} is_vel={is_vel} cons={is_cons} "
          f"ppm={ppm}",
          flush=True)
    wc, wg = pair(**cfg)
    res = []
    mg = gpu_for(wg)
    mg.set_tuning("edge_scal_fused", fused)
    mg.set_tuning("edge_scal_tile", tile)
    for w, m in ((wc, oracle_for(wc)), (wg, mg)):
        # a MAC velocity with filled ghosts from the first two stages
        m.MakeUtrans(0, w.utilde, w.ufull, w.utrans, None, w.w0_cart)
        m.VelPred(0, w.utilde, w.ufull, w.utrans, w.umac, None, w.vel_force, w.w0_cart)
        m.Addw0(0, w.umac, None, w.w0, 1.0)
        bcs_s, bcs_u = m.bc_setup()
        if is_vel:
            # VelocityAdvance: MakeEdgeScal(uold, uedge, umac, vel_force, true, bcs_u, DIM, 0, 0, DIM)
            from maestroex_amd.fabs import FaceMultiFab
            uedge = FaceMultiFab(w.boxes, dim, 0, dim, w.backend)
            bu = (_abi.BCRec * dim)(*[bcs_u[q] for q in range(dim)])
            m.MakeEdgeScal(0, w.utilde, uedge, w.umac, w.vel_force, True, bu, 0, 0, dim, bool(is_cons))
            out = uedge
        else:
            m.MakeEdgeScal(0, w.scal, w.sedge, w.umac, w.scal_force, False, bcs_s, 4, 4, 3,
                           bool(is_cons))
            m.MakeEdgeScal(0, w.scal, w.sedge, w.umac, w.scal_force, False, bcs_s, 0, 0, 2,
                           bool(is_cons))
            out = w.sedge
        m.sync()
        res.append(out)
    comps = list(range(dim)) if is_vel else [0, 1, 4, 5, 6]
    assert_bitexact(compare_face(res[0], res[1], comps=comps), "sedge")


@pytest.mark.parametrize("spt,ept", [(1, 1), (2, 0), (3, 2), (1, 2), (1, 3), (2, 4), (1, 5), (3, 6)])
@pytest.mark.parametrize("geom", ["planar", "spherical", "spherical_exact"])
def test_flux_pred_types(spt, ept, geom):
    """MakeRhoXFlux / MakeRhoHFlux for every species and enthalpy predictor
    type, planar and spherical (face base states, with and without
    use_exact_base_state)."""
    cfg = dict(dim=3, n=16, max_grid=8, ppm_type=1, spherical=geom != "planar",
               use_exact_base_state=1 if geom == "spherical_exact" else 0)
    print(f"[parity] flux {geom} species_pred_type={spt} enthalpy_pred_type={ept}", flush=True)
    wc, wg = pair(**cfg)
    outs = []
    for w, mk in ((wc, oracle_for), (wg, gpu_for)):
        w.params.species_pred_type = spt
        w.params.enthalpy_pred_type = ept
        m = mk(w)
        _run_pass(m, w)
        outs.append(w)
    comps = [0, 1] + list(range(4, wc.nscal))
    assert_bitexact(compare_face(outs[0].sflux, outs[1].sflux, comps=comps), "sflux")
    assert_bitexact(compare_mf(outs[0].etarhoflux, outs[1].etarhoflux, region="all"), "etarhoflux")


@pytest.mark.parametrize("ppm", [1, 2])
@pytest.mark.parametrize("is_umac", [0, 1])
@pytest.mark.parametrize("bc", [P.BC_INT_DIR, P.BC_EXT_DIR, P.BC_HOEXTRAP, P.BC_FOEXTRAP])
def test_ppm_unit(ppm, is_umac, bc):
    """Maestro::PPM on one box next to every side of the domain, with the
    boundary type that selects the special stencils."""
    from maestroex_amd.fabs import MultiFab
    cfg = dict(dim=3, n=12, max_grid=12, ppm_type=ppm)
    wc, wg = pair(**cfg)
    bcs = (_abi.BCRec * 1)()
    for d in range(3):
        bcs[0].lo[d] = bc
        bcs[0].hi[d] = bc
    outs = []
    for w, mk in ((wc, oracle_for), (wg, gpu_for)):
        m = mk(w)
        lo, hi = w.boxes[0]
        Ip = MultiFab(w.boxes, 3, 1, 3, w.backend)
        Im = MultiFab(w.boxes, 3, 1, 3, w.backend)
        bx = ([l - 1 for l in lo], [h + 1 for h in hi])
        if is_umac:
            # face velocities: utrans after MakeUtrans (1 ghost face)
            m.MakeUtrans(0, w.utilde, w.ufull, w.utrans, None, w.w0_cart)
            vel = [(w.utrans[d], 0) for d in range(3)]
        else:
            vel = []
            for d in range(3):
                v = MultiFab(w.boxes, 1, 4, 3, w.backend)
                v.arrays[0] = w.ufull.arrays[0][d:d + 1].clone() if hasattr(w.ufull.arrays[0], "clone") \
                    else w.ufull.arrays[0][d:d + 1].copy()
                vel.append((v, 0))
        m.PPM(0, bx, (w.scal, 0), vel[0], vel[1], vel[2], (Ip, 0), (Im, 0), bcs, is_umac, 0, 0)
        m.sync()
        outs.append((Ip, Im))
    assert_bitexact(compare_mf(outs[0][0], outs[1][0], region="all"), "Ip")
    assert_bitexact(compare_mf(outs[0][1], outs[1][1], region="all"), "Im")


@pytest.mark.parametrize("order", [0, 2, 4])
@pytest.mark.parametrize("bc", [P.BC_INT_DIR, P.BC_EXT_DIR, P.BC_HOEXTRAP])
def test_slope_unit(order, bc):
    from maestroex_amd.fabs import MultiFab
    cfg = dict(dim=3, n=12, max_grid=12, ppm_type=0)
    wc, wg = pair(**cfg)
    bcs = (_abi.BCRec * 7)()
    for q in range(7):
        for d in range(3):
            bcs[q].lo[d] = bc
            bcs[q].hi[d] = bc
    outs = []
    for w, mk in ((wc, oracle_for), (wg, gpu_for)):
        m = mk(w)
        lo, hi = w.boxes[0]
        bx = ([l - 1 for l in lo], [h + 1 for h in hi])
        res = []
        for d in range(3):
            sl = MultiFab(w.boxes, 3, 1, 3, w.backend)
            m.Slope(0, d, bx, (w.scal, 0), (sl, 0), bcs, 3, 4, order)
            res.append(sl)
        m.sync()
        outs.append(res)
    for d in range(3):
        assert_bitexact(compare_mf(outs[0][d], outs[1][d], region="all"), f"slope dir {d}")


def test_cell_to_edge():
    import maestroex_amd as mx
    from maestroex_amd.fabs import BaseState
    rng = np.random.default_rng(7)
    nr = 40
    cell = rng.random(nr)
    outs = []
    for kind in ("cpu", "gpu"):
        w_bk = mx.device_backend() if kind == "gpu" else mx.Backend("numpy")
        c = BaseState(cell, w_bk)
        e = BaseState(np.zeros(nr + 1), w_bk)
        from maestroex_amd.fabs import Geometry, make_params
        g = Geometry(3, [8, 8, nr], [1.0 / nr] * 3, [1, 1, 0], [0, 0, 2], [0, 0, 2])
        p = make_params(g)
        boxes = [((0, 0, 0), (7, 7, nr - 1))]
        m = (gpu_for if kind == "gpu" else oracle_for)(type("W", (), dict(params=p, boxes=boxes))())
        m.CelltoEdge(0, nr, c, e)
        m.sync()
        outs.append(e.to_numpy())
    assert np.array_equal(outs[0], outs[1])
