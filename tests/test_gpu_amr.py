"""GPU parity of the batched hierarchy pass: every box of an uneven AMR
hierarchy (three levels, refinement ratio 2, properly nested) in one pass
call, eagerly, as one graph launch (replayed several times with other device
work between the replays, outputs cleared before each), with one fused
launch per box, and after the autotuner changed the launch shape -- all
bit-exact against the oracle's hierarchy pass, which runs the reference's
stages over all levels with AverageDownFaces and FillPatchUedge between
them.  FillPatchUedge alone is compared on random face data."""
import numpy as np
import pytest

import pyoracle
from harness import assert_bitexact, compare_face, compare_mf
from maestroex_amd.fabs import Backend
from maestroex_amd.workloads import Hierarchy

pytestmark = pytest.mark.gpu

OUTS = ("utrans", "umac", "sedge", "sflux")


def oracle_hierarchy(ppm):
    h = Hierarchy(ppm_type=ppm, backend=Backend("numpy"))
    m = pyoracle.oracle_maestro(h.params, h.boxes)
    m.advance_pass_levels([h.pass_fields(lev) for lev in range(h.nlev)])
    return h


def gpu_hierarchy(ppm):
    import maestroex_amd as mx
    h = Hierarchy(ppm_type=ppm, backend=mx.device_backend())
    m = mx.gpu_maestro(h.params, h.boxes)
    return h, m


def fields(h):
    return [h.pass_fields(lev) for lev in range(h.nlev)]


def clear_outputs(h):
    for lev in range(h.nlev):
        f = h.pass_fields(lev)
        for name in OUTS:
            for d in range(3):
                for a in f[name][d].arrays:
                    a.zero_()
        for a in f["etarhoflux"].arrays:
            a.zero_()


def compare(hc, hg, what):
    comps = [0, 1] + list(range(4, hc.nscal))
    for lev in range(hc.nlev):
        fc, fg = hc.pass_fields(lev), hg.pass_fields(lev)
        assert_bitexact(compare_face(fc["utrans"], fg["utrans"], region="all"),
                        f"{what} lev {lev} utrans")
        assert_bitexact(compare_face(fc["umac"], fg["umac"], region="all"), f"{what} lev {lev} umac")
        assert_bitexact(compare_face(fc["sedge"], fg["sedge"], comps=comps),
                        f"{what} lev {lev} sedge")
        assert_bitexact(compare_face(fc["sflux"], fg["sflux"], comps=comps),
                        f"{what} lev {lev} sflux")
        assert_bitexact(compare_mf(fc["etarhoflux"], fg["etarhoflux"], region="all"),
                        f"{what} lev {lev} etarhoflux")


@pytest.mark.parametrize("ppm", [1, 2])
@pytest.mark.parametrize("mode", ["fused", "staged", "per_box"])
def test_hierarchy_pass(ppm, mode):
    print(f"[parity] hierarchy ppm={ppm} mode={mode}", flush=True)
    hc = oracle_hierarchy(ppm)
    hg, m = gpu_hierarchy(ppm)
    m.set_tuning("edge_scal_fused", 0 if mode == "staged" else 1)
    m.set_tuning("edge_scal_per_box", 1 if mode == "per_box" else 0)
    m.advance_pass_levels(fields(hg))
    m.sync()
    compare(hc, hg, mode)


@pytest.mark.parametrize("ppm", [1, 2])
def test_hierarchy_graph_replays(ppm):
    import torch
    print(f"[parity] hierarchy graph ppm={ppm}", flush=True)
    hc = oracle_hierarchy(ppm)
    hg, m = gpu_hierarchy(ppm)
    g = m.pass_graph(fields(hg))
    m.sync()
    compare(hc, hg, "graph creation")
    scratch = torch.zeros(1 << 20, dtype=torch.float64, device="cuda:0")
    for rep in range(3):
        clear_outputs(hg)
        scratch.add_(1.0)  # other device work between the replays
        torch.cuda.synchronize()
        g.launch()
        m.sync()
        compare(hc, hg, f"replay {rep}")
    g.close()


def test_autotune_keeps_results():
    print("[parity] hierarchy autotune", flush=True)
    hc = oracle_hierarchy(1)
    hg, m = gpu_hierarchy(1)
    cands, best = m.autotune(fields(hg))
    print("[autotune]", cands, "best", best, flush=True)
    # 4 tile shapes x 4 z-chunk lengths
    assert len(cands) == 16 and 0 <= best < len(cands)
    assert sorted({t for t, _, _ in cands}) == [0, 2, 3, 4]
    assert all(ms > 0 for _, _, ms in cands)
    assert cands[best][2] == min(ms for _, _, ms in cands)
    clear_outputs(hg)
    m.advance_pass_levels(fields(hg))
    m.sync()
    compare(hc, hg, "tuned")


@pytest.mark.parametrize("nlev", [1, 3])
def test_graph_replay_follows_set_dt(nlev):
    """dt and rel_eps change every step of the reference's time loop
    (EstDt, MaestroDt.cpp:481-483).  A graph captured at one dt and replayed
    after mx_adv_set_dt must compute the pass at the new dt: each replay is
    compared bit for bit with the oracle's pass at that dt."""
    import torch
    print(f"[parity] graph replay after set_dt nlev={nlev}", flush=True)
    hg, m = gpu_hierarchy(1)
    hc = Hierarchy(ppm_type=1, backend=Backend("numpy"))
    mo = pyoracle.oracle_maestro(hc.params, hc.boxes)
    fg = fields(hg)[:nlev]
    g = m.pass_graph(fg)
    m.sync()
    scratch = torch.zeros(1 << 16, dtype=torch.float64, device="cuda:0")
    for scale in (0.5, 0.75, 0.75, 1.0):
        dt, rel_eps = hg.dt * scale, hg.rel_eps * (1.0 + scale)
        mo.set_dt(dt, rel_eps)
        mo.advance_pass_levels([hc.pass_fields(lev) for lev in range(nlev)])
        m.set_dt(dt, rel_eps)
        clear_outputs(hg)
        scratch.add_(1.0)
        torch.cuda.synchronize()
        g.launch()
        m.sync()
        for lev in range(nlev):
            fc, fgl = hc.pass_fields(lev), hg.pass_fields(lev)
            for name in OUTS:
                assert_bitexact(compare_face(fc[name], fgl[name], region="all" if name in
                                             ("utrans", "umac") else "valid",
                                             comps=None if name in ("utrans", "umac") else
                                             [0, 1] + list(range(4, hc.nscal))),
                                f"dt x{scale} lev {lev} {name}")
    g.close()


@pytest.mark.parametrize("seed", [1, 2])
def test_fill_patch_uedge_parity(seed):
    """FillPatchUedge over the three-level hierarchy on random face data
    (valid faces and ghosts), every face compared bit for bit."""
    import maestroex_amd as mx
    from maestroex_amd.fabs import FaceMultiFab
    print(f"[parity] FillPatchUedge seed={seed}", flush=True)
    h = Hierarchy(ppm_type=1, backend=Backend("numpy"))
    rng = np.random.default_rng(seed)
    host = []
    for lev in range(h.nlev):
        f = FaceMultiFab(h.boxes[lev], 1, 1, 3, Backend("numpy"))
        for d in range(3):
            for b in range(len(h.boxes[lev])):
                f[d].arrays[b][...] = rng.standard_normal(f[d].arrays[b].shape)
        host.append(f)
    dev = []
    bk = mx.device_backend()
    for lev in range(h.nlev):
        f = FaceMultiFab(h.boxes[lev], 1, 1, 3, bk)
        for d in range(3):
            for b in range(len(h.boxes[lev])):
                f[d].arrays[b] = bk.from_numpy(host[lev][d].arrays[b])
        dev.append(f)
    mo = pyoracle.oracle_maestro(h.params, h.boxes)
    mo.FillPatchUedge(host)
    mg = mx.gpu_maestro(h.params, h.boxes)
    mg.FillPatchUedge(dev)
    mg.sync()
    for lev in range(h.nlev):
        assert_bitexact(compare_face(host[lev], dev[lev], region="all"), f"lev {lev}")
