"""GPU parity at the benchmark's own sizes: every workload the benchmark
times is run through the HIP library and through the CPU oracle (the -O3
build, on all host threads of the box's share) on identical inputs, and
every output array is compared bit for bit.

  * 3-D 256^3 as 8 boxes of 128^3, 4 ghosts, 5 scalars, ppm_type 1 and 2,
    eager and as a replayed HIP graph (the 16x11 / 16x12 tiles leave partial
    tiles at a 128-cell box edge; the kernels keep 32-bit plane strides);
  * 2-D 256^2 as one box, periodic in x and outflow in y, ppm_type 1;
  * the batched AMR hierarchy of the benchmark (level 0 256^3 in boxes of
    32-64 cells, 150 level-1 and 200 level-2 boxes);
  * 384^3 spherical in 216 uneven boxes (edges 48-96), ppm_type 2;
  * the differentiable step at 128^3 with 4 scalars: the 20-step forward
    rollout, the dot-product test of the adjoint, and central finite
    differences (eps = 1e-6) of the 20-step loss at 64 cells.

Each comparison prints (and under a GPU run keeps, in
gpurun_out/parity_metrics.jsonl) the max relative error and the count of
differing values per output array.  A limiter- or upwind-branch
disagreement between GPU and CPU changes the value it selects, so a zero
count of differing values bounds the branch disagreements that reach an
output at zero.
"""
import os
import time

import numpy as np
import pytest

import pyoracle
from harness import assert_bitexact, compare_face, compare_mf, record_parity
from maestroex_amd.fabs import Backend

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(900)]

THREADS = 16  # the host-thread share of one GPU on the GPU box


def _oracle_threads():
    n = min(THREADS, len(os.sched_getaffinity(0)))
    pyoracle.set_threads(n, fast=True)
    return n


def _log(msg):
    print(f"[fullsize] {msg}", flush=True)


def _compare_pass(case, wc, wg, nlev=1):
    comps = [0, 1] + list(range(4, wc.nscal))
    worst = 0
    for lev in range(nlev):
        fc = wc.pass_fields(lev) if nlev > 1 else wc.pass_fields()
        fg = wg.pass_fields(lev) if nlev > 1 else wg.pass_fields()
        tag = f"{case} lev {lev}" if nlev > 1 else case
        for name, region, cs in (("utrans", "all", None), ("umac", "all", None),
                                 ("sedge", "valid", comps), ("sflux", "valid", comps)):
            r = record_parity(tag, name, compare_face(fc[name], fg[name], region=region, comps=cs))
            worst = max(worst, r["ndiff"])
        r = record_parity(tag, "etarhoflux",
                          compare_mf(fc["etarhoflux"], fg["etarhoflux"], region="all"))
        worst = max(worst, r["ndiff"])
    return worst


def _clear(w, nlev=1):
    for lev in range(nlev):
        f = w.pass_fields(lev) if nlev > 1 else w.pass_fields()
        for name in ("utrans", "umac", "sedge", "sflux"):
            for d in range(w.dim if hasattr(w, "dim") else 3):
                for a in f[name][d].arrays:
                    a.zero_()
        for a in f["etarhoflux"].arrays:
            a.zero_()


def _single_level(case, cfg, graph_replays=0):
    import maestroex_amd as mx
    from maestroex_amd.workloads import Workload
    t0 = time.perf_counter()
    wc = Workload(backend=Backend("numpy"), **cfg)
    wg = Workload(backend=mx.device_backend(), **cfg)
    _log(f"{case}: {len(wc.boxes)} boxes, {wc.valid_cells()} cells built in "
         f"{time.perf_counter() - t0:.1f}s")
    nth = _oracle_threads()
    t0 = time.perf_counter()
    mo = pyoracle.oracle_maestro(wc.params, [wc.boxes], fast=True)
    mo.advance_pass(0, wc.pass_fields())
    _log(f"{case}: oracle pass {time.perf_counter() - t0:.1f}s on {nth} threads")
    mg = mx.gpu_maestro(wg.params, [wg.boxes])
    mg.advance_pass(0, wg.pass_fields())
    mg.sync()
    nd = _compare_pass(case + " eager", wc, wg)
    assert nd == 0, f"{case}: {nd} values differ"
    if graph_replays:
        import torch
        g = mg.pass_graph([wg.pass_fields()])
        for rep in range(graph_replays):
            _clear(wg)
            torch.cuda.synchronize()
            g.launch()
            mg.sync()
            nd = _compare_pass(f"{case} graph replay {rep}", wc, wg)
            assert nd == 0, f"{case} graph replay {rep}: {nd} values differ"
        g.close()


@pytest.mark.parametrize("ppm", [1, 2])
def test_config1_256c_8x128(ppm):
    _log(f"config 1: 256^3 as 8 x 128^3, ppm_type {ppm}")
    _single_level(f"256^3 8x128^3 ppm{ppm}", dict(dim=3, n=256, max_grid=128, ppm_type=ppm),
                  graph_replays=2)


def test_config0_2d_256():
    _log("config 0: 2-D 256^2, one box, periodic x, outflow y, ppm_type 1")
    _single_level("2-D 256^2 ppm1", dict(dim=2, n=256, max_grid=256, ppm_type=1))


def test_config4_spherical_384():
    from maestroex_amd.workloads import uneven_boxes
    _log("config 4: 384^3 spherical in 216 uneven boxes, ppm_type 2")
    _single_level("384^3 spherical 216 boxes ppm2",
                  dict(dim=3, n=384, ppm_type=2, spherical=True,
                       boxes=uneven_boxes(384, 6, 48, 96, seed=4)))


def test_config2_amr_hierarchy():
    import maestroex_amd as mx
    from maestroex_amd.workloads import Hierarchy
    kw = dict(n0=256, level0_edges=(32, 64), counts=(150, 200), edges=((16, 48), (8, 32)),
              fracs=(0.585, 0.368), ppm_type=1)
    t0 = time.perf_counter()
    hc = Hierarchy(backend=Backend("numpy"), **kw)
    hg = Hierarchy(backend=mx.device_backend(), **kw)
    _log(f"config 2: {[len(b) for b in hc.boxes]} boxes, {hc.valid_cells()} cells built in "
         f"{time.perf_counter() - t0:.1f}s")
    nth = _oracle_threads()
    t0 = time.perf_counter()
    mo = pyoracle.oracle_maestro(hc.params, hc.boxes, fast=True)
    mo.advance_pass_levels([hc.pass_fields(lev) for lev in range(hc.nlev)])
    _log(f"config 2: oracle pass {time.perf_counter() - t0:.1f}s on {nth} threads")
    mg = mx.gpu_maestro(hg.params, hg.boxes)
    fields = [hg.pass_fields(l) for l in range(hg.nlev)]
    cands, best = mg.autotune(fields)
    _log(f"config 2: autotuned {cands[best]}")
    _clear(hg, hg.nlev)
    g = mg.pass_graph(fields)
    _clear(hg, hg.nlev)
    g.launch()
    mg.sync()
    nd = _compare_pass("AMR hierarchy graph", hc, hg, nlev=hc.nlev)
    g.close()
    assert nd == 0, f"hierarchy: {nd} values differ"


def _adjoint_case(n=128, nc=4):
    from maestroex_amd.adjoint import periodic_mac_velocity
    u = periodic_mac_velocity(n, seed=1)
    x = (np.arange(n) + 0.5) / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    rng = np.random.default_rng(3)
    s0 = np.stack([np.exp(-((X - 0.5) ** 2 + (Y - 0.3 - 0.1 * c) ** 2 + (Z - 0.5) ** 2) / 0.02)
                   + 0.05 * rng.standard_normal((n, n, n)) for c in range(nc)])
    tgt = rng.standard_normal(s0.shape)
    return s0, u, tgt, 0.5 / n, 1e-8


def _dev(a):
    import torch
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def test_config3_adjoint_128c_forward_rollout():
    """20 forward steps at 128^3 with 4 scalars, bit-exact against the
    oracle's step at every step."""
    from maestroex_amd.adjoint import ScalStepper
    s0, u, _, dt, eps = _adjoint_case()
    n, nc = s0.shape[1], s0.shape[0]
    _log(f"config 3: 20-step forward rollout {n}^3, {nc} scalars")
    st = ScalStepper(n, nc, dt, eps, ppm_type=1)
    ud = tuple(_dev(x) for x in u)
    sg = _dev(s0)
    so = s0
    for step in range(20):
        sg = st.step(sg, ud)
        so = pyoracle.scal_step(so, u, dt, [1.0 / n] * 3, eps, ppm_type=1, fast=True)
        g = sg.cpu().numpy()
        res = dict(max_abs=float(np.max(np.abs(g - so))), n=g.size,
                   ndiff=int(np.count_nonzero(g != so)))
        res["max_rel"] = res["max_abs"] / float(np.max(np.abs(so)))
        record_parity(f"adjoint 128^3 forward step {step}", "snew", res)
        assert res["ndiff"] == 0, f"step {step}: {res['ndiff']} values differ"


def test_config3_adjoint_128c_dot_product():
    """<v, J w> from the oracle's forward-mode derivative equals <J^T v, w>
    from the GPU adjoint at 128^3, 4 scalars, for w in the state and the
    velocity, to 1e-12 relative."""
    from maestroex_amd.adjoint import ScalStepper
    s0, u, _, dt, eps = _adjoint_case()
    n, nc = s0.shape[1], s0.shape[0]
    _log(f"config 3: dot-product test {n}^3, {nc} scalars")
    rng = np.random.default_rng(11)
    ws = rng.standard_normal(s0.shape)
    wu = [rng.standard_normal(u[0].shape) for _ in range(3)]
    v = rng.standard_normal(s0.shape)
    _, jw = pyoracle.scal_step(s0, u, dt, [1.0 / n] * 3, eps, ws, wu, ppm_type=1, fast=True)
    st = ScalStepper(n, nc, dt, eps, ppm_type=1)
    sb, ub = st.vjp(_dev(s0), tuple(_dev(x) for x in u), _dev(v))
    sb = sb.cpu().numpy()
    ub = [x.cpu().numpy() for x in ub]
    lhs = float(np.sum(v * jw))
    rhs = float(np.sum(sb * ws) + sum(np.sum(ub[d] * wu[d]) for d in range(3)))
    rel = abs(lhs - rhs) / max(abs(lhs), 1.0)
    _log(f"<v, J w> = {lhs!r}  <J^T v, w> = {rhs!r}  rel {rel:.2e}")
    record_parity("adjoint 128^3 dot product", "<v,Jw>-<J^T v,w>",
                  dict(max_rel=rel, max_abs=abs(lhs - rhs), ndiff=int(rel > 1e-12), n=1))
    assert rel <= 1e-12


def test_config3_adjoint_128c_finite_differences():
    """Central differences (eps = 1e-6) of the 20-step rollout loss
    L = |s_20 - target|^2 / 2 at 64 random cells of the initial state,
    against the checkpointed (every 5 steps) adjoint gradient: 1e-6 relative.
    A point whose differences with eps and eps / 2 disagree lies at a limiter
    switch within eps and is reported and left out."""
    import torch
    from maestroex_amd.adjoint import Rollout, ScalStepper
    s0, u, tgt, dt, eps = _adjoint_case()
    n, nc = s0.shape[1], s0.shape[0]
    _log(f"config 3: finite differences of the 20-step loss, {n}^3, {nc} scalars")
    st = ScalStepper(n, nc, dt, eps, ppm_type=1)
    ud = tuple(_dev(x) for x in u)
    tg = _dev(tgt)
    sd = _dev(s0).requires_grad_(True)
    out = Rollout.apply(st, 20, 5, sd, *[x.clone().requires_grad_(True) for x in ud])
    (0.5 * ((out - tg) ** 2).sum()).backward()
    grad = sd.grad.cpu().numpy()

    def final(s):
        x = s
        for _ in range(20):
            x = st.step(x, ud)
        return x

    def ldiff(delta, idx):
        sp = _dev(s0)
        sm = _dev(s0)
        sp[idx] += delta
        sm[idx] -= delta
        ap, am = final(sp), final(sm)
        return float(torch.sum(0.5 * (ap - am) * (ap + am - 2.0 * tg)))

    h = 1e-6
    rng = np.random.default_rng(13)
    checked, kinks, worst = 0, 0, 0.0
    for _ in range(64):
        idx = tuple(int(x) for x in (rng.integers(0, nc), *rng.integers(0, n, size=3)))
        central = ldiff(h, idx) / (2 * h)
        half = ldiff(h / 2, idx) / h
        g = float(grad[idx])
        scale = max(abs(central), abs(g), 1e-3)
        if abs(central - half) > 1e-6 * scale:
            kinks += 1
            _log(f"point {idx} at a limiter switch: D(h) {central} D(h/2) {half}")
            continue
        err = abs(central - g) / scale
        worst = max(worst, err)
        checked += 1
        assert err <= 1e-6, f"{idx}: fd {central} adjoint {g}"
    _log(f"{checked} points checked, worst relative error {worst:.2e}, {kinks} at switches")
    record_parity("adjoint 128^3 finite differences", "grad_s0",
                  dict(max_rel=worst, max_abs=worst, ndiff=0, n=checked))
    assert kinks <= 8 and checked >= 56
