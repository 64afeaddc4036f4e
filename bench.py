"""Benchmark: one FP64 advection pass of MAESTROeX's hot path on one MI355X.

Workload (BASELINE configuration "3-D 256^3 as 8 boxes of 128^3"): 256^3 cells
as 8 boxes of 128^3 with 4 ghost cells, 5 advected scalars (rho', (rho h)',
3 mass fractions) plus 3 velocity components, periodic in x and y, outflow in
z, PPM type 1 (or --ppm 2).  One step is one pass of the path over the whole
level with inputs resident in HBM:
    MakeUtrans -> Addw0 -> VelPred -> Addw0 -> MakeEdgeScal (5 scalars)
      -> MakeRhoXFlux -> MakeRhoHFlux
value = valid cells / pass time, in Gcell-updates/s.

roofline: the dominant kernel's algorithmic HBM bytes per launch over its
average launch time, both measured here with HIP events on the library's
stream (mx_adv_prof_*).  cpu_baseline: the CPU oracle (a line-faithful
OpenMP restatement of the reference) timed on one 128^3 box of the same
workload on this host.
"""
import argparse
import ctypes as C
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "Gcell-updates/s per FP64 advect step, % HBM peak; adjoint/forward time ratio"
HBM_PEAK_GBS = 8000.0
# MI355X FP64 vector rate without FMA (1024 SIMDs x 16 lanes x 2.4 GHz): the
# op count below counts an FMA-able pair as two operations
FP64_PEAK_TOPS = 39.3
# FP64 operations per cell and component of MakeEdgeScal (3-D, advective,
# ppm_type 1) in the formulation with neighbour-shared slopes and face values:
# the oracle's op-counting instantiation of the step on smooth data gives 439
# per cell in the per-point form (PPM and trace 277, predictor 9, corner 78,
# edge 75), less 2 van Leer slopes (13.5 on average) and 1 face value (9) per
# cell and direction shared with the neighbours; tests/test_oracle.py
# (test_step_op_count) recomputes it
EDGE_OPS_PER_CELL_COMP = 331

# Algorithmic HBM bytes per valid cell of one pass (FP64, valid region only,
# compulsory traffic across the reference's stage boundaries): 3-D, nc = 5.
PASS_BYTES_PER_CELL = {3: 488, 2: 352}
# spherical: w0mac read in MakeUtrans and VelPred as well (+6 doubles)
PASS_BYTES_PER_CELL_SPHERICAL = 536


# edge_scal_tile values: tile shape (cell columns) and threads per block
TILE_NAMES = ["32x8", "16x16", "16x12", "16x8", "16x11"]
TILE_THREADS = [384, 384, 256, 192, 256]

def kernel_bytes_per_cell(name, nc, dim):
    """Algorithmic bytes per valid cell of one launch of each kernel: its
    compulsory inputs read once and its outputs written once (FP64)."""
    D = dim
    table = {
        # MakeUtrans faces: read utilde (D) + ufull (D), write utrans (D)
        "k_utrans": 8 * (3 * D),
        # VelPred PPM: read utilde (D) + ufull (D), write Ip/Im (2 D per comp)
        "k_vp_ppm": 8 * (2 * D + 2 * D * D),
        # VelPred transverse: read Ip/Im (2 D D) + utrans (D), write 6 states
        "k_vp_trans": 8 * (2 * D * D + D + 6),
        # VelPred velocities: read states (2 D + 6) + utrans + force, write umac
        "k_vp_vel": 8 * (2 * D + 6 + D + D + D),
        # EdgeScal PPM per launch over nc comps: read s, umac (D); write Ip/Im (2 D)
        "k_es_ppm": 8 * (nc * (1 + 2 * D) + D),
        # EdgeScal transverse: read Ip/Im (2 D) + umac; write 6 states
        "k_es_trans": 8 * (nc * (2 * D + 6) + D),
        # EdgeScal edges: read Ip/Im, states, force, umac; write sedge (D)
        "k_es_edge": 8 * (nc * (2 * D + 6 + 1 + D) + D),
        # fused MakeEdgeScal: read s, force, umac (D); write sedge (D)
        "k_es_fused": 8 * (nc * (2 + D) + D),
        # fused VelPred: read utilde, ufull, utrans, force (D each); write umac (D)
        "k_vp_fused": 8 * (5 * D),
    }
    return table.get(name)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--ppm", type=int, default=1)
    ap.add_argument("--dim", type=int, default=3,
                    help="2: the 2-D workload (256^2 as one box, periodic x, outflow y)")
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--max-grid", type=int, default=128)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--headline-only", action="store_true",
                    help="time only the headline pass (no ppm_type 2, AMR or adjoint fields)")
    ap.add_argument("--autotune", action="store_true",
                    help="let the autotuner pick the fused edge kernel's launch shape (the "
                         "hierarchy mode always does)")
    ap.add_argument("--no-autotune", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--cpu-n", type=int, default=128)
    ap.add_argument("--cpu-reps", type=int, default=5)
    ap.add_argument("--tune", action="append", default=[], metavar="KEY=VALUE",
                    help="kernel choice passed to mx_adv_set_tuning (repeatable)")
    ap.add_argument("--amr", action="store_true",
                    help="the batched AMR hierarchy (3 levels, uneven boxes) instead of 256^3")
    ap.add_argument("--adjoint", action="store_true",
                    help="the differentiable step: 128^3 periodic, 4 scalars, adjoint/forward cost")
    ap.add_argument("--base", action="store_true",
                    help="the base-state <-> Cartesian producers on the 3-D level")
    ap.add_argument("--update", action="store_true",
                    help="UpdateScal + UpdateVel (the update that closes the step) on the 3-D level")
    ap.add_argument("--plotfile", action="store_true",
                    help="write the 3-D level as an AMReX plotfile")
    ap.add_argument("--fill", action="store_true",
                    help="FillPatch (level 0) of the scalars on the 3-D level")
    ap.add_argument("--spherical", action="store_true",
                    help="the spherical star: 384^3 in 216 uneven boxes (edges 48-96), "
                         "ppm_type 2, radial w0, one graph launch per pass")
    ap.add_argument("--calib-copy-mib", type=int, default=0,
                    help="also run a plain device copy of this many MiB (counter calibration)")
    return ap.parse_args()


def committed_traffic(kernel, workload_tag):
    """HBM bytes per launch of `kernel` from the committed counter summary
    profiles/traffic_<kernel>_<workload_tag>.md (written by
    maestroex_amd.profiling from two rocprofv3 --pmc runs), or None."""
    path = os.path.join(ROOT, "profiles", f"traffic_{kernel}_{workload_tag}.md")
    if not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def host_cpu():
    """(nproc, CPU model name) of this host."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return os.cpu_count(), model


def cpu_threads():
    """Host threads of the CPU baseline: the host-thread share of one GPU on
    the GPU box (OMP_NUM_THREADS there, 16), at most the CPUs this process
    may run on; MX_CPU_THREADS overrides."""
    n = int(os.environ.get("MX_CPU_THREADS", "0")) or int(os.environ.get("OMP_NUM_THREADS", "0"))
    n = n or 16
    return max(1, min(n, len(os.sched_getaffinity(0))))


def cpu_baseline(args, ppm=None):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pyoracle
    from maestroex_amd.fabs import Backend
    from maestroex_amd.workloads import Workload
    ppm = args.ppm if ppm is None else ppm
    cores = cpu_threads()
    pyoracle.set_threads(cores, fast=True)
    lib, _ = pyoracle.load(True)
    lib.mxo_alloc_ns.restype = C.c_longlong
    lib.mxo_alloc_ns.argtypes = [C.c_int]
    dim = getattr(args, "dim", 3)
    n = args.cpu_n if dim == 3 else args.n
    w = Workload(dim=dim, n=n, max_grid=n, ppm_type=ppm, backend=Backend("numpy"))
    m = pyoracle.oracle_maestro(w.params, [w.boxes], fast=True)
    m.advance_pass(0, w.pass_fields())  # warm-up (page-in, thread start)
    times, excl = [], []
    for _ in range(args.cpu_reps):
        lib.mxo_alloc_ns(1)
        t0 = time.perf_counter()
        m.advance_pass(0, w.pass_fields())
        t = time.perf_counter() - t0
        times.append(t)
        excl.append(t - lib.mxo_alloc_ns(1) * 1e-9)
    med = sorted(times)[len(times) // 2]
    med_x = sorted(excl)[len(excl) // 2]
    nproc, model = host_cpu()
    cells = n ** dim
    return {
        "value": round(cells / med / 1e9, 6),
        "unit": "Gcell-updates/s",
        "cores": cores,
        "kind": "port",
        "sample": f"one {n}^{dim} box of the same construction (4 ghosts, 5 scalars, ppm_type {ppm}), "
                  f"median of {args.cpu_reps} passes after 1 warm-up, OpenMP {cores} threads "
                  f"(the host-thread share of one GPU), g++ -O3 -fopenmp -ffp-contract=off; "
                  f"{med:.3f} s per pass allocation-inclusive, {med_x:.3f} s allocation-exclusive",
        "alloc_exclusive_value": round(cells / med_x / 1e9, 6),
        "host_nproc": nproc,
        "host_cpu": model,
    }


def copy_bandwidth_gbs(mib=2048, reps=5):
    """Measured device-to-device copy bandwidth (read + write bytes / time)
    of a plain torch copy of `mib` MiB, the best of `reps` copies: the
    achievable HBM rate the kernels are also compared with."""
    import torch
    n = mib * (1 << 20) // 8
    src = torch.ones(n, dtype=torch.float64, device="cuda:0")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    best = 1e30
    for _ in range(reps):
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record()
        dst.copy_(src)
        b.record()
        torch.cuda.synchronize()
        best = min(best, a.elapsed_time(b))
    del src, dst
    return 2.0 * n * 8 / (best * 1e-3) / 1e9


def time_ms(fn, steps, warmup):
    """Mean wall time of fn() over `steps` calls after `warmup`, bracketed by
    device synchronisation (the library works on torch's default stream)."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        fn()
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps


def amr_main(args, emit=True):
    """Batched AMR hierarchy: level 0 256^3 in uneven boxes of 32-64 cells
    per edge, level 1 (ratio 2) 150 boxes of 16-48, level 2 200 boxes of
    8-32, all 5 scalars + 3 velocities.  The hierarchy pass is timed as one
    graph launch, as an eager multi-level call, and with one fused launch per
    box; the per-cell cost is compared with a single 128^3 box.  The AMR
    restriction that follows an update (AverageDown of the scalars,
    AverageDownFaces of umac) is timed on the same hierarchy."""
    import numpy as np
    import maestroex_amd as mx
    from maestroex_amd.workloads import Hierarchy, Workload

    dev = mx.device_backend("cuda:0")
    t0 = time.perf_counter()
    h = Hierarchy(n0=args.n, level0_edges=(32, 64), counts=(150, 200),
                  edges=((16, 48), (8, 32)), fracs=(0.585, 0.368), ppm_type=args.ppm,
                  backend=dev, verbose=True)
    cells = h.valid_cells()
    print(f"[bench] hierarchy {[len(b) for b in h.boxes]} boxes, {cells} cells, built in "
          f"{time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    m = mx.gpu_maestro(h.params, h.boxes)
    fields = [h.pass_fields(l) for l in range(h.nlev)]
    for kv in args.tune:  # keys other than the edge kernel's shape, which the autotuner sets
        key, val = kv.split("=")
        m.set_tuning(key, int(val))
    cands, best = m.autotune(fields)
    tuned = {"tile": TILE_NAMES[cands[best][0]], "tz": cands[best][1]}
    print(f"[bench] autotune {cands} -> {tuned}", file=sys.stderr, flush=True)
    eager_ms = time_ms(lambda: m.advance_pass_levels(fields), args.steps, args.warmup)
    g = m.pass_graph(fields)
    graph_ms = time_ms(g.launch, args.steps, args.warmup)
    m.set_tuning("edge_scal_per_box", 1)
    per_box_ms = time_ms(lambda: m.advance_pass_levels(fields), args.steps, args.warmup)
    m.set_tuning("edge_scal_per_box", 0)
    g.close()

    # AMR restriction after the update: AverageDown of every scalar and
    # AverageDownFaces of umac, finest level first (HIP events per kernel);
    # descriptors built once, as a C++ caller holds its MultiFab views
    nsc = h.params.nscal
    descs = [(lev, fields[lev + 1]["scal"].desc(), fields[lev]["scal"].desc(),
              fields[lev + 1]["umac"].desc(), fields[lev]["umac"].desc())
             for lev in range(h.nlev - 2, -1, -1)]

    def restrict():
        for lev, fs, cs, fu, cu in descs:
            rc = m.fns["average_down"](m.ctx, lev, C.byref(fs), C.byref(cs), 0, nsc)
            rc = rc or m.fns["average_down_faces"](m.ctx, lev, C.byref(fu), C.byref(cu))
            if rc:
                raise RuntimeError(f"restriction: error {rc}")
    sync_ms = time_ms(restrict, args.steps, args.warmup)
    m.fns["prof_enable"](m.ctx, 1)
    restrict()
    nk = m.fns["prof_read"](m.ctx, -1, None, None, None)
    sync_kernel_ms = 0.0
    for idx in range(nk):
        nm, kms, cnt = C.c_char_p(), C.c_double(), C.c_longlong()
        m.fns["prof_read"](m.ctx, idx, C.byref(nm), C.byref(kms), C.byref(cnt))
        if nm.value.decode() == "k_average_down":
            sync_kernel_ms = kms.value
    m.fns["prof_enable"](m.ctx, 0)
    # FillPatch of the level-0 scalars (150 boxes, 4 ghost cells), as after
    # the update: host planning (box intersections) plus the kernels
    s0 = fields[0]["scal"].desc()
    bs, _ = m.bc_setup()

    def fill0():
        rc = m.fns["fill_patch"](m.ctx, 0, C.byref(s0), C.byref(s0), 0, 0, nsc, bs, 0)
        if rc:
            raise RuntimeError(f"fill_patch: error {rc}")
    fill_ms = time_ms(fill0, args.steps, args.warmup)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        fill0()
    fill_host_ms = (time.perf_counter() - t0) * 1e3 / args.steps  # enqueue time per call
    m.sync()
    fine_cells = sum(int(np.prod([hi[d] - lo[d] + 1 for d in range(3)]))
                     for bl in h.boxes[1:] for (lo, hi) in bl)
    # bytes: every fine cell read and an eighth as many coarse cells written
    # (nscal components); half the fine faces of each direction read and a
    # quarter of that written (1 component)
    sync_bytes = 8 * fine_cells * (nsc * (1 + 1 / 8) + 3 * 0.5 * (1 + 1 / 4))

    # the single-box reference: one 128^3 box, its own autotuned edge-kernel
    # shape, one graph launch per pass (as the hierarchy)
    w = Workload(dim=3, n=128, max_grid=128, ppm_type=args.ppm, backend=dev)
    ms1 = mx.gpu_maestro(w.params, [w.boxes])
    f1 = w.pass_fields()
    cands1, best1 = ms1.autotune([f1])
    g1 = ms1.pass_graph([f1])
    single_ms = time_ms(g1.launch, args.steps, args.warmup)
    g1.close()
    per_cell = graph_ms * 1e6 / cells
    per_cell_1 = single_ms * 1e6 / w.valid_cells()
    value = cells / (graph_ms * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(graph_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"AMR hierarchy: level 0 {args.n}^3 in {len(h.boxes[0])} boxes "
                               f"of 32-64, level 1 {len(h.boxes[1])} boxes of 16-48, level 2 "
                               f"{len(h.boxes[2])} boxes of 8-32, 5 scalars + 3 velocities, "
                               f"ppm_type {args.ppm}, one graph launch per pass",
                   "global_batch": cells, "parallelism": "single GPU"},
        "amr": {"graph_ms": round(graph_ms, 4), "eager_ms": round(eager_ms, 4),
                "per_box_launch_ms": round(per_box_ms, 4),
                "ns_per_cell": round(per_cell, 4),
                "single_box_ns_per_cell": round(per_cell_1, 4),
                "per_cell_ratio_vs_single_box": round(per_cell / per_cell_1, 4),
                "single_box_autotuned": {"tile": TILE_NAMES[cands1[best1][0]],
                                         "tz": cands1[best1][1]},
                "autotuned": tuned,
                "autotune_candidates_ms": [[c[0], c[1], round(c[2], 4)] for c in cands],
                "fill_patch_level0": {"ms": round(fill_ms, 4),
                                      "host_enqueue_ms": round(fill_host_ms, 4),
                                      "boxes": len(h.boxes[0]), "ncomp": nsc},
                "restriction": {"ms": round(sync_ms, 4), "kernel_ms": round(sync_kernel_ms, 4),
                                "fine_cells": fine_cells,
                                "kernel_gbs": round(sync_bytes / (sync_kernel_ms * 1e-3) / 1e9, 1)
                                if sync_kernel_ms > 0 else None}},
        "roofline": None,
        "cpu_baseline": None,
    }
    if emit:
        print(json.dumps(out), flush=True)
    return out


def adjoint_main(args, emit=True):
    """Differentiable step (mx_adv_scal_step / _vjp): 128^3 periodic, 4
    scalars, a fixed divergence-free MAC velocity (max |u| = 1, CFL 0.5).
    Times one forward step and one adjoint step (tape + reverse sweeps), and
    a 20-step rollout gradient with checkpoints every 5 steps."""
    import numpy as np
    import torch
    import maestroex_amd as mx
    from maestroex_amd.adjoint import Rollout, ScalStepper, periodic_mac_velocity

    mx.device_backend("cuda:0")
    n, nc = 128, 4
    u = [torch.from_numpy(x).to("cuda:0") for x in periodic_mac_velocity(n, seed=1)]
    x = (np.arange(n) + 0.5) / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    rng = np.random.default_rng(3)
    s0 = np.stack([np.exp(-((X - 0.5) ** 2 + (Y - 0.3 - 0.1 * c) ** 2 + (Z - 0.5) ** 2) / 0.02)
                   + 0.05 * rng.standard_normal((n, n, n)) for c in range(nc)])
    s = torch.from_numpy(s0).to("cuda:0")
    tgt = torch.from_numpy(rng.standard_normal(s0.shape)).to("cuda:0")
    st = ScalStepper(n, nc, 0.5 / n, 1e-8, ppm_type=args.ppm)
    for kv in args.tune:
        key, val = kv.split("=")
        st.m.set_tuning(key, int(val))
    fwd_ms = time_ms(lambda: st.step(s, u), args.steps, args.warmup)
    g = torch.ones_like(s)
    adj_ms = time_ms(lambda: st.vjp(s, u, g), args.steps, args.warmup)
    # per-kernel times of one forward and one adjoint step (HIP events)
    fns = st.fns
    fns["prof_enable"](st.m.ctx, 1)
    st.step(s, u)
    st.vjp(s, u, g)
    nk = fns["prof_read"](st.m.ctx, -1, None, None, None)
    kern = {}
    for idx in range(nk):
        nm, ms, cnt = C.c_char_p(), C.c_double(), C.c_longlong()
        fns["prof_read"](st.m.ctx, idx, C.byref(nm), C.byref(ms), C.byref(cnt))
        kern[nm.value.decode()] = round(ms.value, 4)
    fns["prof_enable"](st.m.ctx, 0)
    print(f"[bench] step kernels ms: {kern}", file=sys.stderr, flush=True)

    def rollout():
        sd = s.clone().requires_grad_(True)
        ud = [t.clone().requires_grad_(True) for t in u]
        out = Rollout.apply(st, 20, 5, sd, *ud)
        (0.5 * ((out - tgt) ** 2).sum()).backward()

    roll_ms = time_ms(rollout, 1, 1)
    cells = n ** 3
    out = {
        "metric": METRIC,
        "value": round(cells / (fwd_ms * 1e-3) / 1e9, 4),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(fwd_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "differentiable scalar step: 128^3 periodic, 4 scalars, fixed "
                               f"divergence-free MAC velocity, ppm_type {args.ppm}",
                   "global_batch": cells, "parallelism": "single GPU"},
        "adjoint": {"forward_ms": round(fwd_ms, 4), "adjoint_ms": round(adj_ms, 4),
                    "adjoint_over_forward": round(adj_ms / fwd_ms, 3),
                    "rollout_20_steps_grad_ms": round(roll_ms, 2),
                    "rollout_over_forward_steps": round(roll_ms / (20 * fwd_ms), 3)},
        "roofline": None,
        "cpu_baseline": None,
    }
    if emit:
        print(json.dumps(out), flush=True)
    return out


def sphere_adjoint(args, w):
    """The spherical star's rollout gradient with the box-list step and its
    adjoint (mx_adv_scal_step_boxes / _vjp_boxes): rho' and (rho h)' of the
    star (384^3 in 216 uneven boxes, outflow sides, ppm_type 2) advected by
    the MAC velocity of the pass just timed (VelPred with w0).  Times one
    forward step, one adjoint step and the gradient of a 10-step rollout
    with checkpoints every 5 steps (its forward, the recomputed segments and
    the adjoint steps)."""
    from maestroex_amd.adjoint import BoxStepper
    from maestroex_amd.fabs import FaceMultiFab, MultiFab
    n, nc, nsteps, every = args.n, 2, 10, 5
    st = BoxStepper(w.boxes, [n] * 3, [0, 0, 0], nc, w.dt, w.rel_eps, ppm_type=2, spherical=1)
    bk, ng = st.bk, st.NG[2]

    def cells(ngrow):
        return MultiFab(w.boxes, nc, ngrow, 3, bk)

    s0 = cells(ng)
    for b in range(len(w.boxes)):
        s0.arrays[b].copy_(w.scal.arrays[b][0:nc])  # the workload's 4 ghosts
    u = w.umac
    a, b_ = cells(ng), cells(ng)
    for x, y in zip(a.arrays, s0.arrays):
        x.copy_(y)
    fwd_ms = time_ms(lambda: st.step_mf(a, u, b_), args.steps, args.warmup)
    g, sb = cells(0), cells(0)
    for x in g.arrays:
        x.fill_(1.0)
    ub = FaceMultiFab(w.boxes, 1, 0, 3, bk)
    adj_ms = time_ms(lambda: st.vjp_mf(a, u, g, sb, ub), args.steps, args.warmup)
    fns = st.fns
    fns["prof_enable"](st.m.ctx, 1)
    st.step_mf(a, u, b_)
    st.vjp_mf(a, u, g, sb, ub)
    nk = fns["prof_read"](st.m.ctx, -1, None, None, None)
    kern = {}
    for idx in range(nk):
        nm, ms, cnt = C.c_char_p(), C.c_double(), C.c_longlong()
        fns["prof_read"](st.m.ctx, idx, C.byref(nm), C.byref(ms), C.byref(cnt))
        kern[nm.value.decode()] = round(ms.value, 4)
    fns["prof_enable"](st.m.ctx, 0)
    print(f"[bench] spherical step kernels ms: {kern}", file=sys.stderr, flush=True)

    # the rollout: states in preallocated MultiFabs, the loss gradient at
    # the end taken as the final state (a zero target)
    ckpt = [s0, cells(ng)]
    seg_states = [cells(ng) for _ in range(every - 1)]
    work = [cells(ng), cells(ng)]
    gbs = [cells(0), cells(0)]

    def rollout():
        cur, nxt = work
        src = s0
        for k in range(nsteps):
            st.step_mf(src, u, nxt)
            cur, nxt = nxt, cur
            src = cur
            if (k + 1) % every == 0 and k + 1 < nsteps:
                for x, y in zip(ckpt[(k + 1) // every].arrays, cur.arrays):
                    x.copy_(y)
        gb, sbn = gbs
        for bi in range(len(w.boxes)):
            gb.arrays[bi].copy_(cur.arrays[bi][(slice(None),) + cur.valid_region(bi)])
        for seg in reversed(range(nsteps // every)):
            states = [ckpt[seg]]
            for q in range(every - 1):
                st.step_mf(states[-1], u, seg_states[q])
                states.append(seg_states[q])
            for sx in reversed(states):
                st.vjp_mf(sx, u, gb, sbn, ub)
                gb, sbn = sbn, gb

    roll_ms = time_ms(rollout, 1, 1)
    out = {"workload": f"rho' and (rho h)' of the 384^3 star in {len(w.boxes)} uneven boxes, "
                       "outflow sides, ppm_type 2, the pass's MAC velocity",
           "forward_ms": round(fwd_ms, 4), "adjoint_ms": round(adj_ms, 4),
           "adjoint_over_forward": round(adj_ms / fwd_ms, 3),
           "rollout_10_steps_grad_ms": round(roll_ms, 2),
           "rollout_over_forward_steps": round(roll_ms / (nsteps * fwd_ms), 3)}
    print(f"[bench] spherical adjoint: {out}", file=sys.stderr, flush=True)
    return out


# algorithmic bytes per cell of the update stages (3-D planar, NumSpec = 3,
# each face counted once): UpdateScal species 2 + 6 NumSpec doubles (old and
# new density and species, three fluxes and the force of each species),
# enthalpy 6, UpdateVel 22 (old and new velocity, force, MAC velocity, nine
# edge velocities, w0)
UPDATE_BYTES_PER_CELL = {"k_update_rhox": 8 * 20, "k_update_rhoh": 8 * 6, "k_update_vel": 8 * 22}


def update_fields(boxes, nspec, backend, fill):
    """Inputs and outputs of UpdateScal and UpdateVel on one level (3-D);
    fill(array, lo, hi) puts uniform random values into an array."""
    from maestroex_amd.fabs import FaceMultiFab, MultiFab
    nc = 4 + nspec
    f = {
        "sold": MultiFab(boxes, nc, 0, 3, backend), "snew": MultiFab(boxes, nc, 0, 3, backend),
        "force": MultiFab(boxes, nc, 0, 3, backend), "sflux": FaceMultiFab(boxes, nc, 0, 3, backend),
        "uold": MultiFab(boxes, 3, 0, 3, backend), "unew": MultiFab(boxes, 3, 0, 3, backend),
        "vforce": MultiFab(boxes, 3, 0, 3, backend), "umac": FaceMultiFab(boxes, 1, 0, 3, backend),
        "uedge": FaceMultiFab(boxes, 3, 0, 3, backend), "w0c": MultiFab(boxes, 3, 1, 3, backend),
    }
    for key in ("force", "uold", "vforce", "w0c"):
        for a in f[key].arrays:
            fill(a, -1.0, 1.0)
    for key in ("sflux", "umac", "uedge"):
        for d in range(3):
            for a in f[key][d].arrays:
                fill(a, -0.01, 0.01)
    for a in f["sold"].arrays:
        fill(a, 0.2, 1.0)
        a[0] = a[4:4 + nspec].sum(0)  # rho: the sum of the species
    return f


def run_update(m, f, nspec, bcd):
    m.UpdateScal(0, f["sold"], f["snew"], f["sflux"], f["force"], 4, nspec, None, bcd)
    m.UpdateScal(0, f["sold"], f["snew"], f["sflux"], f["force"], 1, 1, None, bcd)
    m.UpdateVel(0, f["uold"], f["unew"], f["umac"], f["uedge"], f["vforce"], None, None,
                f["w0c"], False)


def update_caller(m, f, nspec, bcd):
    """run_update with every descriptor built once, as a C++ caller holds its
    MultiFab views: each call is then the three C-ABI calls alone."""
    from maestroex_amd.fabs import empty_face_desc, empty_mf_desc
    fns, ctx = m.fns, m.ctx
    so, sn, fo = f["sold"].desc(), f["snew"].desc(), f["force"].desc()
    sf = f["sflux"].desc()
    uo, un, vf, w0 = f["uold"].desc(), f["unew"].desc(), f["vforce"].desc(), f["w0c"].desc()
    um, ue = f["umac"].desc(), f["uedge"].desc()
    none_mf, none_face = empty_mf_desc(), empty_face_desc()
    r = [C.byref(x) for x in (so, sn, sf, fo, none_mf, uo, un, um, ue, vf, none_face, w0)]
    keep = (so, sn, fo, sf, uo, un, vf, w0, um, ue, none_mf, none_face)

    def call():
        rc = fns["update_scal"](ctx, 0, r[0], r[1], r[2], r[3], 4, nspec, r[4], bcd, 0)
        rc = rc or fns["update_scal"](ctx, 0, r[0], r[1], r[2], r[3], 1, 1, r[4], bcd, 0)
        rc = rc or fns["update_vel"](ctx, 0, r[5], r[6], r[7], r[8], r[9], r[4], r[10], r[11], 0)
        if rc:
            raise RuntimeError(f"update: error {rc}")
    call.keep = keep
    return call


def update_main(args):
    """UpdateScal (species, enthalpy) and UpdateVel, the update that closes
    the step, on the 3-D 256^3 level of 8 boxes of 128^3 with 3 species:
    HBM-bound stages measured against the HBM roofline."""
    import torch
    import maestroex_amd as mx
    from maestroex_amd.fabs import Backend, Geometry, make_params, split_domain

    dev = mx.device_backend("cuda:0")
    n, nspec, bcd = args.n, 3, 1e-3
    geom = Geometry(3, [n] * 3, [1.0 / n] * 3, [1, 1, 1], [0] * 3, [0] * 3)
    params = make_params(geom, nspec=nspec, dt=1e-3)
    boxes = split_domain(geom, args.max_grid)
    gen = torch.Generator(device="cuda:0").manual_seed(5)

    def fill(a, lo, hi):
        a.uniform_(lo, hi, generator=gen)
    f = update_fields(boxes, nspec, dev, fill)
    m = mx.gpu_maestro(params, [boxes])
    call = update_caller(m, f, nspec, bcd)
    ms = time_ms(call, args.steps, args.warmup)
    fns = m.fns
    fns["prof_enable"](m.ctx, 1)
    for _ in range(args.steps):
        call()
    nk = fns["prof_read"](m.ctx, -1, None, None, None)
    kern = {}
    for idx in range(nk):
        nm, kms, cnt = C.c_char_p(), C.c_double(), C.c_longlong()
        fns["prof_read"](m.ctx, idx, C.byref(nm), C.byref(kms), C.byref(cnt))
        kern[nm.value.decode()] = kms.value / max(cnt.value, 1)
    fns["prof_enable"](m.ctx, 0)
    cells = n ** 3
    per_kernel = {}
    for k, avg in sorted(kern.items()):
        bpc = UPDATE_BYTES_PER_CELL.get(k)
        gbs = bpc * cells / (avg * 1e-3) / 1e9 if bpc else None
        per_kernel[k] = {"avg_launch_ms": round(avg, 4), "bytes_per_cell": bpc,
                         "achieved_gbs": round(gbs, 1) if gbs else None,
                         "frac": round(gbs / HBM_PEAK_GBS, 4) if gbs else None}
    print(f"[bench] update kernels: {per_kernel}", file=sys.stderr, flush=True)
    dom = "k_update_vel"
    copy_gbs = copy_bandwidth_gbs()
    d = per_kernel[dom]
    cpu = None
    if not args.no_cpu:
        cpu = update_cpu_baseline(args, nspec, bcd)
    out = {
        "metric": METRIC,
        "value": round(cells / (ms * 1e-3) / 1e9, 4),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"UpdateScal (3 species, then enthalpy) + UpdateVel on 3-D {n}^3 "
                               f"FP64 as {len(boxes)} boxes of {args.max_grid}^3",
                   "global_batch": cells, "parallelism": "single GPU"},
        "update": per_kernel,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": d["achieved_gbs"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": d["frac"], "traffic": None,
                     "avg_launch_ms": d["avg_launch_ms"], "bytes_per_cell": d["bytes_per_cell"],
                     "measured_copy_gbs": round(copy_gbs, 1),
                     "frac_of_measured_copy": round(d["achieved_gbs"] / copy_gbs, 4)},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def fill_main(args):
    """Maestro::FillPatch at level 0 of every scalar (7 components) on the 3-D
    256^3 level of 8 boxes of 128^3 with 4 ghost cells: periodic x and y,
    outflow z (BCSetup's BCRecs).  HBM-bound: one read and one write per
    ghost cell and component."""
    import numpy as np
    import torch
    import maestroex_amd as mx
    from maestroex_amd import _abi
    from maestroex_amd.fabs import Geometry, MultiFab, make_params, split_domain

    dev = mx.device_backend("cuda:0")
    n, ng = args.n, 4
    out_bc = [_abi.PHYS_INTERIOR, _abi.PHYS_INTERIOR, _abi.PHYS_OUTFLOW]
    geom = Geometry(3, [n] * 3, [1.0 / n] * 3, [1, 1, 0], out_bc, out_bc)
    params = make_params(geom, nspec=3)
    boxes = split_domain(geom, args.max_grid)
    nc = params.nscal
    m = mx.gpu_maestro(params, [boxes])
    mf = MultiFab(boxes, nc, ng, 3, dev)
    gen = torch.Generator(device="cuda:0").manual_seed(9)
    for a in mf.arrays:
        a.uniform_(-1.0, 1.0, generator=gen)
    bs, _ = m.bc_setup()
    desc = mf.desc()

    def call():
        rc = m.fns["fill_patch"](m.ctx, 0, C.byref(desc), C.byref(desc), 0, 0, nc, bs, 0)
        if rc:
            raise RuntimeError(f"fill_patch: error {rc}")
    ms = time_ms(call, args.steps, args.warmup)
    m.fns["prof_enable"](m.ctx, 1)
    for _ in range(args.steps):
        call()
    nk = m.fns["prof_read"](m.ctx, -1, None, None, None)
    kern = {}
    for idx in range(nk):
        nm, kms, cnt = C.c_char_p(), C.c_double(), C.c_longlong()
        m.fns["prof_read"](m.ctx, idx, C.byref(nm), C.byref(kms), C.byref(cnt))
        kern[nm.value.decode()] = (kms.value / args.steps, cnt.value // args.steps)
    m.fns["prof_enable"](m.ctx, 0)
    ghost = sum(int(np.prod([hi[d] - lo[d] + 1 + 2 * ng for d in range(3)]))
                - int(np.prod([hi[d] - lo[d] + 1 for d in range(3)])) for lo, hi in boxes)
    slab = sum(int(np.prod([hi[d] - lo[d] + 1 + 2 * ng for d in range(2)])) * ng
               * ((lo[2] == 0) + (hi[2] == n - 1)) for lo, hi in boxes)
    # one read and one write of 8 bytes per ghost cell and component
    bytes_copy = 16 * nc * (ghost - slab)
    bytes_phys = 16 * nc * slab
    kms_total = sum(v[0] for v in kern.values())
    achieved = 16 * nc * ghost / (kms_total * 1e-3) / 1e9
    per_kernel = {k: {"ms_per_call": round(v[0], 4), "launches_per_call": v[1]}
                  for k, v in sorted(kern.items())}
    per_kernel.get("k_fill_copy", {})["achieved_gbs"] = (
        round(bytes_copy / (kern["k_fill_copy"][0] * 1e-3) / 1e9, 1) if "k_fill_copy" in kern
        else None)
    per_kernel.get("k_fill_phys", {})["achieved_gbs"] = (
        round(bytes_phys / (kern["k_fill_phys"][0] * 1e-3) / 1e9, 1) if "k_fill_phys" in kern
        else None)
    print(f"[bench] fill kernels: {per_kernel}", file=sys.stderr, flush=True)
    cpu = None if args.no_cpu else fill_cpu_baseline(args, ng)
    out = {
        "metric": METRIC,
        "value": round(n ** 3 / (ms * 1e-3) / 1e9, 4),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"FillPatch (level 0) of {nc} scalars, {ng} ghost cells, on 3-D "
                               f"{n}^3 FP64 as {len(boxes)} boxes of {args.max_grid}^3, periodic "
                               f"x/y, outflow z",
                   "global_batch": n ** 3, "ghost_cells": ghost, "parallelism": "single GPU"},
        "fill": per_kernel,
        "roofline": {"bound": "hbm", "kernel": "k_fill_copy + k_fill_phys",
                     "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
                     "kernel_ms_per_call": round(kms_total, 4),
                     "bytes_per_ghost_cell": 16 * nc},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


def plotfile_main(args):
    """mx_adv_write_plotfile of 7 components of the 3-D 256^3 level (8 boxes
    of 128^3, 4 ghost cells, valid cells written): device-to-host copies and
    the file writes, into a scratch directory removed after every write."""
    import shutil
    import tempfile
    import torch
    import maestroex_amd as mx
    from maestroex_amd.fabs import Geometry, MultiFab, make_params, split_domain

    dev = mx.device_backend("cuda:0")
    n = args.n
    geom = Geometry(3, [n] * 3, [1.0 / n] * 3, [1, 1, 1], [0] * 3, [0] * 3)
    params = make_params(geom, nspec=3)
    boxes = split_domain(geom, args.max_grid)
    nc = params.nscal
    m = mx.gpu_maestro(params, [boxes])
    mf = MultiFab(boxes, nc, 4, 3, dev)
    gen = torch.Generator(device="cuda:0").manual_seed(11)
    for a in mf.arrays:
        a.uniform_(-1.0, 1.0, generator=gen)
    names = [f"var{c}" for c in range(nc)]
    root = tempfile.mkdtemp(prefix="mx_plt_", dir=os.environ.get("TMPDIR", "/tmp"))
    reps = max(1, min(args.steps, 3))
    times = []
    try:
        for r in range(reps + 1):
            path = os.path.join(root, f"plt{r:05d}")
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            m.write_plotfile(path, [mf], names, 0.0, [r], [0.0] * 3, [1.0] * 3)
            t = time.perf_counter() - t0
            size = sum(os.path.getsize(os.path.join(d, f)) for d, _, fs in os.walk(path)
                       for f in fs)
            shutil.rmtree(path)
            if r > 0:  # the first write is the warm-up
                times.append(t)
            print(f"[bench] plotfile write {r}: {t * 1e3:.1f} ms, {size / 1e6:.1f} MB",
                  file=sys.stderr, flush=True)
    finally:
        shutil.rmtree(root, ignore_errors=True)
    t = min(times)
    out = {
        "metric": METRIC,
        "value": round(n ** 3 / t / 1e9, 4),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": reps,
        "warmup": 1,
        "ms_per_step": round(t * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"AMReX plotfile of {nc} components of 3-D {n}^3 FP64 as "
                               f"{len(boxes)} boxes of {args.max_grid}^3 (device to host, then "
                               f"the files)", "global_batch": n ** 3,
                   "parallelism": "single GPU"},
        "plotfile": {"bytes": size, "write_gbs": round(size / t / 1e9, 3)},
        "roofline": None,
        "cpu_baseline": None if args.no_cpu else plotfile_cpu_baseline(args, nc),
    }
    print(json.dumps(out), flush=True)


def plotfile_cpu_baseline(args, nc):
    """The plotfile restatement (numpy) writing the same components of one
    128^3 box from host memory."""
    import shutil
    import tempfile
    import numpy as np
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from plotfile_oracle import write_plotfile
    n = args.cpu_n
    a = np.random.default_rng(11).uniform(-1.0, 1.0, (nc, n, n, n))
    root = tempfile.mkdtemp(prefix="mx_plt_cpu_", dir=os.environ.get("TMPDIR", "/tmp"))
    times = []
    try:
        for r in range(args.cpu_reps):
            path = os.path.join(root, f"plt{r}")
            t0 = time.perf_counter()
            write_plotfile(path, 3, [f"var{c}" for c in range(nc)], 0.0, [0.0] * 3, [1.0] * 3,
                           [((0, 0, 0), (n - 1,) * 3)], [[1.0 / n] * 3], [0],
                           [[((0, 0, 0), (n - 1,) * 3)]], [[a]])
            times.append(time.perf_counter() - t0)
            shutil.rmtree(path)
    finally:
        shutil.rmtree(root, ignore_errors=True)
    return {"value": round(n ** 3 / min(times) / 1e9, 5), "unit": "Gcell-updates/s",
            "cores": 1, "kind": "port",
            "sample": f"the plotfile restatement writing {nc} components of one {n}^3 box from "
                      f"host memory, best of {args.cpu_reps}"}


def fill_cpu_baseline(args, ng):
    """The oracle's FillPatch of the same 7 scalars on one 128^3 box (periodic
    x/y, outflow z), timed on the host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle
    from maestroex_amd import _abi
    from maestroex_amd.fabs import Backend, Geometry, MultiFab, make_params, split_domain
    cores = cpu_threads()
    pyoracle.set_threads(cores, fast=True)
    n = args.cpu_n
    out_bc = [_abi.PHYS_INTERIOR, _abi.PHYS_INTERIOR, _abi.PHYS_OUTFLOW]
    geom = Geometry(3, [n] * 3, [1.0 / n] * 3, [1, 1, 0], out_bc, out_bc)
    params = make_params(geom, nspec=3)
    boxes = split_domain(geom, n)
    m = pyoracle.oracle_maestro(params, [boxes], fast=True)
    mf = MultiFab(boxes, params.nscal, ng, 3, Backend("numpy"))
    rng = np.random.default_rng(9)
    for a in mf.arrays:
        a[...] = rng.uniform(-1.0, 1.0, size=a.shape)
    bs, _ = m.bc_setup()
    m.FillPatch(0, mf, mf, 0, 0, params.nscal, 0, bs, 0)
    times = []
    for _ in range(args.cpu_reps):
        t0 = time.perf_counter()
        m.FillPatch(0, mf, mf, 0, 0, params.nscal, 0, bs, 0)
        times.append(time.perf_counter() - t0)
    t = min(times)
    return {"value": round(n ** 3 / t / 1e9, 5), "unit": "Gcell-updates/s", "cores": cores,
            "kind": "port", "sample": f"the oracle's FillPatch of {params.nscal} scalars on one "
                                      f"{n}^3 box, best of {args.cpu_reps}"}


def update_cpu_baseline(args, nspec, bcd):
    """The oracle's UpdateScal + UpdateVel on one 128^3 box, timed on the
    host cores."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import numpy as np
    import pyoracle
    from maestroex_amd.fabs import Backend, Geometry, make_params, split_domain
    cores = cpu_threads()
    pyoracle.set_threads(cores, fast=True)
    n = args.cpu_n
    geom = Geometry(3, [n] * 3, [1.0 / n] * 3, [1, 1, 1], [0] * 3, [0] * 3)
    params = make_params(geom, nspec=nspec, dt=1e-3)
    boxes = split_domain(geom, n)
    rng = np.random.default_rng(5)

    def fill(a, lo, hi):
        a[...] = rng.uniform(lo, hi, size=a.shape)
    f = update_fields(boxes, nspec, Backend("numpy"), fill)
    m = pyoracle.oracle_maestro(params, [boxes], fast=True)
    run_update(m, f, nspec, bcd)
    times = []
    for _ in range(args.cpu_reps):
        t0 = time.perf_counter()
        run_update(m, f, nspec, bcd)
        times.append(time.perf_counter() - t0)
    times.sort()
    med = times[len(times) // 2]
    return {"value": round(n ** 3 / med / 1e9, 6), "unit": "Gcell-updates/s", "cores": cores,
            "kind": "port",
            "sample": f"one {n}^3 box, 3 species, median of {args.cpu_reps} updates after 1 "
                      f"warm-up, OpenMP {cores} threads; {med * 1e3:.1f} ms per update"}


def base_main(args):
    """The base-state <-> Cartesian producers on the 3-D 256^3 level of 8
    boxes of 128^3: planar Put1dArrayOnCart, PutInPertForm and
    ConvertRhoXToX, then on a spherical domain Put1dArrayOnCart (quadratic),
    MakeW0mac and MakeS0mac (linear).  Each writes its output once and reads
    little besides (the base state stays in cache): HBM-bound on the writes."""
    import numpy as np
    import maestroex_amd as mx
    from maestroex_amd import _abi
    from maestroex_amd.fabs import BaseState, FaceMultiFab, Geometry, MultiFab, make_params
    from maestroex_amd.fabs import split_domain

    dev = mx.device_backend("cuda:0")
    n = args.n
    cells = n ** 3
    faces = 3 * n * n * (n + 1)
    # algorithmic bytes per point: the outputs written and the cell fields read
    point_bytes = {"k_put1d": 8, "k_pert_form": 16, "k_convert": 8 * 7, "k_put1d_sph": 24,
                   "k_w0mac": 8, "k_s0mac": 8,
                   # rho, gpi, the w0 force, w, w0_cart in; 3 components out
                   "k_vel_force": 8 * 10,
                   # rho, gpi, w0 force, Utilde . n (3 + 3 normal) in; 3 out
                   "k_vel_force_sph": 8 * 16,
                   # force in and out, state, umac, w0_cart, s0_cart
                   "k_scal_force": 8 * 8, "k_scal_force_sph": 8 * 8}
    per_kernel = {}

    def timed(name, fn, npts):
        ms = time_ms(fn, args.steps, args.warmup)
        gbs = point_bytes[name] * npts / (ms * 1e-3) / 1e9
        per_kernel[name] = {"ms": round(ms, 4), "bytes_per_point": point_bytes[name],
                            "achieved_gbs": round(gbs, 1), "frac": round(gbs / HBM_PEAK_GBS, 4)}

    rng = np.random.default_rng(2)
    for sph in (False, True):
        periodic = [0 if sph else 1] * 3
        geom = Geometry(3, [n] * 3, [1.0 / n] * 3, periodic, [0] * 3, [0] * 3)
        params = make_params(geom, nspec=3, spherical=int(sph))
        boxes = split_domain(geom, args.max_grid)
        m = mx.gpu_maestro(params, [boxes])
        if not sph:
            s0 = BaseState(rng.uniform(0.5, 2.0, (1, n + 1)), dev)
            out = MultiFab(boxes, 1, 0, 3, dev)
            scal = MultiFab(boxes, params.nscal, 0, 3, dev)
            for a in scal.arrays:
                a.uniform_(0.5, 2.0)
            d_s0, d_out, d_scal = s0.desc(), out.desc(), scal.desc()
            f = m.fns
            timed("k_put1d", lambda: f["put_1d_array_on_cart"](m.ctx, 0, C.byref(d_s0), C.byref(d_out),
                                                              0, 0, None), cells)
            timed("k_pert_form", lambda: f["put_in_pert_form"](m.ctx, 0, C.byref(d_scal), C.byref(d_s0),
                                                               0, 1, None), cells)
            timed("k_convert", lambda: f["convert_rhox_to_x"](m.ctx, 0, C.byref(d_scal), 1), cells)
            vf, gpi, w0c = (MultiFab(boxes, 3, 0, 3, dev), MultiFab(boxes, 3, 0, 3, dev),
                            MultiFab(boxes, 3, 1, 3, dev))
            ue = FaceMultiFab(boxes, 3, 0, 3, dev)
            s0c = MultiFab(boxes, 1, 1, 3, dev)
            for a in gpi.arrays + w0c.arrays + s0c.arrays:
                a.uniform_(-1.0, 1.0)
            d_vf, d_gpi, d_w0c, d_ue, d_s0c = vf.desc(), gpi.desc(), w0c.desc(), ue.desc(), s0c.desc()
            grav = BaseState(rng.uniform(-2.0, -0.5, (1, n)), dev)
            d_grav, none_mf = grav.desc(), _abi.MultiFabDesc()
            timed("k_vel_force", lambda: f["make_vel_force"](
                m.ctx, 0, C.byref(d_vf), C.byref(d_ue), C.byref(d_scal), C.byref(d_gpi),
                C.byref(d_gpi), C.byref(d_w0c), C.byref(none_mf), C.byref(d_s0), C.byref(d_grav),
                C.byref(d_s0), 0.5, 1.0, 1, 0, None), cells)
            timed("k_scal_force", lambda: f["modify_scal_force"](
                m.ctx, 0, C.byref(d_scal), C.byref(d_scal), C.byref(d_ue), C.byref(d_s0),
                C.byref(d_s0c), C.byref(d_w0c), C.byref(d_s0), 4, 0, None), cells)
        else:
            dx = 1.0 / n
            drf = dx / 5
            nr = int(np.sqrt(3.0) * (0.5 + 3 * dx) / drf) + 6
            rcc = dev.from_numpy((np.arange(nr) + 0.5) * drf)
            redge = dev.from_numpy(np.arange(nr + 1) * drf)
            g = _abi.SphGeom()
            for d in range(3):
                g.center[d], g.prob_lo[d] = 0.5, 0.0
            g.dr_fine, g.nr_fine = drf, nr
            g.r_cc_loc, g.r_edge_loc = dev.ptr(rcc), dev.ptr(redge)
            g.s0_interp_type, g.w0_interp_type, g.s0mac_interp_type, g.w0mac_interp_type = 3, 2, 2, 2
            s0 = BaseState(rng.uniform(0.5, 2.0, (1, nr)), dev)
            w0 = BaseState(rng.uniform(-1.0, 1.0, (1, nr + 1)), dev)
            vec = MultiFab(boxes, 3, 0, 3, dev)
            mac = FaceMultiFab(boxes, 1, 1, 3, dev)
            d_s0, d_w0, d_vec, d_mac = s0.desc(), w0.desc(), vec.desc(), mac.desc()
            none = _abi.MultiFabDesc()
            f = m.fns
            timed("k_put1d_sph", lambda: f["put_1d_array_on_cart"](
                m.ctx, 0, C.byref(d_w0), C.byref(d_vec), 1, 1, C.byref(g)), cells)
            timed("k_w0mac", lambda: f["make_w0mac"](m.ctx, 0, C.byref(d_w0), C.byref(none),
                                                     C.byref(d_mac), C.byref(g)), faces)
            timed("k_s0mac", lambda: f["make_s0mac"](m.ctx, 0, C.byref(d_s0), C.byref(none),
                                                     C.byref(d_mac), C.byref(g)), faces)
            scal = MultiFab(boxes, params.nscal, 0, 3, dev)
            gpi, nrm, s0c = (MultiFab(boxes, 3, 0, 3, dev), MultiFab(boxes, 3, 0, 3, dev),
                             MultiFab(boxes, 1, 1, 3, dev))
            ue = FaceMultiFab(boxes, 1, 0, 3, dev)
            for a in scal.arrays:
                a.uniform_(0.5, 2.0)
            for a in gpi.arrays + nrm.arrays + s0c.arrays:
                a.uniform_(-1.0, 1.0)
            vf = MultiFab(boxes, 3, 0, 3, dev)
            grav = BaseState(rng.uniform(-2.0, -0.5, (1, nr)), dev)
            d_scal, d_gpi, d_nrm, d_s0c, d_ue, d_vf, d_grav = (
                scal.desc(), gpi.desc(), nrm.desc(), s0c.desc(), ue.desc(), vf.desc(), grav.desc())
            timed("k_vel_force_sph", lambda: f["make_vel_force"](
                m.ctx, 0, C.byref(d_vf), C.byref(d_ue), C.byref(d_scal), C.byref(d_gpi),
                C.byref(d_gpi), C.byref(none), C.byref(d_nrm), C.byref(d_s0), C.byref(d_grav),
                C.byref(d_w0), 0.5, 1.0, 1, 0, C.byref(g)), cells)
            timed("k_scal_force_sph", lambda: f["modify_scal_force"](
                m.ctx, 0, C.byref(d_scal), C.byref(d_scal), C.byref(d_ue), C.byref(d_w0),
                C.byref(d_s0c), C.byref(none), C.byref(d_w0), 4, 0, C.byref(g)), cells)
        del m
    print(f"[bench] base-state producers: {per_kernel}", file=sys.stderr, flush=True)
    total_ms = sum(v["ms"] for v in per_kernel.values())
    dom = "k_w0mac"
    d = per_kernel[dom]
    out = {
        "metric": METRIC,
        "value": round(len(per_kernel) * cells / (total_ms * 1e-3) / 1e9, 4),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(total_ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"base-state producers on 3-D {n}^3 FP64 as 8 boxes of "
                               f"{args.max_grid}^3: planar Put1dArrayOnCart, PutInPertForm, "
                               "ConvertRhoXToX, MakeVelForce, ModifyScalForce; spherical "
                               "Put1dArrayOnCart (w0, vector), MakeW0mac and MakeS0mac (linear), "
                               "MakeVelForce, ModifyScalForce; value counts the 10 calls",
                   "global_batch": cells, "parallelism": "single GPU"},
        "base": per_kernel,
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": d["achieved_gbs"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": d["frac"], "traffic": None},
        "cpu_baseline": None,
    }
    print(json.dumps(out), flush=True)


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    if args.amr:
        return amr_main(args)
    if args.adjoint:
        return adjoint_main(args)
    if args.update:
        return update_main(args)
    if args.base:
        return base_main(args)
    if args.fill:
        return fill_main(args)
    if args.plotfile:
        return plotfile_main(args)
    import torch

    import maestroex_amd as mx
    from maestroex_amd.workloads import Workload

    dev = mx.device_backend("cuda:0")
    t0 = time.perf_counter()
    if args.spherical:
        from maestroex_amd.workloads import uneven_boxes
        args.n, args.ppm, args.no_cpu = 384, 2, True
        w = Workload(dim=3, n=args.n, ppm_type=args.ppm, spherical=True, backend=dev,
                     boxes=uneven_boxes(args.n, 6, 48, 96, seed=4))
        workload = (f"3-D spherical star 384^3 FP64 in {len(w.boxes)} uneven boxes (edges 48-96), "
                    f"4 ghosts, 5 scalars + 3 velocities, radial w0 = 0.05 r (1 - r), ppm_type 2")
        bytes_per_cell = PASS_BYTES_PER_CELL_SPHERICAL
    elif args.dim == 2:
        args.max_grid = args.n
        w = Workload(dim=2, n=args.n, max_grid=args.n, ppm_type=args.ppm, backend=dev)
        workload = (f"2-D {args.n}^2 FP64 as one box, 4 ghosts, periodic x, outflow y, "
                    f"5 scalars + 2 velocities, ppm_type {args.ppm}")
        bytes_per_cell = PASS_BYTES_PER_CELL[2]
    else:
        w = Workload(dim=3, n=args.n, max_grid=args.max_grid, ppm_type=args.ppm, backend=dev)
        workload = (f"3-D {args.n}^3 FP64 as {len(w.boxes)} boxes of {args.max_grid}^3, "
                    f"4 ghosts, 5 scalars + 3 velocities, ppm_type {args.ppm}")
        bytes_per_cell = PASS_BYTES_PER_CELL[3]
    print(f"[bench] workload {args.n}^3 in {len(w.boxes)} boxes built in "
          f"{time.perf_counter() - t0:.1f}s", file=sys.stderr, flush=True)
    m = mx.gpu_maestro(w.params, [w.boxes])
    lib, fns = mx.load()
    fields = w.pass_fields()
    tuned = None
    # the library's defaults
    shape = {"tile": "16x11", "tz": 32}
    if args.autotune and not args.no_autotune:
        # the autotuner picks the fused edge kernel's tile and z-chunk on
        # this workload (untimed, before the warm-up)
        cands, best = m.autotune([fields])
        tuned = {"tile": TILE_NAMES[cands[best][0]], "tz": cands[best][1]}
        shape = dict(tuned)
        print(f"[bench] autotune {[(c[0], c[1], round(c[2], 3)) for c in cands]} -> {tuned}",
              file=sys.stderr, flush=True)
    for kv in args.tune:
        key, val = kv.split("=")
        m.set_tuning(key, int(val))
        if key == "edge_scal_tile":
            shape["tile"] = TILE_NAMES[int(val)]
        elif key == "edge_scal_tz":
            shape["tz"] = int(val)

    if args.calib_copy_mib:
        n = args.calib_copy_mib * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda:0")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    # timed regions, each bracketed by events on the default stream (the
    # library's stream): (1) eager passes; (2) the pass as one HIP graph
    # launch (the measured value); (3) eager passes with the library's
    # per-kernel events, for the kernel breakdown and the roofline
    eager_ms = time_ms(lambda: m.advance_pass(0, fields), args.steps, args.warmup)
    graph = m.pass_graph([fields])
    total_ms = time_ms(graph.launch, args.steps, args.warmup) * args.steps
    graph.close()
    rc = fns["prof_enable"](m.ctx, 1)
    assert rc == 0, rc
    for _ in range(args.steps):
        m.advance_pass(0, fields)
    nk = fns["prof_read"](m.ctx, -1, None, None, None)
    kernels = {}
    for idx in range(nk):
        nm = C.c_char_p()
        ms = C.c_double()
        cnt = C.c_longlong()
        fns["prof_read"](m.ctx, idx, C.byref(nm), C.byref(ms), C.byref(cnt))
        kernels[nm.value.decode()] = (ms.value, cnt.value)
    fns["prof_enable"](m.ctx, 0)

    cells = w.valid_cells()
    ms_per_step = total_ms / args.steps
    value = cells / (ms_per_step * 1e-3) / 1e9
    pass_gbs = bytes_per_cell * cells / (ms_per_step * 1e-3) / 1e9

    dom = max(kernels.items(), key=lambda kv: kv[1][0])
    dname, (dms, dcnt) = dom
    avg_ms = dms / dcnt
    # units one launch of the dominant kernel processes: the valid cells of
    # all boxes (x the components the launch covers, folded into bytes/cell)
    ncomp_launch = {"k_es_ppm": None, "k_es_trans": None, "k_es_edge": None, "k_es_fused": None}
    if dname in ncomp_launch:
        # the pass's 5 edge-state components (3 species, rho', (rho h)')
        # over the kernel's launches per pass (1: all in one launch; 3:
        # species, then rho, then rhoh)
        nc = 5.0 / (dcnt / args.steps)
    else:
        nc = 1
    bpc = kernel_bytes_per_cell(dname, nc, args.dim)
    if (dname == "k_es_fused" and not args.spherical and nc == 5 and
            "rhoh_flux_epilogue=0" not in args.tune):
        # the (rho h)' blocks also write the enthalpy flux (the pass's
        # MakeRhoHFlux for enthalpy_pred_type 1 in the kernel's epilogue)
        bpc += 8 * args.dim
    achieved = bpc * cells / (avg_ms * 1e-3) / 1e9 if bpc else None
    roofline = {
        "bound": "hbm",
        "kernel": dname,
        "achieved": round(achieved, 2) if achieved else None,
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
        "traffic": committed_traffic(
            dname, f"{args.n}c_ppm{args.ppm}_{shape['tile']}_tz{shape['tz']}"),
        "avg_launch_ms": round(avg_ms, 4),
        "bytes_per_cell": bpc,
        "pass_achieved": round(pass_gbs, 2),
        "pass_frac": round(pass_gbs / HBM_PEAK_GBS, 4),
    }
    breakdown = {k: round(v[0] / args.steps, 4) for k, v in sorted(kernels.items(),
                                                                    key=lambda kv: -kv[1][0])}
    print(f"[bench] per-pass kernel ms: {breakdown}", file=sys.stderr, flush=True)

    if dname == "k_es_fused" and args.ppm == 1 and not args.spherical:
        # the kernel's FP64 operation rate against the VALU rate (the
        # co-binding roofline)
        ops = EDGE_OPS_PER_CELL_COMP * nc * cells
        tops = ops / (avg_ms * 1e-3) / 1e12
        roofline["fp64_ops_per_launch"] = int(ops)
        roofline["fp64_tops"] = round(tops, 3)
        roofline["fp64_peak_tops"] = FP64_PEAK_TOPS
        roofline["fp64_frac"] = round(tops / FP64_PEAK_TOPS, 4)
    copy_gbs = copy_bandwidth_gbs()
    roofline["measured_copy_gbs"] = round(copy_gbs, 1)
    roofline["frac_of_measured_copy"] = round(achieved / copy_gbs, 4) if achieved else None
    print(f"[bench] device copy {copy_gbs:.0f} GB/s", file=sys.stderr, flush=True)

    extras = {}
    if not args.headline_only and not args.spherical and args.dim == 3:
        extras = rest_of_metric(args, w, fields)
    if not args.headline_only and args.spherical:
        extras["sphere_adjoint"] = sphere_adjoint(args, w)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args)

    out = {
        "metric": METRIC,
        "value": round(value, 4),
        "unit": "Gcell-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": workload,
                   "global_batch": cells, "parallelism": "single GPU",
                   "pass_launch": "one HIP graph launch per pass",
                   "eager_ms_per_step": round(eager_ms, 4),
                   "edge_kernel_launch": dict(shape, threads=TILE_THREADS[TILE_NAMES.index(shape["tile"])]),
                   "autotuned": tuned is not None},
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    out.update(extras)
    print(json.dumps(out), flush=True)


def rest_of_metric(args, w, fields):
    """The other parts of the metric, on the same run: the same 256^3
    workload with ppm_type 2 (graph launch per pass), the batched AMR
    hierarchy's per-cell cost against a single 128^3 box, and the adjoint /
    forward time ratio of the differentiable step (128^3, 4 scalars)."""
    import copy
    import maestroex_amd as mx
    out = {}
    # ppm_type 2: a second context on the same fields (4 ghost cells)
    p2 = mx._abi.Params()
    C.memmove(C.byref(p2), C.byref(w.params), C.sizeof(mx._abi.Params))
    p2.ppm_type = 2
    m2 = mx.gpu_maestro(p2, [w.boxes])
    g2 = m2.pass_graph([fields])
    ms2 = time_ms(g2.launch, args.steps, args.warmup)
    g2.close()
    cells = w.valid_cells()
    out["ppm2"] = {"ms_per_step": round(ms2, 4),
                   "value": round(cells / (ms2 * 1e-3) / 1e9, 4),
                   "pass_frac": round(PASS_BYTES_PER_CELL[3] * cells / (ms2 * 1e-3) / 1e9
                                      / HBM_PEAK_GBS, 4)}
    m2.close()
    print(f"[bench] ppm_type 2: {out['ppm2']}", file=sys.stderr, flush=True)
    a = copy.copy(args)
    a.ppm = 1
    amr = amr_main(a, emit=False)
    out["amr"] = {k: amr["amr"][k] for k in ("graph_ms", "ns_per_cell", "single_box_ns_per_cell",
                                              "per_cell_ratio_vs_single_box")}
    out["amr"]["cells"] = amr["config"]["global_batch"]
    out["amr"]["value"] = amr["value"]
    print(f"[bench] AMR: {out['amr']}", file=sys.stderr, flush=True)
    adj = adjoint_main(a, emit=False)
    out["adjoint"] = dict(adj["adjoint"])
    print(f"[bench] adjoint: {out['adjoint']}", file=sys.stderr, flush=True)
    return out


if __name__ == "__main__":
    main()
