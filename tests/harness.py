"""Shared helpers of the parity and oracle tests.

Builds the same seeded workload on the host (for the CPU oracle) and on the
device (for the HIP library), runs the same call sequence on both and
compares the outputs point by point.
"""
import numpy as np

import pyoracle
from maestroex_amd import _abi
from maestroex_amd.fabs import Backend, FaceMultiFab, MultiFab
from maestroex_amd.workloads import Workload


def oracle_for(w):
    return pyoracle.oracle_maestro(w.params, [w.boxes])


def gpu_for(w):
    import maestroex_amd as mx
    return mx.gpu_maestro(w.params, [w.boxes])


def pair(**cfg):
    """(host workload, device workload) with identical inputs."""
    import maestroex_amd as mx
    wc = Workload(backend=Backend("numpy"), **cfg)
    wg = Workload(backend=mx.device_backend(), **cfg)
    return wc, wg


def mf_arrays(mf, region="valid"):
    out = []
    for b, a in enumerate(mf.to_numpy()):
        out.append(a[(slice(None),) + mf.valid_region(b)] if region == "valid" else a)
    return out


def compare_mf(a, b, region="valid", comps=None):
    """-> dict(max_abs, max_rel, ndiff, n) over all boxes (and comps)."""
    xa = mf_arrays(a, region)
    xb = mf_arrays(b, region)
    max_abs = 0.0
    scale = 0.0
    ndiff = 0
    n = 0
    for u, v in zip(xa, xb):
        if comps is not None:
            u = u[list(comps)]
            v = v[list(comps)]
        d = np.abs(u - v)
        max_abs = max(max_abs, float(np.max(d)) if d.size else 0.0)
        scale = max(scale, float(np.max(np.abs(u))) if u.size else 0.0)
        ndiff += int(np.count_nonzero(u != v))
        n += u.size
    rel = max_abs / max(scale, 1e-300)
    return dict(max_abs=max_abs, max_rel=rel, ndiff=ndiff, n=n)


def compare_face(a, b, region="valid", comps=None):
    res = [compare_mf(a[d], b[d], region, comps) for d in range(a.dim)]
    return dict(max_abs=max(r["max_abs"] for r in res), max_rel=max(r["max_rel"] for r in res),
                ndiff=sum(r["ndiff"] for r in res), n=sum(r["n"] for r in res))


def record_parity(case, array, res):
    """Print and keep the parity metrics of one output array: max relative
    error ||gpu - cpu||_inf / ||cpu||_inf and the count of differing values.
    Lines go to stdout and, when the folder gpurun_out/ exists at the repo
    root (a GPU run), to gpurun_out/parity_metrics.jsonl."""
    import json
    import os
    rec = {"case": case, "array": array, "max_rel": res["max_rel"], "max_abs": res["max_abs"],
           "ndiff": res["ndiff"], "n": res["n"]}
    print(f"[parity-metric] {json.dumps(rec)}", flush=True)
    out = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gpurun_out")
    if os.path.isdir(out):
        with open(os.path.join(out, "parity_metrics.jsonl"), "a") as f:
            f.write(json.dumps(rec) + "\n")
    return res


def assert_bitexact(res, what):
    assert res["ndiff"] == 0, f"{what}: {res['ndiff']} of {res['n']} values differ (max rel {res['max_rel']:.3e})"


def global_face_array(mf, dim, d, n_cells, comp=0):
    """Assemble the valid faces of a face MultiFab (normal d) into one array
    indexed [k, j, i] over the whole domain."""
    shape = [n_cells[q] + (1 if q == d else 0) for q in range(dim)]
    shape = shape + [1] * (3 - dim)
    g = np.full((shape[2], shape[1], shape[0]), np.nan)
    for b, a in enumerate(mf.to_numpy()):
        lo, hi = mf.boxes[b]
        reg = mf.valid_region(b)
        v = a[comp][reg]
        hi2 = list(hi)
        hi2[d] += 1
        sl = tuple(slice(lo[q], hi2[q] + 1) if q < dim else slice(0, 1) for q in (2, 1, 0))
        g[sl] = v
    return g
