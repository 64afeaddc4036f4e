"""Build libmaestroex_amd.so (HIP, gfx950) in-tree.

Compile flags that matter for results:
  -ffp-contract=off   no FMA contraction: every expression rounds as the
                      reference's CPU build does (bit-exact parity);
  -fno-fast-math      IEEE division and square root;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions.

Each source compiles to its own object (in parallel), then one link makes
the library.  build() recompiles every source; build(incremental=True)
recompiles only sources newer than their object (or than any header).
"""
import glob
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = ["mx_ctx.hip", "mx_velocity.hip", "mx_scalar.hip", "mx_edge_fused.hip", "mx_edge1d.hip",
           "mx_graph.hip", "mx_adjoint.hip", "mx_update.hip", "mx_base.hip", "mx_sync.hip",
           "mx_fill.hip", "mx_plotfile.hip", "mx_velpred_fused.hip", "mx_noalloc.hip",
           "mx_amr.hip", "mx_fluxreg.hip"]
LIB = os.path.join(HERE, "lib", "libmaestroex_amd.so")
OBJ = os.path.join(HERE, "_build")
ARCH = os.environ.get("MX_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Werror=return-type",
]


def _headers():
    return (glob.glob(os.path.join(HERE, "csrc", "*.H")) +
            glob.glob(os.path.join(ROOT, "include", "*.h")))


def _compile(src, obj, verbose):
    cmd = ["hipcc", *FLAGS, "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc"),
           "-c", "-o", obj + ".tmp", src]
    if verbose:
        print("[build]", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return f"{os.path.basename(src)}:\n{r.stdout}{r.stderr}"
    os.replace(obj + ".tmp", obj)
    return None


def build(verbose=True, incremental=False, jobs=None):
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    os.makedirs(OBJ, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(HERE, "csrc", s))]
    hdr_t = max([os.path.getmtime(h) for h in _headers()] + [0.0])
    work = []
    objs = []
    for s in srcs:
        src = os.path.join(HERE, "csrc", s)
        obj = os.path.join(OBJ, s.replace(".hip", ".o"))
        objs.append(obj)
        if incremental and os.path.exists(obj):
            t = os.path.getmtime(obj)
            if t >= os.path.getmtime(src) and t >= hdr_t:
                continue
        work.append((src, obj))
    jobs = jobs or min(8, max(1, len(work)))
    errs = []
    if work:
        with ThreadPoolExecutor(max_workers=jobs) as ex:
            for e in ex.map(lambda w: _compile(w[0], w[1], verbose), work):
                if e:
                    errs.append(e)
    if errs:
        sys.stderr.write("\n".join(errs))
        raise RuntimeError("hipcc failed building libmaestroex_amd.so")
    tmp = LIB + ".tmp"
    cmd = ["hipcc", f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs]
    if verbose:
        print("[build]", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed linking libmaestroex_amd.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build(incremental="-i" in sys.argv[1:])
