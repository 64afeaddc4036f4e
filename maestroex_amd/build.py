"""Build libmaestroex_amd.so (HIP, gfx950) in-tree.

Compile flags that matter for results:
  -ffp-contract=off   no FMA contraction: every expression rounds as the
                      reference's CPU build does (bit-exact parity);
  -fno-fast-math      IEEE division and square root;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed FP32 instructions.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SOURCES = ["mx_ctx.hip", "mx_velocity.hip", "mx_scalar.hip", "mx_edge_fused.hip", "mx_edge1d.hip",
           "mx_graph.hip", "mx_adjoint.hip", "mx_update.hip", "mx_base.hip", "mx_sync.hip",
           "mx_fill.hip", "mx_plotfile.hip"]
LIB = os.path.join(HERE, "lib", "libmaestroex_amd.so")
ARCH = os.environ.get("MX_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-shared",
    "-ffp-contract=off",
    "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Werror=return-type",
]


def build(verbose=True):
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    srcs = [os.path.join(HERE, "csrc", s) for s in SOURCES]
    tmp = LIB + ".tmp"
    cmd = ["hipcc", *FLAGS, "-I", os.path.join(ROOT, "include"), "-I", os.path.join(HERE, "csrc"),
           "-o", tmp, *srcs]
    if verbose:
        print("[build]", " ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("hipcc failed building libmaestroex_amd.so")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    build()
