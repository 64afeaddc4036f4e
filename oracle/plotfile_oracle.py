"""CPU restatement of amrex::WriteMultiLevelPlotfile (the Header of
WriteGenericPlotfileHeader, then VisMF::Write of each level with one data
file), as Maestro::WritePlotFile calls it (MaestroPlot.cpp:99-116).

TEST INFRASTRUCTURE ONLY: the parity tests compare the library's plotfile
with this one byte for byte.  AMReX is an un-vendored submodule of the
reference (external/amrex is empty, its pinned commit not recorded), so the
layout is restated from AMReX's published source:
  Header: "HyperCLaw-V1.1", ncomp, the names, dim, time (%.17g), finest
    level, prob_lo and prob_hi (each value followed by a space), the
    refinement ratios, the level domains "((lo) (hi) (0,..))", the level
    steps, dx per level, coordinate system 0, boundary width 0, then per
    level "lev nboxes time", the step, the physical lo/hi of every box per
    direction, and "Level_<lev>/Cell";
  Cell_H: VisMF version 1, how 1 (NFiles), ncomp, ngrow 0, the BoxArray
    "(n 0" ... ")", the count and "FabOnDisk: Cell_D_00000 <offset>" per
    FAB, a blank line, then "n,ncomp" and per FAB the minima, then the
    maxima, each value %.16e followed by a comma, each table followed by a
    blank line;
  Cell_D_00000: per FAB "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (1 2 3 4 5
    6 7 8)))" + box + " ncomp\\n", then the values as little-endian
    doubles, i fastest, component outermost."""
import os

import numpy as np

DESCR = "FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (1 2 3 4 5 6 7 8)))"


def box_str(lo, hi, dim):
    j = lambda v: ",".join(str(int(x)) for x in v[:dim])
    return f"(({j(lo)}) ({j(hi)}) ({','.join('0' * dim)}))"


def g17(v):
    return "%.17g" % v


def write_plotfile(path, dim, varnames, time, prob_lo, prob_hi, domains, dxs, level_steps,
                   boxes, data):
    """boxes[lev]: [(lo, hi)]; data[lev][b]: array (ncomp, nz, ny, nx) of the
    valid cells; domains[lev]: (lo, hi)."""
    nlev = len(boxes)
    ncomp = len(varnames)
    os.makedirs(path, exist_ok=True)
    h = ["HyperCLaw-V1.1", str(ncomp)] + list(varnames)
    h += [str(dim), g17(time), str(nlev - 1)]
    h.append("".join(g17(prob_lo[d]) + " " for d in range(dim)))
    h.append("".join(g17(prob_hi[d]) + " " for d in range(dim)))
    ratios = []
    for lev in range(nlev - 1):
        nc = domains[lev][1][0] - domains[lev][0][0] + 1
        nf = domains[lev + 1][1][0] - domains[lev + 1][0][0] + 1
        ratios.append(nf // nc)
    h.append("".join(f"{r} " for r in ratios))
    h.append("".join(box_str(lo, hi, dim) + " " for lo, hi in domains))
    h.append("".join(f"{s} " for s in level_steps))
    for lev in range(nlev):
        h.append("".join(g17(dxs[lev][d]) + " " for d in range(dim)))
    h += ["0", "0"]
    for lev in range(nlev):
        h.append(f"{lev} {len(boxes[lev])} {g17(time)}")
        h.append(str(level_steps[lev]))
        for lo, hi in boxes[lev]:
            for d in range(dim):
                dx = dxs[lev][d]
                a = lo[d] - domains[lev][0][d]
                b = hi[d] - domains[lev][0][d]
                h.append(g17(prob_lo[d] + dx * a) + " " + g17(prob_lo[d] + dx * (b + 1)))
        h.append(f"Level_{lev}/Cell")
    with open(os.path.join(path, "Header"), "w") as f:
        f.write("\n".join(h) + "\n")
    for lev in range(nlev):
        ldir = os.path.join(path, f"Level_{lev}")
        os.makedirs(ldir, exist_ok=True)
        offs, mins, maxs = [], [], []
        off = 0
        with open(os.path.join(ldir, "Cell_D_00000"), "wb") as f:
            for (lo, hi), a in zip(boxes[lev], data[lev]):
                head = (DESCR + box_str(lo, hi, dim) + f" {ncomp}\n").encode()
                body = np.ascontiguousarray(a[:ncomp], dtype="<f8").tobytes()
                offs.append(off)
                f.write(head)
                f.write(body)
                off += len(head) + len(body)
                flat = a[:ncomp].reshape(ncomp, -1)
                mins.append([min(flat[n].tolist()) for n in range(ncomp)])
                maxs.append([max(flat[n].tolist()) for n in range(ncomp)])
        c = ["1", "1", str(ncomp), "0", f"({len(boxes[lev])} 0"]
        c += [box_str(lo, hi, dim) for lo, hi in boxes[lev]]
        c += [")", str(len(boxes[lev]))]
        c += [f"FabOnDisk: Cell_D_00000 {o}" for o in offs]
        c.append("")
        for t in (mins, maxs):
            c.append(f"{len(boxes[lev])},{ncomp}")
            c += ["".join("%.16e," % v for v in row) for row in t]
            c.append("")
        with open(os.path.join(ldir, "Cell_H"), "w") as f:
            f.write("\n".join(c) + "\n")
