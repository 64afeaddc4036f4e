"""Reader of the AMReX plotfile that mx_adv_write_plotfile writes
(amrex::WriteMultiLevelPlotfile, format HyperCLaw-V1.1, as Maestro's
WritePlotFile calls it, MaestroPlot.cpp:99-116): Header, then per level the
VisMF header Level_<l>/Cell_H and the FAB data it points at."""
import os
import re

import numpy as np

_BOX = re.compile(r"\(\(([-\d,]+)\) \(([-\d,]+)\) \(([\d,]+)\)\)")


def _box(text):
    m = _BOX.search(text)
    if not m:
        raise ValueError(f"not a box: {text!r}")
    lo = tuple(int(v) for v in m.group(1).split(","))
    hi = tuple(int(v) for v in m.group(2).split(","))
    return lo, hi


class PlotFile:
    """Header fields and, per level, the boxes, their data (ncomp, nz, ny, nx)
    and the per-FAB minima and maxima of Cell_H."""

    def __init__(self, path):
        self.path = path
        with open(os.path.join(path, "Header")) as f:
            lines = f.read().split("\n")
        it = iter(lines)
        self.version = next(it)
        if self.version != "HyperCLaw-V1.1":
            raise ValueError(f"unknown plotfile version {self.version!r}")
        ncomp = int(next(it))
        self.varnames = [next(it) for _ in range(ncomp)]
        self.dim = int(next(it))
        self.time = float(next(it))
        self.finest_level = int(next(it))
        self.prob_lo = [float(v) for v in next(it).split()]
        self.prob_hi = [float(v) for v in next(it).split()]
        self.ref_ratio = [int(v) for v in next(it).split()]
        self.domains = [_box(m.group(0)) for m in _BOX.finditer(next(it))]
        self.level_steps = [int(v) for v in next(it).split()]
        self.dx = [[float(v) for v in next(it).split()] for _ in range(self.finest_level + 1)]
        self.coord_sys = int(next(it))
        self.bwidth = int(next(it))
        self.levels = []
        for lev in range(self.finest_level + 1):
            lv, nbox, t = next(it).split()
            if int(lv) != lev:
                raise ValueError("levels out of order")
            step = int(next(it))
            phys = [[tuple(float(v) for v in next(it).split()) for _ in range(self.dim)]
                    for _ in range(int(nbox))]
            prefix = next(it)
            self.levels.append(self._level(lev, prefix, float(t), step, phys))

    def _level(self, lev, prefix, t, step, phys):
        d = os.path.dirname(prefix)
        with open(os.path.join(self.path, prefix + "_H")) as f:
            lines = f.read().split("\n")
        it = iter(lines)
        vers, how, ncomp, ngrow = (int(next(it)) for _ in range(4))
        head = next(it)
        nbox = int(head[1:].split()[0])
        boxes = [_box(next(it)) for _ in range(nbox)]
        if next(it) != ")":
            raise ValueError("bad BoxArray")
        nfod = int(next(it))
        fods = []
        for _ in range(nfod):
            _, name, off = next(it).split()
            fods.append((name, int(off)))
        next(it)

        def table():
            n, m = (int(v) for v in next(it).split(","))
            rows = [[float(v) for v in next(it).rstrip(",").split(",")] for _ in range(n)]
            next(it)
            return np.array(rows).reshape(n, m)
        mins, maxs = table(), table()
        data = []
        for (lo, hi), (name, off) in zip(boxes, fods):
            with open(os.path.join(self.path, d, name), "rb") as f:
                f.seek(off)
                line = f.readline().decode()
                if not line.startswith("FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (1 2 3 4 5 6 7 8)))"):
                    raise ValueError(f"unsupported FAB header {line!r}")
                flo, fhi = _box(line)
                nc = int(line.rsplit(")", 1)[1])
                ext = [fhi[q] - flo[q] + 1 for q in range(self.dim)] + [1] * (3 - self.dim)
                a = np.fromfile(f, dtype="<f8", count=nc * ext[0] * ext[1] * ext[2])
                data.append(a.reshape(nc, ext[2], ext[1], ext[0]))
        return {"time": t, "step": step, "phys": phys, "boxes": boxes, "data": data,
                "ncomp": ncomp, "ngrow": ngrow, "vers": vers, "how": how,
                "min": mins, "max": maxs}


def read_plotfile(path):
    return PlotFile(path)
