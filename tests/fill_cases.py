"""Layouts and fields for the level-0 FillPatch tests: a domain of n cells per
direction split into boxes of max_grid, with periodic and bounded directions;
valid cells hold a field of the global index, ghost cells a sentinel."""
import numpy as np

from maestroex_amd import _abi
from maestroex_amd.fabs import Geometry, MultiFab, make_params, split_domain

SENTINEL = -777.0
FOEXTRAP, HOEXTRAP, EXT_DIR = 2, 4, 3
REFLECT_EVEN, REFLECT_ODD, INT_DIR = 1, -1, 0


def layout(dim=3, n=8, max_grid=4, periodic=(1, 1, 0), nspec=3):
    per = list(periodic)[:dim]
    phys = [_abi.PHYS_INTERIOR if p else _abi.PHYS_OUTFLOW for p in per]
    geom = Geometry(dim, [n] * dim, [1.0 / n] * dim, per, phys, phys)
    params = make_params(geom, nspec=nspec)
    return params, split_domain(geom, max_grid)


def fill_valid(m, fn, sentinel=SENTINEL):
    """valid cells: fn(I, J, K, c); ghost cells: the sentinel"""
    for b, (lo, hi) in enumerate(m.boxes):
        a = m.arrays[b]
        beg = m.begin[b]
        nc, nz, ny, nx = a.shape
        K, J, I = np.meshgrid(np.arange(beg[2], beg[2] + nz), np.arange(beg[1], beg[1] + ny),
                              np.arange(beg[0], beg[0] + nx), indexing="ij")
        inside = np.ones(I.shape, bool)
        for d, X in enumerate((I, J, K)):
            if d < m.dim:
                inside &= (X >= lo[d]) & (X <= hi[d])
        v = np.stack([np.where(inside, fn(I, J, K, c), sentinel) for c in range(nc)])
        m.arrays[b] = m.backend.from_numpy(np.ascontiguousarray(v))


def new_mf(boxes, ncomp, ng, dim, backend, fn, sentinel=SENTINEL):
    m = MultiFab(boxes, ncomp, ng, dim, backend)
    fill_valid(m, fn, sentinel)
    return m


def bcs(ncomp, lo, hi):
    """the same (lo, hi) BCRec for every component"""
    return [(tuple(lo), tuple(hi))] * ncomp


def random_field(seed):
    def fn(I, J, K, c):
        return _hash(seed, c, I, J, K)
    return fn


def _hash(seed, c, I, J, K):
    h = (np.int64(seed) * 1000003 + c * 7919 + I.astype(np.int64) * 73856093
         + J.astype(np.int64) * 19349663 + K.astype(np.int64) * 83492791)
    h = (h ^ (h >> 13)) * 1274126177 % 2147483647
    return (h % 200003) / 100001.5 - 1.0
