"""The oracle's AMR restriction (oracle/oracle_sync.cpp) against closed
forms: the fine field i + 2j + 4k (and its faces) has binary-exact means, so
every covered coarse cell and face must equal the field at the coarse point;
uncovered coarse points and coarse ghosts keep their values."""
import numpy as np
import pytest

import pyoracle
from maestroex_amd.fabs import Backend, FaceMultiFab, MultiFab
from sync_cases import two_levels

NP = Backend("numpy")


def index_field(m, shift=(0.5, 0.5, 0.5), scale=1.0):
    for b, a in enumerate(m.arrays):
        beg = m.begin[b]
        nc, nz, ny, nx = a.shape
        K, J, I = np.meshgrid(np.arange(beg[2], beg[2] + nz), np.arange(beg[1], beg[1] + ny),
                              np.arange(beg[0], beg[0] + nx), indexing="ij")
        v = ((I + shift[0]) + 2 * (J + shift[1]) + 4 * (K + shift[2])) * scale
        a[:] = np.stack([v + c for c in range(nc)])


def covered(point, fine_boxes, d=None):
    """is the coarse cell (or d-face) covered by the coarsened fine boxes"""
    for lo, hi in fine_boxes:
        ok = True
        for q in range(3):
            clo, chi = lo[q] // 2, hi[q] // 2
            if q == d:
                chi += 1
            ok = ok and clo <= point[q] <= chi
        if ok:
            return True
    return False


@pytest.mark.parametrize("dim", [2, 3])
def test_average_down_cells(dim):
    params, coarse, fine = two_levels(dim)
    m = pyoracle.oracle_maestro(params, [coarse, fine])
    f = MultiFab(fine, 2, 1, dim, NP)
    shift = (0.5, 0.5, 0.5) if dim == 3 else (0.5, 0.5, 0.0)
    index_field(f, shift)
    c = MultiFab(coarse, 2, 1, dim, NP, fill=-5.0)
    m.average_down(0, f, c, 0, 2)
    for b, (lo, hi) in enumerate(coarse):
        a = c.arrays[b]
        beg = c.begin[b]
        for idx in np.ndindex(*a.shape[1:]):
            k, j, i = idx
            p = (i + beg[0], j + beg[1], k + beg[2])
            inside = all(lo[q] <= p[q] <= hi[q] for q in range(dim))
            if inside and covered(p, fine):
                # the mean of the fine cells 2p .. 2p + 1: the field at 2p + 1/2
                want = (2 * p[0] + 0.5 + shift[0]) + 2 * (2 * p[1] + 0.5 + shift[1])
                want += 4 * (2 * p[2] + 0.5 + shift[2]) if dim == 3 else 4 * shift[2]
                if dim == 2:
                    want = (2 * p[0] + 1.0) + 2 * (2 * p[1] + 1.0)
                assert a[0][idx] == want and a[1][idx] == want + 1, (b, p)
            else:
                assert a[0][idx] == -5.0 and a[1][idx] == -5.0, (b, p)


@pytest.mark.parametrize("dim", [2, 3])
def test_average_down_faces(dim):
    params, coarse, fine = two_levels(dim)
    m = pyoracle.oracle_maestro(params, [coarse, fine])
    ff = FaceMultiFab(fine, 1, 0, dim, NP)
    cf = FaceMultiFab(coarse, 1, 0, dim, NP, fill=-5.0)
    for d in range(dim):
        index_field(ff[d], (0.0, 0.0, 0.0))
    m.average_down_faces(0, ff, cf)
    for d in range(dim):
        for b, (lo, hi) in enumerate(coarse):
            a = cf[d].arrays[b]
            beg = cf[d].begin[b]
            for idx in np.ndindex(*a.shape[1:]):
                k, j, i = idx
                p = (i + beg[0], j + beg[1], k + beg[2])
                if covered(p, fine, d):
                    # the fine faces at 2p in d, 2p and 2p + 1 across: the
                    # field at 2p (d) and 2p + 1/2 (across)
                    q = [2 * p[t] + (0.0 if t == d else 0.5) for t in range(3)]
                    if dim == 2:
                        q[2] = 0.0
                    want = q[0] + 2 * q[1] + 4 * q[2]
                    assert a[0][idx] == want, (d, b, p)
                else:
                    assert a[0][idx] == -5.0, (d, b, p)
