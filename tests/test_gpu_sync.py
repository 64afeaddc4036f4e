"""AverageDown and AverageDownFaces on the GPU against the oracle, bit for
bit, on two-level layouts (2-D and 3-D, several coarse and fine boxes, fine
boxes across coarse box edges) with random fine data; the uncovered coarse
points keep their values.  Faces shared by two fine boxes get one value in
both fine FABs first, as FillBoundary leaves them."""
import numpy as np
import pytest

import pyoracle
from maestroex_amd.fabs import Backend, FaceMultiFab, MultiFab
from sync_cases import two_levels

pytestmark = pytest.mark.gpu
NP = Backend("numpy")


def global_fill(m, seed, n_glob, nc):
    """random values per global point (one value per shared face)"""
    G = np.random.default_rng(seed).uniform(-1, 1, (nc, n_glob, n_glob, n_glob))
    for q, a in enumerate(m.arrays):
        beg = m.begin[q]
        _, nz, ny, nx = a.shape
        sl = (slice(None), slice(beg[2] + 1, beg[2] + 1 + nz), slice(beg[1] + 1, beg[1] + 1 + ny),
              slice(beg[0] + 1, beg[0] + 1 + nx))
        m.arrays[q] = m.backend.from_numpy(G[sl])


@pytest.mark.parametrize("dim", [2, 3])
def test_average_down_parity(dim):
    import maestroex_amd as mx
    print(f"[parity] average_down dim={dim}", flush=True)
    params, coarse, fine = two_levels(dim)
    outs = []
    for be, mk in ((NP, pyoracle.oracle_maestro), (mx.device_backend(), mx.gpu_maestro)):
        m = mk(params, [coarse, fine])
        f = MultiFab(fine, 3, 1, dim, be)
        global_fill(f, 1, 20, 3)
        c = MultiFab(coarse, 3, 1, dim, be)
        global_fill(c, 2, 12, 3)
        m.average_down(0, f, c, 1, 2)
        ff = FaceMultiFab(fine, 2, 0, dim, be)
        cf = FaceMultiFab(coarse, 2, 0, dim, be)
        for d in range(dim):
            global_fill(ff[d], 10 + d, 20, 2)
            global_fill(cf[d], 20 + d, 12, 2)
        m.average_down_faces(0, ff, cf)
        m.sync()
        outs.append((c, cf))
    for q, (x, y) in enumerate(zip(outs[0][0].to_numpy(), outs[1][0].to_numpy())):
        assert np.array_equal(x, y), f"cells FAB {q}"
    for d in range(dim):
        for q, (x, y) in enumerate(zip(outs[0][1][d].to_numpy(), outs[1][1][d].to_numpy())):
            assert np.array_equal(x, y), f"faces dir {d} FAB {q}"
