"""The plotfile restatement (oracle/plotfile_oracle.py) and the package's
reader (maestroex_amd/plotfile.py): a known-answer Header and Cell_H for a
one-box 2-D plotfile, written out from the published format; the FAB data
bytes against numpy's little-endian doubles; a two-level 3-D round trip."""
import os

import numpy as np

from maestroex_amd.plotfile import read_plotfile
from plotfile_oracle import write_plotfile

HEADER_2D = """HyperCLaw-V1.1
2
rho
temp
2
0.5
0
0 0 
1 0.5 

((0,0) (3,1) (0,0)) 
7 
0.25 0.25 
0
0
0 1 0.5
7
0 1
0 0.5
Level_0/Cell
"""

CELL_H_2D = """1
1
2
0
(1 0
((0,0) (3,1) (0,0))
)
1
FabOnDisk: Cell_D_00000 0

1,2
0.0000000000000000e+00,-8.0000000000000000e+00,

1,2
7.0000000000000000e+00,-1.0000000000000000e+00,

"""


def test_known_answer_2d(tmp_path):
    a = np.arange(16, dtype=float).reshape(2, 1, 2, 4)
    a[1] = -a[0] - 1.0
    path = str(tmp_path / "plt")
    write_plotfile(path, 2, ["rho", "temp"], 0.5, [0.0, 0.0], [1.0, 0.5],
                   [((0, 0, 0), (3, 1, 0))], [[0.25, 0.25]], [7], [[((0, 0, 0), (3, 1, 0))]], [[a]])
    assert open(os.path.join(path, "Header")).read() == HEADER_2D
    assert open(os.path.join(path, "Level_0", "Cell_H")).read() == CELL_H_2D
    raw = open(os.path.join(path, "Level_0", "Cell_D_00000"), "rb").read()
    head = b"FAB ((8, (64 11 52 0 1 12 0 1023)),(8, (1 2 3 4 5 6 7 8)))((0,0) (3,1) (0,0)) 2\n"
    assert raw[:len(head)] == head
    assert raw[len(head):] == a.astype("<f8").tobytes()
    pf = read_plotfile(path)
    assert pf.varnames == ["rho", "temp"] and pf.dim == 2 and pf.time == 0.5
    assert np.array_equal(pf.levels[0]["data"][0], a)


def test_round_trip_two_levels_3d(tmp_path):
    rng = np.random.default_rng(3)
    boxes = [[((0, 0, 0), (7, 7, 3)), ((0, 0, 4), (7, 7, 7))], [((4, 4, 4), (9, 7, 11))]]
    domains = [((0, 0, 0), (7, 7, 7)), ((0, 0, 0), (15, 15, 15))]
    data = [[rng.standard_normal((3, hi[2] - lo[2] + 1, hi[1] - lo[1] + 1, hi[0] - lo[0] + 1))
             for lo, hi in bl] for bl in boxes]
    path = str(tmp_path / "plt3")
    write_plotfile(path, 3, ["a", "b", "c"], 1.25, [0.0] * 3, [2.0] * 3, domains,
                   [[0.25] * 3, [0.125] * 3], [3, 3], boxes, data)
    pf = read_plotfile(path)
    assert pf.finest_level == 1 and pf.ref_ratio == [2]
    assert pf.domains == [((0, 0, 0), (7, 7, 7)), ((0, 0, 0), (15, 15, 15))]
    assert pf.levels[1]["phys"][0] == [(0.5, 1.25), (0.5, 1.0), (0.5, 1.5)]
    for lev in range(2):
        assert pf.levels[lev]["boxes"] == boxes[lev]
        for b, a in enumerate(data[lev]):
            assert np.array_equal(pf.levels[lev]["data"][b], a)
            assert np.array_equal(pf.levels[lev]["min"][b], a.reshape(3, -1).min(axis=1))
            assert np.array_equal(pf.levels[lev]["max"][b], a.reshape(3, -1).max(axis=1))
