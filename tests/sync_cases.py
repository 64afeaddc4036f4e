"""Two-level layouts for the AMR restriction tests: a coarse domain of n
cells per direction in boxes of max_grid, and fine boxes (ratio 2) over part
of it."""
from maestroex_amd.fabs import Geometry, make_params, split_domain


def two_levels(dim=3, n=8, max_grid=4, fine_boxes=None):
    geoms = [Geometry(dim, [n * s] * dim, [1.0 / (n * s)] * dim, [1] * dim, [0] * dim, [0] * dim)
             for s in (1, 2)]
    params = make_params(geoms[0], nspec=3, extra_levels=[geoms[1]])
    coarse = split_domain(geoms[0], max_grid)
    if fine_boxes is None:
        if dim == 3:
            fine_boxes = [((4, 4, 4), (11, 11, 7)), ((4, 4, 8), (11, 11, 11)),
                          ((12, 0, 0), (15, 3, 3))]
        else:
            fine_boxes = [((4, 4, 0), (11, 7, 0)), ((4, 8, 0), (11, 11, 0)), ((12, 0, 0), (15, 3, 0))]
    return params, coarse, list(fine_boxes)
