"""The oracle's base-state <-> Cartesian producers (oracle/oracle_base.cpp)
against what the mathematics fixes:

* planar Put1dArrayOnCart is a copy of row `lev` (cell) or the average of
  two edges, into component dim - 1 of a vector only;
* spherical interpolation of a base state linear in r: the linear types
  reproduce it at every radius to rounding, the quadratic types too, and the
  piecewise-constant types pick the bin int(r / dr) (or the nearer edge);
  vector outputs point along the radius;
* MakeW0mac and MakeS0mac of linear data give the face-centre radius's
  value (types 2, 3), the average of the two cells (type 1) and the average
  of the four nodes (type 4);
* PutInPertForm subtracts and adds back exactly; ConvertRhoXToX and
  ConvertRhoHToH divide and multiply by Rho;
* a base state too short for the domain is an error, not a read past it.
"""
import numpy as np
import pytest

import pyoracle
from base_cases import SphSetup
from maestroex_amd import _abi
from maestroex_amd.fabs import Backend, BaseState, FaceMultiFab, Geometry, MultiFab, make_params
from maestroex_amd.fabs import split_domain

NP = Backend("numpy")


def context(n=8, dim=3, max_grid=4, spherical=False, nspec=3):
    periodic = [0 if spherical else 1] * dim
    geom = Geometry(dim, [n] * dim, [1.0 / n] * dim, periodic, [0] * dim, [0] * dim)
    params = make_params(geom, nspec=nspec, spherical=int(spherical))
    boxes = split_domain(geom, max_grid)
    return params, boxes, pyoracle.oracle_maestro(params, [boxes])


def valid(mf, b):
    return mf.to_numpy()[b][(slice(None),) + mf.valid_region(b)]


def centres(mf, b, h=(0.5, 0.5, 0.5), n=8):
    """coordinates relative to the centre of the points of FAB b"""
    a = mf.arrays[b]
    beg = mf.begin[b]
    _, nz, ny, nx = a.shape
    k, j, i = np.meshgrid(np.arange(beg[2], beg[2] + nz), np.arange(beg[1], beg[1] + ny),
                          np.arange(beg[0], beg[0] + nx), indexing="ij")
    dx = 1.0 / n
    return ((i + h[0]) * dx - 0.5, (j + h[1]) * dx - 0.5, (k + h[2]) * dx - 0.5)


@pytest.mark.parametrize("dim", [2, 3])
def test_put1d_planar(dim):
    params, boxes, m = context(dim=dim)
    n = 8
    s0 = BaseState(np.arange(n + 1, dtype=np.float64)[None] * 0.25 + 1.0, NP)
    out = MultiFab(boxes, dim, 0, dim, NP, fill=-7.0)
    m.Put1dArrayOnCart(0, s0, out, False, True)
    edge = MultiFab(boxes, 1, 0, dim, NP)
    m.Put1dArrayOnCart(0, s0, edge, True, False)
    for b, (lo, hi) in enumerate(boxes):
        a, e = valid(out, b), valid(edge, b)
        r = np.arange(lo[dim - 1], hi[dim - 1] + 1)
        want = 1.0 + 0.25 * r
        shape = [1, 1, 1]
        shape[3 - dim] = len(r)  # the vertical axis of the (z, y, x) array
        if dim == 2:
            shape = [1, len(r), 1]
        assert np.all(a[dim - 1] == want.reshape(shape))
        assert np.all(a[:dim - 1] == -7.0)  # the other components are not touched
        assert np.all(e[0] == (0.5 * (want + (want + 0.25))).reshape(shape))


def linear_s0(sph, a=2.0, b=-3.0, edge=False):
    r = sph.r_edge if edge else sph.r_cc
    return BaseState((a + b * r)[None], NP), (lambda rad: a + b * rad)


@pytest.mark.parametrize("edge,types,tol", [(False, (2, 2, 1, 1), 1e-13), (False, (3, 2, 1, 1), 1e-13),
                                            (True, (3, 2, 1, 1), 1e-13), (True, (3, 3, 1, 1), 1e-13)])
def test_put1d_spherical_linear(edge, types, tol):
    params, boxes, m = context(spherical=True)
    sph = SphSetup(NP, 8, types=types)
    s0, f = linear_s0(sph, edge=edge)
    out = MultiFab(boxes, 3, 0, 3, NP)
    m.Put1dArrayOnCart(0, s0, out, edge, True, sph.geom())
    for b in range(len(boxes)):
        x, y, z = [c[out.valid_region(b)] for c in centres(out, b)]
        rad = np.sqrt(x * x + y * y + z * z)
        a = valid(out, b)
        for comp, q in enumerate((x, y, z)):
            assert np.allclose(a[comp], f(rad) * q / rad, rtol=0, atol=tol)


def test_put1d_spherical_piecewise_constant():
    params, boxes, m = context(spherical=True)
    sph = SphSetup(NP, 8, types=(1, 1, 1, 1))
    vals = np.arange(sph.nr_fine + 1, dtype=np.float64)
    out = MultiFab(boxes, 1, 0, 3, NP)
    m.Put1dArrayOnCart(0, BaseState(vals[None, :-1], NP), out, False, False, sph.geom())
    oute = MultiFab(boxes, 1, 0, 3, NP)
    m.Put1dArrayOnCart(0, BaseState(vals[None], NP), oute, True, False, sph.geom())
    for b in range(len(boxes)):
        x, y, z = [c[out.valid_region(b)] for c in centres(out, b)]
        rad = np.sqrt(x * x + y * y + z * z)
        idx = (rad / sph.drf).astype(int)
        assert np.array_equal(valid(out, b)[0], idx.astype(np.float64))
        rfac = (rad - idx * sph.drf) / sph.drf
        assert np.array_equal(valid(oute, b)[0], np.where(rfac > 0.5, idx + 1, idx).astype(float))


def face_neighbours(shape, d):
    """slices of a 2-ghost cell array holding the cells below and above each
    face of a 1-ghost face array of the given (z, y, x) shape, faces normal
    to d: the face array starts one cell below the box, the cell array two"""
    ax = 2 - d
    below = tuple(slice(0 if q == ax else 1, (0 if q == ax else 1) + shape[q]) for q in range(3))
    above = tuple(slice(1, 1 + shape[q]) for q in range(3))
    return below, above


def node_w0(i, j, k, d, sph, w0, n=8):
    dx = 1.0 / n
    x, y, z = i * dx - 0.5, j * dx - 0.5, k * dx - 0.5
    rad = np.sqrt(x * x + y * y + z * z)
    idx = (rad / sph.drf).astype(int)
    rfac = (rad - idx * sph.drf) / sph.drf
    v = rfac * w0[idx + 1] + (1.0 - rfac) * w0[idx]
    q = (x, y, z)[d]
    with np.errstate(invalid="ignore", divide="ignore"):
        return np.where(rad == 0.0, v, v * q / rad)


def node_average(fmf, b, d, sph, w0):
    """the average of w0 (component d) over the four nodes of each face"""
    beg = fmf.begin[b]
    _, nz, ny, nx = fmf.arrays[b].shape
    k, j, i = np.meshgrid(np.arange(beg[2], beg[2] + nz), np.arange(beg[1], beg[1] + ny),
                          np.arange(beg[0], beg[0] + nx), indexing="ij")
    t = [q for q in range(3) if q != d]
    e = [np.eye(3, dtype=int)[q] for q in t]
    pts = [(0, 0, 0), tuple(e[0]), tuple(e[1]), tuple(e[0] + e[1])]
    return 0.25 * sum(node_w0(i + o[0], j + o[1], k + o[2], d, sph, w0) for o in pts)


@pytest.mark.parametrize("wtype", [2, 3, 4, 1])
def test_make_w0mac(wtype):
    params, boxes, m = context(spherical=True)
    sph = SphSetup(NP, 8, types=(3, 2, 1, wtype))
    w0, f = linear_s0(sph, a=0.5, b=2.0, edge=True)
    w0c = MultiFab(boxes, 3, 2, 3, NP)
    for b in range(len(boxes)):  # w0_cart with its ghost cells: the radial field
        x, y, z = centres(w0c, b)
        rad = np.sqrt(x * x + y * y + z * z)
        w0c.arrays[b][:] = np.stack([f(rad) * q / rad for q in (x, y, z)])
    w0mac = FaceMultiFab(boxes, 1, 1, 3, NP)
    m.MakeW0mac(0, w0, w0c, w0mac, sph.geom())
    for d in range(3):
        for b in range(len(boxes)):
            h = [0.5, 0.5, 0.5]
            h[d] = 0.0
            x, y, z = centres(w0mac[d], b, h)
            rad = np.sqrt(x * x + y * y + z * z)
            want = f(rad) * (x, y, z)[d] / rad
            got = w0mac[d].arrays[b][0]
            if wtype == 1:
                # the average of the two cells beside each face
                c = w0c.arrays[b][d]
                sl_m, sl_0 = face_neighbours(got.shape, d)
                assert np.array_equal(got, 0.5 * (c[sl_m] + c[sl_0]))
            elif wtype == 4:
                # the average of w0 at the face's four nodes; w0 at a node is
                # the linear interpolation in r along the radius, and all of
                # w0 (every component) at the node at r = 0
                assert np.allclose(got, node_average(w0mac[d], b, d, sph, w0.array[0]), rtol=0,
                                   atol=1e-14), (d, b)
            else:
                # linear and quadratic interpolation of linear data
                assert np.allclose(got, want, rtol=0, atol=1e-12), (d, b)


@pytest.mark.parametrize("stype", [2, 3, 1])
def test_make_s0mac(stype):
    params, boxes, m = context(spherical=True)
    sph = SphSetup(NP, 8, types=(3, 2, stype, 1))
    s0, f = linear_s0(sph, a=1.0, b=-0.75)
    s0c = MultiFab(boxes, 1, 2, 3, NP)
    for b in range(len(boxes)):
        x, y, z = centres(s0c, b)
        s0c.arrays[b][0] = np.sin(7 * x) + y * z  # any cell field: type 1 averages it
    s0mac = FaceMultiFab(boxes, 1, 1, 3, NP)
    m.MakeS0mac(0, s0, s0c, s0mac, sph.geom())
    for d in range(3):
        for b in range(len(boxes)):
            h = [0.5, 0.5, 0.5]
            h[d] = 0.0
            x, y, z = centres(s0mac[d], b, h)
            got = s0mac[d].arrays[b][0]
            if stype == 1:
                c = s0c.arrays[b][0]
                sl_m, sl_0 = face_neighbours(got.shape, d)
                assert np.array_equal(got, 0.5 * (c[sl_m] + c[sl_0]))
            else:
                rad = np.sqrt(x * x + y * y + z * z)
                assert np.allclose(got, f(rad), rtol=0, atol=1e-12)


def test_pert_form_and_conversions():
    params, boxes, m = context(dim=3)
    nc = params.nscal
    scal = MultiFab(boxes, nc, 0, 3, NP)
    rng = np.random.default_rng(1)
    for a in scal.arrays:
        a[:] = rng.integers(1, 64, size=a.shape) / 16.0
        a[0] = 2.0  # Rho: division and multiplication are exact
    orig = [a.copy() for a in scal.arrays]
    s0 = BaseState(np.linspace(0.5, 1.5, 9)[None], NP)
    m.PutInPertForm(0, scal, s0, 1, True)
    for b, (lo, hi) in enumerate(boxes):
        r = np.arange(lo[2], hi[2] + 1)
        want = orig[b][1] - s0.array[0, r][:, None, None]
        assert np.array_equal(scal.arrays[b][1], want)
    m.PutInPertForm(0, scal, s0, 1, False)
    for b in range(len(boxes)):
        assert np.allclose(scal.arrays[b][1], orig[b][1], rtol=0, atol=1e-15)
    m.ConvertRhoXToX(0, scal, True)
    m.ConvertRhoHToH(0, scal, True)
    for b in range(len(boxes)):
        assert np.array_equal(scal.arrays[b][4:7], orig[b][4:7] / 2.0)
        assert np.array_equal(scal.arrays[b][2:4], orig[b][2:4])  # Temp, Pi untouched
    m.ConvertRhoXToX(0, scal, False)
    for b in range(len(boxes)):
        assert np.array_equal(scal.arrays[b][4:7], orig[b][4:7])


def test_short_base_state_is_an_error():
    params, boxes, m = context(spherical=True)
    sph = SphSetup(NP, 8)
    short = BaseState(np.ones((1, sph.nr_fine // 4)), NP)  # ends inside the domain
    out = MultiFab(boxes, 1, 0, 3, NP)
    with pytest.raises(_abi.MxError) as e:
        m.Put1dArrayOnCart(0, short, out, False, False, sph.geom())
    assert e.value.code == _abi.MX_ERR_BOX
