"""The oracle's force producers (oracle/oracle_force.cpp) against closed
forms:

* MakeVelForce, planar: -gpi / rho in the horizontal components; the
  buoyancy (rho - rho0) g - gpi over rho minus the w0 force in the vertical,
  with the buoyancy cut off below buoyancy_cutoff_factor base_cutoff_density;
  the Utilde . grad w0 term of linear w0 except in the top cell; all values
  binary-exact;
* MakeVelForce, spherical: constant rho0 and g along the radius and linear w0
  give (rho - rho0) g x/r - gpi over rho - w0 force - (Utilde . n) dw0/dr n;
* ModifyScalForce, planar: linear MAC velocity and w0 give div(U) exactly,
  full form and perturbational form, with the edge state above the domain top
  taken from the top cell (non-periodic) or the bottom (periodic);
* ModifyScalForce, spherical: constant w0 gives div(w0 er) = 2 w0 / r.
"""
import numpy as np
import pytest

import pyoracle
from base_cases import SphSetup
from maestroex_amd.fabs import Backend, BaseState, FaceMultiFab, Geometry, MultiFab, make_params
from maestroex_amd.fabs import split_domain

NP = Backend("numpy")


def context(n=8, max_grid=4, spherical=False, vert_periodic=False):
    periodic = [0] * 3 if spherical else [1, 1, 1 if vert_periodic else 0]
    geom = Geometry(3, [n] * 3, [1 / 8, 1 / 4, 1 / 2][:3] if not spherical else [1.0 / n] * 3,
                    periodic, [0] * 3, [0] * 3)
    params = make_params(geom, nspec=3, spherical=int(spherical))
    boxes = split_domain(geom, max_grid)
    return params, boxes, pyoracle.oracle_maestro(params, [boxes])


def mf(boxes, nc, ng=0, fill=0.0):
    return MultiFab(boxes, nc, ng, 3, NP, fill=fill)


def fill(m, fn):
    for b, a in enumerate(m.arrays):
        beg = m.begin[b]
        _, nz, ny, nx = a.shape
        K, J, I = np.meshgrid(np.arange(beg[2], beg[2] + nz), np.arange(beg[1], beg[1] + ny),
                              np.arange(beg[0], beg[0] + nx), indexing="ij")
        a[:] = np.broadcast_to(fn(I, J, K), a.shape)


def valid(m, b):
    return m.arrays[b][(slice(None),) + m.valid_region(b)]


def kindex(m, b):
    """the vertical index of the valid points of box b, as a (z, y, x) array"""
    lo, hi = m.boxes[b]
    sl = m.valid_region(b)
    return (np.arange(lo[2], hi[2] + 1)[:, None, None] + 0 * m.arrays[b][0][sl])


def test_make_vel_force_planar():
    params, boxes, m = context()
    n = 8
    vf = mf(boxes, 3, fill=9.0)
    rho = mf(boxes, params.nscal)
    fill(rho, lambda I, J, K: np.where(I == 0, 0.25, 2.0)[None] + 0.0 * np.zeros((params.nscal, 1, 1, 1)))
    gpi, wf = mf(boxes, 3), mf(boxes, 3)
    fill(gpi, lambda I, J, K: np.array([0.25, 0.5, 0.75])[:, None, None, None] + 0.0 * I)
    fill(wf, lambda I, J, K: np.array([0.0, 0.0, 0.125])[:, None, None, None] + 0.0 * I)
    w0c = mf(boxes, 3, ng=1)
    fill(w0c, lambda I, J, K: np.stack([0.0 * K, 0.0 * K, 0.5 * K]))  # dw0/dz cell-to-cell 0.5
    ue = FaceMultiFab(boxes, 1, 0, 3, NP)
    for d in range(3):
        fill(ue[d], lambda I, J, K: 0.25 + 0.0 * I[None])
    rho0 = BaseState(np.full((1, n), 1.5), NP)
    grav = BaseState(np.full((1, n), -1.0), NP)
    # buoyancy cut off where rho = 0.25 < 0.5 * 1.0
    m.MakeVelForce(0, vf, ue, rho, gpi, wf, w0c, None, rho0, grav, None, 0.5, 1.0, True)
    dz = 1 / 2
    for b in range(len(boxes)):
        got = valid(vf, b)
        r = valid(rho, b)[0]
        k = kindex(vf, b)
        assert np.array_equal(got[0], -0.25 / r) and np.array_equal(got[1], -0.5 / r)
        rhopert = np.where(r < 0.5, 0.0, r - 1.5)
        want = (rhopert * -1.0 - 0.75) / r - 0.125
        # the Utilde term: (w(k+1) + w(k)) (w0(k+1) - w0(k)) / (2 dz), not in the top cell
        want = np.where(k >= n - 1, want, want - (0.25 + 0.25) * 0.5 / (2.0 * dz))
        assert np.array_equal(got[2], want), b


def test_make_vel_force_spherical():
    params, boxes, m = context(spherical=True)
    n = 8
    sph = SphSetup(NP, n, types=(2, 2, 1, 1))
    vf, rho, gpi, wf = mf(boxes, 3), mf(boxes, params.nscal), mf(boxes, 3), mf(boxes, 3)
    fill(rho, lambda I, J, K: 2.0 + 0.0 * np.zeros((params.nscal, 1, 1, 1)) + 0.0 * I[None])
    fill(gpi, lambda I, J, K: np.array([0.25, 0.5, 0.75])[:, None, None, None] + 0.0 * I)
    fill(wf, lambda I, J, K: np.array([0.1, 0.2, 0.3])[:, None, None, None] + 0.0 * I)
    dx = 1.0 / n
    nrm = mf(boxes, 3)
    ue = FaceMultiFab(boxes, 1, 0, 3, NP)
    U = [0.5, -0.25, 1.0]
    for d in range(3):
        fill(ue[d], lambda I, J, K, d=d: U[d] + 0.0 * I[None])
    fill(nrm, lambda I, J, K: np.stack([(q + 0.5) * dx - 0.5 for q in (I, J, K)]) /
         np.sqrt(sum(((q + 0.5) * dx - 0.5) ** 2 for q in (I, J, K)))[None])
    slope = 0.75
    rho0 = BaseState(np.full((1, sph.nr_fine), 1.5), NP)
    grav = BaseState(np.full((1, sph.nr_fine), -1.0), NP)
    w0 = BaseState((slope * sph.r_edge)[None], NP)
    m.MakeVelForce(0, vf, ue, rho, gpi, wf, None, nrm, rho0, grav, w0, 0.0, 0.0, True, False,
                   sph.geom())
    for b in range(len(boxes)):
        got = valid(vf, b)
        nv = valid(nrm, b)
        ut = sum(U[d] * nv[d] for d in range(3))
        for d in range(3):
            want = (0.5 * -1.0 * nv[d] - (0.25, 0.5, 0.75)[d]) / 2.0 - (0.1, 0.2, 0.3)[d]
            want = want - ut * slope * nv[d]
            assert np.allclose(got[d], want, rtol=0, atol=1e-12), (b, d)


@pytest.mark.parametrize("vert_periodic", [False, True])
@pytest.mark.parametrize("fullform", [True, False])
def test_modify_scal_force_planar(fullform, vert_periodic):
    params, boxes, m = context(vert_periodic=vert_periodic)
    n = 8
    dxs = [1 / 8, 1 / 4, 1 / 2]
    slopes = [0.5, -0.25, 1.0]
    comp = 4
    f = mf(boxes, params.nscal)
    fill(f, lambda I, J, K: 1.0 + 0.0 * np.zeros((params.nscal, 1, 1, 1)) + 0.0 * I[None])
    st = mf(boxes, params.nscal)
    fill(st, lambda I, J, K: 3.0 + 0.0 * np.zeros((params.nscal, 1, 1, 1)) + 0.0 * I[None])
    um = FaceMultiFab(boxes, 1, 0, 3, NP)
    for d in range(3):
        fill(um[d], lambda I, J, K, d=d: (slopes[d] * (I, J, K)[d] * dxs[d])[None])
    w0c = mf(boxes, 3, ng=1)
    fill(w0c, lambda I, J, K: np.stack([0.0 * K, 0.0 * K, 0.25 * K * dxs[2]]))
    s0c = mf(boxes, 1)
    fill(s0c, lambda I, J, K: 2.0 + 0.0 * I[None])
    s0e = BaseState(np.arange(n + 1, dtype=np.float64)[None] * 0.5, NP)
    m.ModifyScalForce(0, f, st, um, s0e, s0c, w0c, None, comp, fullform)
    divu = sum(slopes) + 0.25
    for b in range(len(boxes)):
        got = valid(f, b)[comp]
        k = kindex(f, b)
        if fullform:
            want = 1.0 - 3.0 * divu + 0.0 * k
        else:
            w = lambda kk: slopes[2] * kk * dxs[2]
            e0 = 0.5 * k
            ktop = np.where(k + 1 > n - 1, 0 if vert_periodic else n - 1, k + 1)
            e1 = 0.5 * ktop
            divs0u = 2.0 * (slopes[0] + slopes[1]) + (w(k + 1) * e1 - w(k) * e0) / dxs[2]
            want = 1.0 + (-(3.0 - 2.0) * divu - divs0u)
        assert np.allclose(got, want, rtol=0, atol=1e-13), b


def test_modify_scal_force_spherical_full_form():
    params, boxes, m = context(spherical=True)
    n = 8
    sph = SphSetup(NP, n, types=(2, 2, 1, 1))
    comp = 4
    f, st = mf(boxes, params.nscal), mf(boxes, params.nscal)
    fill(st, lambda I, J, K: 1.0 + 0.0 * np.zeros((params.nscal, 1, 1, 1)) + 0.0 * I[None])
    um = FaceMultiFab(boxes, 1, 0, 3, NP)
    w0 = BaseState(np.full((1, sph.nr_fine + 1), 0.5), NP)
    m.ModifyScalForce(0, f, st, um, None, None, None, w0, comp, True, sph.geom())
    dx = 1.0 / n
    for b in range(len(boxes)):
        lo, hi = boxes[b]
        K, J, I = np.meshgrid(*[np.arange(lo[q], hi[q] + 1) for q in (2, 1, 0)], indexing="ij")
        rad = np.sqrt(sum(((q + 0.5) * dx - 0.5) ** 2 for q in (I, J, K)))
        # div(w0 er) = 2 w0 / r, to the accuracy of the bins and interpolation
        assert np.allclose(valid(f, b)[comp], -2.0 * 0.5 / rad, rtol=2e-2, atol=0), b
