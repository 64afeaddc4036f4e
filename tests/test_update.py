"""The oracle's UpdateScal and UpdateVel (oracle/oracle_update.cpp) against
what the mathematics fixes, chosen so that a dropped term, a wrong sign, a
swapped direction or spacing, or a wrong order of the species passes fails:

* fluxes linear in each direction with different slopes and different cell
  spacings give a flux divergence known exactly (all values binary-exact);
* on a periodic domain the species and enthalpy totals change by dt times
  the total force, and the density update is the sum of the species updates;
* the negative-species redistribution of a hand-computed cell, the density
  floor and the auxiliary rescaling;
* UpdateVel with linear edge velocities, constant MAC velocities and
  constant w0 (planar and spherical), with and without the sponge;
* the w0_cart read one cell above the box (planar) needs a ghost cell.
"""
import numpy as np
import pytest

import pyoracle
from maestroex_amd import _abi
from maestroex_amd.fabs import Backend
from update_cases import UpdateCase

NP = Backend("numpy")


def oracle(case):
    return pyoracle.oracle_maestro(case.params, [case.boxes])


def valid(mf, b):
    return mf.to_numpy()[b][(slice(None),) + mf.valid_region(b)]


def linear_faces(case, mf, slopes):
    """face MultiFab with value slopes[d][c] * (face index) * dx[d] on the
    d-faces (binary-exact for the spacings used here)."""
    for d in range(case.dim):
        sl = np.asarray(slopes[d], dtype=np.float64)

        def fn(I, J, K, d=d, sl=sl):
            idx = (I, J, K)[d]
            return sl[:, None, None, None] * (idx * case.dx[d])[None]
        case.fill(mf[d], fn)


@pytest.mark.parametrize("dim", [2, 3])
def test_update_scal_linear_flux_divergence(dim):
    dx = [1 / 16, 1 / 8, 1 / 4][:dim]
    case = UpdateCase(NP, dim=dim, n=8, max_grid=4, nspec=3, dt=0.25, dx=dx)
    ns, nc = 3, case.nscal
    slopes = [[0.5 * (c + 1) for c in range(nc)], [-1.0 - c for c in range(nc)],
              [2.0 + 0.25 * c for c in range(nc)]][:dim]
    sold, snew, force = case.cell_mf(nc), case.cell_mf(nc), case.cell_mf(nc)
    sflux = case.face_mf(nc)
    linear_faces(case, sflux, slopes)
    sv = np.array([3.0, 2.0, 0.0, 0.0] + [1.0, 0.5, 1.5], dtype=np.float64)
    fv = np.array([0.0, 1.0, 0.0, 0.0] + [0.25, -0.5, 0.75], dtype=np.float64)
    case.fill(sold, lambda I, J, K: sv[:, None, None, None] + 0.0 * I)
    case.fill(force, lambda I, J, K: fv[:, None, None, None] + 0.0 * I)
    m = oracle(case)
    m.UpdateScal(0, sold, snew, sflux, force, 4, ns, None, 0.0)
    m.UpdateScal(0, sold, snew, sflux, force, 1, 1, None, 0.0)
    div = sum(np.asarray(slopes[d]) for d in range(dim))
    want = sv + 0.25 * (-div + fv)
    want[0] = sv[0] + sum(want[4:] - sv[4:])
    for b in range(len(case.boxes)):
        got = valid(snew, b)
        for c in [0, 1, 4, 5, 6]:
            assert np.all(got[c] == want[c]), (b, c, got[c].ravel()[:3], want[c])


def test_update_scal_conserves_on_periodic_domain():
    case = UpdateCase(NP, dim=3, n=8, max_grid=4, nspec=3, dt=0.01)
    rng = np.random.default_rng(3)
    nc = case.nscal
    sold, snew, force = case.cell_mf(nc), case.cell_mf(nc), case.cell_mf(nc)
    sflux = case.face_mf(nc)
    case.random_fill(sold, rng, 1.0, 2.0)
    case.random_fill(force, rng, -0.5, 0.5)
    for d in range(3):
        case.random_fill(sflux[d], rng, -0.1, 0.1)  # periodic: face n is face 0
    m = oracle(case)
    m.UpdateScal(0, sold, snew, sflux, force, 4, 3, None, 0.0)
    m.UpdateScal(0, sold, snew, sflux, force, 1, 1, None, 0.0)
    so = np.concatenate([valid(sold, b).reshape(nc, -1) for b in range(len(case.boxes))], 1)
    sn = np.concatenate([valid(snew, b).reshape(nc, -1) for b in range(len(case.boxes))], 1)
    fo = np.concatenate([valid(force, b).reshape(nc, -1) for b in range(len(case.boxes))], 1)
    for c in [1, 4, 5, 6]:
        assert abs((sn[c] - so[c]).sum() - 0.01 * fo[c].sum()) <= 1e-12
    # density: the sum of the species updates, cell by cell
    assert np.allclose(sn[0] - so[0], (sn[4:7] - so[4:7]).sum(0), rtol=0, atol=1e-14)
    assert np.all(sn[4:7] > 0)


def one_cell_case(species_new, rho_old, bcd, naux=0, aux_old=()):
    """a 1-cell-per-box 2-D case whose species update is species_new exactly
    (zero flux, dt = 1, force = species_new - species_old with species_old 0)"""
    ns = len(species_new)
    case = UpdateCase(NP, dim=2, n=2, max_grid=1, nspec=ns, naux=naux, dt=1.0)
    nc = case.nscal
    sv = np.zeros(nc)
    sv[0] = rho_old
    sv[4 + ns:] = aux_old
    fv = np.zeros(nc)
    fv[4:4 + ns] = species_new
    sold, snew, force = case.cell_mf(nc), case.cell_mf(nc), case.cell_mf(nc)
    case.fill(sold, lambda I, J, K: sv[:, None, None, None] + 0.0 * I)
    case.fill(force, lambda I, J, K: fv[:, None, None, None] + 0.0 * I)
    m = oracle(case)
    m.UpdateScal(0, sold, snew, case.face_mf(nc), force, 4, ns, None, bcd)
    return valid(snew, 0)[:, 0, 0, 0]


def test_update_scal_negative_species_redistribution():
    # x = (-0.125, 0.5, 0.25, 0.25): the deficit 0.125 comes from the others
    # in proportion to their values (sum 1.0)
    out = one_cell_case([-0.125, 0.5, 0.25, 0.25], rho_old=0.0, bcd=0.0)
    assert out[0] == 0.875  # rho: 0 + the species updates
    assert np.array_equal(out[4:8], [0.0, 0.4375, 0.21875, 0.21875])
    # two negative species: handled in component order, the second sees the
    # first's redistribution
    out = one_cell_case([-0.25, 1.0, -0.25, 0.5], rho_old=0.0, bcd=0.0)
    x = np.array([-0.25, 1.0, -0.25, 0.5])
    for c1 in range(4):
        if x[c1] < 0:
            delta = -x[c1]
            others = [c2 for c2 in range(4) if c2 != c1 and x[c2] >= 0]
            s = sum(x[c2] for c2 in others)
            for c2 in others:
                x[c2] -= x[c2] / s * delta
            x[c1] = 0.0
    assert np.allclose(out[4:8], x, rtol=0, atol=1e-15)
    assert abs(out[4:8].sum() - 1.0) <= 1e-15 and np.all(out[4:8] >= 0)


def test_update_scal_density_floor_and_aux():
    # rho after the update 0.0625 + 2 * 0.03125 = 0.125 < 0.5 * bcd = 0.25:
    # the species are scaled by 0.25 / 0.125 = 2 and rho set to 0.25; the
    # auxiliaries scale with the updated (pre-floor) density over the old one
    out = one_cell_case([0.03125, 0.03125], rho_old=0.0625, bcd=0.5, naux=2,
                        aux_old=(1.0, 3.0))
    assert out[0] == 0.25
    assert np.array_equal(out[4:6], [0.0625, 0.0625])
    assert np.array_equal(out[6:8], [2.0, 6.0])
    # above the floor nothing is scaled
    out = one_cell_case([0.5, 0.25], rho_old=0.0, bcd=1.0)
    assert out[0] == 0.75 and np.array_equal(out[4:6], [0.5, 0.25])


def vel_inputs(case, U, B, W, uo, fo, sp=None, w0mac=None):
    dim = case.dim
    uold, unew, force = case.cell_mf(dim), case.cell_mf(dim), case.cell_mf(dim)
    umac, uedge = case.face_mf(1), case.face_mf(dim)
    w0c = case.cell_mf(dim, ngrow=1)
    sponge = case.cell_mf(1)
    case.fill(uold, lambda I, J, K: np.asarray(uo)[:, None, None, None] + 0.0 * I)
    case.fill(force, lambda I, J, K: np.asarray(fo)[:, None, None, None] + 0.0 * I)
    for d in range(dim):
        case.fill(umac[d], lambda I, J, K, d=d: U[d] + 0.0 * I[None])
    linear_faces(case, uedge, B)
    wv = np.zeros(dim)
    wv[dim - 1] = W
    case.fill(w0c, lambda I, J, K: wv[:, None, None, None] + 0.0 * I)
    if sp is not None:
        case.fill(sponge, lambda I, J, K: sp + 0.0 * I[None])
    wm = None
    if w0mac is not None:
        wm = case.face_mf(1)
        for d in range(dim):
            case.fill(wm[d], lambda I, J, K, d=d: w0mac[d] + 0.0 * I[None])
    return uold, unew, umac, uedge, force, sponge, wm, w0c


@pytest.mark.parametrize("dim,sponge", [(2, False), (3, False), (3, True)])
def test_update_vel_planar_closed_form(dim, sponge):
    dx = [1 / 16, 1 / 8, 1 / 4][:dim]
    case = UpdateCase(NP, dim=dim, n=8, max_grid=4, dt=0.125, dx=dx)
    U = [0.5, -0.25, 1.0][:dim]           # MAC velocity per direction
    B = [[1.0, -2.0, 0.5], [0.25, 1.0, -1.0], [2.0, 0.5, 0.75]]
    B = [row[:dim] for row in B[:dim]]  # B[d][n]: d-gradient of component n
    W, uo, fo = 0.5, [1.0, 2.0, -1.0][:dim], [0.25, 0.5, 1.0][:dim]
    sp = 0.75 if sponge else None
    uold, unew, umac, uedge, force, spg, _, w0c = vel_inputs(case, U, B, W, uo, fo, sp)
    oracle(case).UpdateVel(0, uold, unew, umac, uedge, force, spg, None, w0c, sponge)
    dt = 0.125
    for n in range(dim):
        g = sum(U[d] * B[d][n] for d in range(dim))
        want = uo[n] - dt * g + dt * fo[n] - dt * W * B[dim - 1][n]
        if sponge:
            want *= sp
        for b in range(len(case.boxes)):
            assert np.all(valid(unew, b)[n] == want), (n, b)


def test_update_vel_spherical_closed_form():
    case = UpdateCase(NP, dim=3, n=8, max_grid=4, dt=0.125, spherical=True)
    U = [0.5, -0.25, 1.0]
    B = [[1.0, -2.0, 0.5], [0.25, 1.0, -1.0], [2.0, 0.5, 0.75]]
    W0 = [0.25, -0.5, 0.125]  # w0mac per direction
    uo, fo = [1.0, 2.0, -1.0], [0.25, 0.5, 1.0]
    uold, unew, umac, uedge, force, spg, wm, _ = vel_inputs(case, U, B, 0.0, uo, fo, 0.5, W0)
    oracle(case).UpdateVel(0, uold, unew, umac, uedge, force, spg, wm, None, True)
    dt = 0.125
    for n in range(3):
        g = sum(U[d] * B[d][n] for d in range(3))
        w0g = sum(B[d][n] * W0[d] for d in range(3))
        want = (uo[n] - dt * g + dt * fo[n] - dt * w0g) * 0.5
        for b in range(len(case.boxes)):
            assert np.all(valid(unew, b)[n] == want), n


def test_update_vel_needs_w0_ghost_and_rejects_eos():
    case = UpdateCase(NP, dim=3, n=8, max_grid=4, dt=0.125)
    m = oracle(case)
    uold, unew, force = case.cell_mf(3), case.cell_mf(3), case.cell_mf(3)
    w0c = case.cell_mf(3, ngrow=0)  # the cell above the box is missing
    with pytest.raises(_abi.MxError) as e:
        m.UpdateVel(0, uold, unew, case.face_mf(1), case.face_mf(3), force, None, None, w0c,
                    False)
    assert e.value.code == _abi.MX_ERR_BOX
    nc = case.nscal
    with pytest.raises(_abi.MxError) as e:
        m.UpdateScal(0, case.cell_mf(nc), case.cell_mf(nc), case.face_mf(nc), case.cell_mf(nc),
                     1, 1, None, 1.0, True)
    assert e.value.code == _abi.MX_ERR_UNSUPPORTED
