"""MakeVelForce and ModifyScalForce on the GPU against the oracle, bit for
bit, on uneven multi-box levels with random fields: planar (vertical
non-periodic and periodic) and spherical, the Utilde . grad w0 term on and
off, full and perturbational forms, and densities on both sides of the
buoyancy cutoff."""
import numpy as np
import pytest

import pyoracle
from base_cases import SphSetup
from maestroex_amd.fabs import Backend, BaseState, FaceMultiFab, Geometry, MultiFab, make_params
from maestroex_amd.fabs import split_domain

pytestmark = pytest.mark.gpu
NP = Backend("numpy")


def setups(spherical=False, vert_periodic=False, n=16):
    import maestroex_amd as mx
    periodic = [0] * 3 if spherical else [1, 1, 1 if vert_periodic else 0]
    geom = Geometry(3, [n] * 3, [1.0 / n] * 3, periodic, [0] * 3, [0] * 3)
    params = make_params(geom, nspec=3, spherical=int(spherical))
    boxes = split_domain(geom, (8, 4, 6))
    return [(NP, pyoracle.oracle_maestro(params, [boxes])),
            (mx.device_backend(), mx.gpu_maestro(params, [boxes]))], boxes, params


def rand_mf(be, boxes, nc, ng, seed, lo=-1.0, hi=1.0):
    m = MultiFab(boxes, nc, ng, 3, be)
    for q, a in enumerate(m.arrays):
        m.arrays[q] = be.from_numpy(np.random.default_rng(seed + q).uniform(lo, hi, a.shape))
    return m


def rand_faces(be, boxes, seed):
    f = FaceMultiFab(boxes, 1, 0, 3, be)
    for d in range(3):
        for q, a in enumerate(f[d].arrays):
            f[d].arrays[q] = be.from_numpy(np.random.default_rng(seed + 10 * d + q).uniform(-1, 1, a.shape))
    return f


def same(a, b, what):
    for q, (x, y) in enumerate(zip(a.to_numpy(), b.to_numpy())):
        nd = int(np.sum(x != y))
        assert nd == 0, f"{what}: FAB {q}: {nd} values differ"


@pytest.mark.parametrize("spherical,add", [(False, True), (False, False), (True, True), (True, False)])
def test_make_vel_force_parity(spherical, add):
    print(f"[parity] make_vel_force spherical={spherical} add_utilde={add}", flush=True)
    ms, boxes, params = setups(spherical=spherical)
    outs = []
    for be, m in ms:
        sph = SphSetup(be, 16, types=(3, 2, 1, 1)) if spherical else None
        nr = sph.nr_fine if spherical else 16
        rng = np.random.default_rng(4)
        rho0 = BaseState(rng.uniform(0.5, 1.5, (1, nr)), be)
        grav = BaseState(rng.uniform(-2.0, -0.5, (1, nr)), be)
        w0 = BaseState(rng.uniform(-1.0, 1.0, (1, nr + 1)), be)
        vf = MultiFab(boxes, 3, 1, 3, be)
        rho = rand_mf(be, boxes, params.nscal, 0, 1, 0.2, 2.0)
        gpi = rand_mf(be, boxes, 3, 0, 2)
        wf = rand_mf(be, boxes, 3, 0, 3)
        w0c = rand_mf(be, boxes, 3, 1, 5)
        nrm = rand_mf(be, boxes, 3, 0, 6)
        ue = rand_faces(be, boxes, 7)
        m.MakeVelForce(0, vf, ue, rho, gpi, wf, w0c, nrm, rho0, grav, w0, 0.5, 1.2, add, False,
                       sph.geom() if spherical else None)
        m.sync()
        outs.append(vf)
    same(outs[0], outs[1], "vel_force")


@pytest.mark.parametrize("spherical,vper", [(False, False), (False, True), (True, False)])
@pytest.mark.parametrize("fullform", [True, False])
def test_modify_scal_force_parity(spherical, vper, fullform):
    print(f"[parity] modify_scal_force spherical={spherical} vert_periodic={vper} "
          f"fullform={fullform}", flush=True)
    ms, boxes, params = setups(spherical=spherical, vert_periodic=vper)
    outs = []
    for be, m in ms:
        sph = SphSetup(be, 16, types=(2, 2, 1, 1)) if spherical else None
        nr = sph.nr_fine if spherical else 16
        rng = np.random.default_rng(8)
        s0e = BaseState(rng.uniform(0.5, 1.5, (1, nr + 1)), be)
        w0 = BaseState(rng.uniform(-1.0, 1.0, (1, nr + 1)), be)
        f = rand_mf(be, boxes, params.nscal, 0, 11)
        st = rand_mf(be, boxes, params.nscal, 0, 12)
        s0c = rand_mf(be, boxes, 1, 1, 13)
        w0c = rand_mf(be, boxes, 3, 1, 14)
        um = rand_faces(be, boxes, 15)
        for comp in (4, 1):
            m.ModifyScalForce(0, f, st, um, s0e, s0c, w0c, w0, comp, fullform,
                              sph.geom() if spherical else None)
        m.sync()
        outs.append(f)
    same(outs[0], outs[1], "scal_force")
