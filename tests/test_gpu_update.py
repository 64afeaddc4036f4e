"""UpdateScal and UpdateVel on the GPU against the oracle, bit for bit, on
random multi-box levels (2-D and 3-D, uneven boxes, spherical), with inputs
that drive cells through every branch of the species pass: the density floor,
negative species (one and several per cell), auxiliary components, and more
species than the register form holds (the global-memory form)."""
import numpy as np
import pytest

import pyoracle
from maestroex_amd.fabs import Backend
from update_cases import UpdateCase

pytestmark = pytest.mark.gpu
NP = Backend("numpy")


def dev_backend():
    import maestroex_amd as mx
    return mx.device_backend()


def both(make, **kw):
    """the same case on the host (oracle) and on the device"""
    return make(UpdateCase(NP, **kw)), make(UpdateCase(dev_backend(), **kw))


def assert_same(mc, mg, what):
    for b, (a, g) in enumerate(zip(mc.to_numpy(), mg.to_numpy())):
        sl = (slice(None),) + mc.valid_region(b)
        a, g = a[sl], g[sl]
        nd = int(np.sum(~((a == g) | (np.isnan(a) & np.isnan(g)))))
        assert nd == 0, f"{what}: box {b}: {nd} values differ"


def scal_fields(seed):
    def make(case):
        rng = np.random.default_rng(seed)
        nc, ns = case.nscal, case.params.num_spec
        sold, snew, force = case.cell_mf(nc), case.cell_mf(nc), case.cell_mf(nc)
        sflux = case.face_mf(nc)
        G = case.random_fill(sold, rng, 0.2, 1.0)
        # rho = sum of the species; a few cells with a small density
        small = rng.uniform(size=G.shape[1:]) < 0.08
        G[4:4 + ns] *= np.where(small, 0.02, 1.0)
        G[0] = G[4:4 + ns].sum(0)
        p = case.n
        case.fill(sold, lambda I, J, K: G[:, (K % p) if case.dim == 3 else 0 * K, J % p, I % p])
        F = case.random_fill(force, rng, -2.0, 2.0)
        # drive some species negative: a force below -s / dt
        neg = rng.uniform(size=(ns,) + G.shape[1:]) < 0.1
        F[4:4 + ns] = np.where(neg, -1.6 * G[4:4 + ns] / case.params.dt, F[4:4 + ns])
        case.fill(force, lambda I, J, K: F[:, (K % p) if case.dim == 3 else 0 * K, J % p, I % p])
        for d in range(case.dim):
            case.random_fill(sflux[d], rng, -0.01, 0.01)
        return case, sold, snew, sflux, force
    return make


SCAL_CASES = [
    ("3d", dict(dim=3, n=16, max_grid=8, nspec=3)),
    ("3d_uneven_aux", dict(dim=3, n=24, max_grid=(16, 8, 12), nspec=3, naux=2)),
    ("3d_many_species", dict(dim=3, n=12, max_grid=6, nspec=10)),
    ("2d", dict(dim=2, n=32, max_grid=16, nspec=4)),
]


@pytest.mark.parametrize("name,kw", SCAL_CASES, ids=[c[0] for c in SCAL_CASES])
def test_update_scal_parity(name, kw):
    print(f"[parity] update_scal {name} {kw}", flush=True)
    (cc, so_c, sn_c, sf_c, fo_c), (cg, so_g, sn_g, sf_g, fo_g) = both(scal_fields(7), dt=0.05, **kw)
    import maestroex_amd as mx
    mc = pyoracle.oracle_maestro(cc.params, [cc.boxes])
    mg = mx.gpu_maestro(cg.params, [cg.boxes])
    ns = cc.params.num_spec
    bcd = 0.3
    for m, so, sn, sf, fo in ((mc, so_c, sn_c, sf_c, fo_c), (mg, so_g, sn_g, sf_g, fo_g)):
        m.UpdateScal(0, so, sn, sf, fo, 4, ns, None, bcd)
        m.UpdateScal(0, so, sn, sf, fo, 1, 1, None, bcd)
    mg.sync()
    assert_same(sn_c, sn_g, f"{name} snew")
    # the inputs reached the floor and the negative-species branch
    out = np.concatenate([a[(slice(None),) + sn_c.valid_region(b)].reshape(cc.nscal, -1)
                          for b, a in enumerate(sn_c.to_numpy())], 1)
    floored = int(np.sum(out[0] == 0.5 * bcd))
    zeroed = int(np.sum(out[4:4 + ns] == 0.0))
    print(f"[update] {name}: {floored} cells at the floor, {zeroed} species set to zero",
          flush=True)
    assert floored > 0 and zeroed > 0


def vel_fields(seed, sponge):
    def make(case):
        rng = np.random.default_rng(seed)
        dim = case.dim
        uold, unew, force = case.cell_mf(dim), case.cell_mf(dim), case.cell_mf(dim)
        umac, uedge = case.face_mf(1), case.face_mf(dim)
        w0c = case.cell_mf(dim, ngrow=1)
        spg = case.cell_mf(1)
        case.random_fill(uold, rng)
        case.random_fill(force, rng)
        case.random_fill(w0c, rng)
        case.random_fill(spg, rng, 0.5, 1.0)
        for d in range(dim):
            case.random_fill(umac[d], rng)
            case.random_fill(uedge[d], rng)
        wm = None
        if case.params.spherical:
            wm = case.face_mf(1)
            for d in range(dim):
                case.random_fill(wm[d], rng)
        return case, uold, unew, umac, uedge, force, spg, wm, w0c
    return make


VEL_CASES = [
    ("3d", dict(dim=3, n=16, max_grid=8), False),
    ("3d_sponge_uneven", dict(dim=3, n=24, max_grid=(16, 8, 12)), True),
    ("2d_sponge", dict(dim=2, n=32, max_grid=16), True),
    ("3d_spherical_sponge", dict(dim=3, n=16, max_grid=8, spherical=True), True),
]


@pytest.mark.parametrize("name,kw,sponge", VEL_CASES, ids=[c[0] for c in VEL_CASES])
def test_update_vel_parity(name, kw, sponge):
    print(f"[parity] update_vel {name} {kw} sponge={sponge}", flush=True)
    host, dev = both(vel_fields(11, sponge), dt=0.05, **kw)
    import maestroex_amd as mx
    outs = []
    for i, (case, uold, unew, umac, uedge, force, spg, wm, w0c) in enumerate((host, dev)):
        m = (pyoracle.oracle_maestro(case.params, [case.boxes]) if i == 0
             else mx.gpu_maestro(case.params, [case.boxes]))
        m.UpdateVel(0, uold, unew, umac, uedge, force, spg, wm, w0c, sponge)
        m.sync()
        outs.append(unew)
    assert_same(outs[0], outs[1], f"{name} unew")
