"""The base-state <-> Cartesian producers on the GPU against the oracle, bit
for bit: Put1dArrayOnCart (planar 2-D / 3-D; spherical, every interpolation
type, bin- and edge-centred input, scalar and vector output), MakeW0mac
(types 1-4), MakeS0mac (types 1-3), PutInPertForm, ConvertRhoXToX and
ConvertRhoHToH, on uneven multi-box levels with random base states."""
import numpy as np
import pytest

import pyoracle
from base_cases import SphSetup
from maestroex_amd import _abi
from maestroex_amd.fabs import Backend, BaseState, FaceMultiFab, Geometry, MultiFab, make_params
from maestroex_amd.fabs import split_domain

pytestmark = pytest.mark.gpu
NP = Backend("numpy")


def setups(dim=3, n=16, max_grid=(8, 4, 6), spherical=False):
    import maestroex_amd as mx
    periodic = [0 if spherical else 1] * dim
    geom = Geometry(dim, [n] * dim, [1.0 / n] * dim, periodic, [0] * dim, [0] * dim)
    params = make_params(geom, nspec=3, spherical=int(spherical))
    boxes = split_domain(geom, max_grid if dim == 3 else max_grid[:2])
    dev = mx.device_backend()
    return [(NP, pyoracle.oracle_maestro(params, [boxes])),
            (dev, mx.gpu_maestro(params, [boxes]))], boxes, params


def same(a, b, what):
    for q, (x, y) in enumerate(zip(a.to_numpy(), b.to_numpy())):
        nd = int(np.sum(x != y))
        assert nd == 0, f"{what}: FAB {q}: {nd} values differ"


def random_base(backend, length, seed, nlev=1):
    return BaseState(np.random.default_rng(seed).uniform(0.5, 2.0, (nlev, length)), backend)


@pytest.mark.parametrize("dim", [2, 3])
def test_put1d_planar_parity(dim):
    print(f"[parity] put1d planar dim={dim}", flush=True)
    ms, boxes, params = setups(dim=dim)
    outs = []
    for be, m in ms:
        s0 = random_base(be, 17, 3)
        o1 = MultiFab(boxes, dim, 0, dim, be, fill=-1.0)
        o2 = MultiFab(boxes, 1, 0, dim, be)
        m.Put1dArrayOnCart(0, s0, o1, False, True)
        m.Put1dArrayOnCart(0, s0, o2, True, False)
        m.sync()
        outs.append((o1, o2))
    same(outs[0][0], outs[1][0], "cell vector")
    same(outs[0][1], outs[1][1], "edge scalar")


@pytest.mark.parametrize("s0t,w0t", [(1, 1), (2, 2), (3, 3)])
def test_put1d_spherical_parity(s0t, w0t):
    print(f"[parity] put1d spherical s0_interp_type={s0t} w0_interp_type={w0t}", flush=True)
    ms, boxes, params = setups(spherical=True)
    outs = []
    for be, m in ms:
        sph = SphSetup(be, 16, types=(s0t, w0t, 1, 1))
        g = sph.geom()
        s0 = random_base(be, sph.nr_fine, 5)
        w0 = random_base(be, sph.nr_fine + 1, 6)
        a = MultiFab(boxes, 1, 0, 3, be)
        b = MultiFab(boxes, 3, 0, 3, be)
        m.Put1dArrayOnCart(0, s0, a, False, False, g)
        m.Put1dArrayOnCart(0, w0, b, True, True, g)
        m.sync()
        outs.append((a, b))
    same(outs[0][0], outs[1][0], "s0 scalar")
    same(outs[0][1], outs[1][1], "w0 vector")


@pytest.mark.parametrize("wt", [1, 2, 3, 4])
def test_make_w0mac_parity(wt):
    print(f"[parity] make_w0mac type {wt}", flush=True)
    ms, boxes, params = setups(spherical=True)
    outs = []
    for be, m in ms:
        sph = SphSetup(be, 16, types=(3, 2, 1, wt))
        w0 = random_base(be, sph.nr_fine + 1, 7)
        w0c = MultiFab(boxes, 3, 2, 3, be)
        for q, arr in enumerate(w0c.arrays):
            w0c.arrays[q] = be.from_numpy(np.random.default_rng(q).uniform(-1, 1, arr.shape))
        w0mac = FaceMultiFab(boxes, 1, 1, 3, be)
        m.MakeW0mac(0, w0, w0c, w0mac, sph.geom())
        m.sync()
        outs.append(w0mac)
    for d in range(3):
        same(outs[0][d], outs[1][d], f"w0mac dir {d}")


@pytest.mark.parametrize("st", [1, 2, 3])
def test_make_s0mac_parity(st):
    print(f"[parity] make_s0mac type {st}", flush=True)
    ms, boxes, params = setups(spherical=True)
    outs = []
    for be, m in ms:
        sph = SphSetup(be, 16, types=(3, 2, st, 1))
        s0 = random_base(be, sph.nr_fine, 8)
        s0c = MultiFab(boxes, 1, 2, 3, be)
        for q, arr in enumerate(s0c.arrays):
            s0c.arrays[q] = be.from_numpy(np.random.default_rng(10 + q).uniform(-1, 1, arr.shape))
        s0mac = FaceMultiFab(boxes, 1, 1, 3, be)
        m.MakeS0mac(0, s0, s0c, s0mac, sph.geom())
        m.sync()
        outs.append(s0mac)
    for d in range(3):
        same(outs[0][d], outs[1][d], f"s0mac dir {d}")


@pytest.mark.parametrize("spherical", [False, True])
def test_pert_form_and_conversions_parity(spherical):
    print(f"[parity] pert form and conversions spherical={spherical}", flush=True)
    ms, boxes, params = setups(spherical=spherical)
    outs = []
    for be, m in ms:
        nc = params.nscal
        scal = MultiFab(boxes, nc, 0, 3, be)
        for q, arr in enumerate(scal.arrays):
            scal.arrays[q] = be.from_numpy(np.random.default_rng(20 + q).uniform(0.5, 2.0, arr.shape))
        sph = SphSetup(be, 16, types=(2, 2, 1, 1)) if spherical else None
        s0 = random_base(be, sph.nr_fine if spherical else 16, 9)
        g = sph.geom() if spherical else None
        m.PutInPertForm(0, scal, s0, 0, True, g)
        m.ConvertRhoXToX(0, scal, True)
        m.ConvertRhoHToH(0, scal, True)
        m.PutInPertForm(0, scal, s0, 1, False, g)
        m.ConvertRhoXToX(0, scal, False)
        m.sync()
        outs.append(scal)
    same(outs[0], outs[1], "scal")


def test_short_base_state_rejected_on_the_host():
    """The library checks on the host that the base state and r_cc_loc /
    r_edge_loc reach every radius its kernels read; a short one is an error
    and nothing is launched."""
    print("[parity] short base state", flush=True)
    ms, boxes, params = setups(spherical=True)
    be, m = ms[1]
    sph = SphSetup(be, 16)
    short = random_base(be, sph.nr_fine // 4, 1)
    out = MultiFab(boxes, 1, 0, 3, be)
    with pytest.raises(_abi.MxError) as e:
        m.Put1dArrayOnCart(0, short, out, False, False, sph.geom())
    assert e.value.code == _abi.MX_ERR_ARG
