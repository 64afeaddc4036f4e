"""Pins the CPU oracle with checks that do not restate its formulas.

* closed forms: constant data are reproduced exactly; on a linear or a
  quadratic profile PPM type 1 and 2 reconstruct the exact face values
  (4th-order interpolation is exact on cubics and neither limiter fires),
  so the traced states equal the exact average of the profile over the
  characteristic domain of dependence, and with a uniform normal velocity
  and no transverse velocity the 3-D edge state reduces to that 1-D trace;
* invariants: periodic translation (bit-exact), x-mirror with u -> -u
  (bit-exact for PPM type 1: Ip and Im swap), agreement of the two boxes
  that share a periodic face, F_Rho summed in component order;
* behaviour of the reference's own unit test (test_advect): a Gaussian
  advected one period with u = +-1 along x, y, z returns with the same
  error in every direction;
* Maestro::BCSetup's table (MaestroSetup.cpp:276-559) and error paths.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import pyoracle
from harness import global_face_array, oracle_for
from maestroex_amd import _abi
from maestroex_amd.fabs import Backend, FaceMultiFab, MultiFab
from maestroex_amd.workloads import Workload

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NP = Backend("numpy")


def fill_mf(mf, fn, comp):
    """mf[comp] = fn(I, J, K) over every point (valid and ghost) of each FAB."""
    for b, a in enumerate(mf.arrays):
        beg = mf.begin[b]
        nc, nz, ny, nx = a.shape
        K, J, I = np.meshgrid(np.arange(beg[2], beg[2] + nz), np.arange(beg[1], beg[1] + ny),
                              np.arange(beg[0], beg[0] + nx), indexing="ij")
        a[comp] = fn(I, J, K)


def single_box(n=16, ppm=1, dim=3, **kw):
    return Workload(dim=dim, n=n, max_grid=n, ppm_type=ppm, all_periodic=True, backend=NP, **kw)


def ppm_on(w, m, s_comp, vel_faces, bx=None):
    lo, hi = w.boxes[0]
    bx = bx or (lo, hi)
    Ip = MultiFab(w.boxes, 3, 1, 3, NP)
    Im = MultiFab(w.boxes, 3, 1, 3, NP)
    bcs = (_abi.BCRec * 1)()
    m.PPM(0, bx, (w.scal, 0), (vel_faces[0], 0), (vel_faces[1], 0), (vel_faces[2], 0), (Ip, 0),
          (Im, 0), bcs, 1, s_comp, 0)
    return Ip, Im


def const_faces(w, u):
    f = FaceMultiFab(w.boxes, 1, 1, 3, NP)
    for d in range(3):
        f[d].arrays[0][:] = u[d]
    return f


# ---------------------------------------------------------------- closed forms

@pytest.mark.parametrize("ppm", [1, 2])
def test_constant_field_ppm(ppm):
    w = single_box(ppm=ppm)
    m = oracle_for(w)
    w.scal.arrays[0][0] = 3.25
    Ip, Im = ppm_on(w, m, 0, const_faces(w, (0.7, -0.4, 0.3)))
    valid = Ip.valid_region(0)
    assert np.all(Ip.arrays[0][(slice(None),) + valid] == 3.25)
    assert np.all(Im.arrays[0][(slice(None),) + valid] == 3.25)


def _avg_poly(coef, a, b):
    """exact average of sum c_p x^p over [a, b]"""
    tot = 0.0
    for p, c in enumerate(coef):
        tot += c * (b ** (p + 1) - a ** (p + 1)) / (p + 1)
    return tot / (b - a)


@pytest.mark.parametrize("ppm", [1, 2])
@pytest.mark.parametrize("coef", [(0.5, 1.3), (2.0, 0.1, 0.8)], ids=["linear", "quadratic"])
def test_polynomial_trace_exact(ppm, coef):
    """Cell averages of a monotone polynomial in x: Ip(i) is the exact average
    of the polynomial over [x_{i+1/2} - u dt, x_{i+1/2}] (u > 0), Im(i) over
    [x_{i-1/2}, x_{i-1/2} + |u| dt] (u < 0)."""
    n = 16
    w = single_box(n=n, ppm=ppm)
    m = oracle_for(w)
    h = 1.0 / n
    # shift x so the profile is monotone on and around the box
    x0 = 2.0

    def avg(I, J, K):
        return np.vectorize(lambda i: _avg_poly(coef, x0 + i * h, x0 + (i + 1) * h))(I)

    fill_mf(w.scal, avg, 0)
    u = 0.8
    dt = w.params.dt
    for sgn in (1.0, -1.0):
        Ip, Im = ppm_on(w, m, 0, const_faces(w, (sgn * u, 0.0, 0.0)))
        sig = u * dt / h
        for i in range(2, n - 2):
            xl, xr = x0 + i * h, x0 + (i + 1) * h
            a = Ip.arrays[0][0, 5, 5, i + 1]  # Ip has one ghost cell: index = i + 1
            b = Im.arrays[0][0, 5, 5, i + 1]
            if sgn > 0:
                exact = _avg_poly(coef, xr - sig * h, xr)
                assert abs(a - exact) <= 1e-12 * abs(exact), (i, a, exact)
            else:
                exact = _avg_poly(coef, xl, xl + sig * h)
                assert abs(b - exact) <= 1e-12 * abs(exact), (i, b, exact)


@pytest.mark.parametrize("ppm", [1, 2])
def test_uniform_velocity_edge_state_is_1d_trace(ppm):
    """MakeEdgeScal with umac = (u, 0, 0): the transverse corrections vanish
    and sedge_x(i) = Ip_x(i-1) + dt/2 f, the exact characteristic average of
    a quadratic profile plus half a step of forcing."""
    n = 16
    w = single_box(n=n, ppm=ppm)
    m = oracle_for(w)
    h = 1.0 / n
    coef = (1.0, 0.3, 0.5)
    x0 = 1.5
    fill_mf(w.scal, lambda I, J, K: np.vectorize(
        lambda i: _avg_poly(coef, x0 + i * h, x0 + (i + 1) * h))(I), 4)
    f = 0.125
    w.scal_force.arrays[0][4] = f
    u = 0.6
    umac = const_faces(w, (u, 0.0, 0.0))
    bcs_s, _ = m.bc_setup()
    m.MakeEdgeScal(0, w.scal, w.sedge, umac, w.scal_force, False, bcs_s, 4, 4, 1, False)
    dt = w.params.dt
    sig = u * dt / h
    se = w.sedge[0].arrays[0][4]
    for i in range(3, n - 2):
        xr = x0 + i * h  # face i is the right face of cell i-1
        exact = _avg_poly(coef, xr - sig * h, xr) + 0.5 * dt * f
        got = se[5, 5, i]
        assert abs(got - exact) <= 1e-12 * abs(exact), (i, got, exact)


def test_constant_scalars_pass():
    """Constant scalars with no forcing: every edge state equals the value and
    the fluxes are the closed forms of MakeRhoXFlux / MakeRhoHFlux."""
    w = Workload(dim=3, n=16, max_grid=8, ppm_type=1, backend=NP)
    vals = {0: 0.05, 1: -0.2, 4: 0.3, 5: 0.5, 6: 0.2}
    for b in range(len(w.boxes)):
        for c, v in vals.items():
            w.scal.arrays[b][c] = v
        w.scal_force.arrays[b][:] = 0.0
    m = oracle_for(w)
    m.advance_pass(0, w.pass_fields())
    rho0 = w.rho0.to_numpy()[0]
    rho0e = w.rho0_edge.to_numpy()[0]
    rhoh0 = w.rhoh0.to_numpy()[0]
    rhoh0e = w.rhoh0_edge.to_numpy()[0]
    for d in range(3):
        for b, (lo, hi) in enumerate(w.boxes):
            reg = w.sedge[d].valid_region(b)
            se = w.sedge[d].arrays[b]
            for c, v in vals.items():
                assert np.all(se[c][reg] == v)
            um = w.umac[d].arrays[b][0][w.umac[d].valid_region(b)]
            sf = w.sflux[d].arrays[b]
            kk = np.arange(lo[2], hi[2] + 1 + (1 if d == 2 else 0))[:, None, None]
            r0 = (0.5 * (rho0e[kk] + rho0e[kk])) if d == 2 else 0.5 * (rho0[kk] + rho0[kk])
            rh0 = (0.5 * (rhoh0e[kk] + rhoh0e[kk])) if d == 2 else 0.5 * (rhoh0[kk] + rhoh0[kk])
            tot = 0.0
            for c in (4, 5, 6):
                exp = um * (r0 + vals[0]) * vals[c]
                assert np.array_equal(sf[c][reg], exp)
                tot = tot + exp
            assert np.array_equal(sf[0][reg], 0.0 + tot)
            assert np.array_equal(sf[1][reg], um * (rh0 + vals[1]))


@pytest.mark.parametrize("exact", [0, 1])
def test_constant_scalars_spherical_pass(exact):
    """Spherical: constant scalars give constant edge states, and the fluxes
    take the face base states: F_X = u (rho0mac + rho') X (pred type 1),
    F_rhoh = u (rho0mac h0mac + (rho h)') (rhohprime, face base states
    averaged old/new), or u (rh0mac + (rho h)') with use_exact_base_state."""
    w = Workload(dim=3, n=16, max_grid=8, ppm_type=2, backend=NP, spherical=True,
                 use_exact_base_state=exact)
    vals = {0: 0.05, 1: -0.2, 4: 0.3, 5: 0.5, 6: 0.2}
    for b in range(len(w.boxes)):
        for c, v in vals.items():
            w.scal.arrays[b][c] = v
        w.scal_force.arrays[b][:] = 0.0
    m = oracle_for(w)
    m.advance_pass(0, w.pass_fields())
    for d in range(3):
        for b in range(len(w.boxes)):
            reg = w.sedge[d].valid_region(b)
            se = w.sedge[d].arrays[b]
            for c, v in vals.items():
                assert np.all(se[c][reg] == v)
            um = w.umac[d].arrays[b][0][w.umac[d].valid_region(b)]
            r0 = w.r0mac[d].arrays[b][0][w.r0mac[d].valid_region(b)]
            h0 = w.h0mac[d].arrays[b][0][w.h0mac[d].valid_region(b)]
            rh0 = w.rh0mac[d].arrays[b][0][w.rh0mac[d].valid_region(b)]
            sf = w.sflux[d].arrays[b]
            r0e = 0.5 * r0 + 0.5 * r0
            tot = 0.0
            for c in (4, 5, 6):
                exp = um * (r0e + vals[0]) * vals[c]
                assert np.array_equal(sf[c][reg], exp)
                tot = tot + exp
            assert np.array_equal(sf[0][reg], 0.0 + tot)
            if exact:
                exp_h = um * ((0.5 * rh0 + 0.5 * rh0) + vals[1])
            else:
                exp_h = um * (r0e * (0.5 * h0 + 0.5 * h0) + vals[1])
            assert np.array_equal(sf[1][reg], exp_h)
    assert all(np.all(a == 0.0) for a in w.etarhoflux.to_numpy())


# ---------------------------------------------------------------- invariants

def test_translation_invariance():
    """Periodic domain, input shifted by (3, 2) cells in x and y -> every
    output shifted by the same amount, bit for bit."""
    base = dict(dim=3, n=16, max_grid=8, ppm_type=2, all_periodic=True, backend=NP)
    w0 = Workload(**base)
    oracle_for(w0).advance_pass(0, w0.pass_fields())
    w1 = Workload(shift=(3, 2, 0), **base)
    oracle_for(w1).advance_pass(0, w1.pass_fields())
    for name in ("umac", "sedge", "sflux"):
        for d in range(3):
            for comp in ([0] if name == "umac" else [0, 1, 4, 6]):
                a = global_face_array(getattr(w0, name)[d], 3, d, [16] * 3, comp)
                b = global_face_array(getattr(w1, name)[d], 3, d, [16] * 3, comp)
                if d == 0:
                    assert np.array_equal(np.roll(a[:, :, :16], (2, 3), axis=(1, 2)), b[:, :, :16])
                elif d == 1:
                    assert np.array_equal(np.roll(a[:, :16, :], (2, 3), axis=(1, 2)), b[:, :16, :])
                else:
                    assert np.array_equal(np.roll(a, (2, 3), axis=(1, 2)), b)


def test_periodic_shared_faces_agree():
    """The face at x = 1 (stored by the hi boxes) and x = 0 (lo boxes) are the
    same periodic face: computed independently, they agree bit for bit."""
    w = Workload(dim=3, n=16, max_grid=8, ppm_type=1, all_periodic=True, backend=NP)
    oracle_for(w).advance_pass(0, w.pass_fields())
    for name in ("umac", "sedge", "sflux"):
        for d in range(3):
            if name == "sflux" and d == 2:
                continue  # vertical fluxes carry the height-dependent rho0_edge(r)
            comps = [0] if name == "umac" else [0, 1, 4, 5, 6]
            for comp in comps:
                g = global_face_array(getattr(w, name)[d], 3, d, [16] * 3, comp)
                idx = [slice(None)] * 3
                ax = 2 - d
                lo = list(idx)
                hi = list(idx)
                lo[ax] = 0
                hi[ax] = 16
                assert np.array_equal(g[tuple(lo)], g[tuple(hi)]), (name, d, comp)


def test_mirror_symmetry_ppm1():
    """x -> -x with u_x -> -u_x swaps Ip and Im (PPM type 1, bit-exact)."""
    w = single_box(n=16, ppm=1)
    m = oracle_for(w)
    rng = np.random.default_rng(5)
    s = w.scal.arrays[0]
    s[0] = rng.random(s[0].shape)
    vel = FaceMultiFab(w.boxes, 1, 1, 3, NP)
    for d in range(3):
        vel[d].arrays[0][:] = rng.normal(size=vel[d].arrays[0].shape)
    Ip, Im = ppm_on(w, m, 0, vel)
    # mirrored problem
    s[0] = s[0][..., ::-1].copy()
    vel[0].arrays[0][:] = -vel[0].arrays[0][..., ::-1]
    for d in (1, 2):
        vel[d].arrays[0][:] = vel[d].arrays[0][..., ::-1]
    Ip2, Im2 = ppm_on(w, m, 0, vel)
    reg = (slice(1, -1),) * 3
    A, B, A2, B2 = (x.arrays[0] for x in (Ip, Im, Ip2, Im2))
    assert np.array_equal(A2[0][..., ::-1][reg], B[0][reg])
    assert np.array_equal(B2[0][..., ::-1][reg], A[0][reg])
    for d in (1, 2):
        assert np.array_equal(A2[d][..., ::-1][reg], A[d][reg])
        assert np.array_equal(B2[d][..., ::-1][reg], B[d][reg])


def test_rho_flux_is_ordered_species_sum():
    w = Workload(dim=3, n=16, max_grid=8, ppm_type=2, backend=NP)
    oracle_for(w).advance_pass(0, w.pass_fields())
    for d in range(3):
        for b in range(len(w.boxes)):
            reg = w.sflux[d].valid_region(b)
            sf = w.sflux[d].arrays[b]
            acc = np.zeros_like(sf[0][reg])
            for c in (4, 5, 6):
                acc = acc + sf[c][reg]
            assert np.array_equal(sf[0][reg], acc)


@pytest.mark.parametrize("dim", [2, 3])
def test_advect_direction_independence(dim):
    """test_advect (Exec/unit_tests/test_advect/MaestroEvolve.cpp:82-210): a
    Gaussian advected one period by u = +-1 along each axis returns with the
    same L2 error for every direction and sign (no directional bias)."""
    n = 16
    errs = []
    for d in range(dim):
        for sgn in (1.0, -1.0):
            w = Workload(dim=dim, n=n, max_grid=n, ppm_type=1, all_periodic=True, backend=NP,
                         zero_velocity=True)
            m = oracle_for(w)
            h = 1.0 / n
            dt = 0.5 * h
            m.set_dt(dt, 1e-8)
            umac = FaceMultiFab(w.boxes, 1, 1, dim, NP)
            umac[d].arrays[0][:] = sgn
            for b in range(len(w.boxes)):
                w.scal_force.arrays[b][:] = 0.0
            ng = w.ng
            X = (np.arange(n) + 0.5) * h
            grids = np.meshgrid(*([X] * dim), indexing="ij")
            r2 = sum((g - 0.5) ** 2 for g in grids)
            s0 = np.exp(-r2 / 0.1 ** 2)
            s0 = s0.transpose() if dim == 3 else s0.T
            s = s0.copy()
            bcs_s, _ = m.bc_setup()
            for step in range(2 * n):
                full = np.pad(s, ng, mode="wrap")
                if dim == 2:
                    full = full[None]
                w.scal.arrays[0][4] = full
                m.MakeEdgeScal(0, w.scal, w.sedge, umac, w.scal_force, False, bcs_s, 4, 4, 1,
                               False)
                se = w.sedge[d].arrays[0][4]
                if dim == 2:
                    se = se[0]
                F = sgn * se
                ax = dim - 1 - d
                lo = [slice(None)] * dim
                hi = [slice(None)] * dim
                lo[ax] = slice(0, n)
                hi[ax] = slice(1, n + 1)
                s = s - dt / h * (F[tuple(hi)] - F[tuple(lo)])
            errs.append(float(np.sqrt(np.mean((s - s0) ** 2))))
    errs = np.array(errs)
    assert np.all(np.abs(errs - errs[0]) <= 1e-12 * errs[0]), errs
    assert 0 < errs[0] < 0.1


# ---------------------------------------------------------------- BCSetup, ABI

def test_bc_setup_table():
    """MaestroSetup.cpp:276-559: velocity and scalar BCRecs per phys_bc."""
    P = _abi
    w = Workload(dim=3, n=8, max_grid=8, ppm_type=1, backend=NP,
                 phys_bc={0: (P.PHYS_INFLOW, P.PHYS_OUTFLOW), 1: (P.PHYS_SYMMETRY, P.PHYS_SLIPWALL),
                          2: (P.PHYS_NOSLIPWALL, P.PHYS_OUTFLOW)})
    m = oracle_for(w)
    bs, bu = m.bc_setup()
    exp_u = {  # (dir, side) -> (all comps, comp == dir)
        (0, 0): (P.BC_EXT_DIR, P.BC_EXT_DIR), (0, 1): (P.BC_FOEXTRAP, P.BC_FOEXTRAP),
        (1, 0): (P.BC_REFLECT_EVEN, P.BC_REFLECT_ODD), (1, 1): (P.BC_HOEXTRAP, P.BC_EXT_DIR),
        (2, 0): (P.BC_EXT_DIR, P.BC_EXT_DIR), (2, 1): (P.BC_FOEXTRAP, P.BC_FOEXTRAP)}
    exp_s = {  # (dir, side) -> (Rho/RhoH/Temp/species, Pi)
        (0, 0): (P.BC_EXT_DIR, P.BC_FOEXTRAP), (0, 1): (P.BC_FOEXTRAP, P.BC_EXT_DIR),
        (1, 0): (P.BC_REFLECT_EVEN, P.BC_REFLECT_EVEN), (1, 1): (P.BC_HOEXTRAP, P.BC_FOEXTRAP),
        (2, 0): (P.BC_HOEXTRAP, P.BC_FOEXTRAP), (2, 1): (P.BC_FOEXTRAP, P.BC_EXT_DIR)}
    for (d, side), (allc, own) in exp_u.items():
        for comp in range(3):
            v = (bu[comp].lo if side == 0 else bu[comp].hi)[d]
            assert v == (own if comp == d else allc), (d, side, comp)
    for (d, side), (sc, pi) in exp_s.items():
        for comp in range(w.nscal):
            v = (bs[comp].lo if side == 0 else bs[comp].hi)[d]
            assert v == (pi if comp == 3 else sc), (d, side, comp)


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "mx_adv.h")).read()
    return sorted(set(re.findall(r"\b(mx_adv_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_header_symbols():
    import maestroex_amd as mx
    assert os.path.exists(mx.LIB_PATH), "libmaestroex_amd.so not built"
    lib = C.CDLL(mx.LIB_PATH)
    syms = header_symbols()
    assert len(syms) >= 20
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing


def test_library_without_gpu_fails_loudly():
    """No GPU in this container: context creation reports MX_ERR_NOLIB and
    the package raises instead of falling back to the CPU."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible")
    import maestroex_amd as mx
    w = Workload(dim=3, n=8, max_grid=8, ppm_type=1, backend=NP)
    with pytest.raises(_abi.MxError) as ei:
        mx.gpu_maestro(w.params, [w.boxes])
    assert ei.value.code == _abi.MX_ERR_NOLIB
    with pytest.raises(RuntimeError):
        mx.device_backend()


def test_oracle_reports_short_fabs():
    """A state with too few ghost cells for PPM is reported (MX_ERR_BOX)."""
    w = Workload(dim=3, n=8, max_grid=8, ppm_type=2, backend=NP)
    m = oracle_for(w)
    small = MultiFab(w.boxes, w.nscal, 1, 3, NP)
    bcs_s, _ = m.bc_setup()
    with pytest.raises(_abi.MxError) as ei:
        m.MakeEdgeScal(0, small, w.sedge, w.umac, w.scal_force, False, bcs_s, 4, 4, 1, False)
    assert ei.value.code == _abi.MX_ERR_BOX


def test_oracle_rejects_2d_traced_forces():
    """2-D VelPred with ppm_trace_forces reads never-written temporaries in the
    reference (MaestroVelPred.cpp:128-133, 556): reported as unsupported."""
    w = Workload(dim=2, n=16, max_grid=16, ppm_type=1, backend=NP, ppm_trace_forces=1)
    m = oracle_for(w)
    with pytest.raises(_abi.MxError) as ei:
        m.advance_pass(0, w.pass_fields())
    assert ei.value.code == _abi.MX_ERR_UNSUPPORTED


# ---------------------------------------------------------------------------
# MakeEdgeState1d (MaestroMakeEdgeState.cpp)

from edge1d_cases import dummy_boxes, params as e1_params, random_fields, run as e1_run
from edge1d_cases import two_level_geom


def e1_oracle(ppm, sph=0, so=4, **kw):
    p = e1_params(ppm, sph, so, **kw)
    return p, pyoracle.oracle_maestro(p, [dummy_boxes()])


@pytest.mark.parametrize("sph", [0, 1])
@pytest.mark.parametrize("ppm,so", [(0, 0), (0, 2), (0, 4), (1, 4), (2, 4)])
def test_edge_state_1d_constant(ppm, so, sph):
    """Constant s and force: every edge state is s + dt/2 f exactly, for any
    w0 (the slopes, face values and limiters all vanish)."""
    nr = 32
    p, m = e1_oracle(ppm, sph, so)
    rg = _abi.radial_geom(nr, 1.0 / nr)
    s = np.full((1, nr), 1.7)
    f = np.full((1, nr), 0.3)
    w0 = np.random.default_rng(5).uniform(-1, 1, (1, nr + 1))
    e = e1_run(m, rg, s, f, w0, NP)
    dth = 0.5 * p.dt
    np.testing.assert_array_equal(e, np.full_like(e, 1.7 + dth * 0.3))


@pytest.mark.parametrize("sph", [0, 1])
@pytest.mark.parametrize("ppm", [1, 2])
def test_edge_state_1d_zero_w0(ppm, sph):
    """w0 = 0: the trace returns the cell value on both sides, and the
    Riemann solver averages the two neighbours (a = s + dt/2 f):
    sedge(r) = (a(r) + a(r-1)) / 2, a(0) and a(nr-1) at the ends."""
    nr = 40
    p, m = e1_oracle(ppm, sph)
    rg = _abi.radial_geom(nr, 1.0 / nr)
    s, f, _ = random_fields(1, nr, 11)
    w0 = np.zeros((1, nr + 1))
    e = e1_run(m, rg, s, f, w0, NP)[0]
    a = s[0] + (0.5 * p.dt) * f[0]
    ref = np.empty(nr + 1)
    ref[1:nr] = 0.5 * (a[:-1] + a[1:])
    ref[0] = a[0]
    ref[nr] = a[nr - 1]
    np.testing.assert_array_equal(e, ref)


@pytest.mark.parametrize("sph", [0, 1])
@pytest.mark.parametrize("ppm", [1, 2])
@pytest.mark.parametrize("W", [0.7, -0.7])
def test_edge_state_1d_linear_trace(ppm, sph, W):
    """A linear profile s = a + b r with uniform w0 = W: both PPM types
    reconstruct it exactly and no limiter fires, so an interior edge state
    is the exact characteristic average s_upwind +- b (1 - sigma) / 2."""
    nr, dr, dt = 48, 1.0 / 48, 0.01
    p, m = e1_oracle(ppm, sph, dt=dt)
    rg = _abi.radial_geom(nr, dr)
    b = 0.25
    s = (2.0 + b * np.arange(nr))[None, :]
    f = np.zeros((1, nr))
    w0 = np.full((1, nr + 1), W)
    e = e1_run(m, rg, s, f, w0, NP)[0]
    sig = abs(W) * dt / dr
    r = np.arange(5, nr - 4)
    ref = s[0, r - 1] + b * (1 - sig) / 2 if W > 0 else s[0, r] - b * (1 - sig) / 2
    np.testing.assert_allclose(e[r], ref, rtol=1e-13, atol=0)


@pytest.mark.parametrize("ppm", [1, 2])
def test_edge_state_1d_coarse_fine_sync(ppm):
    """Two radial levels, w0 = 0: at the coarse-fine edges each level takes
    its outer state from the other level (MaestroMakeEdgeState.cpp:927-950),
    so both levels hold (a_fine + a_coarse) / 2 there."""
    p, m = e1_oracle(ppm)
    rg = two_level_geom()
    s, f, _ = random_fields(2, 64, 3)
    w0 = np.zeros((2, 65))
    e = e1_run(m, rg, s, f, w0, NP)
    a = s + (0.5 * p.dt) * f
    assert e[1, 16] == 0.5 * (a[1, 16] + a[0, 7])
    assert e[0, 8] == 0.5 * (a[1, 16] + a[0, 7])
    assert e[1, 48] == 0.5 * (a[0, 24] + a[1, 47])
    assert e[0, 24] == 0.5 * (a[0, 24] + a[1, 47])
    # interior of each level: the single-level closed form
    assert e[1, 30] == 0.5 * (a[1, 30] + a[1, 29])
    assert e[0, 3] == 0.5 * (a[0, 3] + a[0, 2])
    # fine entries outside the chunk are untouched
    assert np.isnan(e[1, 10]) and np.isnan(e[1, 60])


def test_edge_state_1d_rejects_blocking_factor_one():
    """A planar ppm chunk starting at r = 2 aborts in the reference
    (MaestroMakeEdgeState.cpp:427-436); the oracle returns MX_ERR_ARG."""
    p, m = e1_oracle(1)
    rg = _abi.radial_geom(64, 1.0 / 64, levels=[(32, 1.0 / 32, [(0, 31)]),
                                                (64, 1.0 / 64, [(2, 40)])])
    s, f, w0 = random_fields(2, 64, 1)
    with pytest.raises(_abi.MxError):
        e1_run(m, rg, s, f, w0, NP)


# ---------------------------------------------------------------------------
# the differentiable scalar step (oracle_step.cpp)

def test_scal_step_jvp_matches_finite_differences():
    """The dual-number derivative of the oracle step equals central finite
    differences of the oracle step along random directions (state and
    velocity), and its primal part is the step itself."""
    from maestroex_amd.adjoint import periodic_mac_velocity
    n = 8
    rng = np.random.default_rng(0)
    u = periodic_mac_velocity(n, seed=4)
    s = rng.random((2, n, n, n))
    dt, dx, eps = 0.5 / n, [1.0 / n] * 3, 1e-8
    ds = rng.standard_normal(s.shape)
    du = [rng.standard_normal((n, n, n)) for _ in range(3)]
    s1 = pyoracle.scal_step(s, u, dt, dx, eps)
    s2, js = pyoracle.scal_step(s, u, dt, dx, eps, ds, du)
    assert np.array_equal(s1, s2)
    h = 1e-6
    fp = pyoracle.scal_step(s + h * ds, [u[d] + h * du[d] for d in range(3)], dt, dx, eps)
    fm = pyoracle.scal_step(s - h * ds, [u[d] - h * du[d] for d in range(3)], dt, dx, eps)
    fd = (fp - fm) / (2 * h)
    assert np.max(np.abs(fd - js)) <= 1e-7 * np.max(np.abs(js))


def test_scal_step_conserves_and_keeps_constants():
    """Periodic, divergence-free velocity: the step conserves the total of
    every component to roundoff and keeps a constant field constant."""
    from maestroex_amd.adjoint import periodic_mac_velocity
    n = 8
    u = periodic_mac_velocity(n, seed=6)
    s = np.random.default_rng(1).random((1, n, n, n))
    out = pyoracle.scal_step(s, u, 0.5 / n, [1.0 / n] * 3, 1e-8)
    assert abs(out.sum() - s.sum()) <= 1e-12 * s.sum()
    c = np.full((1, n, n, n), 0.7)
    outc = pyoracle.scal_step(c, u, 0.5 / n, [1.0 / n] * 3, 1e-8)
    assert np.max(np.abs(outc - 0.7)) <= 1e-14


@pytest.mark.parametrize("ppm", [1, 2])
def test_scal_step_equals_the_pass_stages(ppm):
    """The differentiable step (oracle_step.cpp) is bit-identical to the
    oracle's own stages on the same periodic box: MakeEdgeScal (advective,
    no force) of the species, MakeRhoXFlux with species_pred_type 2
    (F = umac (rho X)_edge) and UpdateScal of the species."""
    from maestroex_amd.adjoint import periodic_mac_velocity
    from maestroex_amd.fabs import Geometry, make_params
    n, nspec, ng = 16, 3, 4
    g = Geometry(3, [n] * 3, [1.0 / n] * 3, [1, 1, 1], [0] * 3, [0] * 3)
    dt, eps = 0.5 / n, 1e-8
    params = make_params(g, ppm_type=ppm, nspec=nspec, species_pred_type=2, dt=dt,
                         rel_eps=eps)
    boxes = [((0, 0, 0), (n - 1, n - 1, n - 1))]
    u = periodic_mac_velocity(n, seed=5)
    rng = np.random.default_rng(9)
    x = (np.arange(n) + 0.5) / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    sp = np.stack([1.5 + 0.3 * np.sin(2 * np.pi * (X + 0.3 * c)) * np.cos(2 * np.pi * (Y - Z))
                   + 0.05 * rng.random((n, n, n)) for c in range(nspec)])
    first = params.first_spec
    nscal = params.nscal
    idx = np.arange(-ng, n + ng) % n
    state = MultiFab(boxes, nscal, ng, 3, NP)
    full = np.zeros((nscal, n, n, n))
    full[first:first + nspec] = sp
    full[params.comp_rho] = sp.sum(axis=0)
    state.arrays[0][...] = full[:, idx][:, :, idx][:, :, :, idx]
    umac = FaceMultiFab(boxes, 1, 1, 3, NP)
    for d in range(3):
        fidx = np.arange(-1, n + 2) % n
        sl = [np.arange(-1, n + 1) % n] * 3
        sl[2 - d] = fidx
        umac[d].arrays[0][0] = u[d][sl[0]][:, sl[1]][:, :, sl[2]]
    force = MultiFab(boxes, nscal, 1, 3, NP)
    sedge = FaceMultiFab(boxes, nscal, 0, 3, NP)
    sflux = FaceMultiFab(boxes, nscal, 0, 3, NP)
    eta = MultiFab(boxes, 1, 1, 3, NP, nodal_dir=2)
    snew = MultiFab(boxes, nscal, 0, 3, NP)
    m = pyoracle.oracle_maestro(params, [boxes])
    bs, _ = m.bc_setup()
    m.MakeEdgeScal(0, state, sedge, umac, force, 0, bs, first, first, nspec, 0)
    from maestroex_amd.fabs import BaseState
    zero = BaseState(np.zeros((1, n + 1)), NP)
    m.MakeRhoXFlux(0, sflux, eta, sedge, umac, zero, zero, None, zero, zero, None, zero, zero,
                   first, nspec)
    m.UpdateScal(0, state, snew, sflux, force, first, nspec)
    got = snew.arrays[0][first:first + nspec]
    want = pyoracle.scal_step(sp, u, dt, [1.0 / n] * 3, eps, ppm_type=ppm)
    assert np.array_equal(got, want), float(np.max(np.abs(got - want)))


def test_step_op_count():
    """The op-counting instantiation of the oracle's step: the phase counts
    are fixed by the formulas (predictor 9, corner 78, edge 75, flux and
    update 17 per cell; one face value 13), a constant field takes the
    shortest slope branch everywhere, and the non-redundant MakeEdgeScal
    figure that bench.py reports against the VALU rate follows from the
    counts on a smooth field."""
    import bench
    from maestroex_amd.adjoint import periodic_mac_velocity
    n = 16
    u = periodic_mac_velocity(n, seed=1)
    x = (np.arange(n) + 0.5) / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    smooth = np.exp(-((X - 0.5) ** 2 + (Y - 0.3) ** 2 + (Z - 0.5) ** 2) / 0.02)
    r = pyoracle.step_op_count(smooth, u, 0.5 / n, [1.0 / n] * 3, 1e-8)
    # per cell and direction: simh is one upwind (u > 0, |u|, |u| > 0: 3);
    # a corner state 2 x (sum, product, difference, product, difference) + 3;
    # an edge state 2 x 2 x 5 + 2 + 3; the flux update 6 products, 3
    # differences, 3 quotients, 2 sums and 3 for the update; a face value
    # 5 + 2 min + 2 max (the dt/h factors are constants, not counted)
    assert (r["predictor"], r["corner"], r["edge"], r["flux_update"], r["face"]) == (
        9.0, 78.0, 75.0, 17.0, 9.0)
    # a van Leer slope costs 8 (differences, product, test) or 15 (with the
    # sign, three |.| and two min, and the product)
    assert 8.0 <= r["slope"] <= 15.0
    const = pyoracle.step_op_count(np.full((n, n, n), 2.5), u, 0.5 / n, [1.0 / n] * 3, 1e-8)
    assert const["slope"] == 8.0
    per_point = r["ppm"] + r["predictor"] + r["corner"] + r["edge"]
    shared = per_point - 3 * (2 * r["slope"] + r["face"])
    print(f"[oracle] step ops per cell: {r}, MakeEdgeScal non-redundant {shared:.1f}")
    assert abs(shared - bench.EDGE_OPS_PER_CELL_COMP) <= 0.01 * bench.EDGE_OPS_PER_CELL_COMP


def test_uneven_boxes_partition_the_domain():
    """uneven_boxes: parts^3 boxes, edges within [lo, hi], covering [0, n)^3
    exactly once."""
    from maestroex_amd.workloads import uneven_boxes
    n, parts, lo, hi = 48, 4, 8, 18
    boxes = uneven_boxes(n, parts, lo, hi, seed=7)
    assert len(boxes) == parts ** 3
    cover = np.zeros((n, n, n), dtype=np.int32)
    for blo, bhi in boxes:
        for d in range(3):
            assert lo <= bhi[d] - blo[d] + 1 <= hi
        cover[blo[2]:bhi[2] + 1, blo[1]:bhi[1] + 1, blo[0]:bhi[0] + 1] += 1
    assert np.all(cover == 1)


def _global_faces(fmf, d, n, nc):
    """The valid faces of direction d of every box, scattered into one global
    (nc, n(+1 in d), ...) array; faces shared by two boxes must agree."""
    shape = [nc, n, n, n]
    shape[3 - d] += 1
    out = np.full(shape, np.nan)
    mf = fmf[d]
    for b, (lo, hi) in enumerate(mf.boxes):
        a = mf.to_numpy()[b][(slice(None),) + mf.valid_region(b)]
        sl = [slice(None)] + [slice(lo[q], hi[q] + 1 + (1 if q == d else 0)) for q in (2, 1, 0)]
        prev = out[tuple(sl)]
        seen = ~np.isnan(prev)
        assert np.array_equal(prev[seen], a[seen]), "a shared face differs between boxes"
        out[tuple(sl)] = a
    return out


@pytest.mark.parametrize("spherical", [False, True])
def test_pass_is_independent_of_box_decomposition(spherical):
    """One level's pass gives the same value on every face whether the domain
    is cut into regular boxes or into uneven ones: the ghost cells carry the
    neighbours' values, so the decomposition only changes where each face is
    computed (spherical: ppm_type 2, all sides outflow)."""
    from maestroex_amd.fabs import Backend
    from maestroex_amd.workloads import Workload, uneven_boxes
    n = 24
    cfg = dict(dim=3, n=n, ppm_type=2 if spherical else 1, spherical=spherical,
               backend=Backend("numpy"))
    results = []
    for boxes in (None, uneven_boxes(n, 3, 6, 10, seed=4)):
        w = Workload(max_grid=12, boxes=boxes, **cfg)
        m = pyoracle.oracle_maestro(w.params, [w.boxes])
        m.advance_pass(0, w.pass_fields())
        results.append([(_global_faces(w.umac, d, n, 1), _global_faces(w.sedge, d, n, w.nscal),
                         _global_faces(w.sflux, d, n, w.nscal)) for d in range(3)])
    comps = [0, 1] + list(range(4, 4 + 3))
    for d in range(3):
        ua, ea, fa = results[0][d]
        ub, eb, fb = results[1][d]
        assert np.array_equal(ua, ub)
        assert np.array_equal(ea[comps], eb[comps])
        assert np.array_equal(fa[comps], fb[comps])


def test_scal_step_ppm2_jvp_matches_finite_differences():
    """ppm_type 2 (Colella-Sekora): the dual-number derivative of the oracle
    step equals central finite differences along a random direction, at the
    cells and faces whose limiter branches do not switch within the step, and
    the step conserves totals and keeps constants."""
    from maestroex_amd.adjoint import periodic_mac_velocity
    n = 8
    rng = np.random.default_rng(2)
    u = periodic_mac_velocity(n, seed=4)
    x = (np.arange(n) + 0.5) / n
    Z, Y, X = np.meshgrid(x, x, x, indexing="ij")
    s = np.exp(-((X - 0.5) ** 2 + (Y - 0.4) ** 2 + (Z - 0.5) ** 2) / 0.03)[None] \
        + 0.02 * rng.standard_normal((1, n, n, n))
    dt, dx, eps = 0.5 / n, [1.0 / n] * 3, 1e-8
    ds = rng.standard_normal(s.shape)
    du = [rng.standard_normal((n, n, n)) for _ in range(3)]
    s1 = pyoracle.scal_step(s, u, dt, dx, eps, ppm_type=2)
    s2, js = pyoracle.scal_step(s, u, dt, dx, eps, ds, du, ppm_type=2)
    assert np.array_equal(s1, s2)
    assert not np.array_equal(s1, pyoracle.scal_step(s, u, dt, dx, eps, ppm_type=1))

    def fd(h):
        fp = pyoracle.scal_step(s + h * ds, [u[d] + h * du[d] for d in range(3)], dt, dx, eps,
                                ppm_type=2)
        fm = pyoracle.scal_step(s - h * ds, [u[d] - h * du[d] for d in range(3)], dt, dx, eps,
                                ppm_type=2)
        return (fp - fm) / (2 * h)

    f1, f2, f3 = fd(1e-6), fd(5e-7), fd(2.5e-7)
    scale = np.max(np.abs(js))
    # no branch switch within h: the three differences agree
    smooth = (np.abs(f1 - f2) <= 1e-8 * scale) & (np.abs(f2 - f3) <= 1e-8 * scale)
    assert smooth.mean() > 0.9  # a direction through every cell crosses a few switches
    assert np.max(np.abs(f3 - js)[smooth]) <= 1e-8 * scale
    # where a switch lies between h/4 and h the smallest difference still agrees
    assert np.mean(np.abs(f3 - js) <= 1e-8 * scale) > 0.95
    assert abs(s1.sum() - s.sum()) <= 1e-12 * abs(s.sum())
    c = np.full((1, n, n, n), 0.3)
    assert np.max(np.abs(pyoracle.scal_step(c, u, dt, dx, eps, ppm_type=2) - 0.3)) <= 1e-14
