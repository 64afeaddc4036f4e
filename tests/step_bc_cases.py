"""Inputs of the differentiable step on a domain with outflow sides, shared by
the CPU and GPU tests: a discretely divergence-free MAC velocity, seeded
scalar fields, box lists, and the MultiFabs of a box list filled from
global arrays the way FillPatch fills them (periodic images, foextrap
copies of the nearest valid cell beyond an outflow side).
"""
import numpy as np

from maestroex_amd import _abi
from maestroex_amd.fabs import FaceMultiFab, Geometry, MultiFab, make_params


def _diff(a, axis, periodic):
    """Forward difference along an axis: a[x + 1] - a[x], wrapping when
    periodic (n values from n nodes), else n - 1 values from n nodes."""
    if periodic:
        return np.roll(a, -1, axis=axis) - a
    sl1 = [slice(None)] * a.ndim
    sl0 = [slice(None)] * a.ndim
    sl1[axis] = slice(1, None)
    sl0[axis] = slice(None, -1)
    return a[tuple(sl1)] - a[tuple(sl0)]


def mac_velocity(n, periodic, seed=0, nmodes=12, umax=1.0):
    """The discrete curl of a random smooth vector potential on cell edges:
    divergence free on every cell to roundoff.  n = (nx, ny, nz); direction
    d periodic (integer wavenumbers, faces 0..n[d]-1) or not (faces
    0..n[d]).  Returns (ux, uy, uz) indexed [k, j, i], scaled to max|u| =
    umax."""
    rng = np.random.default_rng(seed)
    h = [1.0 / n[d] for d in range(3)]
    nodes = [np.arange(n[d] + (0 if periodic[d] else 1)) * h[d] for d in range(3)]
    cells = [(np.arange(n[d]) + 0.5) * h[d] for d in range(3)]
    modes = []
    for _ in range(nmodes):
        k = np.array([rng.integers(-3, 4) if periodic[d] else rng.uniform(-2.5, 2.5)
                      for d in range(3)], dtype=float)
        if not k.any():
            k[0] = 1.0
        modes.append((k, rng.normal(size=3), rng.uniform(0, 2 * np.pi, size=3)))
    A = []
    for comp in range(3):
        # component comp: cell centres along comp, nodes elsewhere
        xs = [cells[q] if q == comp else nodes[q] for q in range(3)]
        Z, Y, X = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")
        a = np.zeros(Z.shape)
        for k, amp, ph in modes:
            a += amp[comp] * np.cos(2 * np.pi * (k[0] * X + k[1] * Y + k[2] * Z) + ph[comp])
        A.append(a)
    Ax, Ay, Az = A
    # axes of a [k, j, i] array: x -> 2, y -> 1, z -> 0
    ux = _diff(Az, 1, periodic[1]) / h[1] - _diff(Ay, 0, periodic[2]) / h[2]
    uy = _diff(Ax, 0, periodic[2]) / h[2] - _diff(Az, 2, periodic[0]) / h[0]
    uz = _diff(Ay, 2, periodic[0]) / h[0] - _diff(Ax, 1, periodic[1]) / h[1]
    s = umax / max(np.abs(ux).max(), np.abs(uy).max(), np.abs(uz).max())
    return ux * s, uy * s, uz * s


def divergence(u, periodic):
    n = [u[1].shape[2], u[0].shape[1], u[0].shape[0]]
    return (_diff(u[0], 2, periodic[0]) * n[0] + _diff(u[1], 1, periodic[1]) * n[1]
            + _diff(u[2], 0, periodic[2]) * n[2])


def scalar_fields(n, ncomp, seed=0, noise=0.05):
    """Gaussian blobs plus seeded noise, (ncomp, nz, ny, nx)."""
    rng = np.random.default_rng(seed)
    xs = [(np.arange(n[d]) + 0.5) / n[d] for d in range(3)]
    Z, Y, X = np.meshgrid(xs[2], xs[1], xs[0], indexing="ij")
    out = []
    for c in range(ncomp):
        ctr = 0.3 + 0.4 * rng.random(3)
        b = np.exp(-((X - ctr[0]) ** 2 + (Y - ctr[1]) ** 2 + (Z - ctr[2]) ** 2) / 0.03)
        out.append(1.0 + b + noise * rng.standard_normal(X.shape))
    return np.stack(out)


def geometry(n, periodic):
    lo = [_abi.PHYS_INTERIOR if periodic[d] else _abi.PHYS_OUTFLOW for d in range(3)]
    return Geometry(3, list(n), [1.0 / n[d] for d in range(3)], [1 if p else 0 for p in periodic],
                    lo, lo)


def step_params(n, periodic, ppm, ncomp, dt, rel_eps, spherical=0):
    g = geometry(n, periodic)
    return make_params(g, ppm_type=ppm, nspec=ncomp, dt=dt, rel_eps=rel_eps,
                       species_pred_type=2, spherical=spherical)


def _map(x, nn, periodic, faces):
    """Global index -> stored index: periodic image, else the nearest stored
    point (cells 0..nn-1, faces 0..nn)."""
    if periodic:
        return np.mod(x, nn)
    return np.clip(x, 0, nn if faces else nn - 1)


def cell_mf(boxes, data, ngrow, periodic, backend):
    """A cell MultiFab with ngrow ghosts from the global (ncomp, nz, ny, nx)
    array: ghosts from periodic images or the nearest valid cell."""
    nc = data.shape[0]
    n = [data.shape[3], data.shape[2], data.shape[1]]
    mf = MultiFab(boxes, nc, ngrow, 3, backend)
    for b, (lo, hi) in enumerate(boxes):
        idx = [_map(np.arange(lo[d] - ngrow, hi[d] + ngrow + 1), n[d], periodic[d], False)
               for d in range(3)]
        mf.arrays[b] = backend.from_numpy(data[:, idx[2]][:, :, idx[1]][:, :, :, idx[0]])
    return mf


def face_mf(boxes, u, ngrow, periodic, backend):
    """A face MultiFab (one component) with ngrow ghosts from the global face
    arrays u[d] (faces 0..n[d] along an outflow d, 0..n[d]-1 along a
    periodic one)."""
    n = [u[1].shape[2], u[0].shape[1], u[0].shape[0]]
    fm = FaceMultiFab(boxes, 1, ngrow, 3, backend)
    for d in range(3):
        for b, (lo, hi) in enumerate(boxes):
            idx = []
            for q in range(3):
                top = hi[q] + (1 if q == d else 0)
                idx.append(_map(np.arange(lo[q] - ngrow, top + ngrow + 1), n[q], periodic[q],
                                q == d))
            a = u[d][idx[2]][:, idx[1]][:, :, idx[0]]
            fm[d].arrays[b] = backend.from_numpy(a[None])
    return fm


def gather_cells(mf, n):
    """The valid cells of a cell MultiFab as one (ncomp, nz, ny, nx) array."""
    out = None
    for b, a in enumerate(mf.to_numpy()):
        lo, hi = mf.boxes[b]
        if out is None:
            out = np.full((a.shape[0], n[2], n[1], n[0]), np.nan)
        out[:, lo[2]:hi[2] + 1, lo[1]:hi[1] + 1, lo[0]:hi[0] + 1] = a[(slice(None),)
                                                                       + mf.valid_region(b)]
    return out


def gather_faces(fm, d, shape):
    """The valid faces of direction d of a face MultiFab (component 0) as one
    global array of the given [k, j, i] shape (periodic: face n is face 0,
    dropped)."""
    out = np.full(shape, np.nan)
    for b, a in enumerate(fm[d].to_numpy()):
        lo, hi = fm[d].boxes[b]
        v = a[0][fm[d].valid_region(b)]
        top = [hi[q] + (1 if q == d else 0) for q in range(3)]
        sl = [slice(lo[q], min(top[q], shape[2 - q] - 1) + 1) for q in range(3)]
        out[sl[2], sl[1], sl[0]] = v[:sl[2].stop - sl[2].start, :sl[1].stop - sl[1].start,
                                     :sl[0].stop - sl[0].start]
    return out
